"""Rocking-curve products after the slice loop.

- `domainsum`: the weighted sum of a multi-domain ``.md`` file into a single-domain ``.s``
  (sim-trhepd-rheed/src/domainsum.f90:7-121).  Beams of every other domain are rotated by the
  difference of the domain angles and added onto the beam of the domain with the most beams
  that they coincide with.
- `convolution_beam`: the angular convolution of one beam column of a ``.s`` file with a
  Gaussian of full width `omega` at half maximum (sim-trhepd-rheed/script/make_convolution.py),
  output ``"%f %.9f"`` rows.
- `convolution`: the same for every beam (str-orig/tool/make_convolution.py), output rows of
  ``%.5e`` values: the angle, then the convolved beams.

These are host-side text transforms of a few thousand numbers, as in the reference (whose
tools are Fortran and Python); the device is not involved.
"""
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import deckio

_RAD = math.atan(1.0) / 45.0


def _rows(text: str) -> List[str]:
    return text.splitlines()


def _floats(line: str) -> List[float]:
    return [deckio._parse_fnum(t) for t in line.replace(",", " ").split()]


def domainsum(md_text: str, wdom: Optional[Sequence[float]] = None) -> str:
    """The single-domain ``.s`` text of a multi-domain ``.md`` text (domainsum.f90).  `wdom`
    replaces the weights stored in the file, as the reference's interactive entry does.
    Raises ValueError on a single-domain file."""
    lines = _rows(md_text)
    it = iter(lines)
    if next(it).strip() != "#azimuths,g-angles,domains,nh,nk,gam":
        raise ValueError("single domain")
    h = next(it).split()
    naz, ng, ndom, nh, nk = (int(v) for v in h[:5])
    gam = float(h[5])
    nrep1, nrep2 = (naz, ng) if ng > naz else (ng, naz)
    next(it)   # "#beams,domain angle,domain weight"
    nb, rdom, w = [], [], []
    for _ in range(ndom):
        v = next(it).split()
        nb.append(int(v[0]))
        rdom.append(float(v[1]))
        w.append(deckio._parse_fnum(v[2]))
    if wdom is not None:
        w = [float(x) for x in wdom]
    # the reference domain: the first with the most beams
    idomr = max(range(ndom), key=lambda d: (nb[d], -d))
    mb = nb[idomr]
    next(it)   # "#ih,ik"
    ih: List[int] = []
    ik: List[int] = []
    for d in range(ndom):
        vals: List[int] = []
        while len(vals) < 2 * nb[d]:
            vals += [int(x) for x in next(it).split()]
        ih += vals[0::2]
        ik += vals[1::2]
    ibtr = sum(nb[:idomr])

    def gxy(j):
        gx, gy = ih[j] / nh, ik[j] / nk
        if abs(gam - 120.0) < 1e-4:
            gx = (gx + gx + gy) / math.sqrt(3.0)
        return gx, gy

    s2 = np.zeros((nrep1, nrep2, mb))
    angle = [0.0] * nrep2
    ibt = 0
    for d in range(ndom):
        target = None
        if d != idomr:
            dr = rdom[idomr] - rdom[d]
            cdr, sdr = math.cos(dr * _RAD), math.sin(dr * _RAD)
            target = []
            for ib in range(nb[d]):
                gx, gy = gxy(ib + ibt)
                gxr, gyr = cdr * gx - sdr * gy, sdr * gx + cdr * gy
                hit = None
                for ibr in range(mb):
                    rx, ry = gxy(ibr + ibtr)
                    if abs(gxr - rx) < 1e-4 and abs(gyr - ry) < 1e-4:
                        hit = ibr
                        break
                target.append(hit)
        for r1 in range(nrep1):
            for r2 in range(nrep2):
                vals: List[float] = []
                while len(vals) < nb[d] + 1:
                    vals += _floats(next(it))
                angle[r2] = vals[0]
                f2 = vals[1:nb[d] + 1]
                for ib in range(nb[d]):
                    t = ib if target is None else target[ib]
                    if t is not None:
                        s2[r1, r2, t] += f2[ib] * w[d]
            next(it)   # blank line after each irep1
        ibt += nb[d]
    out = ["#azimuths,g-angles,beams", "%4d%4d%4d" % (naz, ng, mb), "#ih,ik"]
    pairs = [v for j in range(ibtr, ibtr + mb) for v in (ih[j], ik[j])]
    for i in range(0, len(pairs), 400):
        out.append("".join("%4d" % v for v in pairs[i:i + 400]))
    for r1 in range(nrep1):
        for r2 in range(nrep2):
            out.append(deckio.format_s_row(angle[r2], s2[r1, r2], 12, 4))
        out.append("")   # the empty record of write (3,*), as gfortran writes it
    return "\n".join(out) + "\n"


def _gauss_fwhm(omega: float):
    # make_convolution.py (sim-trhepd-rheed/script): the Gaussian of full width omega at
    # half maximum with its published constants
    return lambda x: (0.939437 / omega) * np.exp(-2.77259 * (x ** 2.0 / omega ** 2.0))


def convolution_beam(s_text: str, first_line: int = 5, last_line: int = 74, row_number: int = 2,
                     omega: float = 0.5, step: float = 0.1) -> Tuple[List[float], List[float], str]:
    """Convolution of column `row_number` (1 = the angle) of data lines first_line..last_line
    (1-based) of a ``.s`` text, with the script's weight `step` per sample.  Returns the
    angles, the convolved values and the output text."""
    g = _gauss_fwhm(omega)
    lines = s_text.splitlines()
    deg, col = [], []
    for line in lines[first_line - 1:last_line]:
        data = line.replace(",", "").split()
        deg.append(float(data[0]))
        col.append(float(data[row_number - 1]))
    out = []
    for i in range(len(col)):
        acc = 0.0
        for j in range(len(col)):
            acc += col[j] * g(deg[i] - deg[j]) * step
        out.append(acc)
    text = "".join("%f %.9f\n" % (deg[i], out[i]) for i in range(len(out)))
    return deg, out, text


def convolution(s_text: str, omega: float = 0.5) -> Tuple[np.ndarray, str]:
    """Convolution of every beam of a single-azimuth ``.s`` text (str-orig/tool/
    make_convolution.py): a normalised Gaussian of full width omega at half maximum, the
    sample spacing taken from the first two angles.  Returns the array (angle, beams...) per
    glancing angle and the ``%.5e`` text."""
    sigma = 0.5 * omega / (np.sqrt(2.0 * np.log(2.0)))

    def g(x):
        return (1.0 / (sigma * np.sqrt(2.0 * np.pi))) * np.exp(-0.5 * x ** 2.0 / sigma ** 2.0)

    lines = s_text.splitlines()
    h = lines[1].replace(",", "").split()
    ng, nb = int(h[1]), int(h[2])
    org = np.zeros((ng, nb + 1))
    cnv = np.zeros((ng, nb + 1))
    for a in range(ng):
        data = lines[4 + a].replace(",", "").split()
        org[a, 0] = cnv[a, 0] = float(data[0])
        for b in range(nb):
            org[a, b + 1] = float(data[b + 1])
    step = org[1, 0] - org[0, 0]
    for b in range(nb):
        for i in range(ng):
            acc = 0.0
            for j in range(ng):
                acc += org[j, b + 1] * g(org[i, 0] - org[j, 0]) * step
            cnv[i, b + 1] = acc
    text = "".join(" ".join("%.5e" % v for v in row) + "\n" for row in cnv)
    return cnv, text
