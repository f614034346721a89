"""Readers and writers for the sim-trhepd-rheed file formats.

The formats are the reference's own and stay byte-compatible:

* ``bulk.txt`` / ``surf.txt`` -- Fortran list-directed decks, read in the order of
  ``bulkio.f90:34-117`` and ``surfio.f90:125-195`` (sim-trhepd-rheed/src).
* ``bulkP.b`` -- Fortran unformatted sequential file (4-byte little-endian record
  markers) written by ``bulkio.f90:126-141,173-176`` and ``bulk.f90:122``.
* ``*.s`` -- rocking-curve output written by ``surfio.f90:270-274`` and
  ``surflf.f90:154-165`` with the format ``(E12.4,200(",",E23.15))``.
"""
from __future__ import annotations

import dataclasses
import math
import re
import struct
from typing import List, Optional, Sequence

import numpy as np

# --------------------------------------------------------------------------
# list-directed input

_SEP = re.compile(r"\s*,\s*|\s+")


class ListDirected:
    """Minimal Fortran list-directed READ: each ``read(n)`` consumes whole records
    until ``n`` values are gathered; the rest of the last record is discarded."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def read(self, n: int) -> List[str]:
        vals: List[str] = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise EOFError("deck ended while reading %d values" % n)
            line = self.lines[self.pos].strip()
            self.pos += 1
            if not line:
                continue
            for tok in _SEP.split(line):
                if len(vals) == n:
                    break
                if tok == "":
                    continue
                if tok.startswith("/"):
                    break
                vals.append(tok)
        return vals

    def ints(self, n: int) -> List[int]:
        return [int(float(_fnum(t))) for t in self.read(n)]

    def floats(self, n: int) -> List[float]:
        return [float(_fnum(t)) for t in self.read(n)]


def _fnum(tok: str) -> str:
    return tok.replace("d", "e").replace("D", "e")


# --------------------------------------------------------------------------
# decks


@dataclasses.dataclass
class Element:
    iz: int
    da1: float
    sap: float
    bh: float
    bk: float
    bz: float


@dataclasses.dataclass
class Atom:
    ielm: int
    ocr: float
    x: float
    y: float
    z: float


@dataclasses.dataclass
class BulkDeck:
    """``bulk.txt`` (read order of bulkio.f90:34-117)."""
    nh: int
    nk: int
    ndom: int
    nb: List[int]
    rdom: List[float]
    ih: List[int]
    ik: List[int]
    be: float
    azi: float
    azf: float
    daz: float
    gi: float
    gf: float
    dg: float
    dz: float
    ml: int
    elements: List[Element]
    nsg: int
    aa: float
    bb: float
    gam: float
    cc: float
    dx: float
    dy: float
    atoms: List[Atom]

    @staticmethod
    def parse(text: str) -> "BulkDeck":
        r = ListDirected(text)
        nh, nk, ndom = r.ints(3)
        nb = r.ints(ndom)
        rdom = r.floats(ndom)
        ih: List[int] = []
        ik: List[int] = []
        for d in range(ndom):
            v = r.ints(2 * nb[d])
            ih += v[0::2]
            ik += v[1::2]
        be, azi, azf, daz, gi, gf, dg = r.floats(7)
        dzv = r.read(2)
        dz, ml = float(_fnum(dzv[0])), int(float(_fnum(dzv[1])))
        (nelm,) = r.ints(1)
        els = []
        for _ in range(nelm):
            t = r.read(3)
            iz, da1, sap = int(float(t[0])), float(_fnum(t[1])), float(_fnum(t[2]))
            bh, bk, bz = r.floats(3)
            els.append(Element(iz, da1, sap, bh, bk, bz))
        t = r.read(7)
        nsg = int(float(t[0]))
        aa, bb, gam, cc, dx, dy = (float(_fnum(x)) for x in t[1:])
        (natm,) = r.ints(1)
        atoms = []
        for _ in range(natm):
            t = r.read(5)
            atoms.append(Atom(int(float(t[0])), *(float(_fnum(x)) for x in t[1:])))
        return BulkDeck(nh, nk, ndom, nb, rdom, ih, ik, be, azi, azf, daz, gi, gf, dg, dz, ml,
                        els, nsg, aa, bb, gam, cc, dx, dy, atoms)

    def _beam_lines(self) -> List[str]:
        # one record per domain: each domain's beams are one list-directed read (bulkio.f90)
        out, ib = [], 0
        for n in self.nb:
            out.append(",".join("%d,%d" % (self.ih[j], self.ik[j]) for j in range(ib, ib + n)) + "\t,(IH(I),IK(I)")
            ib += n
        return out

    def format(self) -> str:
        o = ["%d,%d,%d\t\t,NH,NK,NDOM" % (self.nh, self.nk, self.ndom),
             ",".join(str(n) for n in self.nb) + "\t\t,NB",
             ",".join(_g(x) for x in self.rdom) + "\t\t,RDOM",
             *self._beam_lines(),
             ",".join(_g(x) for x in (self.be, self.azi, self.azf, self.daz, self.gi, self.gf, self.dg))
             + "\t,BE,AZI,AZF,DAZ,GI,GF,DG",
             "%s,%d\t\t,DZ,ML" % (_g(self.dz), self.ml),
             "%d\t\t,NELM" % len(self.elements)]
        for e in self.elements:
            o.append("%d,%s,%s\t\t,Z,da1,sap" % (e.iz, _g(e.da1), _g(e.sap)))
            o.append("%s,%s,%s\t\t,BH(I),BK(I),BZ(I)" % (_g(e.bh), _g(e.bk), _g(e.bz)))
        o.append(",".join([str(self.nsg)] + [_g(x) for x in (self.aa, self.bb, self.gam, self.cc, self.dx, self.dy)])
                 + "\t,NSG,AA,BB,GAM,CC,DX,DY")
        o.append("%d\t\t,NATM" % len(self.atoms))
        for a in self.atoms:
            o.append("%d,%s,%s,%s,%s\t,IELM(I),ocr(I),X(I),Y(I),Z(I)" % (a.ielm, _g(a.ocr), _g(a.x), _g(a.y), _g(a.z)))
        return "\n".join(o) + "\n"


@dataclasses.dataclass
class SurfDeck:
    """``surf.txt`` (read order of surfio.f90:125-195)."""
    elements: List[Element]
    nsgs: int
    msa: int
    msb: int
    nsa: int
    nsb: int
    dthick: float
    dxs: float
    dys: float
    atoms: List[Atom]
    wdom: List[float]

    @staticmethod
    def parse(text: str, ndom: int = 1) -> "SurfDeck":
        r = ListDirected(text)
        (nelms,) = r.ints(1)
        els = []
        for _ in range(nelms):
            t = r.read(3)
            iz, da1, sap = int(float(t[0])), float(_fnum(t[1])), float(_fnum(t[2]))
            bh, bk, bz = r.floats(3)
            els.append(Element(iz, da1, sap, bh, bk, bz))
        t = r.read(8)
        nsgs, msa, msb, nsa, nsb = (int(float(x)) for x in t[:5])
        dthick, dxs, dys = (float(_fnum(x)) for x in t[5:])
        (natms,) = r.ints(1)
        atoms = []
        for _ in range(natms):
            t = r.read(5)
            atoms.append(Atom(int(float(t[0])), *(float(_fnum(x)) for x in t[1:])))
        wdom = r.floats(ndom) if ndom > 1 else [1.0]
        return SurfDeck(els, nsgs, msa, msb, nsa, nsb, dthick, dxs, dys, atoms, wdom)

    def format(self) -> str:
        o = ["%d\t\t,NELMS" % len(self.elements)]
        for e in self.elements:
            o.append("%d,%s,%s\t\t,Z,da1,sap" % (e.iz, _g(e.da1), _g(e.sap)))
            o.append("%s,%s,%s\t\t,BH(I),BK(I),BZ(I)" % (_g(e.bh), _g(e.bk), _g(e.bz)))
        o.append(",".join(str(v) for v in (self.nsgs, self.msa, self.msb, self.nsa, self.nsb))
                 + ",%s,%s,%s\t,NSGS,msa,msb,nsa,nsb,dthick,DXS,DYS" % (_g(self.dthick), _g(self.dxs), _g(self.dys)))
        o.append("%d\t\t,NATM" % len(self.atoms))
        for a in self.atoms:
            o.append("%d,%s,%s,%s,%s" % (a.ielm, _g(a.ocr), _g(a.x), _g(a.y), _g(a.z)))
        if len(self.wdom) > 1:
            # read only for several domains (surfio.f90:191-194)
            o.append(",".join(_g(w) for w in self.wdom) + "\t\t,WDOM")
        return "\n".join(o) + "\n"


def _g(x: float) -> str:
    return repr(float(x))


# --------------------------------------------------------------------------
# bulkP.b


class _Rec:
    def __init__(self, payload: bytes):
        self.b = payload
        self.o = 0

    def i(self, n=1):
        v = struct.unpack_from("<%di" % n, self.b, self.o)
        self.o += 4 * n
        return v if n > 1 else v[0]

    def d(self, n=1):
        v = struct.unpack_from("<%dd" % n, self.b, self.o)
        self.o += 8 * n
        return v if n > 1 else v[0]


def _records(data: bytes):
    o = 0
    while o < len(data):
        (n,) = struct.unpack_from("<i", data, o)
        payload = data[o + 4:o + 4 + n]
        (n2,) = struct.unpack_from("<i", data, o + 4 + n)
        if n2 != n:
            raise ValueError("corrupt unformatted record at byte %d" % o)
        yield payload
        o += 8 + n


def _record(fmt: str, *vals) -> bytes:
    p = struct.pack("<" + fmt, *vals)
    return struct.pack("<i", len(p)) + p + struct.pack("<i", len(p))


@dataclasses.dataclass
class BulkElement:
    a: np.ndarray       # nabr+nabi raw coefficients (before asfcrr)
    b: np.ndarray
    dth: float
    dtk: float
    dtz: float
    elm: float
    dchg: float


@dataclasses.dataclass
class BulkDomain:
    iord: np.ndarray    # (nb,) 1-based: sorted position -> original beam index
    nbg: np.ndarray     # (ngr,) group sizes
    igh: np.ndarray     # (nvb,)
    igk: np.ndarray
    blocks: List[List[List[np.ndarray]]]   # [irep1][irep2][group] -> (n_g, n_g) complex


@dataclasses.dataclass
class BulkP:
    """Contents of ``bulkP.b`` (record layout of bulkio.f90:126-141,173-176; bulk.f90:122)."""
    ndom: int
    inegpos: int
    isym: int
    nh: int
    nk: int
    iabr: int
    iabp: int
    iabe: int
    nabr: int
    nabi: int
    idiag: int
    nb: List[int]
    rdom: List[float]
    ih: List[int]
    ik: List[int]
    be: float
    azi: float
    daz: float
    naz: int
    gi: float
    dg: float
    ng: int
    tem: float
    dz: float
    ml: int
    epsb: float
    nsg: int
    mp: int
    aa: float
    bb: float
    gam: float
    cc: float
    dx: float
    dy: float
    zoffset: float
    elements: List[BulkElement]
    atoms: List[Atom]
    domains: List[BulkDomain]

    @property
    def nrep(self):
        return (self.naz, self.ng) if self.ng > self.naz else (self.ng, self.naz)

    @staticmethod
    def read(path_or_bytes) -> "BulkP":
        data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
        recs = iter(_records(data))
        nx = lambda: _Rec(next(recs))
        ndom = nx().i()
        inegpos, isym, nh, nk, iabr, iabp, iabe, nabr, nabi, idiag = nx().i(10)
        r = nx()
        nb, rdom = [], []
        for _ in range(ndom):
            nb.append(r.i())
            rdom.append(r.d())
        nbt = sum(nb)
        r = nx()
        hk = r.i(2 * nbt) if nbt > 0 else ()
        hk = hk if isinstance(hk, tuple) else (hk,)
        ih, ik = list(hk[0::2]), list(hk[1::2])
        r = nx()
        be, azi, daz = r.d(3)
        naz = r.i()
        gi, dg = r.d(2)
        ng = r.i()
        r = nx()
        tem, dz = r.d(2)
        ml = r.i()
        epsb = r.d()
        r = nx()
        nsg, mp = r.i(2)
        aa, bb, gam, cc, dx, dy, zoffset = r.d(7)
        nelm, natm = nx().i(2)
        els = []
        na = nabr + nabi
        for _ in range(nelm):
            r = nx()
            ab = np.array(r.d(2 * na))
            r = nx()
            dth, dtk, dtz, elm, dchg = r.d(5)
            els.append(BulkElement(ab[:na].copy(), ab[na:].copy(), dth, dtk, dtz, elm, dchg))
        atoms = []
        for _ in range(natm):
            r = nx()
            ielm = r.i()
            ocr, x, y, z = r.d(4)
            atoms.append(Atom(ielm, ocr, x, y, z))
        nrep1, nrep2 = (naz, ng) if ng > naz else (ng, naz)
        doms = []
        for d in range(ndom):
            r = nx()
            iord = np.array(r.i(nb[d]) if nb[d] > 1 else (r.i(),), dtype=np.int32)
            ngr, nvb = nx().i(2)
            r = nx()
            nbg = np.array(r.i(ngr) if ngr > 1 else (r.i(),), dtype=np.int32)
            r = nx()
            g = r.i(2 * nvb)
            g = g if isinstance(g, tuple) else (g,)
            igh = np.array(g[0::2], dtype=np.int32)
            igk = np.array(g[1::2], dtype=np.int32)
            blocks = []
            for _ in range(nrep1):
                row = []
                for _ in range(nrep2):
                    grp = []
                    for n in nbg:
                        rb = next(recs)
                        a = np.frombuffer(rb, dtype="<c16", count=int(n) * int(n))
                        grp.append(a.reshape((int(n), int(n)), order="F").copy())
                    row.append(grp)
                blocks.append(row)
            doms.append(BulkDomain(iord, nbg, igh, igk, blocks))
        return BulkP(ndom, inegpos, isym, nh, nk, iabr, iabp, iabe, nabr, nabi, idiag, nb, rdom, ih, ik,
                     be, azi, daz, naz, gi, dg, ng, tem, dz, ml, epsb, nsg, mp, aa, bb, gam, cc, dx, dy,
                     zoffset, els, atoms, doms)

    def to_bytes(self) -> bytes:
        o = [_record("i", self.ndom),
             _record("10i", self.inegpos, self.isym, self.nh, self.nk, self.iabr, self.iabp, self.iabe,
                     self.nabr, self.nabi, self.idiag)]
        o.append(_record("id" * self.ndom, *[v for p in zip(self.nb, self.rdom) for v in p]))
        o.append(_record("%di" % (2 * len(self.ih)), *[v for p in zip(self.ih, self.ik) for v in p]))
        o.append(_record("3di2di", self.be, self.azi, self.daz, self.naz, self.gi, self.dg, self.ng))
        o.append(_record("2did", self.tem, self.dz, self.ml, self.epsb))
        o.append(_record("2i7d", self.nsg, self.mp, self.aa, self.bb, self.gam, self.cc, self.dx, self.dy, self.zoffset))
        o.append(_record("2i", len(self.elements), len(self.atoms)))
        for e in self.elements:
            na = len(e.a)
            o.append(_record("%dd" % (2 * na), *e.a, *e.b))
            o.append(_record("5d", e.dth, e.dtk, e.dtz, e.elm, e.dchg))
        for a in self.atoms:
            o.append(_record("i4d", a.ielm, a.ocr, a.x, a.y, a.z))
        for d in self.domains:
            o.append(_record("%di" % len(d.iord), *d.iord))
            o.append(_record("2i", len(d.nbg), len(d.igh)))
            o.append(_record("%di" % len(d.nbg), *d.nbg))
            o.append(_record("%di" % (2 * len(d.igh)), *[v for p in zip(d.igh, d.igk) for v in p]))
            for row in d.blocks:
                for grp in row:
                    for blk in grp:
                        p = np.asarray(blk, dtype="<c16").reshape(-1, order="F").tobytes()
                        o.append(struct.pack("<i", len(p)) + p + struct.pack("<i", len(p)))
        return b"".join(o)


# --------------------------------------------------------------------------
# .s rocking-curve files


def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran ``Ew.d`` edit descriptor (0.dddE+xx, exponent letter dropped for |exp|>99)."""
    if not math.isfinite(x):
        return ("NaN" if x != x else ("Infinity" if x > 0 else "-Infinity")).rjust(w)
    if x == 0.0:
        return ("0." + "0" * d + "E+00").rjust(w)
    digits, e10 = ("%.*e" % (d - 1, abs(x))).split("e")
    ex = int(e10) + 1
    es = ("E%+03d" % ex) if abs(ex) <= 99 else ("%+04d" % ex)
    return (("-" if x < 0 else "") + "0." + digits.replace(".", "") + es).rjust(w)


def format_s_row(angle_deg: float, f2: Sequence[float], w: int = 23, d: int = 15) -> str:
    """One angle row of ``(E12.4,200(",",Ew.d))`` (E23.15 for srfref, E12.4 for domainsum)
    with format reversion every 200 values.  Output stops at the first data edit descriptor
    without a list item, so a record that ends inside the 200-group carries the group's
    leading "," at its end."""
    lines = []
    cur = fortran_e(angle_deg, 12, 4)
    for i, v in enumerate(f2):
        if i > 0 and i % 200 == 0:
            lines.append(cur)
            cur = ""
        cur += "," + fortran_e(float(v), w, d)
    if len(f2) % 200 != 0 or len(f2) == 0:
        cur += ","
    lines.append(cur)
    return "\n".join(lines)


def write_s(path: str, naz: int, ng: int, ih: Sequence[int], ik: Sequence[int],
            rows: Sequence[Sequence[tuple]]) -> None:
    """Write a single-domain ``.s`` file. ``rows[irep1]`` is a list of ``(angle_deg, f2)``."""
    nb = len(ih)
    out = ["#azimuths,g-angles,beams", "%4d%4d%4d" % (naz, ng, nb), "#ih,ik"]
    pairs = [v for p in zip(ih, ik) for v in p]
    for i in range(0, len(pairs), 400):
        out.append("".join("%4d" % v for v in pairs[i:i + 400]))
    for block in rows:
        for ang, f2 in block:
            out.append(format_s_row(ang, f2))
        out.append("")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def format_md(naz: int, ng: int, nh: int, nk: int, gam_deg: float, nb: Sequence[int], rdom_deg: Sequence[float],
              wdom: Sequence[float], ih: Sequence[int], ik: Sequence[int], rows_by_domain) -> str:
    """Text of a multi-domain ``.md`` file (surfio.f90:276-288 header, then each domain's rows
    as srfref writes them, surflf.f90:154-165).  ``rows_by_domain[idom][irep1]`` is a list of
    ``(angle_deg, f2)``; ih/ik list every domain's beams in turn."""
    ndom = len(nb)
    out = ["#azimuths,g-angles,domains,nh,nk,gam", "%4d%4d%4d%4d%4d%10.3f" % (naz, ng, ndom, nh, nk, gam_deg),
           "#beams,domain angle,domain weight"]
    for d in range(ndom):
        out.append("%4d%10.3f%s" % (nb[d], rdom_deg[d], fortran_e(wdom[d], 12, 3)))
    out.append("#ih,ik")
    ib = 0
    for d in range(ndom):
        pairs = [v for j in range(ib, ib + nb[d]) for v in (ih[j], ik[j])]
        for i in range(0, len(pairs), 400):
            out.append("".join("%4d" % v for v in pairs[i:i + 400]))
        ib += nb[d]
    for rows in rows_by_domain:
        for block in rows:
            for ang, f2 in block:
                out.append(format_s_row(ang, f2))
            out.append("")
    return "\n".join(out) + "\n"


_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD]?[-+]\d+)?")


def _parse_fnum(tok: str) -> float:
    tok = tok.strip()
    m = re.match(r"^([-+]?\d*\.\d+)([-+]\d{3})$", tok)
    if m:   # Ew.d with a 3-digit exponent drops the letter
        return float(m.group(1) + "e" + m.group(2))
    return float(tok.replace("D", "E").replace("d", "e"))


@dataclasses.dataclass
class RockingCurve:
    naz: int
    ng: int
    nb: int
    ih: List[int]
    ik: List[int]
    angles: np.ndarray      # (npoints,) first column as printed
    f2: np.ndarray          # (npoints, nb)


def read_s(path: str) -> RockingCurve:
    """Read a single-domain ``.s`` file the way test/test23/compare.py does."""
    return read_s_text(open(path).read())


def read_s_text(s_text: str) -> RockingCurve:
    """`read_s` of the file's text."""
    text = s_text.splitlines()
    it = iter(text)

    def nextline():
        while True:
            ln = next(it)
            if ln.startswith("#"):
                continue
            return ln

    hdr = nextline().split()
    naz, ng, nb = int(hdr[0]), int(hdr[1]), int(hdr[2])
    vals: List[float] = []
    while len(vals) < 2 * nb:
        vals += [int(x) for x in nextline().replace(",", " ").split()]
    ih, ik = vals[0::2], vals[1::2]
    rows = []
    npts = naz * ng
    for _ in range(npts):
        cur: List[float] = []
        while len(cur) < 1 + nb:
            ln = nextline()
            if not ln.strip():
                continue
            cur += [_parse_fnum(t) for t in ln.split(",") if t.strip()]
        rows.append(cur)
    arr = np.array(rows)
    return RockingCurve(naz, ng, nb, ih, ik, arr[:, 0], arr[:, 1:])
