"""Seeded synthetic decks for the benchmark configurations.

Common to all: a 1x1 hexagonal cell a = b = 3.84 A, gamma = 120 deg, unit-layer spacing
cc = 3.13 A with ABC stacking (DX = -1/3, DY = 1/3), one Si bulk atom per unit, 10 keV
positrons, Debye-Waller B entered as BH = BK = BZ, 4-Gaussian scattering factors, sap =
0.15, da1 = 1.0, bulk slices DZ = 0.02 A.  Beams are the N shortest reciprocal vectors of
the cell (ties broken by (h, k)).  Surface atoms sit at seeded in-plane positions uniform
in the cell and heights uniform in 1.0-2.5 A above the top bulk atom; the surface region
(the slab the slice loop integrates) has a stated height.

The decks are written in the reference's own formats (bulk.txt / surf.txt) and the
bulkP.b header comes from ``driver.bulk_header``.  The bulk reflection matrices the
surface loop starts from are those of the deck's ML unit layers (6) on a non-reflecting
substrate -- what blkref's recursion computes from R = 0 -- obtained on the device by
integrating that slab of bulk unit layers (``bulk_slab_deck``) from F = 0 with the
renormalised ratio-form propagator.  blkref's own route (the transfer matrix of a unit
layer, trmat / bulk.f90:69-93) is ill conditioned on these 1x1 decks: every beam is a bulk
beam, the strongly evanescent ones (|Im gamma| up to ~19 per A at N = 23, more at larger N)
make the unit layer's transfer matrix grow like exp(|Im gamma| cc) ~ 1e25, and two FP64
builds of blkref (the reference's bulk program and the numpy restatement) return reflection
matrices that differ at O(1) and rocking curves that differ by O(10) on the N = 23 deck.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import deckio, driver, lib

A_CELL = 3.84
CC = 3.13
ZMAP = {14: 0.15, 8: 0.15, 1: 0.15}


def shortest_beams(n: int, nh: int = 1, nk: int = 1, aa: float = A_CELL, gam_deg: float = 120.0):
    """The n shortest in-plane reciprocal vectors (h, k), (0, 0) first."""
    pi = math.pi
    gam = math.radians(gam_deg)
    ghx = 2 * pi / (aa * nh)
    ghy = -2 * pi / (aa * math.tan(gam) * nh)
    gky = 2 * pi / (aa * math.sin(gam) * nk)
    r = int(math.ceil(math.sqrt(n))) + 4
    cand = []
    for h in range(-3 * r, 3 * r + 1):
        for k in range(-3 * r, 3 * r + 1):
            g2 = (ghx * h) ** 2 + (ghy * h + gky * k) ** 2
            cand.append((round(g2, 9), h, k))
    cand.sort()
    pick = cand[:n]
    return [p[1] for p in pick], [p[2] for p in pick]


def bulk_deck(nb: int, gi: float, gf: float, dg: float, azi: float = 0.0, b_dw: float = 0.5,
              ml: int = 6) -> deckio.BulkDeck:
    ih, ik = shortest_beams(nb)
    el = deckio.Element(14, 1.0, ZMAP[14], b_dw, b_dw, b_dw)
    return deckio.BulkDeck(1, 1, 1, [nb], [0.0], ih, ik, 10.0, azi, azi, 0.0, gi, gf, dg, 0.02, ml, [el],
                           14, A_CELL, A_CELL, 120.0, CC, -1.0 / 3, 1.0 / 3,
                           [deckio.Atom(1, 1.0, 0.0, 0.0, 0.0)])


def surf_deck(rng: np.random.Generator, zs: Sequence[int], natoms: int, height: float,
              b_range: Tuple[float, float] = (0.5, 0.5)) -> deckio.SurfDeck:
    """Seeded surface: natoms adatoms of elements zs (cycled), p1 symmetry, slab `height` A."""
    els = []
    for z in zs:
        b = float(rng.uniform(*b_range)) if b_range[0] != b_range[1] else b_range[0]
        els.append(deckio.Element(z, 1.0, ZMAP.get(z, 0.15), b, b, b))
    atoms = []
    for i in range(natoms):
        x, y = rng.uniform(0.0, 1.0, 2)
        h = rng.uniform(1.0, 2.5)
        # surface z shares the bulk frame shifted up one unit layer: h above the top bulk atom
        atoms.append(deckio.Atom(1 + i % len(zs), 1.0, float(x), float(y), float(h - CC)))
    zoff = 0.5 * CC   # zoffset of the one-atom bulk unit
    zz = max([0.0] + [a.z + zoff for a in atoms])
    dthick = height - zz - CC
    return deckio.SurfDeck(els, 1, 1, 0, 0, 1, dthick, 0.0, 0.0, atoms, [1.0])


def bulk_slab_deck(bd: deckio.BulkDeck, layers: int) -> deckio.SurfDeck:
    """`layers` periodic bulk unit layers stacked above the top bulk unit (each one cc
    higher and shifted by (dx, dy)), as a surface deck with bulk element indices."""
    atoms = []
    for m in range(1, layers + 1):
        for a in bd.atoms:
            atoms.append(deckio.Atom(-a.ielm, a.ocr, a.x + m * bd.dx, a.y + m * bd.dy, a.z + (m - 1) * bd.cc))
    # the slab ends half a unit layer above the last atom plane
    return deckio.SurfDeck([], 1, 1, 0, 0, 1, 0.5 * bd.cc, 0.0, 0.0, atoms, [1.0])


@dataclasses.dataclass
class Structure:
    bulk: deckio.BulkDeck
    surf: deckio.SurfDeck
    bp: deckio.BulkP


def config(cfg: int, seed: int = 1234, index: int = 0) -> Structure:
    """Structure `index` of configuration cfg (1..5), deterministic in (seed, index)."""
    rng = np.random.default_rng([seed, cfg, index])
    if cfg == 1:
        bd = bulk_deck(23, 0.5, 6.0, 0.1)
        sd = surf_deck(rng, [14], 3, 8.0)
    elif cfg == 2:
        bd = bulk_deck(47, 0.5, 6.0, 0.1)
        sd = surf_deck(rng, [14, 8], 5, 12.0, (0.3, 1.0))
    elif cfg == 3:
        bd = bulk_deck(521, 0.5, 4.4, 0.1)
        sd = surf_deck(rng, [14, 8, 1], 12, 9.0)
    elif cfg in (4, 5):
        sizes = [23, 47, 97, 199, 307] + ([521] if cfg == 5 else [])
        nb = int(rng.choice(sizes))
        nat = int(rng.integers(3, 13))
        height = float(rng.uniform(8.0, 20.0))
        nang = int(rng.integers(28 if cfg == 4 else 20, 57))
        bd = bulk_deck(nb, 0.5, 0.5 + 0.1 * (nang - 1) + 1e-6, 0.1)
        sd = surf_deck(rng, [14, 8, 1], nat, height, (0.3, 1.0))
    else:
        raise ValueError("configs are 1..5")
    return Structure(bd, sd, driver.bulk_header(bd))


def bulk_reflection(eng: "lib.Engine", st: Structure, layers: Optional[int] = None, dz2: float = 0.02,
                    method: int = lib.SP4):
    """Bulk reflection blocks [irep1][irep2][group] for a synthetic structure: the reflection
    of `layers` bulk unit layers (the deck's ML by default) integrated from F = 0 on the
    device."""
    slab = bulk_slab_deck(st.bulk, st.bulk.ml if layers is None else layers)
    job = driver.SurfaceJob(eng, st.bp, slab, dz2, method)
    try:
        blocks = []
        for r1 in range(job.nrep[0]):
            G, F, W, _ = job.inputs(r1, None)
            fo, _ = eng.propagate_batch(job.table_id, job.iv, G, F, job.nb0, None)
            row = []
            for p in range(fo.shape[0]):
                grp, ib = [], 0
                for n in job.nbg:
                    idx = job.iord[ib:ib + n] - 1
                    grp.append(fo[p][np.ix_(idx, idx)].copy())
                    ib += n
                row.append(grp)
            blocks.append(row)
    finally:
        job.release()
    return blocks
