"""A two-domain deck and the reference's own output for it (no multi-domain run ships with
the reference).

The deck is test23's (Si(111), 23 beams) with a second domain turned by +60 deg on the
hexagonal lattice: its beams are domain 1's rotated (h, k), RDOM = 0, 60, domain weights
0.6 / 0.4 in surf.txt.  The surface's space group (NSGS 14, p3m1) makes a 120 deg turn a
symmetry, so the 60 deg domain is the twin whose curves differ from the first domain's.  The reference's own programs, built from its sources by
oracle/ref.mk, produce the fixtures:

- bulkP.b: the header records of bulk.exe's two-domain run, with each domain's reflection
  blocks from a bulk.exe run of a one-domain deck holding only that domain (same beams, same
  RDOM).  A two-domain run cannot be used for the blocks of the second domain: bulkio.f90
  allocates the potential arrays v, vi per domain (bulkio.f90:184-185) and scpot adds into
  them (scpot.f90:183-187) without clearing them first, so a domain after the first starts
  from whatever the heap held.  Run with glibc's MALLOC_PERTURB_ set, the same program writes
  NaN blocks; the first domain of a run gets fresh zeroed pages, so one-domain runs are
  well defined.
- surf-bulkP.md / surf-bulkP.s: surf.exe (sp4, dz 0.1, integrator code 5) on that bulkP.b:
  every domain's curve (surfio.f90:276-288) and, through domainsum (domainsum.f90), the
  summed curve.
All are kept in tests/golden/ref/test23dom/ together with the two decks.

    python tests/golden/make_domain_fixture.py
"""
import copy
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "trhepd-opt_amd")]
import refrun  # noqa: E402
from trhepd import deckio  # noqa: E402

SRC = os.path.join(HERE, "ref", "test23")
DST = os.path.join(HERE, "ref", "test23dom")
TURN = 60.0


def rotate(h, k, deg):
    """(h, k) of the hexagonal reciprocal lattice (a = b, gamma = 120 deg) turned by +deg
    (a multiple of 60)"""
    a1 = np.array([1.0, 0.0])
    a2 = np.array([math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)])
    B = 2 * math.pi * np.linalg.inv(np.array([a1, a2])).T
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    g = np.array([[c, -s], [s, c]]) @ (h * B[0] + k * B[1])
    hk = np.linalg.solve(B.T, g)
    assert np.allclose(hk, np.round(hk), atol=1e-9)
    return int(round(hk[0])), int(round(hk[1]))


def decks():
    bd = deckio.BulkDeck.parse(open(os.path.join(SRC, "bulk.txt")).read())
    two = copy.deepcopy(bd)
    rot = [rotate(h, k, TURN) for h, k in zip(bd.ih, bd.ik)]
    two.ndom, two.nb, two.rdom = 2, [bd.nb[0], bd.nb[0]], [0.0, TURN]
    two.ih = list(bd.ih) + [r[0] for r in rot]
    two.ik = list(bd.ik) + [r[1] for r in rot]
    sd = deckio.SurfDeck.parse(open(os.path.join(SRC, "surf.txt")).read(), 1)
    sd.wdom = [0.6, 0.4]
    return two.format(), sd.format()


def one_domain(two, d):
    """the deck of domain d alone"""
    one = copy.deepcopy(two)
    ib = sum(two.nb[:d])
    one.ndom, one.nb, one.rdom = 1, [two.nb[d]], [two.rdom[d]]
    one.ih, one.ik = two.ih[ib:ib + two.nb[d]], two.ik[ib:ib + two.nb[d]]
    return one


def main():
    os.makedirs(DST, exist_ok=True)
    bulk_txt, surf_txt = decks()
    two_raw, _ = refrun.bulk(bulk_txt.encode(), threads=8)
    two = deckio.BulkP.read(two_raw)
    deck = deckio.BulkDeck.parse(bulk_txt)
    for d in range(deck.ndom):
        one = deckio.BulkP.read(refrun.bulk(one_domain(deck, d).format().encode(), threads=8)[0])
        a, b = one.domains[0], two.domains[d]
        assert all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("iord", "nbg", "igh", "igk"))
        b.blocks = a.blocks
    bulkp = two.to_bytes()
    md, s = refrun.surf_domains(surf_txt.encode(), bulkp, 0.1, 5, threads=8)
    for name, data in (("bulk.txt", bulk_txt.encode()), ("surf.txt", surf_txt.encode()), ("bulkP.b", bulkp),
                       ("surf-bulkP.md", md.encode()), ("surf-bulkP.s", s.encode())):
        with open(os.path.join(DST, name), "wb") as fh:
            fh.write(data)
    print("wrote", DST, "bulkP.b", len(bulkp), "bytes")


if __name__ == "__main__":
    main()
