"""The reference's own programs, built from its sources by oracle/ref.mk, against the
curves and bulk blocks its authors recorded (so that the build can serve as a checker and
as the CPU baseline), and the oracle against that build where no recorded curve exists
(the Magnus integrators and rk4gs, int_method 1-3 and 8).  Skipped where oracle/_ref has
not been built."""
import os

import numpy as np
import pytest

import bulk as OB
import oracle as O
import refrun
from conftest import REF
from trhepd import deckio

pytestmark = pytest.mark.skipif(not (refrun.available(refrun.SURF) and refrun.available(refrun.BULK)),
                                reason="oracle/_ref not built (make -f oracle/ref.mk)")


def _files(case):
    d = os.path.join(REF, case)
    return open(os.path.join(d, "surf.txt"), "rb").read(), open(os.path.join(d, "bulkP.b"), "rb").read()


def _max_rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    mx = np.max(np.abs(b), axis=1, keepdims=True)
    return float((np.abs(a - b) / np.maximum(np.abs(b), 1e-12 * mx)).max())


@pytest.mark.parametrize("case,dz,method", [("test23", "0.68", 5), ("test23", "0.1", 6), ("test47", "0.68", 4),
                                            ("test47", "0.1", 7)])
def test_built_reference_reproduces_recorded_curves(case, dz, method):
    surf_txt, bulkp = _files(case)
    s, _, _ = refrun.surf(surf_txt, bulkp, float(dz), method, threads=4)
    got = deckio.read_s_text(s).f2
    ref = deckio.read_s(os.path.join(REF, case, "perf-%d-%s.txt" % (method, dz))).f2
    e = _max_rel(got, ref)
    print(case, dz, method, "built vs recorded %.3e" % e)
    assert e < 1e-10


def test_built_bulk_program_reproduces_recorded_blocks():
    """bulk.exe on test23's bulk.txt writes the reference's own bulkP.b: every record but the
    reflection blocks byte for byte, the blocks to FP64 rounding."""
    d = os.path.join(REF, "test23")
    mine, _ = refrun.bulk(open(os.path.join(d, "bulk.txt"), "rb").read(), threads=4)
    a, b = deckio.BulkP.read(mine), deckio.BulkP.read(os.path.join(d, "bulkP.b"))
    worst = 0.0
    for ra, rb in zip(a.domains[0].blocks, b.domains[0].blocks):
        for ga, gb in zip(ra, rb):
            for x, y in zip(ga, gb):
                worst = max(worst, float(np.abs(np.asarray(x) - y).max() / np.abs(y).max()))
    print("bulk blocks %.3e" % worst)
    assert worst < 1e-10
    b.domains[0].blocks = a.domains[0].blocks
    assert b.to_bytes() == mine


@pytest.mark.parametrize("method", [1, 2, 3, 8])
@pytest.mark.parametrize("case,dz", [("test23", 0.68), ("test47", 0.1)])
def test_oracle_magnus_and_rk4gs_vs_built_reference(case, dz, method):
    """No recorded curve exists for int_method 1-3 and 8: the built reference pins the
    oracle's restatement of them (surf_magnus.f90:46-339, surf_rk.f90:93-156)."""
    surf_txt, bulkp = _files(case)
    s, _, _ = refrun.surf(surf_txt, bulkp, dz, method, threads=4)
    ref = deckio.read_s_text(s).f2
    surf = O.load_case(os.path.join(REF, case), dz, method)
    pts = list(range(0, len(surf.angles()), 7))
    rows = O.surface_curve(surf, points=pts)
    e = max(O.rel_err(rows[i][1], ref[p]) for i, p in enumerate(pts))
    print(case, dz, "method", method, "oracle vs built reference %.3e" % e)
    assert e < 1e-10


def test_oracle_bulk_vs_built_reference_on_reference_deck():
    """oracle/bulk.py against bulk.exe on the test47 deck (seven groups of 6-7 beams)."""
    d = os.path.join(REF, "test47")
    mine, _ = refrun.bulk(open(os.path.join(d, "bulk.txt"), "rb").read(), threads=4)
    ref = deckio.BulkP.read(mine)
    pts = [0, 34, 68]
    blocks = OB.Bulk(ref).blocks(points=pts)
    worst = 0.0
    for i, p in enumerate(pts):
        for x, y in zip(blocks[0][p], ref.domains[0].blocks[0][p]):
            worst = max(worst, float(np.abs(np.asarray(x) - y).max() / np.abs(y).max()))
    print("oracle bulk vs built reference %.3e" % worst)
    assert worst < 1e-8


def test_two_domain_bulk_run_reads_uncleared_memory():
    """Why the two-domain fixture takes each domain's blocks from a one-domain run: in a
    two-domain run of the reference's bulk program the potential of the second domain is
    summed into arrays that are allocated but never cleared (bulkio.f90:184-198,
    scpot.f90:183-187), so that domain's blocks depend on what the heap held.  Its first
    domain gets fresh zeroed pages and equals the one-domain run; its second domain does not,
    and with glibc's MALLOC_PERTURB_ filling new allocations the run writes NaN."""
    d = os.path.join(REF, "test23dom")
    two = deckio.BulkP.read(refrun.bulk(open(os.path.join(d, "bulk.txt"), "rb").read(), threads=4)[0])
    fixture = deckio.BulkP.read(os.path.join(d, "bulkP.b"))    # blocks of one-domain runs
    err = []
    for k in range(2):
        err.append(max(np.abs(a - b).max() / np.abs(b).max()
                       for ra, rb in zip(two.domains[k].blocks, fixture.domains[k].blocks)
                       for ga, gb in zip(ra, rb) for a, b in zip(ga, gb)))
    print("two-domain run vs one-domain runs: domain 1 %.3e, domain 2 %.3e" % tuple(err))
    assert err[0] == 0.0 and err[1] > 1e-3
    bad = deckio.BulkP.read(refrun.bulk(open(os.path.join(d, "bulk.txt"), "rb").read(), threads=4,
                                        extra_env={"MALLOC_PERTURB_": "165"})[0])
    assert not np.all(np.isfinite(np.asarray(bad.domains[1].blocks[0][0][0])))
    # the oracle, which clears its potential, gives the one-domain runs' blocks for both
    pts = [0, 40]
    n2 = fixture.nrep[1]
    for k in range(2):
        ob = OB.Bulk(fixture, k).blocks(points=pts)
        e = max(np.abs(a - b).max() / np.abs(b).max() for p in pts
                for a, b in zip(ob[p // n2][p % n2], fixture.domains[k].blocks[p // n2][p % n2]))
        assert e < 1e-10, (k, e)
