"""The small-beam benchmark configurations at their configured shapes, against the reference's
own program (sim-trhepd-rheed surf.exe, built from its sources by oracle/ref.mk) at test time.

C1: the seeded N = 23 structure, all 56 angles, sp4 at dz 0.02 (401 slices).
C2: 16 seeded N = 47 structures x 56 angles = 896 solves run as ONE launch of the small-N
    kernel (batch_run_many fuses batches with the same beams, integrator and step), sp4 at
    dz 0.02 (600 slices).
Both sides start from the device's bulk reflection (synth.bulk_reflection), written into the
reference's bulkP.b, so their inputs are bit-identical.  Bar per point: max(1e-10, 10 x) the
reference's own response when the bulk blocks are multiplied by (1 + 2^-52), from a second
run of the reference (these decks' evanescent beams make the lowest angles sensitive to
rounding at the 1e-10 level).
"""
import copy
import os

import numpy as np
import pytest

import oracle as O
import refrun
from trhepd import deckio, driver, lib, synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not refrun.available(), reason="oracle/_ref not built (make -f oracle/ref.mk)")]

THREADS = 16


def _reference(st, blocks):
    """the reference's curve on (st, blocks), and the same with the blocks times (1 + 2^-52)"""
    bp = copy.deepcopy(st.bp)
    surf_txt = st.surf.format().encode()
    bp.domains[0].blocks = blocks
    base = deckio.read_s_text(refrun.surf(surf_txt, bp.to_bytes(), 0.02, lib.SP4, THREADS)[0]).f2
    bp.domains[0].blocks = [[[np.asarray(x) * (1 + 2.0 ** -52) for x in grp] for grp in row] for row in blocks]
    pert = deckio.read_s_text(refrun.surf(surf_txt, bp.to_bytes(), 0.02, lib.SP4, THREADS)[0]).f2
    return np.asarray(base), np.asarray(pert)


def _check(f2, base, pert, label):
    worst = (0.0, 0, 0.0)
    for p in range(len(base)):
        e = O.rel_err(f2[p], base[p])
        resp = O.rel_err(pert[p], base[p])
        bar = max(1e-10, 10 * resp)
        assert e < bar, "%s angle %d: %.3e vs reference (bar %.1e)" % (label, p, e, bar)
        if e > worst[0]:
            worst = (e, p, resp)
        assert np.all(f2[p] >= 0)
    return worst


def test_c1_all_angles_vs_reference(engine):
    st = synth.config(1, seed=1234)
    print("C1: N", st.bp.nb[0], flush=True)
    blocks = synth.bulk_reflection(engine, st)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.02, lib.SP4)
    try:
        G, F, W, A = job.inputs(0, blocks)
        _, f2 = engine.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)
        ns = job.ns
    finally:
        job.release()
    base, pert = _reference(st, blocks)
    assert base.shape == f2.shape == (56, 23)
    e, p, resp = _check(f2, base, pert, "C1")
    print("C1 ns %d: worst %.3e at angle %d (reference one-ulp response there %.2e)" % (ns, e, p, resp))


def test_c2_896_solves_in_one_launch_vs_reference(engine):
    sts = [synth.config(2, seed=1234, index=i) for i in range(16)]
    assert all(s.bulk.format() == sts[0].bulk.format() for s in sts)   # one bulk deck
    blocks = synth.bulk_reflection(engine, sts[0])
    jobs, bids = [], []
    try:
        for st in sts:
            job = driver.SurfaceJob(engine, st.bp, st.surf, 0.02, lib.SP4)
            jobs.append(job)
            G, F, W, A = job.inputs(0, blocks)
            bids.append(engine.batch_create(job.table_id, job.iv, G, F, job.nb0, W))
        print("C2: %d structures x %d angles, N %d, ns %d" % (len(sts), len(A), job.nb, job.ns), flush=True)
        engine.batch_run_many(bids)
        st_all = engine.last_stats()
        f2 = [engine.batch_fetch(b, want_f=False)[1] for b in bids]
        # the fused launch does per angle exactly what a batch run alone does
        engine.batch_run(bids[5])
        alone = engine.batch_fetch(bids[5], want_f=False)[1]
        assert np.array_equal(alone, f2[5])
    finally:
        for b in bids:
            engine.batch_release(b)
        for job in jobs:
            job.release()
    print("C2 device: %d solves in %.1f ms, %d renormalisations" % (16 * 56, st_all["device_ms"], st_all["n_renorm"]))
    worst = (0.0, 0, 0, 0.0)
    for i, st in enumerate(sts):
        base, pert = _reference(st, blocks)
        assert base.shape == f2[i].shape == (56, 47)
        e, p, resp = _check(f2[i], base, pert, "C2 structure %d" % i)
        if e > worst[0]:
            worst = (e, i, p, resp)
    print("C2 896 solves: worst %.3e (structure %d angle %d, reference one-ulp response %.2e)" % worst)
