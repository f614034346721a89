"""GPU parity of rk4_intgs (int_method 8, surf_rk.f90:93-156) on the batched device path:
RK4 on [q; p] with left products, the Cholesky-QR renormalisation of bcgs
(matgs.f90:4-37) when the power estimate of ||[q; p]||^2 exceeds 1e2, strans and the final
transposed solve.

The reference ships no curves for this method, so the oracle is the checker (pinned by its
convergence order and its agreement with rk4, test_oracle_math.py).  Tolerance: the FP64 bar,
per-intensity relative error <= 1e-10 (floor 1e-12 * max).
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import driver, lib, synth
from test_gpu_surface import _inputs

pytestmark = pytest.mark.gpu


def _device(eng, surf, period=0):
    angs, G, F, W = _inputs(surf)
    tid = eng.table_build(surf.pot, surf.dz, surf.ns, 8)
    eng.set_renorm_period(period)
    try:
        _, f2 = eng.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
        st = eng.last_stats()
    finally:
        eng.set_renorm_period(0)
        eng.table_release(tid)
    return f2, st


@pytest.mark.parametrize("case,dz", [("test23", "0.68"), ("test23", "0.1"), ("test47", "0.68"), ("test47", "0.1")])
def test_rk4gs_vs_oracle(engine, case, dz):
    print("case", case, dz, flush=True)
    surf = O.load_case(os.path.join(REF, case), float(dz), 8)
    f2, st = _device(engine, surf)
    stats = []
    rows = O.surface_curve(surf, stats=stats)
    e = max(O.rel_err(f2[p], r[1]) for p, r in enumerate(rows))
    ren = int(sum(s[0] for s in stats))
    print("rel vs oracle %.3e, renorms device %d oracle %d" % (e, st["n_renorm"], ren))
    assert e < 1e-10
    assert abs(st["n_renorm"] - ren) <= max(1, ren // 100)


def test_rk4gs_periodic_matches_rule(engine):
    """Cholesky-QR renormalisation is an exact change of variables: every 4th slice gives
    the same curve as the power-estimate rule."""
    surf = O.load_case(os.path.join(REF, "test47"), 0.1, 8)
    f0, _ = _device(engine, surf, 0)
    f4, st = _device(engine, surf, 4)
    e = max(O.rel_err(f4[p], f0[p]) for p in range(len(f0)))
    print("period 4 vs rule %.3e, renorms %d" % (e, st["n_renorm"]))
    assert e < 1e-10


def test_rk4gs_n521_vs_oracle(engine):
    """N = 521 at dz 0.1 (99 slices, 52 renormalisations in the oracle): every Cholesky QR
    runs the blocked factorisation over 33 column blocks with several row-panel tiles per
    step.  RK4 is stable at this dz (a one-ulp change of the potential table moves the
    oracle's curve by 1e-12 per beam), so the N = 521 FP64 bar applies: 1e-8 per beam,
    1e-10 normalised L2, and the same renormalisation count."""
    print("case test521 rk4gs 0.1", flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), 0.1, 8)
    f2, st = _device(engine, surf)
    stats = []
    rows = O.surface_curve(surf, stats=stats)
    ren = int(sum(s[0] for s in stats))
    e = O.rel_err(f2[0], rows[0][1])
    l2 = np.linalg.norm(f2[0] - rows[0][1]) / np.linalg.norm(rows[0][1])
    print("rel %.3e L2 %.3e, renorms device %d oracle %d" % (e, l2, st["n_renorm"], ren))
    assert e < 1e-8 and l2 < 1e-10
    assert st["n_renorm"] == ren


@pytest.mark.parametrize("method", [8, 2])
def test_dense_methods_n97_vs_oracle(engine, method):
    """A seeded synthetic structure with N = 97 at dz 0.02 (4 of its angles)."""
    st = synth.config(4, seed=1234, index=0)
    blocks = synth.bulk_reflection(engine, st, layers=2)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.02, method)
    sel = [0, 9, 18, 27]
    try:
        G, F, W, A = job.inputs(0, blocks)
        sel = [p for p in sel if p < len(A)]
        _, f2 = engine.propagate_batch(job.table_id, job.iv, G[sel], F[sel], job.nb0, W[sel])
        ren = engine.last_stats()["n_renorm"]
    finally:
        job.release()
    bp = st.bp
    bp.domains[0].blocks = blocks
    surf = O.setup_surface(bp, st.surf, 0.02, method)
    rows = O.surface_curve(surf, points=sel)
    err = max(O.rel_err(f2[i], rows[i][1]) for i in range(len(sel)))
    print("N=97 method %d: max rel err vs oracle %.3e, renorms %d" % (method, err, ren))
    assert err < 1e-10


def test_rk4gs_in_mixed_context_is_fp64(engine, mixed_engine):
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, 8)
    a, _ = _device(engine, surf)
    b, _ = _device(mixed_engine, surf)
    assert np.array_equal(a, b)
