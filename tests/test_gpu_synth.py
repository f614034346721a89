"""GPU parity on the seeded synthetic decks of the benchmark configurations.

The bulk reflection that starts the surface loop is computed on the device (a two-layer
bulk slab integrated from F = 0) and handed to both the HIP path and the oracle, so the
comparison isolates the surface slice loop.  Tolerance: 1e-10 per intensity (floored at
1e-12 of the per-angle maximum)."""
import numpy as np
import pytest

import oracle as O
from trhepd import driver, lib, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,points", [(1, None), (2, [0, 9, 18, 27, 36, 45, 55])])
def test_synthetic_config_vs_oracle(engine, cfg, points):
    print("synthetic config", cfg, flush=True)
    st = synth.config(cfg, seed=7)
    blocks = synth.bulk_reflection(engine, st, layers=2)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.02, lib.SP4)
    try:
        G, F, W, A = job.inputs(0, blocks)
        sel = list(range(len(A))) if points is None else points
        _, f2 = engine.propagate_batch(job.table_id, job.iv, G[sel], F[sel], job.nb0, W[sel])
    finally:
        job.release()
    bp = st.bp
    bp.domains[0].blocks = blocks
    surf = O.setup_surface(bp, st.surf, 0.02, lib.SP4)
    assert surf.ns == job.ns and surf.nv == job.nv
    rows = O.surface_curve(surf, points=sel)
    err = max(O.rel_err(f2[i], rows[i][1]) for i in range(len(sel)))
    print("config", cfg, "max rel err vs oracle %.3e" % err)
    assert err < 1e-10
    assert np.all(f2 >= 0)


def test_resident_batch_rerun_is_deterministic(engine):
    st = synth.config(1, seed=3)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.1, lib.SP4)
    try:
        G, F, W, A = job.inputs(0, None)
        bid = engine.batch_create(job.table_id, job.iv, G[:8], F[:8], job.nb0, W[:8])
        engine.batch_run(bid)
        f_a, f2_a = engine.batch_fetch(bid)
        engine.batch_run(bid)
        f_b, f2_b = engine.batch_fetch(bid)
        engine.batch_release(bid)
    finally:
        job.release()
    assert np.array_equal(f2_a, f2_b) and np.array_equal(f_a, f_b)


def test_c3_periodic_schedule_vs_reference_rule(engine):
    """The bench configuration (N = 521, synthetic deck): the automatic periodic
    renormalisation agrees with the reference's Gershgorin rule to the FP64 bar."""
    st = synth.config(3)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.02, lib.SP4)
    try:
        G, F, W, A = job.inputs(0, None)
        sel = [0, 13, 26, 39]
        out = {}
        for period in (0, -1):
            engine.set_renorm_period(period)
            _, f2 = engine.propagate_batch(job.table_id, job.iv, G[sel], F[sel], job.nb0, W[sel])
            out[period] = (f2, engine.last_stats()["n_renorm"])
    finally:
        engine.set_renorm_period(0)
        job.release()
    import oracle as O
    err = max(O.rel_err(out[-1][0][i], out[0][0][i]) for i in range(len(sel)))
    l2 = max(np.linalg.norm(out[-1][0][i] - out[0][0][i]) / np.linalg.norm(out[0][0][i]) for i in range(len(sel)))
    print("renorms reference rule %d, periodic %d; rel %.3e L2 %.3e" % (out[0][1], out[-1][1], err, l2))
    assert l2 < 1e-10 and err < 1e-8


@pytest.mark.parametrize("nb", [49, 64, 600])
def test_edge_beam_counts_renormalisation_invariance(engine, nb):
    """The batched path at its smallest N, at an N with no row padding (NP = N) and at the
    largest N (the LU panel's third row per thread, 38 back-substitution blocks): the
    renormalisation is an exact change of variables, so a renormalisation after every
    slice, every fourth slice and the reference's Gershgorin rule give the same curve to the
    FP64 bar (a mistake in the LU would not cancel between the schedules)."""
    print("N", nb, flush=True)
    rng = np.random.default_rng([11, nb])
    bd = synth.bulk_deck(nb, 0.5, 0.8 + 1e-6, 0.1)
    st = synth.Structure(bd, synth.surf_deck(rng, [14, 8], 3, 6.0), driver.bulk_header(bd))
    blocks = synth.bulk_reflection(engine, st, layers=1)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.04, lib.SP4)
    out = {}
    try:
        G, F, W, A = job.inputs(0, blocks)
        assert job.nb0 >= 1 and len(A) == 4
        for period in (1, 4, 0):
            engine.set_renorm_period(period)
            _, f2 = engine.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)
            out[period] = (f2, engine.last_stats()["n_renorm"])
    finally:
        engine.set_renorm_period(0)
        job.release()
    ref = out[0][0]
    for period in (1, 4):
        f2, nren = out[period]
        err = max(O.rel_err(f2[i], ref[i]) for i in range(len(A)))
        l2 = max(np.linalg.norm(f2[i] - ref[i]) / np.linalg.norm(ref[i]) for i in range(len(A)))
        print("period %d: %d renormalisations (rule: %d); rel %.3e L2 %.3e" % (period, nren, out[0][1], err, l2))
        assert np.all(np.isfinite(f2)) and l2 < 1e-10 and err < 1e-8
    assert out[1][1] > out[4][1]


@pytest.mark.parametrize("nb", [64, 600])
@pytest.mark.parametrize("method", [lib.M2, lib.RK4GS])
def test_dense_methods_edge_beam_counts_renormalisation_invariance(engine, nb, method):
    """m2 (Gershgorin rule) and rk4gs (power-estimate rule, Cholesky-QR renormalisation) at
    an N with no padding and at the largest N (38 column blocks in the blocked Cholesky and
    triangular solves, 38 row blocks in the split power estimate): renormalising after every
    slice, every fourth slice and by the method's own rule gives the same curve to the FP64
    bar."""
    print("N", nb, "method", method, flush=True)
    rng = np.random.default_rng([13, nb])
    bd = synth.bulk_deck(nb, 0.5, 0.8 + 1e-6, 0.1)
    st = synth.Structure(bd, synth.surf_deck(rng, [14, 8], 3, 6.0), driver.bulk_header(bd))
    blocks = synth.bulk_reflection(engine, st, layers=1)
    job = driver.SurfaceJob(engine, st.bp, st.surf, 0.04, method)
    out = {}
    try:
        G, F, W, A = job.inputs(0, blocks)
        for period in (1, 4, 0):
            engine.set_renorm_period(period)
            _, f2 = engine.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)
            out[period] = (f2, engine.last_stats()["n_renorm"])
    finally:
        engine.set_renorm_period(0)
        job.release()
    ref = out[0][0]
    for period in (1, 4):
        f2, nren = out[period]
        err = max(O.rel_err(f2[i], ref[i]) for i in range(len(A)))
        l2 = max(np.linalg.norm(f2[i] - ref[i]) / np.linalg.norm(ref[i]) for i in range(len(A)))
        print("period %d: %d renormalisations (rule: %d); rel %.3e L2 %.3e" % (period, nren, out[0][1], err, l2))
        assert np.all(np.isfinite(f2)) and l2 < 1e-10 and err < 1e-8
    assert out[1][1] > out[4][1]


@pytest.mark.parametrize("nb", [49, 600])
def test_mixed_edge_beam_counts_vs_fp64(engine, nb):
    """The int8 digit path at the smallest and the largest batched N (column-block and k
    padding of the digit fragments) against the FP64 path, at 6 digits: within the mixed
    bar of 1e-6 per intensity."""
    print("N", nb, flush=True)
    rng = np.random.default_rng([12, nb])
    bd = synth.bulk_deck(nb, 0.5, 0.8 + 1e-6, 0.1)
    st = synth.Structure(bd, synth.surf_deck(rng, [14, 8], 3, 6.0), driver.bulk_header(bd))
    blocks = synth.bulk_reflection(engine, st, layers=1)
    out = {}
    for prec in (lib.PREC_FP64, lib.PREC_MIXED):
        with lib.Engine(0, precision=prec) as eng:
            eng.set_renorm_period(-1)
            if prec == lib.PREC_MIXED:
                eng.set_mixed_digits(6)
            job = driver.SurfaceJob(eng, st.bp, st.surf, 0.04, lib.SP4)
            try:
                G, F, W, A = job.inputs(0, blocks)
                out[prec] = eng.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)[1]
            finally:
                job.release()
    f64, fmx = out[lib.PREC_FP64], out[lib.PREC_MIXED]
    err = max(O.rel_err(fmx[i], f64[i]) for i in range(len(f64)))
    print("mixed vs fp64 %.3e" % err)
    assert np.all(np.isfinite(fmx)) and err < 1e-6
