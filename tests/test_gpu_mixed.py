"""The mixed-precision path (int8 digit-sliced stage products, FP64 state / LU / table)
against the FP64 path and the reference's FP64 output.

Gate: per-intensity relative error <= 1e-6, floored at 1e-12 of the per-angle maximum
(the accuracy metric of the rocking-curve comparison), on the reference's N = 521 deck and
on the full N = 521 synthetic bench curve (40 angles)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import deckio, driver, lib, synth
from test_gpu_surface import _inputs

pytestmark = pytest.mark.gpu


def _run(eng, surf, method, G, F, W, period=0):
    tid = eng.table_build(surf.pot, surf.dz, surf.ns, method)
    eng.set_renorm_period(period)
    try:
        _, f2 = eng.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
        st = eng.last_stats()
    finally:
        eng.set_renorm_period(0)
        eng.table_release(tid)
    return f2, st


@pytest.mark.parametrize("dz,method", [("0.68", 5), ("0.47", 5), ("0.68", 6)])
def test_mixed_n521_vs_fp64_and_reference(engine, mixed_engine, dz, method):
    print("case test521", dz, method, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), float(dz), method)
    angs, G, F, W = _inputs(surf)
    f64, _ = _run(engine, surf, method, G, F, W)
    fmx, st = _run(mixed_engine, surf, method, G, F, W)
    ref = deckio.read_s(os.path.join(REF, "test521", "perf-%d-%s.txt" % (method, dz)))
    e64 = max(O.rel_err(fmx[p], f64[p]) for p in range(len(angs)))
    eref = max(O.rel_err(fmx[p], ref.f2[p]) for p in range(len(angs)))
    l2 = np.linalg.norm(fmx - f64) / np.linalg.norm(f64)
    print("mixed vs fp64 %.3e (L2 %.3e)  vs reference %.3e  stats %s" % (e64, l2, eref, st))
    assert np.all(np.isfinite(fmx))
    assert e64 < 1e-6 and eref < 1e-6


def test_mixed_small_n_is_the_fp64_path(engine, mixed_engine):
    """Below the large-N threshold the mixed context runs the FP64 small-N kernels."""
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, 5)
    angs, G, F, W = _inputs(surf)
    f64, _ = _run(engine, surf, 5, G, F, W)
    fmx, _ = _run(mixed_engine, surf, 5, G, F, W)
    assert np.array_equal(f64, fmx)


@pytest.mark.parametrize("digits", [6, 7])
def test_mixed_c3_full_curve_vs_fp64(engine, digits):
    """The bench workload: N = 521, 40 angles, sp4 at dz 0.02, periodic renormalisation;
    6 digits is what bench.py runs, 7 the library default."""
    st = synth.config(3, seed=1234)
    blocks = synth.bulk_reflection(engine, st, layers=2)
    out = {}
    mixed_engine = lib.Engine(0, precision=lib.PREC_MIXED)
    mixed_engine.set_mixed_digits(digits)
    for name, eng in (("fp64", engine), ("mixed", mixed_engine)):
        job = driver.SurfaceJob(eng, st.bp, st.surf, 0.02, lib.SP4)
        eng.set_renorm_period(-1)
        try:
            G, F, W, A = job.inputs(0, blocks)
            _, f2 = eng.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)
        finally:
            eng.set_renorm_period(0)
            job.release()
        out[name] = f2
        b00 = job.nb0 - 1
    mixed_engine.close()
    f64, fmx = out["fp64"], out["mixed"]
    errs = np.array([O.rel_err(fmx[p], f64[p]) for p in range(len(f64))])
    l2 = np.linalg.norm(fmx - f64, axis=1) / np.linalg.norm(f64, axis=1)
    i00 = np.abs(fmx[:, b00] - f64[:, b00]) / np.abs(f64[:, b00])
    print("C3 mixed (%d digits) vs fp64: max rel %.3e (median %.3e), max L2 %.3e, (00) curve %.3e"
          % (digits, errs.max(), np.median(errs), l2.max(), i00.max()))
    assert np.all(np.isfinite(fmx))
    assert errs.max() < 1e-6


def test_mixed_leapfrog_c3_vs_fp64(engine):
    """Leapfrog (int_method 4) on the bench structure, 8 of its 40 angles, 6 digits, periodic
    renormalisation (so Q is not the identity between renormalisations): within 1e-6 of FP64."""
    st = synth.config(3, seed=1234)
    blocks = synth.bulk_reflection(engine, st, layers=2)
    out = {}
    mixed_engine = lib.Engine(0, precision=lib.PREC_MIXED)
    mixed_engine.set_mixed_digits(6)
    for name, eng in (("fp64", engine), ("mixed", mixed_engine)):
        job = driver.SurfaceJob(eng, st.bp, st.surf, 0.02, lib.LF)
        eng.set_renorm_period(-1)
        try:
            G, F, W, A = job.inputs(0, blocks)
            G, F, W = G[::5], F[::5], W[::5]
            _, f2 = eng.propagate_batch(job.table_id, job.iv, G, F, job.nb0, W)
            print(name, eng.last_stats())
        finally:
            eng.set_renorm_period(0)
            job.release()
        out[name] = f2
    mixed_engine.close()
    errs = np.array([O.rel_err(out["mixed"][p], out["fp64"][p]) for p in range(len(out["fp64"]))])
    print("C3 leapfrog mixed vs fp64: max rel %.3e" % errs.max())
    assert np.all(np.isfinite(out["mixed"]))
    assert errs.max() < 1e-6


def test_every_stage_tile_gives_the_same_bits(monkeypatch):
    """Each int8 stage-product tile (column blocks x row blocks x Q prefetch depth), pinned one
    at a time, gives bit-identical
    intensities: the digit sums are exact integers and every tile combines them in the same
    order.  Five angles of the N = 521 deck, so that several row tiles run."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    angs, G, F, W = _inputs(surf)
    scale = np.array([1.0, 0.999, 1.001, 0.998, 1.002])
    Gb, Fb, Wb = np.array([G[0] * s for s in scale]), np.array([F[0]] * 5), np.array([W[0]] * 5)
    out = {}
    for v in range(12):
        monkeypatch.setenv("TRHEPD_STAGE_VARIANT", str(v))
        with lib.Engine(0, precision=lib.PREC_MIXED) as eng:
            eng.set_mixed_digits(6)
            try:
                out[v], _ = _run(eng, surf, 5, Gb, Fb, Wb)
            except lib.TrhepdError as e:   # a tile the registers do not hold
                assert e.code == lib.E_ARG
                continue
        print("variant", v, "done", flush=True)
    assert len(out) >= 8
    for v, f2 in out.items():
        assert np.array_equal(f2, out[0]), v
