"""GPU parity of the batched large-N path (N = 521, the reference's test521 deck).

Tolerance: per beam and normalised L2, max(1e-10, 10 x) the larger of the reference's own
one-ulp response and the spread between two builds of the reference, both measured per
(method, dz) by tests/golden/make_sensitivity.py (test_gpu_fine._bars)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import deckio
from test_gpu_fine import _bars
from test_gpu_surface import _inputs, _oracle_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dz,method", [("0.68", 5), ("0.68", 6), ("0.47", 5), ("0.47", 6)])
def test_surface_n521_vs_oracle_and_reference(engine, dz, method):
    print("case test521", dz, method, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), float(dz), method)
    angs, G, F, W = _inputs(surf)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, method)
    try:
        fo, f2 = engine.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
        st = engine.last_stats()
    finally:
        engine.table_release(tid)
    of, of2 = _oracle_rows(surf, angs, G, F)
    ref = deckio.read_s(os.path.join(REF, "test521", "perf-%d-%s.txt" % (method, dz)))
    e_or = O.rel_err(f2[0], of2[0])
    e_ref = O.rel_err(f2[0], ref.f2[0])
    l2_ref = np.linalg.norm(f2[0] - ref.f2[0]) / np.linalg.norm(ref.f2[0])
    l2_or = np.linalg.norm(f2[0] - of2[0]) / np.linalg.norm(of2[0])
    print("stats", st, "rel vs oracle %.3e (L2 %.3e)  vs reference %.3e (L2 %.3e)" % (e_or, l2_or, e_ref, l2_ref))
    assert st["n_renorm"] >= 1
    bar_b, bar_l2 = _bars("%d-%s" % (method, dz))
    print("bars %.1e / %.1e" % (bar_b, bar_l2))
    assert e_or < bar_b and l2_or < bar_l2
    assert e_ref < bar_b and l2_ref < bar_l2


def test_batched_many_angles_same_as_single(engine):
    """The batch is in lock-step but every angle is independent: solving 6 angles in one
    call gives the same intensities as solving each alone."""
    surf5 = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    a5, G5, F5, W5 = _inputs(surf5)
    scale = np.array([1.0, 0.999, 1.001])
    Gb = np.array([G5[0] * s for s in scale])
    Fb = np.array([F5[0]] * 3)
    Wb = np.array([W5[0]] * 3)
    tid = engine.table_build(surf5.pot, surf5.dz, surf5.ns, 5)
    try:
        _, f2b = engine.propagate_batch(tid, surf5.iv, Gb, Fb, surf5.nb0, Wb)
        singles = [engine.propagate_batch(tid, surf5.iv, Gb[i:i + 1], Fb[i:i + 1], surf5.nb0, Wb[i:i + 1])[1][0]
                   for i in range(3)]
    finally:
        engine.table_release(tid)
    for i in range(3):
        assert O.rel_err(f2b[i], singles[i]) < 1e-12


def test_concurrent_batches_match_single_batch(engine):
    """Angles split over concurrent batches (separate streams, interleaved slices) give
    bit-identical intensities to one batch holding all of them."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    a5, G5, F5, W5 = _inputs(surf)
    scale = np.linspace(0.995, 1.005, 6)
    Gb = np.array([G5[0] * s for s in scale])
    Fb = np.array([F5[0]] * 6)
    Wb = np.array([W5[0]] * 6)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, 5)
    try:
        whole = engine.batch_create(tid, surf.iv, Gb, Fb, surf.nb0, Wb)
        parts = [engine.batch_create(tid, surf.iv, Gb[i:i + 2], Fb[i:i + 2], surf.nb0, Wb[i:i + 2]) for i in (0, 2, 4)]
        engine.batch_run(whole)
        _, f2w = engine.batch_fetch(whole, want_f=False)
        engine.batch_run_many(parts)
        f2p = np.concatenate([engine.batch_fetch(p, want_f=False)[1] for p in parts])
        for b in [whole] + parts:
            engine.batch_release(b)
    finally:
        engine.table_release(tid)
    assert np.array_equal(f2w, f2p)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_phase_offset_batches_match_single_batch(engine, mixed_engine, prec):
    """Concurrent batches under a fixed renormalisation period (every 3rd of 15 slices),
    with phase offsets on (the second and third batch start after the first one's slices 1
    and 2) and off: bit-identical intensities to one batch holding all six angles."""
    e = engine if prec == "fp64" else mixed_engine
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    a5, G5, F5, W5 = _inputs(surf)
    scale = np.linspace(0.995, 1.005, 6)
    Gb = np.array([G5[0] * s for s in scale])
    Fb = np.array([F5[0]] * 6)
    Wb = np.array([W5[0]] * 6)
    tid = e.table_build(surf.pot, surf.dz, surf.ns, 5)
    e.set_renorm_period(3)
    res = {}
    try:
        whole = e.batch_create(tid, surf.iv, Gb, Fb, surf.nb0, Wb)
        parts = [e.batch_create(tid, surf.iv, Gb[i:i + 2], Fb[i:i + 2], surf.nb0, Wb[i:i + 2]) for i in (0, 2, 4)]
        e.batch_run(whole)
        res["whole"] = e.batch_fetch(whole, want_f=False)[1]
        for mode in (1, 0):
            e.set_phase_offsets(mode)
            e.batch_run_many(parts)
            assert e.last_stats()["n_renorm"] == 6 * 5   # each angle at slices 3, 6, 9, 12, 15
            res[mode] = np.concatenate([e.batch_fetch(p, want_f=False)[1] for p in parts])
        for b in [whole] + parts:
            e.batch_release(b)
    finally:
        e.set_phase_offsets(-1)
        e.set_renorm_period(0)
        e.table_release(tid)
    assert np.all(np.isfinite(res["whole"]))
    assert np.array_equal(res[1], res[0])
    assert np.array_equal(res["whole"], res[1])


def test_phase_offsets_with_batches_of_different_tables(mixed_engine):
    """Phase offsets when the batches run different tables: the lead (dz 0.68, 15 slices)
    is shorter than the batch that starts one slice behind it (dz 0.47, 21 slices); a third
    (dz 0.68) starts three slices behind.  Each batch's intensities equal its run alone."""
    e = mixed_engine
    cases = [O.load_case(os.path.join(REF, "test521"), dz, 5) for dz in (0.68, 0.47, 0.68)]
    tids, bids, alone = [], [], []
    e.set_renorm_period(4)
    try:
        for k, surf in enumerate(cases):
            a5, G5, F5, W5 = _inputs(surf)
            G = np.array([G5[0] * (1.0 + 0.002 * k), G5[0] * (0.999 - 0.002 * k)])
            F = np.array([F5[0]] * 2)
            W = np.array([W5[0]] * 2)
            tids.append(e.table_build(surf.pot, surf.dz, surf.ns, 5))
            bids.append(e.batch_create(tids[-1], surf.iv, G, F, surf.nb0, W))
        e.set_phase_offsets(1)
        for b in bids:
            e.batch_run(b)
            alone.append(e.batch_fetch(b, want_f=False)[1])
        e.batch_run_many(bids)
        together = [e.batch_fetch(b, want_f=False)[1] for b in bids]
        for b in bids:
            e.batch_release(b)
    finally:
        e.set_phase_offsets(-1)
        e.set_renorm_period(0)
        for t in tids:
            e.table_release(t)
    for a, t in zip(alone, together):
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, t)


def test_auto_periodic_renormalisation_n521(engine):
    """Automatic period at a coarse dz: the evanescent growth per slice is so fast that the
    derived period is one slice, as the reference's rule does on this deck."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.47, 5)
    angs, G, F, W = _inputs(surf)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, 5)
    engine.set_renorm_period(-1)
    try:
        _, f2 = engine.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
    finally:
        engine.set_renorm_period(0)
        engine.table_release(tid)
    ref = deckio.read_s(os.path.join(REF, "test521", "perf-5-0.47.txt"))
    l2 = np.linalg.norm(f2[0] - ref.f2[0]) / np.linalg.norm(ref.f2[0])
    bar_b, bar_l2 = _bars("5-0.47")
    assert O.rel_err(f2[0], ref.f2[0]) < bar_b and l2 < bar_l2


def _device_curve(eng, surf, method, period=0):
    angs, G, F, W = _inputs(surf)
    tid = eng.table_build(surf.pot, surf.dz, surf.ns, method)
    eng.set_renorm_period(period)
    try:
        _, f2 = eng.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
        st = eng.last_stats()
    finally:
        eng.set_renorm_period(0)
        eng.table_release(tid)
    return f2, st


@pytest.mark.parametrize("dz", ["0.68", "0.47"])
@pytest.mark.parametrize("period", [0, -1])
def test_leapfrog_n521(engine, dz, period):
    """Stormer-Verlet (int_method 4) on the batched path against the reference's perf-4
    curves and the oracle, at the measured N = 521 FP64 bars."""
    print("case test521 leapfrog", dz, period, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), float(dz), 4)
    f2, st = _device_curve(engine, surf, 4, period)
    ref = deckio.read_s(os.path.join(REF, "test521", "perf-4-%s.txt" % dz))
    angs, G, F, W = _inputs(surf)
    _, of2 = _oracle_rows(surf, angs, G, F)
    for other in (ref.f2[0], of2[0]):
        e = O.rel_err(f2[0], other)
        l2 = np.linalg.norm(f2[0] - other) / np.linalg.norm(other)
        bar_b, bar_l2 = _bars("4-%s" % dz)
        print("renorms", st["n_renorm"], "rel %.3e L2 %.3e (bars %.1e / %.1e)" % (e, l2, bar_b, bar_l2))
        assert e < bar_b and l2 < bar_l2


def _rk4_sensitivity(surf):
    """How far the oracle's own RK4 curve moves when the potential table is perturbed by
    one rounding unit: RK4 at dz >= 0.47 is outside its stability region for the strongly
    evanescent beams of this deck, so every rounding difference is amplified."""
    v, vi = O.table(surf)
    r1, r2, ga, az = surf.angles()[0]
    g = O.gamma(surf, ga, az)
    f = surf.scatter_f(r1, r2)
    a = O.intensity(surf, O.integrate(surf, f.copy(), g, v, vi)[0], g, ga)
    b = O.intensity(surf, O.integrate(surf, f.copy(), g, v * (1 + 2.0 ** -52), vi)[0], g, ga)
    return a, O.rel_err(b, a), np.linalg.norm(a - b) / np.linalg.norm(a)


@pytest.mark.parametrize("dz", ["0.68", "0.47"])
def test_rk4_n521(engine, dz):
    """Classical RK4 (int_method 7) on the batched path at steps outside its stability region
    for this deck's strongly evanescent beams.  Bar: within 10x of the oracle's own one-ulp
    sensitivity (measured here: ~1e-3 per beam, ~2e-4 L2 at dz 0.68) against both the oracle
    and the reference's perf-7 curve.  (The stable steps 0.047 and 0.015 are in
    test_gpu_fine.py, at the reference's own noise level.)"""
    print("case test521 rk4", dz, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), float(dz), 7)
    f2, st = _device_curve(engine, surf, 7)
    orc, s_rel, s_l2 = _rk4_sensitivity(surf)
    ref = deckio.read_s(os.path.join(REF, "test521", "perf-7-%s.txt" % dz))
    for other in (ref.f2[0], orc):
        e = O.rel_err(f2[0], other)
        l2 = np.linalg.norm(f2[0] - other) / np.linalg.norm(other)
        print("renorms", st["n_renorm"], "rel %.3e L2 %.3e (one-ulp sensitivity %.3e / %.3e)" % (e, l2, s_rel, s_l2))
        assert e < 10 * s_rel and l2 < 10 * s_l2
    assert st["n_renorm"] == surf.ns


@pytest.mark.parametrize("method", [4, 7])
def test_leapfrog_rk4_many_angles_same_as_single(engine, method):
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, method)
    a, G, F, W = _inputs(surf)
    scale = np.array([1.0, 0.9995, 1.0005, 0.999])
    Gb = np.array([G[0] * s for s in scale])
    Fb = np.array([F[0]] * 4)
    Wb = np.array([W[0]] * 4)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, method)
    try:
        _, f2b = engine.propagate_batch(tid, surf.iv, Gb, Fb, surf.nb0, Wb)
        singles = [engine.propagate_batch(tid, surf.iv, Gb[i:i + 1], Fb[i:i + 1], surf.nb0, Wb[i:i + 1])[1][0]
                   for i in range(4)]
    finally:
        engine.table_release(tid)
    for i in range(4):
        assert O.rel_err(f2b[i], singles[i]) < 1e-12


def test_leapfrog_mixed_n521(engine, mixed_engine):
    """Leapfrog on the int8 digit-sliced products: within 1e-6 of the FP64 device curve."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, 4)
    f64, _ = _device_curve(engine, surf, 4)
    fmx, _ = _device_curve(mixed_engine, surf, 4)
    e = O.rel_err(fmx[0], f64[0])
    print("leapfrog mixed vs FP64 rel %.3e" % e)
    assert e < 1e-6
