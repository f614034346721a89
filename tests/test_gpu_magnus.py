"""GPU parity of the Magnus integrators (int_method 1-3: m2/m4/m6_int_spa,
surf_magnus.f90:114-339) on the batched device path.

The reference ships no curves for these methods, so the oracle is the checker (pinned by
convergence order, test_oracle_math.py).  Tolerance: the FP64 bar, per-intensity relative
error <= 1e-10 (floor 1e-12 * max) at N = 23 / 47; at N = 521 <= 1e-8 per beam and 1e-10
normalised L2 (the level at which FP64 summation order shows there).
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import lib
from test_gpu_surface import _inputs

pytestmark = pytest.mark.gpu


def _device(eng, surf, method, period=0):
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


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("case,dz", [("test23", "0.68"), ("test23", "0.1"), ("test47", "0.68")])
def test_magnus_vs_oracle(engine, case, dz, method):
    print("case", case, dz, "method", method, flush=True)
    surf = O.load_case(os.path.join(REF, case), float(dz), method)
    f2, st = _device(engine, surf, method)
    stats = []
    rows = O.surface_curve(surf, stats=stats)
    e = max(O.rel_err(f2[p], r[1]) for p, r in enumerate(rows))
    ren = int(sum(s[0] for s in stats))
    print("rel vs oracle %.3e, renorms device %d oracle %d" % (e, st["n_renorm"], ren))
    assert e < 1e-10
    # the same Gershgorin rule decides; a decision on the 1e3 edge may flip with rounding
    assert abs(st["n_renorm"] - ren) <= max(1, ren // 100)


@pytest.mark.parametrize("method", [1, 2, 3])
def test_magnus_periodic_matches_rule(engine, method):
    """A fixed renormalisation period is an exact change of variables for these methods too."""
    surf = O.load_case(os.path.join(REF, "test47"), 0.1, method)
    f0, _ = _device(engine, surf, method, 0)
    f4, st = _device(engine, surf, method, 4)
    e = max(O.rel_err(f4[p], f0[p]) for p in range(len(f0)))
    print("period 4 vs rule %.3e" % e)
    assert e < 1e-10


def _one_ulp_response(surf):
    """How far the oracle's own curve moves when the potential table is scaled by
    1 + 2^-52 (per-beam relative, normalised L2): the rounding amplification of the method
    at this deck and dz."""
    v, vi = O.table(surf)
    r1, r2, ga, az = surf.angles()[0]
    g = O.gamma(surf, ga, az)
    f = surf.scatter_f(r1, r2)
    a = O.intensity(surf, O.integrate(surf, f.copy(), g, v, vi)[0], g, ga)
    b = O.intensity(surf, O.integrate(surf, f.copy(), g, v * (1 + 2.0 ** -52), vi)[0], g, ga)
    return O.rel_err(b, a), np.linalg.norm(a - b) / np.linalg.norm(a)


@pytest.mark.parametrize("method,dz", [(1, "0.68"), (2, "0.68"), (3, "0.68"), (2, "0.1")])
def test_magnus_n521_vs_oracle(engine, method, dz):
    """m2/m4 at the N = 521 FP64 bar (1e-8 per beam, 1e-10 normalised L2).  m6 at dz 0.68
    amplifies rounding about 75x more than m4 (a one-ulp change of the table moves the
    oracle's own curve by ~1.5e-8 per beam, 2e-11 L2, against 1.9e-10 / 7e-14 for m4), so
    its bar is 10x that one-ulp response, as for RK4."""
    print("case test521", dz, "method", method, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), float(dz), method)
    f2, st = _device(engine, surf, method)
    stats = []
    rows = O.surface_curve(surf, stats=stats)
    ren = int(sum(s[0] for s in stats))
    e = O.rel_err(f2[0], rows[0][1])
    l2 = np.linalg.norm(f2[0] - rows[0][1]) / np.linalg.norm(rows[0][1])
    print("N=521 method %d dz %s: rel %.3e L2 %.3e renorms device %d oracle %d" % (method, dz, e, l2, st["n_renorm"], ren))
    bar_e, bar_l2 = 1e-8, 1e-10
    if method == 3:
        s_e, s_l2 = _one_ulp_response(surf)
        print("one-ulp response rel %.3e L2 %.3e" % (s_e, s_l2))
        bar_e, bar_l2 = max(bar_e, 10 * s_e), max(bar_l2, 10 * s_l2)
    assert e < bar_e and l2 < bar_l2
    assert abs(st["n_renorm"] - ren) <= max(1, ren // 100)


def test_magnus_in_mixed_context_is_fp64(engine, mixed_engine):
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, 2)
    a, _ = _device(engine, surf, 2)
    b, _ = _device(mixed_engine, surf, 2)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("method", [2, 8])
def test_dense_methods_many_angles_n521_same_as_single(engine, method):
    """Eight angles at N = 521 in one batch: every step of these methods is per angle (the
    products, the renormalisation decision and factorisation), so each angle of the batch
    gives, bit for bit, the intensities it gives alone (checked to 1e-12)."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.68, method)
    a, G, F, W = _inputs(surf)
    scale = np.linspace(0.999, 1.001, 8)
    Gb = np.array([G[0] * s for s in scale])
    Fb = np.array([F[0]] * 8)
    Wb = np.array([W[0]] * 8)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, method)
    try:
        _, f2b = engine.propagate_batch(tid, surf.iv, Gb, Fb, surf.nb0, Wb)
        singles = [engine.propagate_batch(tid, surf.iv, Gb[i:i + 1], Fb[i:i + 1], surf.nb0, Wb[i:i + 1])[1][0]
                   for i in (0, 3, 7)]
    finally:
        engine.table_release(tid)
    for k, i in enumerate((0, 3, 7)):
        print("method %d angle %d: rel %.3e, bit-identical %s" % (method, i, O.rel_err(f2b[i], singles[k]),
                                                                   np.array_equal(f2b[i], singles[k])))
        assert O.rel_err(f2b[i], singles[k]) < 1e-12
