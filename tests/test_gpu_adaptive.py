"""Adaptive slice thickness on the device: tables over slice lists (trhepd_table_build_steps)
on both slice loops, against the oracle's integration over the same list, and the pilot that
chooses the list (trhepd_adaptive_steps) on the reference's N = 521 deck against the
reference's converged curve (sp4 at dz 0.001).

Bars: N = 23 / 47 1e-10 per intensity (floor 1e-12 of the angle's maximum), as for the uniform
schedule; N = 521: 10x the oracle's own response to a one-ulp change of the table on the same
slice list, at least 1e-10."""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import deckio, lib
from test_gpu_surface import _inputs

pytestmark = pytest.mark.gpu


def _err(a, b):
    return O.rel_err(a, b), float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _slice_list(zmax, n, seed):
    """n slices with thicknesses spread over [0.6, 1.4] of the mean, summing to zmax"""
    w = np.random.default_rng(seed).uniform(0.6, 1.4, n)
    return w * (zmax / w.sum())


def _oracle_f2(surf, steps, G, F, angs, pts, perturb=0.0):
    v, vi = O.table_steps(surf, steps)
    if perturb:
        v = v * (1 + perturb)
    out = []
    for p in pts:
        fo, _ = O.integrate_steps(surf, steps, F[p].copy(), G[p], v, vi)
        out.append(O.intensity(surf, fo, G[p], angs[p][2]))
    return np.array(out)


@pytest.mark.parametrize("case,method", [("test23", 5), ("test23", 6), ("test47", 5)])
def test_slice_list_small_n_vs_oracle(engine, case, method):
    surf = O.load_case(os.path.join(REF, case), 0.1, method)
    steps = _slice_list(surf.dz * surf.ns, 60, seed=7)
    angs, G, F, W = _inputs(surf)
    pts = list(range(0, len(angs), 8))
    tid = engine.table_build_steps(surf.pot, steps, method)
    try:
        _, f2 = engine.propagate_batch(tid, surf.iv, G[pts], F[pts], surf.nb0, W[pts])
    finally:
        engine.table_release(tid)
    of2 = _oracle_f2(surf, steps, G, F, angs, pts)
    e = max(_err(f2[i], of2[i])[0] for i in range(len(pts)))
    print("%s method %d, 60 slices %.3f..%.3f: device vs oracle %.3e" % (case, method, steps.min(), steps.max(), e))
    assert e < 1e-10


def test_slice_list_n521_vs_oracle(engine, mixed_engine):
    """test521's slab (one angle) in 21 slices of uneven thickness (the uniform schedule at
    dz 0.47 has 21), on the FP64 batched path against the oracle, and the mixed path against
    FP64."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.47, 5)
    steps = _slice_list(surf.dz * surf.ns, surf.ns, seed=11)
    angs, G, F, W = _inputs(surf)
    pts = [0]
    tid = engine.table_build_steps(surf.pot, steps, 5)
    try:
        _, f2 = engine.propagate_batch(tid, surf.iv, G[pts], F[pts], surf.nb0, W[pts])
        _, f64 = engine.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
    finally:
        engine.table_release(tid)
    tidm = mixed_engine.table_build_steps(surf.pot, steps, 5)
    try:
        _, fmx = mixed_engine.propagate_batch(tidm, surf.iv, G, F, surf.nb0, W)
    finally:
        mixed_engine.table_release(tidm)
    of2 = _oracle_f2(surf, steps, G, F, angs, pts)
    of2p = _oracle_f2(surf, steps, G, F, angs, pts, perturb=2.0 ** -52)
    for i, p in enumerate(pts):
        e, l2 = _err(f2[i], of2[i])
        re, rl2 = _err(of2p[i], of2[i])
        print("angle %d: device vs oracle %.3e (L2 %.3e); oracle one-ulp response %.2e (L2 %.2e)" % (p, e, l2, re, rl2))
        assert e < max(1e-10, 10 * re) and l2 < max(1e-10, 10 * rl2)
    emx = max(O.rel_err(fmx[p], f64[p]) for p in range(len(angs)))
    print("mixed vs fp64 on the slice list: %.3e" % emx)
    assert np.all(np.isfinite(fmx)) and emx < 1e-6


def test_equal_slice_list_matches_uniform_table(engine):
    """A list of ns equal slices runs the uniform schedule: the two differ only by the
    rounding of the summed slice starts (bar: 10x the reference's own one-ulp response at
    this step, tests/golden/ref/test521/sensitivity.json)."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.47, 5)
    angs, G, F, W = _inputs(surf)
    tu = engine.table_build(surf.pot, surf.dz, surf.ns, 5)
    tl = engine.table_build_steps(surf.pot, [surf.dz] * surf.ns, 5)
    try:
        _, fu = engine.propagate_batch(tu, surf.iv, G, F, surf.nb0, W)
        _, fl = engine.propagate_batch(tl, surf.iv, G, F, surf.nb0, W)
    finally:
        engine.table_release(tu)
        engine.table_release(tl)
    runs = json.load(open(os.path.join(REF, "test521", "sensitivity.json")))["runs"]
    resp = max(runs["5-0.47"]["response"][0], runs["5-0.47"]["build_vs_recorded"][0])
    e = max(O.rel_err(fl[p], fu[p]) for p in range(len(angs)))
    print("equal slice list vs uniform table: %.3e (bar %.2e)" % (e, max(1e-10, 10 * resp)))
    assert e < max(1e-10, 10 * resp)


def test_adaptive_pilot_n521(engine):
    """The pilot on test521's angle (tolerance 1e-7 on the step-doubling estimate): the list
    covers the slab, respects its bounds, is reproducible, and the intensities over it are
    compared with the reference's converged curve beside the uniform schedule with the same
    number of slices."""
    surf = O.load_case(os.path.join(REF, "test521"), 0.1, 5)
    zmax = surf.dz * surf.ns
    angs, G, F, W = _inputs(surf)
    pil = [0]
    hmin, hmax = 0.01, 1.0
    steps, nforced = engine.adaptive_steps(surf.pot, 5, zmax, 1e-7, 0.1, hmin, hmax, surf.iv, G[pil], F[pil],
                                           surf.nb0)
    steps2, _ = engine.adaptive_steps(surf.pot, 5, zmax, 1e-7, 0.1, hmin, hmax, surf.iv, G[pil], F[pil], surf.nb0)
    n = len(steps)
    print("pilot: %d slices, %.4f..%.4f A, %d forced at hmin" % (n, steps.min(), steps.max(), nforced))
    assert np.array_equal(steps, steps2)
    assert abs(steps.sum() - zmax) < 1e-9 * zmax
    assert np.all(steps[:-1] >= hmin * (1 - 1e-12)) and np.all(steps <= hmax)
    pts = list(range(len(angs)))
    acc = deckio.read_s(os.path.join(REF, "test521", "sp4-0.001.s"))
    out = {}
    for name, lst in (("adaptive", steps), ("uniform", np.full(n, zmax / n))):
        tid = engine.table_build_steps(surf.pot, lst, 5)
        try:
            _, f2 = engine.propagate_batch(tid, surf.iv, G[pts], F[pts], surf.nb0, W[pts])
        finally:
            engine.table_release(tid)
        errs = [_err(f2[i], acc.f2[p]) for i, p in enumerate(pts)]
        out[name] = (max(e[0] for e in errs), max(e[1] for e in errs))
        print("%s (%d slices) vs converged: max rel %.3e, L2 %.3e" % (name, n, out[name][0], out[name][1]))
    # a list chosen from local estimates is not guaranteed to beat equal slices of the same
    # count; it must stay of the same accuracy
    assert out["adaptive"][1] < 3 * out["uniform"][1] + 1e-9


def test_refusals(engine):
    """Slice lists are for sp4/sp6 with positive finite thicknesses; the pilot needs the
    batched path (N > 48) and consistent bounds."""
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, 5)
    for steps, method in (([0.1, 0.0, 0.2], 5), ([0.1, -0.1], 5), ([0.1, float("nan")], 6), ([0.1, 0.2], 4),
                          ([0.1, 0.2], 7)):
        with pytest.raises(lib.TrhepdError) as e:
            engine.table_build_steps(surf.pot, steps, method)
        assert e.value.code == lib.E_ARG
    angs, G, F, W = _inputs(surf)
    with pytest.raises(lib.TrhepdError) as e:   # N = 23: the pilot needs the batched path
        engine.adaptive_steps(surf.pot, 5, 1.0, 1e-6, 0.1, 0.01, 1.0, surf.iv, G[:1], F[:1], surf.nb0)
    assert e.value.code == lib.E_ARG
    big = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    ab, Gb, Fb, Wb = _inputs(big)
    for args in ((4, 1.0, 1e-6, 0.1, 0.01, 1.0), (5, 1.0, 1e-6, 0.1, 0.2, 1.0), (5, 1.0, 0.0, 0.1, 0.01, 1.0),
                 (5, -1.0, 1e-6, 0.1, 0.01, 1.0)):
        with pytest.raises(lib.TrhepdError) as e:
            engine.adaptive_steps(big.pot, args[0], args[1], args[2], args[3], args[4], args[5], big.iv, Gb, Fb,
                                  big.nb0)
        assert e.value.code == lib.E_ARG
    with pytest.raises(lib.TrhepdError) as e:   # more slices than max_steps
        engine.adaptive_steps(big.pot, 5, big.dz * big.ns, 1e-6, 0.05, 0.05, 0.05, big.iv, Gb, Fb, big.nb0,
                              max_steps=3)
    assert e.value.code == lib.E_ARG
