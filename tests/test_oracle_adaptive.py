"""Adaptive slice thickness: the reference's step-size controller and integration over slice
lists, on the CPU.

- h211b_limit / h211b_filter (sim-trhepd-rheed/src/h211b_control.f90:1-16) as the oracle and
  the library restate them, against the reference's own functions compiled from that file
  (oracle/_ref/libh211b.so, built by oracle/ref.mk) and against the closed forms they obey;
- the depth schedule of a table over a slice list, and the oracle's RKN integration over it:
  a list of equal slices gives the uniform integration, and a non-uniform list converges to
  the reference's converged curve (sp4 at dz 0.001) at the integrator's order.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF, REPO
from trhepd import deckio, lib

REFLIB = os.path.join(REPO, "oracle", "_ref", "libh211b.so")

US = [-50.0, -3.0, -1.0, 0.0, 0.25, 0.9, 1.0, 1.0 + 1e-9, 1.1, 2.0, 7.5, 1e3]
CASES = [(2.0, 2.0, 1.0, 5), (0.5, 3.0, 1.3, 5), (1e-3, 0.2, 0.7, 7), (40.0, 1e4, 2.5, 7), (1.0, 1.0, 1.0, 5),
         (0.9, 1.7, 0.95, 3)]


def _reference_lib():
    if not os.path.exists(REFLIB):
        pytest.skip("oracle/_ref/libh211b.so is built by oracle/ref.mk from the reference sources")
    L = C.CDLL(REFLIB)
    dp = C.POINTER(C.c_double)
    L.h211b_limit_.argtypes = [dp]
    L.h211b_limit_.restype = C.c_double
    L.h211b_filter_.argtypes = [dp, dp, dp, C.POINTER(C.c_int)]
    L.h211b_filter_.restype = C.c_double
    return L


def test_controller_matches_the_reference_build():
    L = _reference_lib()
    for u in US:
        ref = L.h211b_limit_(C.byref(C.c_double(u)))
        assert O.h211b_limit(u) == pytest.approx(ref, rel=2e-16, abs=0), u
        assert lib.h211b_limit(u) == pytest.approx(ref, rel=2e-16, abs=0), u
    for c1, c0, rho, k in CASES:
        ref = L.h211b_filter_(C.byref(C.c_double(c1)), C.byref(C.c_double(c0)), C.byref(C.c_double(rho)),
                              C.byref(C.c_int(k)))
        assert O.h211b_filter(c1, c0, rho, k) == pytest.approx(ref, rel=4e-16, abs=0), (c1, c0, rho, k)
        assert lib.h211b_filter(c1, c0, rho, k) == pytest.approx(ref, rel=4e-16, abs=0), (c1, c0, rho, k)


@pytest.mark.parametrize("impl", [O, lib])
def test_controller_closed_forms(impl):
    limit, filt = impl.h211b_limit, impl.h211b_filter
    # the limiter: fixed point 1, odd about 1, slope 1 there, asymptotes 1 +- pi/2
    assert limit(1.0) == 1.0
    for x in (1e-3, 0.3, 2.0, 40.0):
        assert limit(1.0 + x) + limit(1.0 - x) == pytest.approx(2.0, abs=1e-15)
    assert (limit(1.0 + 1e-6) - limit(1.0 - 1e-6)) / 2e-6 == pytest.approx(1.0, abs=1e-9)
    assert limit(1e12) == pytest.approx(1.0 + math.pi / 2, abs=1e-11)
    assert limit(-1e12) == pytest.approx(1.0 - math.pi / 2, abs=1e-11)
    # the filter: exponent 1/(4k) on each error ratio, -1/k on the step ratio
    for k in (3, 5, 7):
        for c in (0.01, 0.7, 3.0, 1e5):
            assert filt(c, c, 1.0, k) == pytest.approx(c ** (1.0 / (2 * k)), rel=1e-15)
            assert filt(c, 1.0, 1.0, k) == pytest.approx(c ** (1.0 / (4 * k)), rel=1e-15)
        for rho in (0.5, 1.0, 1.7):
            assert filt(1.0, 1.0, rho, k) == pytest.approx(rho ** (-1.0 / k), rel=1e-15)


def test_slice_list_depths():
    """Equal slices give the uniform schedule (to the rounding of summing the thicknesses),
    z0 shifts every depth, the last column ends the slab."""
    for method, s in ((5, 6), (6, 11)):
        dz, ns = 0.1001455481567265, 17
        uni = O.table_depths(method, dz, ns)
        lst = O.table_depths_steps(method, [dz] * ns)
        assert len(lst) == len(uni) == s * ns + 1
        assert np.max(np.abs(lst - uni)) < 1e-14
        steps = np.array([0.05, 0.2, 0.11, 0.3, 0.07])
        d0 = O.table_depths_steps(method, steps)
        d1 = O.table_depths_steps(method, steps, z0=2.5)
        assert np.max(np.abs(d1 - d0 - 2.5)) < 1e-14
        assert d0[-1] == pytest.approx(steps.sum(), abs=1e-15)
        # slice l's first column is the start of the slice
        assert np.allclose(d0[:-1:s], np.concatenate([[0.0], np.cumsum(steps)[:-1]]), atol=1e-15)


def _curve_steps(surf, steps, pts):
    return [r[1] for r in O.surface_curve_steps(surf, steps, pts)]


def test_equal_slices_give_the_uniform_curve():
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, 5)
    pts = list(range(0, 69, 17))
    uni = [r[1] for r in O.surface_curve(surf, pts)]
    lst = _curve_steps(surf, [surf.dz] * surf.ns, pts)
    err = max(O.rel_err(a, b) for a, b in zip(lst, uni))
    print("equal slices vs uniform: %.3e" % err)
    assert err < 1e-11


def test_non_uniform_slices_converge_at_fourth_order():
    """Alternating thicknesses 0.7 h / 1.3 h over test23's slab, then each slice halved: the
    error against the reference's sp4 curve at dz 0.001 falls by about 2^4."""
    surf = O.load_case(os.path.join(REF, "test23"), 0.1, 5)
    zmax = surf.dz * surf.ns
    acc = deckio.read_s(os.path.join(REF, "test23", "sp4-0.001.s"))
    pts = list(range(0, 69, 8))
    errs = []
    for n in (20, 40):
        h = zmax / n
        steps = np.array([0.7 * h if i % 2 == 0 else 1.3 * h for i in range(n)])
        steps *= zmax / steps.sum()
        rows = _curve_steps(surf, steps, pts)
        errs.append(max(np.linalg.norm(r - acc.f2[p]) / np.linalg.norm(acc.f2[p]) for r, p in zip(rows, pts)))
    print("non-uniform slices: L2 error %.3e (20 slices) %.3e (40 slices), ratio %.1f" %
          (errs[0], errs[1], errs[0] / errs[1]))
    assert errs[0] / errs[1] > 10.0
