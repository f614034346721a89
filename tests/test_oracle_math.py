"""Pins for the oracle functions that no reference golden vector covers directly.

- Plane-group orbits (strfac.f90:88-527; the decks use only groups 6, 9 and 14): the
  general-position orbit size equals the International Tables multiplicity, the affine
  operations recovered from the orbit form a group modulo lattice translations, and every
  special-position branch returns exactly the distinct images of the general operations.
- matAbdivB (matcomp.f90:53-69): residual of X^{-1} Y against numpy.linalg.solve.
- cond_est (matcomp.f90:498-528): Varah's bound makes the Gershgorin estimate an upper
  bound on cond_1 for column-diagonally-dominant matrices; 1e100 when dominance fails.
- Integrators without a reference vector (Magnus m2/m4/m6, rk4_intgs) and the N = 1 path
  (hp_int_spa1d): observed convergence order against sp6 at fine dz, agreement of rk4_intgs
  with rk4 (which perf-7-*.txt pins), and agreement of the N = 1 sp6 path with the general
  N x N sp4 path run at N = 1.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF

# International Tables for Crystallography vol. A, general-position multiplicities of the
# 17 plane groups in the numbering of strfac.f90:15-65
ITA_MULT = {1: 1, 2: 2, 3: 2, 4: 2, 5: 4, 6: 4, 7: 4, 8: 4, 9: 8, 10: 4, 11: 8, 12: 8,
            13: 3, 14: 6, 15: 6, 16: 6, 17: 12}


def _orbit(nsg, x, y):
    xs = np.zeros(16)
    ys = np.zeros(16)
    n = O.lib().orc_orbit(nsg, float(x), float(y), xs, ys)
    return np.stack([xs[:n], ys[:n]], axis=1)


def _mod1(p):
    q = p - np.floor(p + 1e-7)
    q[q > 1 - 1e-7] = 0.0
    return q


def _same_set(a, b, tol=1e-6):
    a, b = _mod1(np.asarray(a)), _mod1(np.asarray(b))
    if len(a) != len(b):
        return False
    used = np.zeros(len(b), bool)
    for p in a:
        d = np.abs(b - p)
        d = np.minimum(d, 1 - d).max(axis=1)
        k = np.where((d < tol) & ~used)[0]
        if len(k) == 0:
            return False
        used[k[0]] = True
    return True


def _dedup(pts, tol=1e-6):
    out = []
    for p in _mod1(np.asarray(pts)):
        if not any(np.minimum(np.abs(q - p), 1 - np.abs(q - p)).max() < tol for q in out):
            out.append(p)
    return np.array(out)


def _general_ops(nsg):
    """Affine operations (M, t) of the general-position branch, recovered from the orbits
    of three generic points (the branch is affine in (x, y))."""
    p0 = np.array([0.1234, 0.3721])
    e = 1e-3
    o0, ox, oy = _orbit(nsg, *p0), _orbit(nsg, p0[0] + e, p0[1]), _orbit(nsg, p0[0], p0[1] + e)
    ops = []
    for k in range(len(o0)):
        M = np.stack([(ox[k] - o0[k]) / e, (oy[k] - o0[k]) / e], axis=1)
        Mi = np.rint(M)
        assert np.allclose(M, Mi, atol=1e-6)
        ops.append((Mi, o0[k] - Mi @ p0))
    return ops


@pytest.mark.parametrize("nsg", range(1, 18))
def test_orbit_general_multiplicity(nsg):
    orb = _orbit(nsg, 0.1234, 0.3721)
    assert len(orb) == ITA_MULT[nsg]
    assert len(_dedup(orb)) == len(orb)


@pytest.mark.parametrize("nsg", range(1, 18))
def test_orbit_operations_form_a_group(nsg):
    ops = _general_ops(nsg)
    assert any(np.array_equal(M, np.eye(2)) and np.allclose(_mod1(t[None])[0], 0) for M, t in ops)
    for M, t in ops:
        assert abs(abs(np.linalg.det(M)) - 1) < 1e-12
    # closure: the composition of any two operations is one of them modulo the lattice
    for M1, t1 in ops:
        for M2, t2 in ops:
            M, t = M1 @ M2, M1 @ t2 + t1
            assert any(np.array_equal(M, N) and np.allclose(_mod1((t - u)[None])[0] % 1.0, 0, atol=1e-9) or
                       (np.array_equal(M, N) and np.allclose(np.abs(((t - u) + 0.5) % 1.0 - 0.5), 0, atol=1e-9))
                       for N, u in ops), (nsg, M, t)


SPECIAL = [(0, 0), (0.5, 0), (0, 0.5), (0.5, 0.5), (1 / 3, 2 / 3), (2 / 3, 1 / 3), (0.25, 0.25),
           (0.25, 0.75), (0.75, 0.25), (0.3, 0), (0.3, 0.5), (0, 0.3), (0.5, 0.3), (0.3, 0.3),
           (0.3, 0.7), (0.3, 0.25), (0.3, 0.75), (0.25, 0.3), (0.2, 0.4), (0.4, 0.2), (0.2, 0.7),
           (0.15, 0.65), (0.65, 0.15), (0.7, 0.2), (0.3, 0.8), (0.45, 0.55)]


def _p6mm_unlisted_mirror(x, y):
    """Mirror lines of p6mm that strfac.f90:486-527 does not test (it tests the axes and
    the mirrors x = 0, y = 0, x = y): x + y = 1, y = 2x, x = 2y.  On them it keeps all
    twelve general images."""
    near = lambda a, b: abs(a - b) < 2e-4
    axes = [(0, 0), (1 / 3, 2 / 3), (2 / 3, 1 / 3), (0.5, 0), (0, 0.5), (0.5, 0.5)]
    if any(near(x, a) and near(y, b) for a, b in axes) or near(x, 0) or near(y, 0) or near(x, y):
        return False
    return near(x + y, 1) or near(y, 2 * x) or near(x, 2 * y)


@pytest.mark.parametrize("nsg", range(1, 18))
def test_orbit_special_positions_are_distinct_general_images(nsg):
    ops = _general_ops(nsg)
    for x, y in SPECIAL:
        p = np.array([x, y])
        images = [M @ p + t for M, t in ops]
        got = _orbit(nsg, x, y)
        if nsg == 17 and _p6mm_unlisted_mirror(x, y):
            # the reference's own behaviour: the general branch, duplicates included
            assert _same_set(got, np.array(images)) and len(_dedup(images)) == 6, (x, y)
        else:
            assert _same_set(got, _dedup(images)), (nsg, x, y, got, images)


@pytest.mark.parametrize("n", [1, 5, 23, 64])
def test_matAbdivB_solves(n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 3 * np.eye(n)
    Y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    t1 = np.asfortranarray(np.concatenate([X, Y], axis=1))
    buf = np.ascontiguousarray(t1.T.reshape(-1))   # column-major (n, 2n)
    assert O.lib().orc_matAbdivB(n, buf, n) == 0
    Z = buf.reshape(2 * n, n).T[:, n:]
    ref = np.linalg.solve(X, Y)
    assert np.max(np.abs(Z - ref)) / np.max(np.abs(ref)) < 1e-12


def _cond_est(A):
    n = A.shape[0]
    buf = np.ascontiguousarray(np.asfortranarray(A).T.reshape(-1))
    return O.lib().orc_cond_est(n, buf, n)


def test_cond_est_bounds_cond1():
    rng = np.random.default_rng(5)
    for n in (3, 10, 40):
        for _ in range(5):
            A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
            d = np.abs(A).sum(axis=0) * (1.2 + rng.random(n))   # column dominance
            A[np.arange(n), np.arange(n)] = d * np.exp(2j * np.pi * rng.random(n))
            assert _cond_est(A) >= np.linalg.cond(A, 1) * (1 - 1e-12)
    D = np.diag([2.0, 5.0, 0.5, 1.0]).astype(complex)
    assert abs(_cond_est(D) - 10.0) < 1e-12
    B = np.ones((3, 3), complex)
    assert _cond_est(B) == 1e100


def _curve(case_dir, m, dz, pts):
    s = O.load_case(case_dir, dz, m)
    return np.array([r[1] for r in O.surface_curve(s, points=list(pts))])


@pytest.mark.parametrize("method,order", [(1, 2), (2, 4), (3, 6), (8, 4)])
def test_convergence_order_without_reference_vector(method, order):
    d = os.path.join(REF, "test23")
    pts = (3, 20, 40)
    ref = _curve(d, 6, 0.005, pts)
    e = [np.max(np.abs(_curve(d, method, dz, pts) - ref)) / np.max(np.abs(ref)) for dz in (0.08, 0.04, 0.02)]
    orders = [np.log2(e[i] / e[i + 1]) for i in range(2)]
    assert all(abs(o - order) < 0.15 for o in orders), (e, orders)


def test_rk4_gram_schmidt_matches_rk4():
    d = os.path.join(REF, "test23")
    a = _curve(d, 7, 0.04, (3, 20, 40))
    b = _curve(d, 8, 0.04, (3, 20, 40))
    mx = np.max(np.abs(a), axis=1, keepdims=True)
    assert np.max(np.abs(a - b) / np.maximum(np.abs(a), 1e-12 * mx)) < 1e-9


def test_one_beam_path_matches_general_path():
    """sp6 at N = 1 runs hp_int_spa1d; sp4 runs the general N x N routine at N = 1.  Both
    solve the same slice equations: sp4's gap to sp6 at fine dz falls as dz^4."""
    from trhepd import deckio
    D = os.path.join(REF, "ge001")
    bp = deckio.BulkP.read(os.path.join(D, "bulkP.b"))
    sd = deckio.SurfDeck.parse(open(os.path.join(D, "surf.txt")).read())

    def curve(m, dz):
        s = O.setup_surface(bp, sd, dz, m)
        assert s.nb == 1
        return np.array([r[1][0] for r in O.surface_curve(s, points=[5, 20, 40, 60])])

    ref = curve(6, 0.002)
    e = [np.max(np.abs(curve(5, dz) / ref - 1)) for dz in (0.04, 0.02, 0.01)]
    assert e[2] < 1e-9
    assert all(abs(np.log2(e[i] / e[i + 1]) - 4) < 0.2 for i in range(2)), e
    e6 = [np.max(np.abs(curve(6, dz) / ref - 1)) for dz in (0.04, 0.02)]
    assert e6[1] < 1e-11 and np.log2(e6[0] / e6[1]) > 5.5, e6
