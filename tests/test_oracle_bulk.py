"""Pins of the bulk-stage oracle (oracle/bulk.py).

- Golden: the reflection blocks the reference's build wrote into bulkP.b (Ge(001) sample,
  test23, test47, and the well-conditioned groups of test521).
- Mathematics: the slice matrix of trmat equals Cm exp(dz i[[0, I], [H, 0]]) Cm^-1 with
  Cm = [[Gamma, I], [Gamma, -I]] (scipy's Pade expm, an independent route to the eigen form);
  the one-beam case has the closed form of a scalar wave; the layer recursion's converged
  R is a fixed point of R = (T11 R + T12)(T21 R + T22)^-1 and its first layer is T12 T22^-1.
"""
import os

import numpy as np
import pytest
import scipy.linalg

import bulk as OB
from conftest import REF
from trhepd import deckio


def _err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _worst(case, pts):
    bp = deckio.BulkP.read(os.path.join(REF, case, "bulkP.b"))
    B = OB.Bulk(bp)
    bl = B.blocks(points=pts)
    _, n1, n2 = B.angles()
    ref = bp.domains[0].blocks
    return max(_err(a, b) for p in pts for a, b in zip(bl[p // n2][p % n2], ref[p // n2][p % n2]))


@pytest.mark.parametrize("case,pts,tol", [
    ("ge001", list(range(70)), 1e-12),
    ("test23", [0, 20, 44, 68], 1e-11),
    ("test47", [0, 30, 68], 5e-9),
])
def test_oracle_bulk_matches_reference_file(case, pts, tol):
    assert _worst(case, pts) <= tol


def test_oracle_bulk_521_groups():
    bp = deckio.BulkP.read(os.path.join(REF, "test521", "bulkP.b"))
    bl = OB.Bulk(bp).blocks()[0][0]
    ref = bp.domains[0].blocks[0][0]
    errs = sorted(_err(a, b) for a, b in zip(bl, ref))
    # the well-conditioned groups agree to rounding; the strongly evanescent ones do not
    assert errs[6] <= 2e-9


def _slice_expm(h, g, dz):
    n = len(g)
    I = np.eye(n)
    A = 1j * np.block([[np.zeros((n, n)), I], [h, np.zeros((n, n))]])
    Cm = np.block([[np.diag(g), I], [np.diag(g), -I]])
    return Cm @ scipy.linalg.expm(dz * A) @ np.linalg.inv(Cm)


def test_trmat_equals_exponential():
    rng = np.random.default_rng(7)
    for n in (2, 3, 6):
        g = rng.uniform(0.5, 3.0, n) + 1j * rng.uniform(0.0, 0.5, n)
        h = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 0.5 + np.diag(g * g)
        t = OB.trmat(h, g, 0.05)
        assert _err(t, _slice_expm(h, g, 0.05)) < 1e-12


def test_trmat_one_beam_closed_form():
    g, h, dz = 1.7 + 0.1j, 2.3 + 0.4j, 0.11
    k = np.sqrt(h)
    c, s = np.cos(k * dz), np.sin(k * dz)
    E = np.array([[c, 1j * s / k], [1j * k * s, c]])
    Cm = np.array([[g, 1], [g, -1]])
    ref = Cm @ E @ np.linalg.inv(Cm)
    assert _err(OB.trmat(np.array([[h]]), np.array([g]), dz), ref) < 1e-14


class _Rec(OB.Bulk):
    def __init__(self, ml, epsb):
        class _P:
            pass
        self.bp = _P()
        self.bp.ml, self.bp.epsb = ml, epsb


def test_recursion_fixed_point_and_first_layer():
    rng = np.random.default_rng(3)
    n = 4
    # a contracting layer: T22 dominant so the recursion converges
    T = np.zeros((2 * n, 2 * n), dtype=complex)
    T[:n, :n] = 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    T[:n, n:] = 0.5 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    T[n:, :n] = 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / n
    T[n:, n:] = 2.0 * np.eye(n) + 0.1 * rng.standard_normal((n, n))
    r1 = _Rec(1, 1e-10).reflection(T)
    assert _err(r1, T[:n, n:] @ np.linalg.inv(T[n:, n:])) < 1e-14
    r = _Rec(200, 1e-14).reflection(T)
    fp = (T[:n, :n] @ r + T[:n, n:]) @ np.linalg.inv(T[n:, :n] @ r + T[n:, n:])
    assert _err(fp, r) < 1e-12
