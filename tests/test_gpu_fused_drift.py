"""The RKN slice loop runs each slice's first drift (Q += c_a1 dz P, the start of hp_int_spa's
stage loop, surf_prkn.f90:296-303) inside the previous slice's end kick for the angles that do
not renormalise there.  The fused form does the same arithmetic as the separate first-stage
pass, so the intensities are bit-identical to it (TRHEPD_FUSE_DRIFT=0), in both precisions,
under the reference's renormalisation rule and under a periodic schedule."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import lib
from test_gpu_surface import _inputs

pytestmark = pytest.mark.gpu


def _curve(precision, method, period, fuse, monkeypatch, surf, G, F, W):
    monkeypatch.setenv("TRHEPD_FUSE_DRIFT", "1" if fuse else "0")
    with lib.Engine(0, precision=precision) as eng:
        if precision == lib.PREC_MIXED:
            eng.set_mixed_digits(6)
        tid = eng.table_build(surf.pot, surf.dz, surf.ns, method)
        eng.set_renorm_period(period)
        _, f2 = eng.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
        st = eng.last_stats()
        eng.table_release(tid)
    return f2, st


@pytest.mark.parametrize("precision", [lib.PREC_FP64, lib.PREC_MIXED])
@pytest.mark.parametrize("method,period", [(5, 0), (5, 3), (6, 0)])
def test_fused_first_drift_same_bits(monkeypatch, precision, method, period):
    print("case test521 dz 0.1 method", method, "period", period, "precision", precision, flush=True)
    surf = O.load_case(os.path.join(REF, "test521"), 0.1, method)
    angs, G, F, W = _inputs(surf)
    G, F, W = G[:5], F[:5], W[:5]
    sep, st0 = _curve(precision, method, period, False, monkeypatch, surf, G, F, W)
    fus, st1 = _curve(precision, method, period, True, monkeypatch, surf, G, F, W)
    print("separate", st0, "fused", st1)
    assert st0["n_renorm"] == st1["n_renorm"] > 0
    assert np.all(np.isfinite(fus))
    assert np.array_equal(sep, fus)
