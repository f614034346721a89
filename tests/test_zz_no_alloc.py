"""trhepd_propagate_batch_noalloc (device arrays in, nothing allocated by the library)
against trhepd_propagate_batch on the same inputs: identical kernels, so identical bits.
Two calls in a row, the second smaller (fewer beams and angles), so that the fixed work area
is reused with the previous call's contents still in it."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF
from trhepd import lib

pytestmark = pytest.mark.gpu


def _inputs(case, dz, method, nang):
    surf = O.load_case(os.path.join(REF, case), dz, method)
    angs = surf.angles()[:nang]
    G = np.array([O.gamma(surf, ga, az) for (_, _, ga, az) in angs])
    F = np.array([surf.scatter_f(r1, r2) for (r1, r2, _, _) in angs])
    W = np.array([surf.wn * np.sin(ga) ** 2 for (_, _, ga, _) in angs])
    return surf, G, F, W


def _noalloc(eng, tid, surf, G, F, W):
    nang, nb = G.shape
    arrs = [lib.DeviceArray(np.ascontiguousarray(np.asarray(surf.iv, dtype=np.int32).T)),
            lib.DeviceArray(np.ascontiguousarray(G, dtype=np.complex128)),
            lib.DeviceArray(np.ascontiguousarray(F.transpose(0, 2, 1), dtype=np.complex128)),
            lib.DeviceArray(np.zeros((nang, nb))),
            lib.DeviceArray(np.ascontiguousarray(W, dtype=np.float64))]
    d_iv, d_g, d_f, d_f2, d_w = arrs
    try:
        eng.propagate_batch_noalloc(tid, nb, nang, d_iv.ptr, d_g.ptr, d_f.ptr, surf.nb0, d_f2.ptr, d_w.ptr)
        return d_f.get().transpose(0, 2, 1), d_f2.get(), eng.last_stats()
    finally:
        for a in arrs:
            a.free()


@pytest.mark.parametrize("period", [0, -1])
def test_noalloc_matches_propagate_batch_twice(period):
    cases = [("test47", 0.68, lib.SP4, 12), ("test23", 0.68, lib.SP6, 5)]   # second call smaller
    with lib.Engine(0) as eng:
        eng.set_renorm_period(period)
        for case, dz, method, nang in cases:
            print("noalloc vs propagate_batch:", case, "dz", dz, "method", method, "angles", nang, flush=True)
            surf, G, F, W = _inputs(case, dz, method, nang)
            tid = eng.table_build(surf.pot, surf.dz, surf.ns, method)
            f_ref, f2_ref = eng.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
            st_ref = eng.last_stats()
            held = eng.bytes_held
            f_na, f2_na, st_na = _noalloc(eng, tid, surf, G, F, W)
            assert eng.bytes_held == held
            assert np.array_equal(f2_na, f2_ref)
            assert np.array_equal(f_na, f_ref)
            assert st_na["n_renorm"] == st_ref["n_renorm"] and st_na["n_products"] == st_ref["n_products"]
            eng.table_release(tid)


def test_noalloc_refusals():
    surf, G, F, W = _inputs("test23", 0.68, lib.SP4, 3)
    with lib.Engine(0) as eng:
        tid = eng.table_build(surf.pot, surf.dz, surf.ns, lib.SP4)
        iv = np.asarray(surf.iv, dtype=np.int32).copy()
        iv[5, 2] = 10 ** 6                      # outside 1..nv: refused before f is touched
        d_iv = lib.DeviceArray(np.ascontiguousarray(iv.T))
        d_g = lib.DeviceArray(np.ascontiguousarray(G))
        f0 = np.ascontiguousarray(F.transpose(0, 2, 1))
        d_f = lib.DeviceArray(f0)
        try:
            with pytest.raises(lib.TrhepdError) as e:
                eng.propagate_batch_noalloc(tid, 23, 3, d_iv.ptr, d_g.ptr, d_f.ptr, surf.nb0)
            assert e.value.code == lib.E_ARG
            assert np.array_equal(d_f.get(), f0)
            with pytest.raises(lib.TrhepdError) as e:      # N above the work area's 48
                eng.propagate_batch_noalloc(tid, 49, 3, d_iv.ptr, d_g.ptr, d_f.ptr, surf.nb0)
            assert e.value.code == lib.E_ARG
        finally:
            for a in (d_iv, d_g, d_f):
                a.free()
        eng.table_release(tid)
