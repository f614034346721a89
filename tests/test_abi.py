"""The C ABI: the shared library loads, exports every entry point include/trhepd.h
declares, and the host-only entry points (no device needed) agree with the oracle."""
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import REF, REPO
from trhepd import deckio, lib


def test_library_exports_header_symbols():
    hdr = open(os.path.join(REPO, "include", "trhepd.h")).read()
    names = set(re.findall(r"\b(trhepd_[a-z0-9_]+)\s*\(", hdr))
    assert names, "no declarations parsed"
    L = lib.load()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(lib.SYMBOLS)


def test_srfghk_matches_oracle():
    bp = deckio.BulkP.read(os.path.join(REF, "test521", "bulkP.b"))
    s = O.load_case(os.path.join(REF, "test521"), 0.68, 5)
    nb = bp.nb[0]
    nvm = 1 + nb * (nb - 1) // 2
    nv, igh, igk, iv = lib.srfghk(nvm, bp.domains[0].igh, bp.domains[0].igk, bp.ih, bp.ik)
    assert nv == s.nv
    assert np.array_equal(igh, s.pot["igh"]) and np.array_equal(igk, s.pot["igk"])
    assert np.array_equal(iv, s.iv)


def test_asfcrr_matches_oracle():
    rng = np.random.default_rng(3)
    nelm = 3
    a = rng.uniform(-2, 2, (nelm, 6))
    b = rng.uniform(0.3, 40, (nelm, 6))
    c = np.zeros((nelm, 6))
    dchg = np.array([0.0, 0.5, 0.0])
    dth, dtk, dtz = (rng.uniform(0.3, 1.0, nelm) for _ in range(3))
    elm = np.array([28.0855, 15.9994, 1.00794])
    for tem in (0.0, 300.0, -1.0):
        out = lib.asfcrr(10.0, nelm, 4, 1, a, b, c, dchg, dth, dtk, dtz, elm, tem, 3.0)
        a2, b2, c2, d2, h2, k2 = (np.array(x, dtype=float).copy() for x in (a, b, c, dchg, dth, dtk))
        import ctypes as C
        wn = C.c_double()
        O.lib().orc_asfcrr(10.0, C.byref(wn), nelm, 4, 1, a2.reshape(-1), b2.reshape(-1), c2.reshape(-1),
                           d2, h2, k2, np.array(dtz, dtype=float), elm, tem, 3.0)
        assert out[0] == pytest.approx(wn.value, rel=1e-15)
        for x, y in zip(out[1:], (a2, b2, c2, d2, h2, k2)):
            np.testing.assert_allclose(x, y, rtol=1e-14, atol=0)


def test_bad_arguments_are_rejected():
    with pytest.raises(lib.TrhepdError):
        lib.srfghk(2, [0], [0], [0, 1, 2], [0, 0, 0])     # nv would exceed nvm


FORT = os.path.join(REPO, "trhepd-opt_amd", "build", "fortran", "abi_check")


@pytest.mark.skipif(not os.path.exists(FORT), reason="amdflang not available: Fortran bindings not built")
def test_fortran_bindings_link_and_report_status():
    """The iso_c_binding module links against the library; without a device the context
    call returns TRHEPD_E_HIP through the Fortran interface (with one, TRHEPD_OK)."""
    import subprocess
    r = subprocess.run([FORT], capture_output=True, text=True, timeout=120)
    first = r.stdout.splitlines()[0]
    assert first.startswith("trhepd_ctx_create status ")
    assert int(first.split()[-1]) in (0, 2)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FORT), reason="amdflang not available: Fortran bindings not built")
def test_fortran_bindings_on_device():
    import subprocess
    r = subprocess.run([FORT], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0
    assert all(line.endswith("status 0") for line in r.stdout.splitlines() if "status" in line)


PROP = os.path.join(REPO, "trhepd-opt_amd", "build", "fortran", "propagate_check")


def _write_fortran_input(path, surf, method, G, F, W):
    """the input file of propagate_check: header, descriptor arrays, the angle batch (the
    byte layouts the C ABI takes, which are the Fortran driver's own arrays)"""
    p = surf.pot
    nang, nb = G.shape
    hdr = np.array([p["nv"], p["nvb"], p["natms"], p["natmb"], p["nelm"], p["nabr"], p["nabi"], surf.ns, method,
                    nb, nang, surf.nb0], dtype=np.int32)
    with open(path, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(np.array([p["cc"], p["smesh"], surf.dz], dtype=np.float64).tobytes())
        fh.write(np.ascontiguousarray(p["ielm"], dtype=np.int32).tobytes())
        for k in ("z", "dchg", "ghk", "ocr", "b", "asf"):
            fh.write(np.ascontiguousarray(p[k], dtype=np.float64).tobytes())
        if p["natms"]:
            fh.write(np.ascontiguousarray(p["st0"], dtype=np.complex128).tobytes())
        if p["natmb"]:
            fh.write(np.ascontiguousarray(p["st1"], dtype=np.complex128).tobytes())
            fh.write(np.ascontiguousarray(p["st2"], dtype=np.complex128).tobytes())
        fh.write(np.ascontiguousarray(np.asarray(surf.iv, dtype=np.int32).T).tobytes())
        fh.write(np.ascontiguousarray(G, dtype=np.complex128).tobytes())
        fh.write(np.ascontiguousarray(np.asarray(F, dtype=np.complex128).transpose(0, 2, 1)).tobytes())
        fh.write(np.ascontiguousarray(W, dtype=np.float64).tobytes())


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(PROP), reason="amdflang not available: Fortran bindings not built")
def test_fortran_drives_table_build_and_propagate_batch(tmp_path, engine):
    """The test23 deck (69 angles, sp4, dz 0.68) through trhepd_table_build and
    trhepd_propagate_batch called from Fortran: the same bits as the same calls from C
    (through the Python binding), and the reference's recorded curve to 1e-10."""
    import subprocess
    surf = O.load_case(os.path.join(REF, "test23"), 0.68, lib.SP4)
    angs = surf.angles()
    G = np.array([O.gamma(surf, ga, az) for (_, _, ga, az) in angs])
    F = np.array([surf.scatter_f(r1, r2) for (r1, r2, _, _) in angs])
    W = np.array([surf.wn * np.sin(ga) ** 2 for (_, _, ga, _) in angs])
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_fortran_input(fin, surf, lib.SP4, G, F, W)
    r = subprocess.run([PROP, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    nang, nb = G.shape
    raw = np.fromfile(fout, dtype=np.uint8)
    nf = nb * nb * nang * 16
    fo = np.frombuffer(raw[:nf].tobytes(), dtype=np.complex128).reshape(nang, nb, nb).transpose(0, 2, 1)
    f2 = np.frombuffer(raw[nf:].tobytes(), dtype=np.float64).reshape(nang, nb)
    tid = engine.table_build(surf.pot, surf.dz, surf.ns, lib.SP4)
    try:
        fc, f2c = engine.propagate_batch(tid, surf.iv, G, F, surf.nb0, W)
    finally:
        engine.table_release(tid)
    assert np.array_equal(f2, f2c) and np.array_equal(fo, fc)
    ref = deckio.read_s(os.path.join(REF, "test23", "perf-5-0.68.txt"))
    e = max(O.rel_err(f2[p], ref.f2[p]) for p in range(nang))
    print("Fortran-driven test23 sp4 dz 0.68 vs recorded %.3e" % e)
    assert e < 1e-10
