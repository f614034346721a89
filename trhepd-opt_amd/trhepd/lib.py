"""ctypes binding of libtrhepd_hip.so (the C ABI declared in include/trhepd.h).

The library is the product: every numerical stage runs in its HIP kernels.  This module
only marshals numpy arrays (Fortran order, complex128, 1-based index maps) across the
ABI and turns non-zero status codes into exceptions.  If the shared library is missing
it raises -- there is no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # trhepd-opt_amd/
LIB_PATH = os.path.join(ROOT, "lib", "libtrhepd_hip.so")

OK, E_ARG, E_HIP, E_NOMEM, E_BUDGET, E_SINGULAR, E_NONFINITE, E_NOTFOUND, E_PENDING = range(9)
M2, M4, M6, LF, SP4, SP6, RK4, RK4GS = range(1, 9)
PREC_FP64, PREC_MIXED = 0, 1

# exported symbols (must match include/trhepd.h)
SYMBOLS = ("trhepd_ctx_create", "trhepd_ctx_destroy", "trhepd_last_error", "trhepd_ctx_bytes",
           "trhepd_asfcrr", "trhepd_srfghk", "trhepd_asfcef", "trhepd_scpot_stf",
           "trhepd_table_build", "trhepd_table_shape", "trhepd_table_fetch", "trhepd_table_release",
           "trhepd_propagate_batch", "trhepd_batch_create", "trhepd_batch_run", "trhepd_batch_run_many", "trhepd_batch_fetch",
           "trhepd_batch_release", "trhepd_last_stats", "trhepd_set_renorm_period", "trhepd_set_phase_offsets", "trhepd_set_mixed_digits", "trhepd_profile",
           "trhepd_kernel_stats", "trhepd_serve_submit", "trhepd_serve_step", "trhepd_serve_result",
           "trhepd_graph_create", "trhepd_graph_launch", "trhepd_graph_release", "trhepd_bulk_reflect",
           "trhepd_propagate_batch_noalloc", "trhepd_set_bulk_batched_above", "trhepd_serve_last_timing",
           "trhepd_serve_set_merge", "trhepd_table_build_steps", "trhepd_adaptive_steps", "trhepd_h211b_limit",
           "trhepd_h211b_filter")


def h211b_limit(u: float) -> float:
    """The library's h211b_limit (h211b_control.f90:1-7)."""
    return load().trhepd_h211b_limit(float(u))


def h211b_filter(c1: float, c0: float, rho: float, k: int) -> float:
    """The library's h211b_filter (h211b_control.f90:9-16)."""
    return load().trhepd_h211b_filter(float(c1), float(c0), float(rho), int(k))


class TrhepdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("trhepd status %d: %s" % (code, msg))
        self.code = code


class _Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("hbm_budget", C.c_size_t)]


class Potential(C.Structure):
    """trhepd_potential: the mod_scpot3 globals (surfio.f90:18-27)."""
    _fields_ = [("nv", C.c_int), ("nvb", C.c_int), ("natms", C.c_int), ("natmb", C.c_int),
                ("nelm", C.c_int), ("nabr", C.c_int), ("nabi", C.c_int),
                ("cc", C.c_double), ("smesh", C.c_double),
                ("ielm", C.c_void_p), ("z", C.c_void_p), ("dchg", C.c_void_p), ("ghk", C.c_void_p),
                ("ocr", C.c_void_p), ("b", C.c_void_p), ("asf", C.c_void_p),
                ("st0", C.c_void_p), ("st1", C.c_void_p), ("st2", C.c_void_p)]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TrhepdError(E_HIP, "%s is not built (run __graft_entry__.build() or make -C trhepd-opt_amd)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.trhepd_ctx_create.argtypes = [C.POINTER(_Opts), C.POINTER(vp)]
    L.trhepd_ctx_destroy.argtypes = [vp]
    L.trhepd_last_error.argtypes = [vp]
    L.trhepd_last_error.restype = C.c_char_p
    L.trhepd_ctx_bytes.argtypes = [vp]
    L.trhepd_ctx_bytes.restype = sz
    L.trhepd_asfcrr.argtypes = [d, dp, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, d, d]
    L.trhepd_srfghk.argtypes = [i, i, vp, vp, i, vp, vp, vp, ip]
    L.trhepd_asfcef.argtypes = [vp, i, i, vp, vp, d, d, d, i, i, vp, vp, vp, vp, vp, vp]
    L.trhepd_scpot_stf.argtypes = [vp, i, i, i, i, i, i, i, vp, vp, vp, vp, d, d, vp]
    L.trhepd_table_build.argtypes = [vp, C.POINTER(Potential), d, i, i, ip]
    L.trhepd_table_shape.argtypes = [vp, i, ip, ip]
    L.trhepd_table_fetch.argtypes = [vp, i, vp, vp]
    L.trhepd_table_release.argtypes = [vp, i]
    L.trhepd_propagate_batch.argtypes = [vp, i, i, i, vp, vp, vp, vp, i, vp]
    L.trhepd_propagate_batch_noalloc.argtypes = [vp, i, i, i, vp, vp, vp, vp, i, vp]
    L.trhepd_last_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]
    L.trhepd_batch_create.argtypes = [vp, i, i, i, vp, vp, vp, i, vp, ip]
    L.trhepd_batch_run.argtypes = [vp, i]
    L.trhepd_batch_run_many.argtypes = [vp, i, vp]
    L.trhepd_batch_fetch.argtypes = [vp, i, vp, vp]
    L.trhepd_batch_release.argtypes = [vp, i]
    L.trhepd_profile.argtypes = [vp, i]
    L.trhepd_set_renorm_period.argtypes = [vp, i]
    L.trhepd_set_phase_offsets.argtypes = [vp, i]
    L.trhepd_set_mixed_digits.argtypes = [vp, i]
    L.trhepd_set_bulk_batched_above.argtypes = [vp, i]
    L.trhepd_kernel_stats.argtypes = [vp, C.POINTER(C.c_longlong), dp]
    L.trhepd_serve_submit.argtypes = [vp, C.POINTER(Potential), d, i, i, i, i, vp, vp, vp, i, vp, ip]
    L.trhepd_serve_step.argtypes = [vp, ip, ip, ip]
    L.trhepd_graph_create.argtypes = [vp, i, vp, ip]
    L.trhepd_bulk_reflect.argtypes = [vp, i, i, vp, i, vp, i, vp, vp, i, d, vp, vp]
    L.trhepd_graph_launch.argtypes = [vp, i]
    L.trhepd_graph_release.argtypes = [vp, i]
    L.trhepd_serve_result.argtypes = [vp, i, vp, vp, dp]
    L.trhepd_serve_last_timing.argtypes = [vp, dp, dp]
    L.trhepd_serve_set_merge.argtypes = [vp, i, d]
    L.trhepd_table_build_steps.argtypes = [vp, C.POINTER(Potential), i, i, vp, d, ip]
    L.trhepd_adaptive_steps.argtypes = [vp, C.POINTER(Potential), i, d, d, d, d, d, i, i, vp, vp, vp, i, i, vp, ip, ip]
    L.trhepd_h211b_limit.argtypes = [d]
    L.trhepd_h211b_limit.restype = d
    L.trhepd_h211b_filter.argtypes = [d, d, d, i]
    L.trhepd_h211b_filter.restype = d
    for name in SYMBOLS:
        getattr(L, name).restype = getattr(L, name).restype or C.c_int
    L.trhepd_last_error.restype = C.c_char_p
    L.trhepd_ctx_bytes.restype = sz
    _lib = L
    return L


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.float64), requirements=["C", "A", "W"]).copy()


def asfcrr(be: float, nelm: int, nabr: int, nabi: int, a, b, c, dchg, dth, dtk, dtz, elm, tem: float,
           smesh: float):
    """asfcrr (scpot.f90:11-58) on (nelm, 6)-shaped C arrays == Fortran (6, nelm); returns wn
    and the updated arrays."""
    L = load()
    a, b, c, dchg, dth, dtk = (_f64(x) for x in (a, b, c, dchg, dth, dtk))
    dtz, elm = _f64(dtz), _f64(elm)
    wn = C.c_double(0.0)
    rc = L.trhepd_asfcrr(be, C.byref(wn), nelm, nabr, nabi, _p(a), _p(b), _p(c), _p(dchg), _p(dth), _p(dtk),
                         _p(dtz), _p(elm), tem, smesh)
    if rc:
        raise TrhepdError(rc, "asfcrr: bad arguments")
    return wn.value, a, b, c, dchg, dth, dtk


def srfghk(nvm: int, igh_b, igk_b, ih, ik) -> Tuple[int, np.ndarray, np.ndarray, np.ndarray]:
    """srfghk (surfsub.f90:9-41): returns nv, igh, igk (length nv) and iv (nb, nb) Fortran order."""
    L = load()
    nvb = len(igh_b)
    nb = len(ih)
    igh = np.zeros(max(nvm, nvb), dtype=np.int32)
    igk = np.zeros(max(nvm, nvb), dtype=np.int32)
    igh[:nvb], igk[:nvb] = igh_b, igk_b
    ihh = np.ascontiguousarray(ih, dtype=np.int32)
    ikk = np.ascontiguousarray(ik, dtype=np.int32)
    iv = np.zeros((nb, nb), dtype=np.int32, order="F")
    nv = C.c_int(0)
    rc = L.trhepd_srfghk(nvm, nvb, _p(igh), _p(igk), nb, _p(ihh), _p(ikk), _p(iv), C.byref(nv))
    if rc:
        raise TrhepdError(rc, "srfghk: nv exceeds nvm or bad arguments")
    return nv.value, igh[:nv.value].copy(), igk[:nv.value].copy(), iv


class Engine:
    """A device context (trhepd_ctx): owns HBM buffers, a HIP stream and the tables."""

    def __init__(self, device: int = 0, precision: int = PREC_FP64, hbm_budget: int = 0):
        L = load()
        self._L = L
        self._h = C.c_void_p()
        rc = L.trhepd_ctx_create(C.byref(_Opts(device, precision, hbm_budget)), C.byref(self._h))
        if rc:
            raise TrhepdError(rc, "ctx_create failed (device %d)" % device)
        self.precision = precision

    def close(self):
        if self._h:
            self._L.trhepd_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc:
            raise TrhepdError(rc, self._L.trhepd_last_error(self._h).decode())

    @property
    def bytes_held(self) -> int:
        return int(self._L.trhepd_ctx_bytes(self._h))

    # -- per-structure setup -------------------------------------------------
    def asfcef(self, igh, igk, ghx, ghy, gky, nabr, nabi, a, c, dth, dtk, nelm):
        nv = len(igh)
        igh = np.ascontiguousarray(igh, dtype=np.int32)
        igk = np.ascontiguousarray(igk, dtype=np.int32)
        a, c, dth, dtk = (np.ascontiguousarray(x, dtype=np.float64) for x in (a, c, dth, dtk))
        asf = np.zeros(nv * 6 * nelm)
        ghk = np.zeros(nv)
        self._chk(self._L.trhepd_asfcef(self._h, nv, nelm, _p(igh), _p(igk), ghx, ghy, gky, nabr, nabi,
                                        _p(a), _p(c), _p(dth), _p(dtk), _p(asf), _p(ghk)))
        return asf, ghk

    def scpot_stf(self, nsg, ma, mb, na, nb, gh, gk, x, y, dx, dy, nv=None):
        natm = len(x)
        gh = np.ascontiguousarray(gh, dtype=np.float64)
        gk = np.ascontiguousarray(gk, dtype=np.float64)
        nv = len(gh) if nv is None else nv
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        st = np.zeros((natm, nv), dtype=np.complex128)   # == Fortran (nv, natm)
        if natm:
            self._chk(self._L.trhepd_scpot_stf(self._h, natm, nv, nsg, ma, mb, na, nb, _p(gh), _p(gk),
                                               _p(x), _p(y), dx, dy, _p(st)))
        return st

    # -- table ---------------------------------------------------------------
    @staticmethod
    def _potential(pot: dict):
        """The trhepd_potential descriptor of a potential dict, and the arrays it points into."""
        keep = {}

        def arr(k, dt):
            keep[k] = np.ascontiguousarray(pot[k], dtype=dt)
            return _p(keep[k])

        P = Potential(int(pot["nv"]), int(pot["nvb"]), int(pot["natms"]), int(pot["natmb"]), int(pot["nelm"]),
                      int(pot["nabr"]), int(pot["nabi"]), float(pot["cc"]), float(pot["smesh"]),
                      arr("ielm", np.int32), arr("z", np.float64), arr("dchg", np.float64),
                      arr("ghk", np.float64), arr("ocr", np.float64), arr("b", np.float64),
                      arr("asf", np.float64), arr("st0", np.complex128), arr("st1", np.complex128),
                      arr("st2", np.complex128))
        return P, keep

    def table_build(self, pot: dict, dz: float, ns: int, method: int) -> int:
        P, keep = self._potential(pot)
        tid = C.c_int(0)
        self._chk(self._L.trhepd_table_build(self._h, C.byref(P), dz, ns, method, C.byref(tid)))
        return tid.value

    def table_build_steps(self, pot: dict, steps, method: int, z0: float = 0.0) -> int:
        """Table of int_method 5 or 6 over a list of slice thicknesses, the first slice
        starting at depth z0."""
        P, keep = self._potential(pot)
        h = np.ascontiguousarray(steps, dtype=np.float64)
        tid = C.c_int(0)
        self._chk(self._L.trhepd_table_build_steps(self._h, C.byref(P), method, len(h), _p(h), float(z0),
                                                   C.byref(tid)))
        return tid.value

    def adaptive_steps(self, pot: dict, method: int, zmax: float, tol: float, h0: float, hmin: float, hmax: float,
                       iv: np.ndarray, gamma: np.ndarray, f: np.ndarray, nb0: int, max_steps: int = 100000):
        """Slice thicknesses for the slab [0, zmax] chosen by the H211b controller from
        step-doubling estimates on the pilot angles (conventions of propagate_batch).
        Returns (steps, number of slices accepted at hmin above tol)."""
        P, keep = self._potential(pot)
        nang, nb = gamma.shape
        ivf = np.asfortranarray(iv, dtype=np.int32)
        g = np.ascontiguousarray(gamma, dtype=np.complex128)
        ff = np.ascontiguousarray(np.asarray(f, dtype=np.complex128).transpose(0, 2, 1))
        out = np.zeros(max_steps)
        n, nf = C.c_int(0), C.c_int(0)
        self._chk(self._L.trhepd_adaptive_steps(self._h, C.byref(P), method, zmax, tol, h0, hmin, hmax, nb, nang,
                                                _p(ivf), _p(g), _p(ff), nb0, max_steps, _p(out), C.byref(n),
                                                C.byref(nf)))
        return out[:n.value].copy(), nf.value

    # -- bulk stage -------------------------------------------------------------
    def bulk_reflect(self, tid: int, nbg, iv: np.ndarray, gamma: np.ndarray, phase, ml: int, epsb: float):
        """Bulk reflection blocks: gamma (nang, nb) complex in group order, iv blkghk's
        (nbgm, nb) map (Fortran order), phase (nb) or None.  Returns ([angle][group] blocks,
        layers (nang, ngr))."""
        nbg = np.ascontiguousarray(nbg, dtype=np.int32)
        ivf = np.asfortranarray(iv, dtype=np.int32)
        nbgm = ivf.shape[0]
        g = np.ascontiguousarray(gamma, dtype=np.complex128)
        nang = g.shape[0]
        ph = None if phase is None else np.ascontiguousarray(phase, dtype=np.complex128)
        out = np.zeros(nang * int(np.sum(nbg.astype(np.int64) ** 2)), dtype=np.complex128)
        layers = np.zeros((nang, len(nbg)), dtype=np.int32)
        self._chk(self._L.trhepd_bulk_reflect(self._h, tid, len(nbg), _p(nbg), nbgm, _p(ivf), nang, _p(g), _p(ph),
                                              ml, epsb, _p(out), _p(layers)))
        blocks, q = [], 0
        for a in range(nang):
            row = []
            for n in nbg:
                row.append(out[q:q + n * n].reshape((n, n), order="F").copy())
                q += n * n
            blocks.append(row)
        return blocks, layers

    # -- captured multi-batch runs ---------------------------------------------
    def graph_create(self, bids) -> int:
        """Captures the complete run of these resident batches into one graph."""
        ids = np.ascontiguousarray(bids, dtype=np.int32)
        gid = C.c_int(0)
        self._chk(self._L.trhepd_graph_create(self._h, len(ids), _p(ids), C.byref(gid)))
        return gid.value

    def graph_launch(self, gid: int):
        self._chk(self._L.trhepd_graph_launch(self._h, gid))

    def graph_release(self, gid: int):
        self._chk(self._L.trhepd_graph_release(self._h, gid))

    # -- serving ---------------------------------------------------------------
    def serve_submit(self, pot: dict, dz: float, ns: int, method: int, iv, gamma, f, nb0: int, wnsga2) -> int:
        """Queues one rocking-curve request (conventions of propagate_batch); returns its id."""
        P, keep = self._potential(pot)
        nang, nb = gamma.shape
        ivf = np.asfortranarray(iv, dtype=np.int32)
        g = np.ascontiguousarray(gamma, dtype=np.complex128)
        ff = np.ascontiguousarray(np.asarray(f, dtype=np.complex128).transpose(0, 2, 1))
        w = np.ascontiguousarray(wnsga2, dtype=np.float64)
        rid = C.c_int(0)
        self._chk(self._L.trhepd_serve_submit(self._h, C.byref(P), dz, ns, method, nb, nang, _p(ivf), _p(g), _p(ff),
                                              nb0, _p(w), C.byref(rid)))
        self._serve_shape = getattr(self, "_serve_shape", {})
        self._serve_shape[rid.value] = (nang, nb)
        return rid.value

    def serve_set_merge(self, max_angles: int, window_ms: float = 0.0):
        """Admit queued requests of one structure together as one batch of at most
        `max_angles` angles, the head of the queue waiting `window_ms` after its submission
        for others to join (0 angles: no merging)."""
        self._chk(self._L.trhepd_serve_set_merge(self._h, int(max_angles), float(window_ms)))

    def serve_step(self):
        """One step of the request loop; returns (admitted, queued, finished-uncollected)."""
        a, q, d = C.c_int(), C.c_int(), C.c_int()
        self._chk(self._L.trhepd_serve_step(self._h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def serve_result(self, rid: int, want_f: bool = False):
        """(f or None, f2, latency_ms, queued_ms, service_ms) of a finished request, or None
        while it is pending (latency = queued + service)."""
        nang, nb = self._serve_shape[rid]
        f = np.zeros((nang, nb, nb), dtype=np.complex128) if want_f else None
        f2 = np.zeros((nang, nb))
        lat = C.c_double()
        rc = self._L.trhepd_serve_result(self._h, rid, _p(f), _p(f2), C.byref(lat))
        if rc == E_PENDING:
            return None
        del self._serve_shape[rid]
        qd, sv = C.c_double(), C.c_double()
        self._L.trhepd_serve_last_timing(self._h, C.byref(qd), C.byref(sv))
        self._chk(rc)
        return (None if f is None else f.transpose(0, 2, 1)), f2, lat.value, qd.value, sv.value

    def table_shape(self, tid: int):
        nv, nc = C.c_int(), C.c_int()
        self._chk(self._L.trhepd_table_shape(self._h, tid, C.byref(nv), C.byref(nc)))
        return nv.value, nc.value

    def table_fetch(self, tid: int):
        nv, nc = self.table_shape(tid)
        v = np.zeros((nc, nv), dtype=np.complex128)
        vi = np.zeros((nc, nv), dtype=np.complex128)
        self._chk(self._L.trhepd_table_fetch(self._h, tid, _p(v), _p(vi)))
        return v, vi

    def table_release(self, tid: int):
        self._chk(self._L.trhepd_table_release(self._h, tid))

    # -- slice loop ----------------------------------------------------------
    def propagate_batch(self, tid: int, iv: np.ndarray, gamma: np.ndarray, f: np.ndarray, nb0: int,
                        wnsga2: Optional[np.ndarray] = None):
        """gamma (nang, nb) complex; f (nang, nb, nb) complex with f[a][i, j] = F_a(i, j).
        Returns (f_out, f2): f_out in the same indexing, f2 (nang, nb) or None."""
        nang, nb = gamma.shape
        ivf = np.asfortranarray(iv, dtype=np.int32)
        g = np.ascontiguousarray(gamma, dtype=np.complex128)
        ff = np.ascontiguousarray(np.asarray(f, dtype=np.complex128).transpose(0, 2, 1))
        f2 = np.zeros((nang, nb)) if wnsga2 is not None else None
        w = None if wnsga2 is None else np.ascontiguousarray(wnsga2, dtype=np.float64)
        self._chk(self._L.trhepd_propagate_batch(self._h, tid, nb, nang, _p(ivf), _p(g), _p(ff), _p(f2), nb0,
                                                 _p(w)))
        return ff.transpose(0, 2, 1), f2

    def propagate_batch_noalloc(self, tid: int, nb: int, nang: int, d_iv: int, d_gamma: int, d_f: int,
                                nb0: int, d_f2: Optional[int] = None, d_wnsga2: Optional[int] = None):
        """trhepd_propagate_batch_noalloc on caller-owned device memory (addresses as ints, the
        library's layouts: iv (nb, nb) int32 column-major, gamma (nb, nang), f (nb, nb, nang)
        complex column-major, f2 (nb, nang), wnsga2 (nang)).  N <= 48; allocates nothing."""
        self._chk(self._L.trhepd_propagate_batch_noalloc(self._h, tid, nb, nang, C.c_void_p(d_iv),
                                                         C.c_void_p(d_gamma), C.c_void_p(d_f),
                                                         None if d_f2 is None else C.c_void_p(d_f2), nb0,
                                                         None if d_wnsga2 is None else C.c_void_p(d_wnsga2)))

    # -- device-resident batches ----------------------------------------------
    def batch_create(self, tid: int, iv, gamma, f, nb0: int, wnsga2=None) -> int:
        """Upload one batch (same conventions as propagate_batch); returns its id."""
        nang, nb = gamma.shape
        self._keep_shape = getattr(self, "_keep_shape", {})
        ivf = np.asfortranarray(iv, dtype=np.int32)
        g = np.ascontiguousarray(gamma, dtype=np.complex128)
        ff = np.ascontiguousarray(np.asarray(f, dtype=np.complex128).transpose(0, 2, 1))
        w = None if wnsga2 is None else np.ascontiguousarray(wnsga2, dtype=np.float64)
        bid = C.c_int(0)
        self._chk(self._L.trhepd_batch_create(self._h, tid, nb, nang, _p(ivf), _p(g), _p(ff), nb0, _p(w),
                                              C.byref(bid)))
        self._keep_shape[bid.value] = (nang, nb, w is not None)
        return bid.value

    def batch_run(self, bid: int):
        self._chk(self._L.trhepd_batch_run(self._h, bid))

    def batch_run_many(self, bids):
        ids = np.ascontiguousarray(bids, dtype=np.int32)
        self._chk(self._L.trhepd_batch_run_many(self._h, len(ids), _p(ids)))

    def batch_fetch(self, bid: int, want_f: bool = True):
        nang, nb, has_f2 = self._keep_shape[bid]
        ff = np.zeros((nang, nb, nb), dtype=np.complex128) if want_f else None
        f2 = np.zeros((nang, nb)) if has_f2 else None
        self._chk(self._L.trhepd_batch_fetch(self._h, bid, _p(ff), _p(f2)))
        return (None if ff is None else ff.transpose(0, 2, 1)), f2

    def batch_release(self, bid: int):
        self._chk(self._L.trhepd_batch_release(self._h, bid))
        self._keep_shape.pop(bid, None)

    def set_renorm_period(self, period: int):
        """0: the reference's Gershgorin rule; K > 0: every angle after every K-th slice."""
        self._chk(self._L.trhepd_set_renorm_period(self._h, int(period)))

    def set_phase_offsets(self, mode: int):
        """Start concurrent large-N batches a fraction of the renormalisation period apart:
        1 on, 0 off, -1 the default (on for the mixed precision)."""
        self._chk(self._L.trhepd_set_phase_offsets(self._h, int(mode)))

    def set_mixed_digits(self, digits: int):
        """Digits per operand of the mixed-precision stage products (5, 6 or 7)."""
        self._chk(self._L.trhepd_set_mixed_digits(self._h, int(digits)))

    def set_bulk_batched_above(self, n: int):
        """Bulk groups of more than n beams run batched over angles (default 64; 0: all)."""
        self._chk(self._L.trhepd_set_bulk_batched_above(self._h, int(n)))

    def profile(self, enable: bool):
        self._chk(self._L.trhepd_profile(self._h, 1 if enable else 0))

    def kernel_stats(self):
        n, ms = C.c_longlong(), C.c_double()
        self._chk(self._L.trhepd_kernel_stats(self._h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def last_stats(self):
        r, p, ms = C.c_longlong(), C.c_longlong(), C.c_double()
        self._chk(self._L.trhepd_last_stats(self._h, C.byref(r), C.byref(p), C.byref(ms)))
        return {"n_renorm": r.value, "n_products": p.value, "device_ms": ms.value}


class DeviceArray:
    """A device copy of a numpy array, allocated through the HIP runtime the library links
    (for callers that hand device memory to the library, e.g. trhepd_propagate_batch_noalloc)."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            load()                               # resolves libamdhip64 the way the library does
            H = C.CDLL("libamdhip64.so.7")
            H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            H.hipFree.argtypes = [C.c_void_p]
            H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            H.hipSetDevice.argtypes = [C.c_int]
            cls._hip = H
        return cls._hip

    def __init__(self, a: np.ndarray, device: int = 0):
        H = self._rt()
        self.a = np.ascontiguousarray(a)
        self.nbytes = self.a.nbytes
        p = C.c_void_p()
        if H.hipSetDevice(device) or H.hipMalloc(C.byref(p), max(self.nbytes, 1)):
            raise TrhepdError(E_NOMEM, "hipMalloc of %d bytes failed" % self.nbytes)
        self.ptr = p.value
        if H.hipMemcpy(p, self.a.ctypes.data_as(C.c_void_p), self.nbytes, 1):   # host to device
            raise TrhepdError(E_HIP, "hipMemcpy to the device failed")

    def get(self) -> np.ndarray:
        out = np.empty_like(self.a)
        if self._rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, 2):
            raise TrhepdError(E_HIP, "hipMemcpy from the device failed")
        return out

    def free(self):
        if self.ptr:
            self._rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0
