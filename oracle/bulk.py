"""CPU restatement of the bulk stage of sim-trhepd-rheed (numpy).

TEST INFRASTRUCTURE ONLY: the checker of the device bulk stage; only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.

Follows, for one domain of a bulk deck (everything it needs is in the bulkP.b header):
  - the bulk potential, bulkio.f90:172-196: slice midpoints z_i = dz (i - 1/2), i = 1..ns
    (ns = int(cc/dz) + 1, dz = cc/ns), summed over the unit layer (structure factors at
    shift 0), the unit below (z_i + cc, shift -(dx, dy)) and the unit above (z_i - cc,
    shift +(dx, dy)) -- evaluated here with the surface-loop potential routine
    (orc_comp_scpot, scpot.f90:115-193) on a descriptor whose three parts are those units;
  - blkghk's in-group index map, bulksub.f90:297-336;
  - blkref, bulk.f90:7-124: per angle and beam group, the slice transfer matrices (trmat),
    their product over the unit layer, the column phases of the (dx, dy) shift, and the
    layer recursion R_l = (T11 R + T12)(T21 R + T22)^-1 with the |R_ii|^2 convergence test;
  - trmat, trmatg.f90:6-59: H = v + vi + diag(gamma^2) (upper triangle conj(v - vi)), its
    right eigenvectors U and eigenvalues lambda (zgeev -> numpy.linalg.eig, which calls the
    same LAPACK routine), kappa = sqrt(lambda), tau = (gamma_i + kappa_j) U_ij,
    rho = (gamma_i - kappa_j) U_ij, T = [[tau e, rho/e], [rho e, tau/e]] [[tau, rho], [rho, tau]]^-1
    with e = exp(i kappa dz);
  - gcmi3, gcmi.f90:5-15: A B^-1 through an LU solve of the transpose.
Pinned by the reference's own bulkP.b files (test/test{23,47,521}), which carry the bulk
reflection blocks its authors' build wrote.
"""
import ctypes as C
import math
from typing import List

import numpy as np

import oracle as O


def _gcmi3(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """a inv(b) (gcmi.f90:5-15)."""
    return np.linalg.solve(b.T, a.T).T


def trmat(h: np.ndarray, g: np.ndarray, dz: float) -> np.ndarray:
    """Transfer matrix of one slice (trmatg.f90:6-59), 2n x 2n."""
    n = len(g)
    if n == 1:
        lam, u = h.reshape(1).copy(), np.ones((1, 1), dtype=complex)
    else:
        lam, u = np.linalg.eig(h)
    kap = np.sqrt(lam)
    ph = np.exp(1j * dz * kap)
    tau = (g[:, None] + kap[None, :]) * u
    rou = (g[:, None] - kap[None, :]) * u
    t1 = np.block([[tau * ph[None, :], rou / ph[None, :]], [rou * ph[None, :], tau / ph[None, :]]])
    t3 = np.block([[tau, rou], [rou, tau]])
    return _gcmi3(t1, t3)


class Bulk:
    """Bulk potential and beam groups of one domain."""

    def __init__(self, bp, idom: int = 0):
        L = O.lib()
        pi = math.atan(1.0) * 4.0
        self.bp = bp
        nabr, nabi = bp.nabr, bp.nabi
        nelm = len(bp.elements)
        a = np.zeros((nelm, O.MAB))
        b = np.zeros((nelm, O.MAB))
        c = np.zeros((nelm, O.MAB))
        dth, dtk, dtz, elm, dchg = (np.zeros(nelm) for _ in range(5))
        for i, e in enumerate(bp.elements):
            na = nabr + nabi
            a[i, :na], b[i, :na] = e.a[:na], e.b[:na]
            dth[i], dtk[i], dtz[i], elm[i], dchg[i] = e.dth, e.dtk, e.dtz, e.elm, e.dchg
        natm = len(bp.atoms)
        ielm = np.array([at.ielm for at in bp.atoms], dtype=np.int32)
        ocr = np.array([at.ocr for at in bp.atoms])
        x = np.array([at.x for at in bp.atoms])
        y = np.array([at.y for at in bp.atoms])
        z = np.array([at.z for at in bp.atoms])
        wn = C.c_double(0.0)
        L.orc_asfcrr(bp.be, C.byref(wn), nelm, nabr, nabi, a.reshape(-1), b.reshape(-1), c.reshape(-1),
                     dchg, dth, dtk, dtz, elm, bp.tem, bp.aa * bp.bb * math.sin(bp.gam))
        self.wn = wn.value
        self.ghx = (pi + pi) / (bp.aa * bp.nh)
        self.ghy = -(pi + pi) / (bp.aa * math.tan(bp.gam) * bp.nh)
        self.gky = (pi + pi) / (bp.bb * math.sin(bp.gam) * bp.nk)
        dom = bp.domains[idom]
        ib0 = sum(bp.nb[:idom])
        nb = bp.nb[idom]
        iord = np.asarray(dom.iord)
        # beams in the sorted (grouped) order of blkibg
        ih0 = np.array(bp.ih[ib0:ib0 + nb])
        ik0 = np.array(bp.ik[ib0:ib0 + nb])
        self.ih, self.ik = ih0[iord - 1], ik0[iord - 1]
        self.nbg = list(dom.nbg)
        self.igh = np.array(dom.igh, dtype=np.int32)
        self.igk = np.array(dom.igk, dtype=np.int32)
        nv = len(self.igh)
        self.nv = nv
        # blkghk's iv: in-group pair (k < l) -> component index (1-based, from 2)
        first = {}
        for m in range(1, nv):
            first.setdefault((int(self.igh[m]), int(self.igk[m])), m + 1)
        self.iv = []
        ib = 0
        for n in self.nbg:
            iv = np.zeros((n, n), dtype=np.int64)
            for k in range(n):
                for l in range(k + 1, n):
                    iv[l, k] = first[(int(self.ih[ib + l] - self.ih[ib + k]), int(self.ik[ib + l] - self.ik[ib + k]))]
            self.iv.append(iv)
            ib += n
        asf = np.zeros(nv * O.MAB * nelm)
        ghk = np.zeros(nv)
        L.orc_asfcef(nv, nelm, self.igh, self.igk, self.ghx, self.ghy, self.gky, nabr, nabi, a.reshape(-1),
                     c.reshape(-1), asf, dth, dtk, ghk)
        gh = (-(pi + pi) / bp.nh) * self.igh.astype(np.float64)
        gk = (-(pi + pi) / bp.nk) * self.igk.astype(np.float64)
        st = np.zeros((3, natm, nv), dtype=np.complex128)
        for j in range(natm):
            L.orc_strfac(bp.nsg, nv, 1, 0, 0, 1, gh, gk, x[j], y[j], 0.0, 0.0, st[0, j])
            L.orc_strfac(bp.nsg, nv, 1, 0, 0, 1, gh, gk, x[j], y[j], -bp.dx, -bp.dy, st[1, j])
            L.orc_strfac(bp.nsg, nv, 1, 0, 0, 1, gh, gk, x[j], y[j], bp.dx, bp.dy, st[2, j])
        # the three units as the three parts of the surface-loop potential: "surface" =
        # unit above at z - cc, "top bulk" = this unit at z, "second bulk" = unit below at z + cc
        pot = dict(nv=nv, nvb=nv, natms=natm, natmb=natm, nelm=nelm, nabr=nabr, nabi=nabi, cc=bp.cc, smesh=1.0,
                   ielm=np.concatenate([ielm, ielm]), z=np.concatenate([z, z]), dchg=dchg, ghk=ghk,
                   ocr=np.concatenate([ocr, ocr]), b=b.reshape(-1).copy(), asf=asf,
                   st0=st[2].reshape(-1), st1=st[0].reshape(-1), st2=st[1].reshape(-1))
        self.pot = pot
        self.ns = int(round(bp.cc / bp.dz))
        self.dz = bp.dz
        s, keep = O._pot_struct(pot)
        v = np.zeros((self.ns, nv), dtype=np.complex128)
        vi = np.zeros((self.ns, nv), dtype=np.complex128)
        for i in range(self.ns):
            L.orc_comp_scpot(C.byref(s), self.dz * (i + 0.5), v[i], vi[i])
        self.v, self.vi = v, vi
        self.azi = bp.azi + bp.rdom[idom]

    def angles(self):
        """(irep1, irep2) -> (glancing, azimuth) in the loop order of bulk.f90:25-50."""
        bp = self.bp
        n1, n2 = (bp.naz, bp.ng) if bp.ng > bp.naz else (bp.ng, bp.naz)
        out = []
        for r1 in range(n1):
            for r2 in range(n2):
                if bp.ng > bp.naz:
                    out.append((bp.gi + r2 * bp.dg, self.azi + r1 * bp.daz))
                else:
                    out.append((bp.gi + r1 * bp.dg, self.azi + r2 * bp.daz))
        return out, n1, n2

    def hamiltonian(self, l: int, grp: int, g: np.ndarray) -> np.ndarray:
        """trmat's H of slice l for group grp (trmatg.f90:20-29)."""
        v, vi = self.v[l], self.vi[l]
        iv = self.iv[grp]
        n = len(g)
        h = np.zeros((n, n), dtype=complex)
        for i in range(n):
            h[i, i] = v[0] + vi[0] + g[i] * g[i]
            for j in range(i + 1, n):
                m = iv[j, i] - 1
                h[j, i] = v[m] + vi[m]
                h[i, j] = np.conj(v[m] - vi[m])
        return h

    def gamma(self, grp: int, ga: float, az: float) -> np.ndarray:
        ib = sum(self.nbg[:grp])
        n = self.nbg[grp]
        wnx = self.wn * math.cos(ga) * math.cos(az)
        wny = self.wn * math.cos(ga) * math.sin(az)
        g = np.zeros(n, dtype=complex)
        for i in range(n):
            wgx = self.ghx * self.ih[ib + i] + wnx
            wgy = self.ghy * self.ih[ib + i] + self.gky * self.ik[ib + i] + wny
            s = self.wn * self.wn - wgx * wgx - wgy * wgy
            g[i] = math.sqrt(s) if s >= 0 else 1j * math.sqrt(-s)
        return g

    def unit_transfer(self, grp: int, g: np.ndarray) -> np.ndarray:
        """Product of the slice transfer matrices with the column phases (bulk.f90:66-93)."""
        n = len(g)
        t4 = None
        for l in range(self.ns):
            t1 = trmat(self.hamiltonian(l, grp, g), g, self.dz)
            t4 = t1 if t4 is None else t1 @ t4
        bp = self.bp
        if abs(bp.dx) > 1e-4 or abs(bp.dy) > 1e-4:
            ib = sum(self.nbg[:grp])
            pi2 = math.atan(1.0) * 8.0
            s = (bp.dx * pi2 / bp.nh) * self.ih[ib:ib + n] + (bp.dy * pi2 / bp.nk) * self.ik[ib:ib + n]
            ph = np.exp(1j * s)
            t4[:, :n] *= ph[None, :]
            t4[:, n:] *= ph[None, :]
        return t4

    def reflection(self, t3: np.ndarray) -> np.ndarray:
        """Layer recursion with the convergence test (bulk.f90:95-119)."""
        n = t3.shape[0] // 2
        r = None
        r20 = None
        for l in range(1, self.bp.ml + 1):
            t1 = t3[:n, n:].copy()
            t2 = t3[n:, n:].copy()
            if l > 1:
                t1 += t3[:n, :n] @ r
                t2 += t3[n:, :n] @ r
            r = _gcmi3(t1, t2)
            d2 = np.abs(np.diag(r)) ** 2
            if l == 20:
                r20 = d2
            elif l > 20:
                rmax = np.max(np.abs(d2 - r20))
                r20 = d2
                if rmax < self.bp.epsb:
                    break
        return r

    def blocks(self, points=None) -> List[List[List[np.ndarray]]]:
        """Reflection blocks [irep1][irep2][group], optionally for a subset of flat angle indices."""
        angs, n1, n2 = self.angles()
        out = [[None] * n2 for _ in range(n1)]
        for p, (ga, az) in enumerate(angs):
            if points is not None and p not in points:
                continue
            row = []
            for grp in range(len(self.nbg)):
                g = self.gamma(grp, ga, az)
                row.append(self.reflection(self.unit_transfer(grp, g)))
            out[p // n2][p % n2] = row
        return out
