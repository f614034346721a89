"""Host driver above the C ABI: the deck setup of the reference's programs and the angle
loop of srfref, with every numerical stage on the device.

* ``bulk_header``  -- the header / domain records bulkio.f90:31-176 writes to bulkP.b
  (element parameters, zoffset, scan ranges, beam grouping blkibg bulksub.f90:247-293 and
  bulk components blkghk bulksub.f90:297-336).  Integer / scalar host work.
* ``SurfaceJob``   -- surfio.f90:60-341: element corrections (trhepd_asfcrr), the
  beam-pair map (trhepd_srfghk), scattering factors and structure factors on the device
  (trhepd_asfcef, trhepd_scpot_stf), then the potential table (trhepd_table_build).
* ``SurfaceJob.run`` -- the angle loop of srfref (surflf.f90:55-166): gamma per angle,
  bulk blocks scattered by iord, one trhepd_propagate_batch per irep1 row.
"""
from __future__ import annotations

import json
import copy
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import deckio, lib

MAB = 6
_ELEM = None


def _elements():
    global _ELEM
    if _ELEM is None:
        _ELEM = json.load(open(os.path.join(os.path.dirname(__file__), "data", "elements.json")))
    return _ELEM


def asfparam(iz: int, iabr: int):
    """asfparam (asf.f90:7-433): Doyle-Turner when iabr == 0 and the element is tabulated,
    otherwise Peng with the Gaussians reversed so that b(1) is the largest."""
    E = _elements()
    if iabr == 0 and str(iz) in E["doyle_turner"]:
        e = E["doyle_turner"][str(iz)]
        return 4, list(e["a"]), list(e["b"])
    if 1 <= iz <= 98:
        e = E["peng"][str(iz)]
        return 4, list(reversed(e["a"])), list(reversed(e["b"]))
    return 0, [], []


def amass(iz: int) -> float:
    m = _elements()["mass"]
    return m[iz - 1] if 1 <= iz <= 110 else 1000.0


# --------------------------------------------------------------------------
# bulk-side setup


def scanrg(azi, azf, daz, gi, gf, dg):
    """scanrg (bulksub.f90:210-243): step counts and radians."""
    rad = math.atan(1.0) / 45.0
    if abs(daz) < 1e-4:
        naz = 1
    else:
        naz = min(max(int((azf - azi) / daz + 0.1) + 1, 1), 200)
    if abs(dg) < 1e-4:
        ng = 1
    else:
        ng = min(max(int((gf - gi) / dg + .1) + 1, 1), 200)
    return azi * rad, daz * rad, naz, gi * rad, dg * rad, ng


def blkibg(ih: List[int], ik: List[int], nh: int, nk: int):
    """blkibg (bulksub.f90:247-293): groups of beams coupled in the bulk; sorts ih/ik in
    place and returns (iord, nbg)."""
    nb = len(ih)
    iord = list(range(1, nb + 1))
    if nh == 1 and nk == 1:
        return iord, [nb]
    nbg = []
    j = 0
    while True:
        i = j
        j = i + 1
        h0, k0 = ih[i], ik[i]
        for k in range(i + 1, nb):
            if (ih[k] - h0) % nh == 0 and (ik[k] - k0) % nk == 0:
                if k > j:
                    l, n, m = ih[k], ik[k], iord[k]
                    for jj in range(k, j, -1):
                        ih[jj], ik[jj], iord[jj] = ih[jj - 1], ik[jj - 1], iord[jj - 1]
                    ih[j], ik[j], iord[j] = l, n, m
                j += 1
        nbg.append(j - i)
        if j >= nb:
            return iord, nbg


def blkghk(nbg: List[int], ih: List[int], ik: List[int]):
    """blkghk (bulksub.f90:297-336): unique in-group (dh, dk) components; returns igh, igk."""
    igh, igk = [0], [0]
    seen = set()        # components 2.. (a (0,0) difference is matched from m = 2 on too)
    ibas = 0
    for n in nbg:
        iend = ibas + n
        for k in range(ibas, iend - 1):
            for l in range(k + 1, iend):
                key = (ih[l] - ih[k], ik[l] - ik[k])
                if key not in seen:
                    seen.add(key)
                    igh.append(key[0])
                    igk.append(key[1])
        ibas = iend
    return igh, igk


def bulk_header(bd: deckio.BulkDeck, inegpos: int = 1, idiag: int = 3) -> deckio.BulkP:
    """bulkP.b header and domain records for a bulk.txt (bulkio.f90:31-176); the
    reflection blocks are left empty."""
    pi = math.atan(1.0) * 4.0
    rad = pi / 180.0
    ml = bd.ml if 1 <= bd.ml <= 1000 else 1000
    iabr, iabp, iabe = 0, 1, 0
    els = []
    nabr, nabi = 0, 1
    for e in bd.elements:
        nabr, pa, pb = asfparam(e.iz, iabr)
        if nabr <= 0:
            raise ValueError("no scattering factor for Z=%d" % e.iz)
        a = list(pa) + [0.0]
        b = list(pb) + [0.0]
        a[0] -= e.da1
        if inegpos > 0:
            a[:nabr] = [-x for x in a[:nabr]]
        a[nabr] = e.sap
        b[nabr] = 0.0
        els.append(deckio.BulkElement(np.array(a), np.array(b), abs(e.bh), abs(e.bk), abs(e.bz), amass(e.iz), 0.0))
    zs = [a.z for a in bd.atoms]
    zoffset = 0.5 * (bd.cc - max(zs) - min(zs))
    atoms = [deckio.Atom(a.ielm, a.ocr, a.x, a.y, a.z + zoffset) for a in bd.atoms]
    azi, daz, naz, gi, dg, ng = scanrg(bd.azi, bd.azf, bd.daz, bd.gi, bd.gf, bd.dg)
    ns = int(bd.cc / bd.dz) + 1
    dz = bd.cc / ns
    doms = []
    ib = 0
    for d in range(bd.ndom):
        ih = list(bd.ih[ib:ib + bd.nb[d]])
        ik = list(bd.ik[ib:ib + bd.nb[d]])
        iord, nbg = blkibg(ih, ik, bd.nh, bd.nk)
        igh, igk = blkghk(nbg, ih, ik)
        doms.append(deckio.BulkDomain(np.array(iord, dtype=np.int32), np.array(nbg, dtype=np.int32),
                                      np.array(igh, dtype=np.int32), np.array(igk, dtype=np.int32), []))
        ib += bd.nb[d]
    return deckio.BulkP(bd.ndom, inegpos, 0, bd.nh, bd.nk, iabr, iabp, iabe, nabr, nabi, idiag, list(bd.nb),
                        [r * rad for r in bd.rdom], list(bd.ih), list(bd.ik), bd.be, azi, daz, naz, gi, dg, ng,
                        0.0, dz, ml, 1e-10, bd.nsg, 0, bd.aa, bd.bb, bd.gam * rad, bd.cc, bd.dx, bd.dy, zoffset,
                        els, atoms, doms)


# --------------------------------------------------------------------------
# bulk stage (bulkio.f90:118-196, bulk.f90:7-124) on the device


class BulkJob:
    """Bulk reflection of one domain of a bulk deck, from its bulkP.b header."""

    def __init__(self, eng: "lib.Engine", bp: deckio.BulkP, idom: int = 0):
        self.eng, self.bp = eng, bp
        pi = math.atan(1.0) * 4.0
        nabr, nabi = bp.nabr, bp.nabi
        nelm = len(bp.elements)
        a = np.zeros((nelm, MAB))
        b = np.zeros((nelm, MAB))
        c = np.zeros((nelm, MAB))
        dth, dtk, dtz, elm, dchg = (np.zeros(nelm) for _ in range(5))
        for i, e in enumerate(bp.elements):
            na = nabr + nabi
            a[i, :na], b[i, :na] = e.a[:na], e.b[:na]
            dth[i], dtk[i], dtz[i], elm[i], dchg[i] = e.dth, e.dtk, e.dtz, e.elm, e.dchg
        ielm = np.array([at.ielm for at in bp.atoms], dtype=np.int32)
        ocr = np.array([at.ocr for at in bp.atoms])
        x = np.array([at.x for at in bp.atoms])
        y = np.array([at.y for at in bp.atoms])
        z = np.array([at.z for at in bp.atoms])
        self.wn, a, b, c, dchg, dth, dtk = lib.asfcrr(bp.be, nelm, nabr, nabi, a, b, c, dchg, dth, dtk, dtz, elm,
                                                      bp.tem, bp.aa * bp.bb * math.sin(bp.gam))
        self.ghx = (pi + pi) / (bp.aa * bp.nh)
        self.ghy = -(pi + pi) / (bp.aa * math.tan(bp.gam) * bp.nh)
        self.gky = (pi + pi) / (bp.bb * math.sin(bp.gam) * bp.nk)
        dom = bp.domains[idom]
        ib0 = sum(bp.nb[:idom])
        nb = bp.nb[idom]
        iord = np.asarray(dom.iord)
        self.ih = np.array(bp.ih[ib0:ib0 + nb])[iord - 1]     # group order of blkibg
        self.ik = np.array(bp.ik[ib0:ib0 + nb])[iord - 1]
        self.nbg = [int(n) for n in dom.nbg]
        igh = np.array(dom.igh, dtype=np.int32)
        igk = np.array(dom.igk, dtype=np.int32)
        nv = len(igh)
        # blkghk's iv(nbgm, nb): row = l - first beam of the group, column = beam k (k < l)
        first = {}
        for m in range(1, nv):
            first.setdefault((int(igh[m]), int(igk[m])), m + 1)
        nbgm = max(self.nbg)
        iv = np.zeros((nbgm, nb), dtype=np.int32, order="F")
        ib = 0
        for n in self.nbg:
            for k in range(n):
                for l in range(k + 1, n):
                    iv[l, ib + k] = first[(int(self.ih[ib + l] - self.ih[ib + k]), int(self.ik[ib + l] - self.ik[ib + k]))]
            ib += n
        self.iv = iv
        asf, ghk = eng.asfcef(igh, igk, self.ghx, self.ghy, self.gky, nabr, nabi, a.reshape(-1), c.reshape(-1),
                              dth, dtk, nelm)
        gh = (-(pi + pi) / bp.nh) * igh.astype(np.float64)
        gk = (-(pi + pi) / bp.nk) * igk.astype(np.float64)
        natm = len(bp.atoms)
        st_in = eng.scpot_stf(bp.nsg, 1, 0, 0, 1, gh, gk, x, y, 0.0, 0.0)
        st_below = eng.scpot_stf(bp.nsg, 1, 0, 0, 1, gh, gk, x, y, -bp.dx, -bp.dy)
        st_above = eng.scpot_stf(bp.nsg, 1, 0, 0, 1, gh, gk, x, y, bp.dx, bp.dy)
        # the unit above (z - cc), this unit (z) and the unit below (z + cc) are the three
        # parts of the table descriptor
        self.pot = dict(nv=nv, nvb=nv, natms=natm, natmb=natm, nelm=nelm, nabr=nabr, nabi=nabi, cc=bp.cc,
                        smesh=1.0, ielm=np.concatenate([ielm, ielm]), z=np.concatenate([z, z]), dchg=dchg, ghk=ghk,
                        ocr=np.concatenate([ocr, ocr]), b=b.reshape(-1), asf=asf, st0=st_above.reshape(-1),
                        st1=st_in.reshape(-1), st2=st_below.reshape(-1))
        self.ns = int(round(bp.cc / bp.dz))
        self.dz = bp.dz
        self.azi = bp.azi + bp.rdom[idom]
        self.phase = None
        if abs(bp.dx) > 1e-4 or abs(bp.dy) > 1e-4:
            pi2 = math.atan(1.0) * 8.0
            s = (bp.dx * pi2 / bp.nh) * self.ih + (bp.dy * pi2 / bp.nk) * self.ik
            self.phase = np.exp(1j * s)
        self.table_id = eng.table_build(self.pot, self.dz, self.ns, lib.M2)

    def release(self):
        if self.table_id is not None:
            self.eng.table_release(self.table_id)
            self.table_id = None

    def angles(self) -> List[Tuple[float, float]]:
        """(glancing, azimuth) in the loop order of bulk.f90:25-50."""
        bp = self.bp
        n1, n2 = (bp.naz, bp.ng) if bp.ng > bp.naz else (bp.ng, bp.naz)
        out = []
        for r1 in range(n1):
            for r2 in range(n2):
                if bp.ng > bp.naz:
                    out.append((bp.gi + r2 * bp.dg, self.azi + r1 * bp.daz))
                else:
                    out.append((bp.gi + r1 * bp.dg, self.azi + r2 * bp.daz))
        return out

    def gamma(self, ga: float, az: float) -> np.ndarray:
        wnx = self.wn * math.cos(ga) * math.cos(az)
        wny = self.wn * math.cos(ga) * math.sin(az)
        wgx = self.ghx * self.ih + wnx
        wgy = self.ghy * self.ih + self.gky * self.ik + wny
        s = self.wn * self.wn - wgx * wgx - wgy * wgy
        return np.where(s >= 0, np.sqrt(np.abs(s)) + 0j, 1j * np.sqrt(np.abs(s)))

    def run(self, points: Optional[Sequence[int]] = None):
        """Reflection blocks [angle][group] for the flat angle indices `points` (all by
        default) and the layer counts used."""
        angs = self.angles()
        pts = list(range(len(angs))) if points is None else list(points)
        G = np.array([self.gamma(*angs[p]) for p in pts])
        return self.eng.bulk_reflect(self.table_id, self.nbg, self.iv, G, self.phase, self.bp.ml, self.bp.epsb)

    def rows(self):
        """This domain's reflection blocks as bulkP.b holds them: [irep1][irep2][group]."""
        blocks, _ = self.run()
        n1, n2 = self.bp.nrep
        return [[blocks[r1 * n2 + r2] for r2 in range(n2)] for r1 in range(n1)]


def bulk_program(eng: "lib.Engine", bd: deckio.BulkDeck) -> deckio.BulkP:
    """What the reference's bulk program writes to bulkP.b for a bulk.txt (bulkio.f90,
    bulk.f90): the header, then every domain's reflection blocks from the device."""
    bp = bulk_header(bd)
    out = copy.deepcopy(bp)
    for d in range(bp.ndom):
        job = BulkJob(eng, bp, d)
        try:
            out.domains[d].blocks = job.rows()
        finally:
            job.release()
    return out


# --------------------------------------------------------------------------
# surface setup and the angle loop


class SurfaceJob:
    """One domain of one structure, set up on an Engine (surfio.f90:60-341)."""

    def __init__(self, eng: "lib.Engine", bp: deckio.BulkP, sd: deckio.SurfDeck, dz2: float, method: int,
                 idom: int = 0, build_table: bool = True):
        self.eng = eng
        self.method = method
        pi = math.atan(1.0) * 4.0
        nabr, nabi = bp.nabr, bp.nabi
        nelmb, nelms = len(bp.elements), len(sd.elements)
        nelm = nelmb + nelms
        a = np.zeros((nelm, MAB))
        b = np.zeros((nelm, MAB))
        c = np.zeros((nelm, MAB))
        dth, dtk, dtz, elm, dchg = (np.zeros(nelm) for _ in range(5))
        for i, e in enumerate(bp.elements):
            na = nabr + nabi
            a[i, :na], b[i, :na] = e.a[:na], e.b[:na]
            dth[i], dtk[i], dtz[i], elm[i], dchg[i] = e.dth, e.dtk, e.dtz, e.elm, e.dchg
        for j, e in enumerate(sd.elements):
            i = nelmb + j
            elm[i] = amass(e.iz)
            nr, pa, pb = asfparam(e.iz, bp.iabr)
            if nr <= 0:
                raise ValueError("no scattering factor for Z=%d" % e.iz)
            nabr = nr
            a[i, :nr], b[i, :nr] = pa, pb
            a[i, 0] -= e.da1
            if bp.inegpos > 0:
                a[i, :nabr] = -a[i, :nabr]
            nabi = 1
            a[i, nabr], b[i, nabr] = e.sap, 0.0
            dth[i], dtk[i], dtz[i] = abs(e.bh), abs(e.bk), abs(e.bz)
        natmb, natms = len(bp.atoms), len(sd.atoms)
        natm = natmb + natms
        ielm = np.zeros(natm, dtype=np.int32)
        ocr, x, y, z = (np.zeros(natm) for _ in range(4))
        for i, at in enumerate(bp.atoms):
            ielm[i], ocr[i], x[i], y[i], z[i] = at.ielm, at.ocr, at.x, at.y, at.z
        for j, at in enumerate(sd.atoms):
            i = natmb + j
            if -nelmb <= at.ielm <= -1:
                ielm[i] = -at.ielm
            elif 1 <= at.ielm <= nelms:
                ielm[i] = at.ielm + nelmb
            else:
                raise ValueError("bad element index %d" % at.ielm)
            ocr[i], x[i], y[i], z[i] = at.ocr, at.x, at.y, at.z + bp.zoffset
        if sd.nsgs < 1:
            raise ValueError("nsgs < 1 (bulk-only run) is not a surface job")
        zz = max([0.0] + [z[i] for i in range(natmb, natm)])
        ns2 = int((zz + sd.dthick + bp.cc) / bp.dz) + 1
        self.ns = max(int(np.rint((bp.dz * ns2) / dz2)), 1)
        self.dz = (bp.dz * ns2) / self.ns
        smesh = float(abs(sd.msa * sd.nsb - sd.msb * sd.nsa))
        self.wn, a, b, c, dchg, dth, dtk = lib.asfcrr(bp.be, nelm, nabr, nabi, a, b, c, dchg, dth, dtk, dtz, elm,
                                                      bp.tem, bp.aa * bp.bb * math.sin(bp.gam) * smesh)
        self.ghx = (pi + pi) / (bp.aa * bp.nh)
        self.ghy = -(pi + pi) / (bp.aa * math.tan(bp.gam) * bp.nh)
        self.gky = (pi + pi) / (bp.bb * math.sin(bp.gam) * bp.nk)
        dom = bp.domains[idom]
        ib0 = sum(bp.nb[:idom])
        nb = bp.nb[idom]
        self.nb = nb
        self.ih = np.array(bp.ih[ib0:ib0 + nb], dtype=np.int32)
        self.ik = np.array(bp.ik[ib0:ib0 + nb], dtype=np.int32)
        self.iord = np.asarray(dom.iord)
        self.nbg = np.asarray(dom.nbg)
        nvm = 1 + nb * (nb - 1) // 2
        nvb = len(dom.igh)
        self.nv, igh, igk, self.iv = lib.srfghk(max(nvm, nvb), dom.igh, dom.igk, self.ih, self.ik)
        nv = self.nv
        asf, ghk = eng.asfcef(igh, igk, self.ghx, self.ghy, self.gky, nabr, nabi, a.reshape(-1), c.reshape(-1),
                              dth, dtk, nelm)
        gh = (-(pi + pi) / bp.nh) * igh.astype(np.float64)
        gk = (-(pi + pi) / bp.nk) * igk.astype(np.float64)
        st0 = eng.scpot_stf(sd.nsgs, sd.msa, sd.msb, sd.nsa, sd.nsb, gh, gk, x[natmb:], y[natmb:],
                            sd.dxs + bp.dx, sd.dys + bp.dy)
        st1 = eng.scpot_stf(bp.nsg, 1, 0, 0, 1, gh[:nvb], gk[:nvb], x[:natmb], y[:natmb], bp.dx, bp.dy)
        st2 = eng.scpot_stf(bp.nsg, 1, 0, 0, 1, gh[:nvb], gk[:nvb], x[:natmb], y[:natmb], 0.0, 0.0)
        self.pot = dict(nv=nv, nvb=nvb, natms=natms, natmb=natmb, nelm=nelm, nabr=nabr, nabi=nabi, cc=bp.cc,
                        smesh=smesh, ielm=ielm, z=z, dchg=dchg, ghk=ghk, ocr=ocr, b=b.reshape(-1), asf=asf,
                        st0=st0.reshape(-1), st1=st1.reshape(-1), st2=st2.reshape(-1))
        self.azi = bp.azi + bp.rdom[idom]
        self.daz, self.naz, self.gi, self.dg, self.ng = bp.daz, bp.naz, bp.gi, bp.dg, bp.ng
        self.nb0 = int(np.nonzero((self.ih == 0) & (self.ik == 0))[0][0]) + 1
        self.table_id = eng.table_build(self.pot, self.dz, self.ns, method) if build_table else None

    def release(self):
        if self.table_id is not None:
            self.eng.table_release(self.table_id)
            self.table_id = None

    # angle loop (surflf.f90:43-103)
    @property
    def nrep(self) -> Tuple[int, int]:
        return (self.naz, self.ng) if self.ng > self.naz else (self.ng, self.naz)

    def angle(self, r1: int, r2: int) -> Tuple[float, float]:
        if self.ng > self.naz:
            return self.gi + r2 * self.dg, self.azi + r1 * self.daz
        return self.gi + r1 * self.dg, self.azi + r2 * self.daz

    def printed_angle(self, ga: float, az: float) -> float:
        return (ga if self.ng > self.naz else az) * 180.0 / 3.141592654

    def gamma(self, ga: float, az: float) -> np.ndarray:
        cga, caz, saz = math.cos(ga), math.cos(az), math.sin(az)
        wnx, wny = self.wn * cga * caz, self.wn * cga * saz
        wgx = self.ghx * self.ih + wnx
        wgy = self.ghy * self.ih + self.gky * self.ik + wny
        s = self.wn * self.wn - wgx * wgx - wgy * wgy
        return np.where(s >= 0.0, np.sqrt(np.abs(s)) + 0j, 1j * np.sqrt(np.abs(s)))

    def scatter(self, blocks: Sequence[np.ndarray]) -> np.ndarray:
        f = np.zeros((self.nb, self.nb), dtype=np.complex128)
        ib = 0
        for blk in blocks:
            n = blk.shape[0]
            idx = self.iord[ib:ib + n] - 1
            f[np.ix_(idx, idx)] = blk
            ib += n
        return f

    def inputs(self, r1: int, bulk_blocks) -> Tuple[np.ndarray, np.ndarray, np.ndarray, List[float]]:
        """gamma, scattered bulk F, wnsga2 and printed angles for one irep1 row."""
        nrep2 = self.nrep[1]
        G, F, W, A = [], [], [], []
        for r2 in range(nrep2):
            ga, az = self.angle(r1, r2)
            G.append(self.gamma(ga, az))
            F.append(self.scatter(bulk_blocks[r1][r2]) if bulk_blocks is not None
                     else np.zeros((self.nb, self.nb), dtype=np.complex128))
            W.append(self.wn * math.sin(ga) ** 2)
            A.append(self.printed_angle(ga, az))
        return np.array(G), np.array(F), np.array(W), A

    def run(self, bulk_blocks) -> List[List[Tuple[float, np.ndarray]]]:
        """The rocking curve: rows[irep1] = [(printed angle, f2), ...]."""
        out = []
        for r1 in range(self.nrep[0]):
            G, F, W, A = self.inputs(r1, bulk_blocks)
            _, f2 = self.eng.propagate_batch(self.table_id, self.iv, G, F, self.nb0, W)
            out.append([(A[i], f2[i]) for i in range(len(A))])
        return out

    def write_s(self, path: str, rows) -> None:
        deckio.write_s(path, self.naz, self.ng, list(self.ih), list(self.ik), rows)


def surface_program(eng: "lib.Engine", bp: deckio.BulkP, sd: deckio.SurfDeck, dz2: float, method: int,
                    wdom: Optional[Sequence[float]] = None) -> Dict[str, Optional[str]]:
    """What the reference's surface program writes for a bulkP.b and a surf.txt
    (surfm.f90:55-90): the ``.s`` text of a single domain; for several domains the ``.md``
    text (every domain's curve, surfio.f90:276-288) and its domain sum as the ``.s``
    (domainsum.f90).  The weights are the deck's unless `wdom` is given."""
    from . import post
    per_dom = []
    jobs = []
    for d in range(bp.ndom):
        job = SurfaceJob(eng, bp, sd, dz2, method, idom=d)
        try:
            per_dom.append(job.run(bp.domains[d].blocks))
        finally:
            job.release()
        jobs.append(job)
    j0 = jobs[0]
    if bp.ndom == 1:
        nb = bp.nb[0]
        out = ["#azimuths,g-angles,beams", "%4d%4d%4d" % (j0.naz, j0.ng, nb), "#ih,ik"]
        pairs = [v for i in range(nb) for v in (bp.ih[i], bp.ik[i])]
        for i in range(0, len(pairs), 400):
            out.append("".join("%4d" % v for v in pairs[i:i + 400]))
        for block in per_dom[0]:
            for ang, f2 in block:
                out.append(deckio.format_s_row(ang, f2))
            out.append("")
        return {"s": "\n".join(out) + "\n", "md": None}
    w = list(wdom) if wdom is not None else list(sd.wdom)
    deg = 180.0 / math.pi
    md = deckio.format_md(j0.naz, j0.ng, bp.nh, bp.nk, bp.gam * deg, bp.nb, [r * deg for r in bp.rdom], w,
                          bp.ih, bp.ik, per_dom)
    return {"s": post.domainsum(md), "md": md}
