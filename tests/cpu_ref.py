"""ctypes bindings of the two CPU checkers (oracle C restatement and the flang
build of the reference routines) with a common calling convention.

Every function takes the ``Layout`` and numpy arrays in the reference storage
(``v`` is ``(nvar, ip11)`` C-contiguous, i.e. the Fortran ``v(ip11, nvar)``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

import oracle
from oracle import ptr

D = ctypes.c_double
L_ = ctypes.c_long


class OrcBlock(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("npc", "id1", "id2", "jd1", "jd2", "kd1", "kd2", "ii1", "ii2", "jj1", "jj2", "kk1",
                 "kk2", "nnn", "pad0", "pad1")]


class OrcPhys(ctypes.Structure):
    _fields_ = [("gam", D), ("gam1", D), ("pinfl", D), ("pr", D), ("prt", D), ("cp", D),
                ("reynz", D), ("tnz", D), ("iflu", ctypes.c_int32), ("pad", ctypes.c_int32)]


class OrcSch(ctypes.Structure):
    _fields_ = [("muscl", ctypes.c_int32), ("ilim", ctypes.c_int32), ("xk", D)]


def orc_block(b):
    return OrcBlock(*b.ext16())


def orc_phys(gam=1.4, pinfl=0.0, pr=0.72, prt=1.0, reynz=1e5, tnz=288.15, iflu=1):
    return OrcPhys(gam, gam - 1.0, pinfl, pr, prt, 1.0 / (gam - 1.0), reynz, tnz, iflu, 0)


def sn_block(sn, b):
    return sn[9 * b.npc: 9 * b.npc + 9 * b.nnn]


class Ref:
    """The reference routines (flang build), one layout at a time."""

    def __init__(self, layout, phys: OrcPhys, muscl=1, xk=1.0 / 3.0, ilim=0, kfmg=0, ns=True, lib=None):
        self.lib = lib if lib is not None else oracle.ref_lib()
        if self.lib is None:
            raise RuntimeError("oracle/_ref/libcodens_ref.so not built")
        self.lay = layout
        ext = np.ascontiguousarray(layout.ext_array())
        self.lib.ref_set_layout(len(layout.blocks), layout.ip11, layout.ip00, ptr(ext))
        self.lib.ref_set_phys(D(phys.gam), D(phys.pinfl), D(phys.pr), D(phys.prt), D(phys.reynz),
                              D(phys.tnz), phys.iflu)
        self.lib.ref_set_scheme(muscl, D(xk), ilim, kfmg, 1 if ns else 0)
        self.t = [np.zeros(layout.ip00) for _ in range(10)]

    def zpres(self, l, v, ps, temp, cson):
        self.lib.ref_zpres(l + 1, ptr(v), ptr(ps), ptr(temp), ptr(cson))

    # on-disk formats (writda / readda / writdg / readdg over all blocks)
    def writda_all(self, path, equat, v, mut):
        p = os.fsencode(path)
        self.lib.ref_writda_all(ctypes.c_char_p(p), len(p), len(self.lay.blocks), equat.encode().ljust(7)[:7],
                                ptr(v), ptr(mut))

    def readda_all(self, path, equat, v, mut):
        p = os.fsencode(path)
        self.lib.ref_readda_all(ctypes.c_char_p(p), len(p), len(self.lay.blocks), equat.encode().ljust(7)[:7],
                                ptr(v), ptr(mut))

    def writdg_all(self, path, x, y, z):
        p = os.fsencode(path)
        self.lib.ref_writdg_all(ctypes.c_char_p(p), len(p), len(self.lay.blocks), ptr(x), ptr(y), ptr(z))

    def readdg_all(self, path, x, y, z):
        p = os.fsencode(path)
        self.lib.ref_readdg_all(ctypes.c_char_p(p), len(p), len(self.lay.blocks), ptr(x), ptr(y), ptr(z))

    def zvismo(self, l, mu, v, temp):
        self.lib.ref_zvismo(l + 1, ptr(mu), ptr(v), ptr(temp))

    def chronos(self, l, eta, mu, mut, v, dt, sn, vol, cson):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.ref_chronos(l + 1, D(eta), ptr(mu), ptr(mut), ptr(v), ptr(dt), ptr(snb), b.nnn,
                             ptr(vol), *[ptr(a) for a in self.t[:4]], ptr(cson))

    def sch_roe_euler(self, l, ityprk, u, v, ff, sn, ps):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.ref_sch_roe_euler(l + 1, ityprk, ptr(u), ptr(v), ptr(ff), ptr(snb), b.nnn,
                                   *[ptr(a) for a in self.t], ptr(ps))

    def sch_roe(self, l, ityprk, u, v, ff, tq, sn, ps):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.ref_sch_roe(l + 1, ityprk, ptr(u), ptr(v), ptr(ff), *[ptr(tq[c]) for c in range(9)],
                             ptr(snb), b.nnn, *[ptr(a) for a in self.t], ptr(ps))

    def zgrad2(self, l, sn, vol, v, temp):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        g = np.zeros((12, self.lay.ip00))
        cmu = np.ones(self.lay.ip11)
        self.lib.ref_zgrad2(l + 1, ptr(snb), b.nnn, ptr(vol), ptr(v), ptr(temp), ptr(g), ptr(cmu))
        return g

    def zfluto(self, l, mu, mut, tq, g):
        self.lib.ref_zfluto(l + 1, ptr(mu), ptr(mut), ptr(tq), ptr(g))

    def sch_expli(self, l, u, v, dt, vol):
        self.lib.ref_sch_expli(l + 1, ptr(u), ptr(v), ptr(dt), ptr(vol))

    def implimf_3d(self, l, u, dt, v, d, ff, mu, mut, lmx, sn, vol, ps, cson, dtpas=1.0, ityprk=0):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.ref_implimf_3d(l + 1, ptr(u), ptr(dt), ptr(v), ptr(d), ptr(ff), ptr(mu), ptr(mut), lmx,
                                ptr(snb), b.nnn, ptr(vol), D(dtpas), ityprk, *[ptr(a) for a in self.t],
                                ptr(ps), ptr(cson))

    def implimf_eu(self, l, u, dt, v, d, ff, lmx, sn, vol, ps, cson, dt1min=1.0, ityprk=0):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.ref_set_dt1min(D(dt1min))
        self.lib.ref_implimf_eu(l + 1, ptr(u), ptr(dt), ptr(v), ptr(d), ptr(ff), lmx, ptr(snb), b.nnn, ptr(vol),
                                D(1.0), ityprk, *[ptr(a) for a in self.t], ptr(ps), ptr(cson))


class Orc:
    """The C restatement, same calling convention as :class:`Ref`."""

    def __init__(self, layout, phys: OrcPhys, muscl=1, xk=1.0 / 3.0, ilim=0, kfmg=0, ns=True):
        self.lib = oracle.oracle_lib()
        self.lay = layout
        self.ph = phys
        self.sc = OrcSch(muscl, ilim, xk)
        self.ns = ns
        self.kfmg = kfmg
        self.blocks = [orc_block(b) for b in layout.blocks]
        ld = layout.ip11
        self.ld = L_(ld)

    def B(self, l):
        return ctypes.byref(self.blocks[l])

    def zpres(self, l, v, ps, temp, cson):
        self.lib.orc_zpres(self.B(l), ctypes.byref(self.ph), ptr(v), self.ld, ptr(ps), ptr(temp), ptr(cson))

    def zvismo(self, l, mu, v, temp):
        return self.lib.orc_zvismo(self.B(l), ctypes.byref(self.ph), ptr(temp), ptr(mu))

    def chronos(self, l, eta, mu, mut, v, dt, sn, vol, cson):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.orc_chronos(self.B(l), ctypes.byref(self.ph), D(eta), 1 if self.ns else 0, ptr(v), self.ld,
                             ptr(mu), ptr(mut), ptr(snb), ptr(vol), ptr(cson), ptr(dt))

    def chrono_global(self, dt1min, dt):
        arr = (OrcBlock * len(self.blocks))(*self.blocks)
        out = D(0.0)
        stop = self.lib.orc_chrono_global(arr, len(self.blocks), D(dt1min), ptr(dt), ctypes.byref(out))
        return stop, out.value

    def sch_roe_euler(self, l, ityprk, u, v, ff, sn, ps):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.orc_sch_roe_euler(self.B(l), ctypes.byref(self.ph), ctypes.byref(self.sc), ityprk, ptr(u),
                                   ptr(v), self.ld, ptr(ff), ptr(snb), ptr(ps))

    def sch_roe(self, l, ityprk, u, v, ff, tq, sn, ps):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.orc_sch_roe(self.B(l), ctypes.byref(self.ph), ctypes.byref(self.sc), ityprk, ptr(u), ptr(v),
                             self.ld, ptr(ff), ptr(tq), ptr(snb), ptr(ps))

    def zgrad2(self, l, sn, vol, v, temp):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        g = np.zeros((12, b.nnn))
        self.lib.orc_zgrad2(self.B(l), ptr(snb), ptr(vol), ptr(v), self.ld, ptr(temp), ptr(g))
        return g

    def zfluto(self, l, mu, mut, tq, g):
        self.lib.orc_zfluto(self.B(l), ctypes.byref(self.ph), ptr(mu), ptr(mut), ptr(tq), self.ld,
                            ptr(np.ascontiguousarray(g)))

    def sch_expli(self, l, u, v, dt, vol):
        self.lib.orc_sch_expli(self.B(l), ptr(u), ptr(v), self.ld, ptr(dt), ptr(vol))

    def rk_stage(self, l, alpha, v0, u, v, dt, vol):
        nv = min(v.shape[0], u.shape[0])
        self.lib.orc_rk_stage(self.B(l), D(alpha), ptr(v0), ptr(u), ptr(v), self.ld, nv, ptr(dt), ptr(vol))

    def implimf_3d(self, l, u, dt, v, d, ff, mu, mut, lmx, sn, vol, ps, cson, dtpas=1.0, ityprk=0):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.orc_implimf_3d(self.B(l), ctypes.byref(self.ph), 1 if self.ns else 0, ptr(u), ptr(dt), ptr(v),
                                self.ld, ptr(d), ptr(ff), ptr(mu), ptr(mut), lmx, ptr(snb), ptr(vol), D(dtpas),
                                self.kfmg, ityprk, ptr(ps), ptr(cson))

    def implimf_eu(self, l, u, dt, v, d, ff, lmx, sn, vol, ps, cson, dt1min=1.0, ityprk=0):
        b = self.lay.blocks[l]
        snb = np.ascontiguousarray(sn_block(sn, b))
        self.lib.orc_implimf_eu(self.B(l), ctypes.byref(self.ph), ptr(u), ptr(dt), ptr(v), self.ld, v.shape[0], ptr(d),
                                ptr(ff),
                                lmx, ptr(snb), ptr(vol), D(dt1min), self.kfmg, ityprk, ptr(ps), ptr(cson))

    # boundary ghosts (face lists as int64 arrays)
    def bc_coincident(self, g, i, r, v, ps, temp, cson):
        self.lib.orc_bc_coincident(L_(len(g)), ptr(g), ptr(i), ptr(r), ptr(v), self.ld, ptr(ps), ptr(temp),
                                   ptr(cson))

    def bc_wall(self, g, i, v, ps, temp, cson):
        self.lib.orc_bc_wall(L_(len(g)), ptr(g), ptr(i), ptr(v), self.ld, ptr(ps), ptr(temp), ptr(cson))

    def bc_farfield(self, g, i, nx, ny, nz, w0, v, ps, temp, cson):
        self.lib.orc_bc_farfield(L_(len(g)), ptr(g), ptr(i), ptr(nx), ptr(ny), ptr(nz), ptr(w0),
                                 ctypes.byref(self.ph), ptr(v), self.ld, ptr(ps), ptr(temp), ptr(cson))

    def bc_tq_copy(self, g, i, tq):
        self.lib.orc_bc_tq_copy(L_(len(g)), ptr(g), ptr(i), ptr(tq), self.ld)

    def bc_tq_coincident(self, g, i, r, tq):
        self.lib.orc_bc_tq_coincident(L_(len(g)), ptr(g), ptr(i), ptr(r), ptr(tq), self.ld)


def _ref_dissip(self, l, v, d, sn, ps, cson, ki2, ki4):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    snu = np.zeros(self.lay.ip00)
    self.lib.ref_dissip_jameson(l + 1, ptr(v), ptr(d), ptr(snb), b.nnn, ptr(snu), ptr(ps), ptr(cson), D(ki2), D(ki4))


def _orc_dissip(self, l, v, d, sn, ps, cson, ki2, ki4):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    self.lib.orc_dissip_jameson(self.B(l), D(ki2), D(ki4), ptr(v), self.ld, ptr(d), ptr(snb), ptr(ps), ptr(cson))


Ref.dissip_jameson = _ref_dissip
Orc.dissip_jameson = _orc_dissip


def _ref_roe2o(self, l, t, d, sn):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    w1, w2 = np.zeros(self.lay.ip00), np.zeros(self.lay.ip00)
    vol = np.zeros(self.lay.ip11)
    self.lib.ref_met_roe2o(l + 1, ptr(t), ptr(d), ptr(snb), b.nnn, ptr(vol), ptr(w1), ptr(w2))


def _ref_roe1(self, l, t, d, sn):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    vol = np.zeros(self.lay.ip11)
    self.lib.ref_met_roe(l + 1, ptr(t), ptr(d), ptr(snb), b.nnn, ptr(vol))


def _orc_roe2o(self, l, t, d, sn):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    self.lib.orc_met_roe2o(self.B(l), ptr(t), self.ld, ptr(d), ptr(snb))


def _orc_roe1(self, l, t, d, sn):
    b = self.lay.blocks[l]
    snb = np.ascontiguousarray(sn_block(sn, b))
    self.lib.orc_met_roe(self.B(l), ptr(t), self.ld, ptr(d), ptr(snb))


def met_roe_outside_entries(lay):
    """(row, entry) pairs of d that the reference's met_roe zeroes outside the
    block it works on: its shifted index (met_roe.f90:184-188) starts the
    zeroing one storage entry early, i.e. at the previous block's last storage
    entry, or, for the first block, at d(0,6) = d(ip11,5) and d(0,7) =
    d(ip11,6) by column-major sequence association.  None of these entries belongs to a cell; the GPU
    kernel and the oracle write inside the block only."""
    out = []
    for b in lay.blocks:
        if b.npc == 0:
            out += [(4, lay.ip11 - 1), (5, lay.ip11 - 1)]
        else:
            out += [(5, b.npc - 1), (6, b.npc - 1)]
    return out


Ref.met_roe2o = _ref_roe2o
Ref.met_roe = _ref_roe1
Orc.met_roe2o = _orc_roe2o
Orc.met_roe = _orc_roe1


class OrcTurb(ctypes.Structure):
    _fields_ = [(n, D) for n in ("cb1", "cb2", "sigma", "kappa", "cw1", "cw2", "cw3", "cv1", "ct3", "ct4",
                                 "betas", "beta", "sigmk", "sigmw", "alfak", "alfae")]


def turb_constants():
    """Spalart-Allmaras (atlecdon.f90:183-204) and Wilcox k-omega
    (atlecdon.f90:353-366) constants; alfak/alfae of a k-eps model."""
    cb1, sigma, cb2, kappa = 0.1355, 2.0 / 3.0, 0.622, 0.41
    cw1 = cb1 / kappa ** 2 + (1 + cb2) / sigma
    return OrcTurb(cb1, cb2, sigma, kappa, cw1, 0.3, 2.0, 7.1, 1.1, 2.0, 0.09, 0.075, 0.5, 0.5, 1.0, 1.3)


def _tc_arr(tc):
    return np.array([getattr(tc, f) for f, _ in OrcTurb._fields_])


def _ref_set_turb(self, tc):
    a = _tc_arr(tc)
    self.lib.ref_set_turb(ptr(a))


def _ref_turb_call(self, name, l, *args):
    return getattr(self.lib, name)(l + 1, *args)


Ref.set_turb = _ref_set_turb


def _o_teq_gradv(self, l, sn, vol, v):
    b = self.lay.blocks[l]
    g = np.zeros((9, b.nnn))
    self.lib.orc_teq_gradv(self.B(l), ptr(np.ascontiguousarray(sn_block(sn, b))), ptr(vol), ptr(v), self.ld, ptr(g))
    return g


def _o_met_prod(self, l, mut, v, g9):
    b = self.lay.blocks[l]
    tp = np.zeros(b.nnn)
    self.lib.orc_met_prod(self.B(l), ptr(mut), ptr(v), self.ld, ptr(np.ascontiguousarray(g9)), ptr(tp))
    return tp


def _o_met_gradtr(self, l, sn, vol, v, fd):
    b = self.lay.blocks[l]
    self.lib.orc_met_gradtr(self.B(l), ptr(np.ascontiguousarray(sn_block(sn, b))), ptr(vol), ptr(v), self.ld, ptr(fd))


def _o_met_smsa(self, l, tc, sn, vol, v, mu, dist, g9, fd, cfke, ts6, ts7):
    b = self.lay.blocks[l]
    self.lib.orc_met_smsa(self.B(l), ctypes.byref(tc), ptr(np.ascontiguousarray(sn_block(sn, b))), ptr(vol), ptr(v),
                          self.ld, ptr(mu), ptr(dist), ptr(np.ascontiguousarray(g9)), ptr(fd), ptr(cfke), ptr(ts6),
                          ptr(ts7))


def _o_met_smko(self, l, tc, v, tprod, cfke, ts6, ts7, cson):
    self.lib.orc_met_smko(self.B(l), ctypes.byref(tc), ptr(v), self.ld, ptr(np.ascontiguousarray(tprod)), ptr(cfke),
                          ptr(ts6), ptr(ts7), ptr(cson))


def _o_met_flud(self, l, kfl, tc, v, mu, mut, fd):
    self.lib.orc_met_flud(self.B(l), kfl, ctypes.byref(tc), ptr(v), self.ld, ptr(mu), ptr(mut), ptr(fd))


def _o_sch_turb(self, l, u, v, d, fd, ts6, ts7, sn, vol):
    b = self.lay.blocks[l]
    self.lib.orc_sch_turb(self.B(l), ptr(u), ptr(v), self.ld, ptr(d), ptr(fd), ptr(ts6), ptr(ts7),
                          ptr(np.ascontiguousarray(sn_block(sn, b))), ptr(vol))


Orc.teq_gradv = _o_teq_gradv
Orc.met_prod = _o_met_prod
Orc.met_gradtr = _o_met_gradtr
Orc.met_smsa = _o_met_smsa
Orc.met_smko = _o_met_smko
Orc.met_flud = _o_met_flud
Orc.sch_turb = _o_sch_turb


def orc_utnrd(lib, ph, mach, alpha, beta, roa1, aa1, pa1):
    """The far-field state of utnrd.f90 (oracle restatement), 5 doubles."""
    w0 = np.zeros(5)
    lib.orc_utnrd(D(mach), D(alpha), D(beta), D(roa1), D(aa1), D(pa1), ctypes.byref(ph), ptr(w0))
    return w0


def orc_rfve(lib, g, i, v, ps, temp, cson):
    lib.orc_rfve(L_(len(g)), ptr(g), ptr(i), ptr(v), L_(v.shape[1]), ptr(ps), ptr(temp), ptr(cson))


def orc_bc_sym(lib, g, i, nx, ny, nz, v):
    lib.orc_bc_sym(L_(len(g)), ptr(g), ptr(i), ptr(nx), ptr(ny), ptr(nz), ptr(v), L_(v.shape[1]))


class RefFrontiers:
    """The reference's boundary routines (rbvc, clpara, utnrd + rbve + clnrd,
    clsym, rbte, rbtc, met_rbse, met_rbsc) over a list of frontiers, each a
    FaceSet-like (ghost, inner) slice in the reference's listing order; the
    flang build sees them as frontiers 1..nfr, face m of frontier f at
    mpb(f)+m.  partner[f] (0-based, -1 for none) is the coincident frontier."""

    def __init__(self, ref: Ref, ghosts, inners, partner):
        self.lib = ref.lib
        self.nfr = len(ghosts)
        self.mt = np.array([len(g) for g in ghosts], dtype=np.int32)
        self.off = np.concatenate([[0], np.cumsum(self.mt)]).astype(np.int64)
        self.ncbd = np.ascontiguousarray(np.concatenate(ghosts).astype(np.int32) + 1)
        self.ncin = np.ascontiguousarray(np.concatenate(inners).astype(np.int32) + 1)
        part = np.array([p + 1 for p in partner], dtype=np.int32)
        self.lib.ref_set_boundaries(self.nfr, ptr(self.mt), ptr(part))

    def _list(self, frs):
        return np.ascontiguousarray(np.array([f + 1 for f in frs], dtype=np.int32))

    def rbvc(self, frs, v, ps, temp, cson):
        fr = self._list(frs)
        self.lib.ref_rbvc(len(fr), ptr(fr), ptr(v), ptr(self.ncbd), ptr(self.ncin), ptr(ps), ptr(temp), ptr(cson))

    def rfve(self, frs, v, ps, temp, cson):
        fr = self._list(frs)
        self.lib.ref_rfve(len(fr), ptr(fr), ptr(v), ptr(ps), ptr(temp), ptr(cson), ptr(self.ncbd), ptr(self.ncin))

    def clpara(self, f, v, ps, temp, cson):
        self.lib.ref_clpara(f + 1, ptr(self.ncbd), ptr(v), ptr(self.ncin), ptr(ps), ptr(temp), ptr(cson))

    def _normals(self, f, nrm):
        n = [np.zeros(len(self.ncbd)) for _ in range(3)]
        for c in range(3):
            n[c][self.off[f]:self.off[f + 1]] = nrm[c]
        return n

    def farfield(self, f, l, mach, alpha, beta, roa1, aa1, pa1, nrm, v, ps, temp, cson, y=None, z=None):
        nx, ny, nz = self._normals(f, nrm)
        n = v.shape[1]
        y = np.zeros(n) if y is None else y
        z = np.zeros(n) if z is None else z
        w0 = np.zeros(5)
        self.lib.ref_farfield(f + 1, l + 1, D(mach), D(alpha), D(beta), D(roa1), D(aa1), D(pa1), ptr(nx), ptr(ny),
                              ptr(nz), ptr(self.ncbd), ptr(self.ncin), ptr(v), ptr(ps), ptr(temp), ptr(cson),
                              ptr(y), ptr(z), ptr(w0))
        return w0

    def clsym(self, f, nrm, v):
        nx, ny, nz = self._normals(f, nrm)
        self.lib.ref_clsym(f + 1, ptr(nx), ptr(ny), ptr(nz), ptr(self.ncin), ptr(self.ncbd), ptr(v))

    def met_bord(self, kinds, kparoi, rokinf, roeinf, v):
        """met_bord.f90:4 over every frontier: kinds[f] in 'rc', 'para', 'nrd', 'sym'."""
        code = np.ascontiguousarray(np.array([["rc", "para", "nrd", "sym"].index(k) for k in kinds], dtype=np.int32))
        self.lib.ref_met_bord(self.nfr, ptr(code), int(kparoi), D(rokinf), D(roeinf), ptr(v), ptr(self.ncbd),
                              ptr(self.ncin))

    def rbte(self, frs, tq):
        fr = self._list(frs)
        self.lib.ref_rbte(len(fr), ptr(fr), ptr(tq), ptr(self.ncbd), ptr(self.ncin))

    def rbtc(self, frs, tq):
        fr = self._list(frs)
        self.lib.ref_rbtc(len(fr), ptr(fr), ptr(tq), ptr(self.ncbd), ptr(self.ncin))

    def zvisqc(self, frn, frc, v, mu, mut, tq, sn, vol, temp):
        """zvisqc.f90:4 over every block, then rbte over the frontiers frn and
        rbtc over the coincident frontiers frc; sn: the whole metric array."""
        a, c = self._list(frn), self._list(frc)
        self.lib.ref_zvisqc(len(a), ptr(a), len(c), ptr(c), ptr(v), ptr(mu), ptr(mut), ptr(tq), ptr(self.ncbd),
                            ptr(self.ncin), ptr(sn), ptr(vol), ptr(temp))

    def normals(self, nrm):
        """nxn, nyn, nzn over every face: nrm[f], the (3, faces) inward normals
        of frontier f, or None (zeros)."""
        n = [np.zeros(len(self.ncbd)) for _ in range(3)]
        for f, x in enumerate(nrm):
            if x is not None:
                for q in range(3):
                    n[q][self.off[f]:self.off[f + 1]] = x[q]
        return n

    def rbord(self, kinds, blocks, coin, n, v, ps, temp, cson, mach, alpha, beta, roa1, aa1, pa1, y, z):
        """rbord.f90:4 over every frontier (drop-in build only): kinds[f] in
        'rc', 'para', 'nrd', 'sym', blocks[f] 0-based, coin the coincident
        frontiers (nfbc), n = normals(...), one far-field state for every
        'nrd' frontier."""
        code = np.ascontiguousarray(np.array([["rc", "para", "nrd", "sym"].index(k) for k in kinds], dtype=np.int32))
        blk = np.ascontiguousarray(np.array([b_ + 1 for b_ in blocks], dtype=np.int32))
        c = self._list(coin)
        self.lib.ref_rbord(self.nfr, ptr(code), ptr(blk), len(c), ptr(c), ptr(self.ncbd), ptr(self.ncin), ptr(n[0]),
                           ptr(n[1]), ptr(n[2]), ptr(v), ptr(ps), ptr(temp), ptr(cson), D(mach), D(alpha), D(beta),
                           D(roa1), D(aa1), D(pa1), ptr(y), ptr(z))

    def met_rbse(self, frs, fd):
        fr = self._list(frs)
        self.lib.ref_met_rbse(len(fr), ptr(fr), ptr(fd), ptr(self.ncbd), ptr(self.ncin))

    def met_rbsc(self, frs, fd):
        fr = self._list(frs)
        self.lib.ref_met_rbsc(len(fr), ptr(fr), ptr(fd), ptr(self.ncbd), ptr(self.ncin))

    def initns(self, f, l, face, b, x, y, z):
        """Inward unit normals of frontier f = face `face` of block l (initns.f90)."""
        nx, ny, nz = (np.zeros(len(self.ncbd)) for _ in range(3))
        rng = {"i": (b.ii1, b.ii2), "j": (b.jj1, b.jj2), "k": (b.kk1, b.kk2)}
        lim = []
        for d in "ijk":
            lo, hi = rng[d]
            if d == face[0]:
                lo = hi = (lo if face[1] == "1" else hi)
            lim += [lo, hi]
        iface = "ijk".index(face[0]) * 2 + (0 if face[1] == "1" else 1) + 1
        self.lib.ref_initns(f + 1, l + 1, iface, *lim, ptr(x), ptr(y), ptr(z), ptr(self.ncbd), ptr(nx), ptr(ny),
                            ptr(nz))
        s = slice(self.off[f], self.off[f + 1])
        return nx[s].copy(), ny[s].copy(), nz[s].copy()
