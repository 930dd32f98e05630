"""CPU stage runner with the same sequence as codens_amd.solver.Solver, built
on the oracle (or the reference build): the checker of multi-step runs."""
from __future__ import annotations

import numpy as np

import cpu_ref as C


class CpuStage:
    def __init__(self, case, use_ref=False):
        self.case = case
        lay = case.layout
        n = lay.ip11
        ph = C.orc_phys(**case.phys)
        sc = case.scheme
        self.O = C.Orc(lay, ph, muscl=sc["muscl"], xk=sc["xk"], ilim=sc["ilim"], ns=case.ns)
        self.K = C.Ref(lay, ph, muscl=sc["muscl"], xk=sc["xk"], ilim=sc["ilim"], ns=case.ns) if use_ref else self.O
        self.v = np.ascontiguousarray(case.v0.copy())
        z = lambda *s: np.zeros(s)
        self.u = z(5, n); self.d = z(5, n); self.ff = z(5, n)
        self.ps = z(n); self.temp = z(n); self.cson = z(n); self.mu = z(n); self.mut = z(n); self.dt = z(n)
        self.tq = z(9, n)

    def stage(self):
        cs, K, O = self.case, self.K, self.O
        nb = len(cs.layout.blocks)
        for l in range(nb):
            K.zpres(l, self.v, self.ps, self.temp, self.cson)
        if cs.ns:
            for l in range(nb):
                K.zvismo(l, self.mu, self.v, self.temp)
        for l in range(nb):
            K.chronos(l, cs.eta, self.mu, self.mut, self.v, self.dt, cs.sn, cs.vol, self.cson)
        if cs.kdtl == 0:
            stop, dtmin = O.chrono_global(cs.dt1min, self.dt)
            assert not stop, f"global step {dtmin} below dt1min {cs.dt1min}"
        c, w, f = cs.coincident, cs.wall, cs.farfield
        if len(c.ghost):
            O.bc_coincident(c.ghost, c.inner, c.remote, self.v, self.ps, self.temp, self.cson)
        if len(w.ghost):
            O.bc_wall(w.ghost, w.inner, self.v, self.ps, self.temp, self.cson)
        if len(f.ghost):
            nrm = np.ascontiguousarray(f.normals)
            O.bc_farfield(f.ghost, f.inner, nrm[0].copy(), nrm[1].copy(), nrm[2].copy(),
                          np.ascontiguousarray(cs.winf, dtype=np.float64), self.v, self.ps, self.temp, self.cson)
        if len(c.ghost):
            O.bc_coincident(c.ghost, c.inner, c.remote, self.v, self.ps, self.temp, self.cson)
        if cs.ns:
            for l in range(nb):
                g = K.zgrad2(l, cs.sn, cs.vol, self.v, self.temp)
                K.zfluto(l, self.mu, self.mut, self.tq, g)
            gg = np.concatenate([w.ghost, f.ghost]); ii = np.concatenate([w.inner, f.inner])
            if len(gg):
                O.bc_tq_copy(gg, ii, self.tq)
            if len(c.ghost):
                O.bc_tq_coincident(c.ghost, c.inner, c.remote, self.tq)
        for l in range(nb):
            if cs.ns:
                K.sch_roe(l, 0, self.u, self.v, self.ff, self.tq, cs.sn, self.ps)
            else:
                K.sch_roe_euler(l, 0, self.u, self.v, self.ff, cs.sn, self.ps)
        for l in range(nb):
            if cs.kmf == 0:
                K.sch_expli(l, self.u, self.v, self.dt, cs.vol)
            else:
                K.implimf_3d(l, self.u, self.dt, self.v, self.d, self.ff, self.mu, self.mut, cs.lmax, cs.sn,
                             cs.vol, self.ps, self.cson)
