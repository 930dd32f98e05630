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
        nv = self.v.shape[0]
        self.u = z(nv, n); self.d = z(nv, n); self.ff = z(nv, n)
        t = case.turb
        if t is not None:
            self.fd = z(6, n); self.cfke = z(n); self.ts6 = z(n); self.ts7 = z(n)
            self.tc = C.OrcTurb(*t.tc)
        # with the reference build, its own boundary routines over the
        # case's frontiers: coincident pairs first, then walls, far fields
        # and symmetry planes, in listing order
        self.RF = None
        self.first = True
        # zvisqc in one call (stresses and their ghosts), if set: replaces the
        # per-block zgrad2 + zfluto and the stress ghosts below
        self.zvisqc = None
        if use_ref:
            self.RF, self.coin_fr, self.other_fr = self.frontiers(self.K)

    def frontiers(self, K):
        """The case's frontiers as the reference's boundary routines of the
        build behind K see them: (RefFrontiers, coincident frontier numbers,
        the others)."""
        case = self.case
        fr, part = [], []
        c = case.coincident
        for f in range(len(c.sizes)):
            s_ = c.frontier(f)
            fr.append(("rc", c, s_, c.blocks[f]))
            part.append(f ^ 1)
        for kind, fs in (("para", case.wall), ("nrd", case.farfield), ("sym", case.sym)):
            for f in range(len(fs.sizes)):
                fr.append((kind, fs, fs.frontier(f), fs.blocks[f]))
                part.append(-1)
        self.fr = fr
        RF = C.RefFrontiers(K, [x[1].ghost[x[2]] for x in fr], [x[1].inner[x[2]] for x in fr], part)
        return (RF, [k for k, x in enumerate(fr) if x[0] == "rc"], [k for k, x in enumerate(fr) if x[0] != "rc"])

    def stage(self):
        """One stage of Solver: thermodynamics, time step, ghosts, [turbulence],
        stresses, residual, then the explicit or the implicit update."""
        self.thermo()
        self.time_step()
        self.residual()
        cs, K = self.case, self.K
        for l in range(len(cs.layout.blocks)):
            if cs.kmf == 0:
                K.sch_expli(l, self.u, self.v, self.dt, cs.vol)
            else:
                K.implimf_3d(l, self.u, self.dt, self.v, self.d, self.ff, self.mu, self.mut, cs.lmax, cs.sn,
                             cs.vol, self.ps, self.cson)

    RK_ALPHA = (0.25, 1.0 / 3.0, 0.5, 1.0)

    def rk4_step(self):
        """One step of the 4-stage explicit scheme, in Solver.rk4_step's order:
        the step's dt from the initial state, then per stage the residual of
        the previous stage's state and v = v0 - alpha dt/vol u (oracle
        orc_rk_stage, real cells)."""
        cs = self.case
        self.thermo()
        self.time_step()
        v0 = self.v.copy()
        for s, a in enumerate(self.RK_ALPHA):
            if s:
                self.thermo()
            self.residual()
            for l in range(len(cs.layout.blocks)):
                self.O.rk_stage(l, a, v0, self.u, self.v, self.dt, cs.vol)

    def thermo(self):
        cs, K = self.case, self.K
        nb = len(cs.layout.blocks)
        for l in range(nb):
            K.zpres(l, self.v, self.ps, self.temp, self.cson)
        if cs.ns:
            for l in range(nb):
                K.zvismo(l, self.mu, self.v, self.temp)

    def time_step(self):
        cs, K, O = self.case, self.K, self.O
        for l in range(len(cs.layout.blocks)):
            K.chronos(l, cs.eta, self.mu, self.mut, self.v, self.dt, cs.sn, cs.vol, self.cson)
        if cs.kdtl == 0:
            stop, dtmin = O.chrono_global(cs.dt1min, self.dt)
            assert not stop, f"global step {dtmin} below dt1min {cs.dt1min}"

    def residual(self):
        """Ghosts (rfve in the first cycle), turbulence, stresses and the
        residual u of the current state."""
        cs, K, O = self.case, self.K, self.O
        nb = len(cs.layout.blocks)
        if self.first:
            # atsch_num.f90:450-458: rfve on the non-coincident frontiers in
            # the first cycle
            self.first = False
            if self.RF is not None:
                if self.other_fr:
                    self.RF.rfve(self.other_fr, self.v, self.ps, self.temp, self.cson)
            else:
                w, f, y = cs.wall, cs.farfield, cs.sym
                gg = np.concatenate([w.ghost, f.ghost, y.ghost]); ii = np.concatenate([w.inner, f.inner, y.inner])
                if len(gg):
                    C.orc_rfve(O.lib, gg, ii, self.v, self.ps, self.temp, self.cson)
        if self.RF is not None:
            self.rbord_ref()
        else:
            self.rbord()
        c, w, f, y = cs.coincident, cs.wall, cs.farfield, cs.sym
        if cs.turb is not None:
            self.turb_stage()
        if cs.ns and self.zvisqc is not None:
            self.zvisqc()
        elif cs.ns:
            for l in range(nb):
                g = K.zgrad2(l, cs.sn, cs.vol, self.v, self.temp)
                K.zfluto(l, self.mu, self.mut, self.tq, g)
            if self.RF is not None:  # zvisqc.f90:182-205
                if self.other_fr:
                    self.RF.rbte(self.other_fr, self.tq)
                if self.coin_fr:
                    self.RF.rbtc(self.coin_fr, self.tq)
            else:
                gg = np.concatenate([w.ghost, f.ghost, y.ghost]); ii = np.concatenate([w.inner, f.inner, y.inner])
                if len(gg):
                    O.bc_tq_copy(gg, ii, self.tq)
                if len(c.ghost):
                    O.bc_tq_coincident(c.ghost, c.inner, c.remote, self.tq)
        for l in range(nb):
            if cs.ns:
                K.sch_roe(l, 0, self.u, self.v, self.ff, self.tq, cs.sn, self.ps)
            else:
                K.sch_roe_euler(l, 0, self.u, self.v, self.ff, cs.sn, self.ps)

    def rbord(self):
        """rbord.f90:137-429 on the oracle: coincident ghosts, walls, far
        fields, symmetry planes, coincident ghosts again."""
        cs, O = self.case, self.O
        c, w, f, y = cs.coincident, cs.wall, cs.farfield, cs.sym
        if len(c.ghost):
            O.bc_coincident(c.ghost, c.inner, c.remote, self.v, self.ps, self.temp, self.cson)
        if len(w.ghost):
            O.bc_wall(w.ghost, w.inner, self.v, self.ps, self.temp, self.cson)
        if len(f.ghost):
            nrm = np.ascontiguousarray(f.normals)
            O.bc_farfield(f.ghost, f.inner, nrm[0].copy(), nrm[1].copy(), nrm[2].copy(),
                          np.ascontiguousarray(cs.winf, dtype=np.float64), self.v, self.ps, self.temp, self.cson)
        if len(y.ghost):
            nrm = np.ascontiguousarray(y.normals)
            C.orc_bc_sym(O.lib, y.ghost, y.inner, nrm[0].copy(), nrm[1].copy(), nrm[2].copy(), self.v)
        if len(c.ghost):
            O.bc_coincident(c.ghost, c.inner, c.remote, self.v, self.ps, self.temp, self.cson)

    def rbord_ref(self):
        """rbord.f90:137-429 with the reference's own routines: rbvc, then
        every frontier (clpara; utnrd, rbve and clnrd; clsym), then rbvc."""
        cs, RF = self.case, self.RF
        a = (self.v, self.ps, self.temp, self.cson)
        if self.coin_fr:
            RF.rbvc(self.coin_fr, *a)
        spec = cs.farspec
        for k, (kind, fs, s_, l) in enumerate(self.fr):
            if kind == "para":
                RF.clpara(k, *a)
            elif kind == "nrd":
                RF.farfield(k, l, spec.mach, spec.alpha, spec.beta, spec.roa1, spec.aa1, spec.pa1,
                            np.ascontiguousarray(fs.normals[:, s_]), *a)
            elif kind == "sym":
                RF.clsym(k, np.ascontiguousarray(fs.normals[:, s_]), self.v)
        if self.coin_fr:
            RF.rbvc(self.coin_fr, *a)

    def turb_stage(self):
        """met_num.f90:162-930 for SA / Wilcox k-omega without wall laws, the
        same sequence as Solver.turb_stage, on the oracle."""
        import ctypes
        from oracle import ptr
        cs, O = self.case, self.O
        t, lay = cs.turb, cs.layout
        n = lay.ip11
        sa = t.model == "1SA"
        L_, D = C.L_, C.D
        # met_num.f90:145-159: met_bord (the reference's own with its build),
        # then met_parko for k-omega (the oracle's, pinned bit for bit against
        # the reference's in tests/test_oracle_metbord.py: the reference's
        # needs the wall tables of atindnor, which this driver does not set)
        w, f, c, y = cs.wall, cs.farfield, cs.coincident, cs.sym
        if self.RF is not None:
            self.RF.met_bord([x[0] for x in self.fr], t.kparoi, t.rokinf, t.roeinf, self.v)
        else:
            lists = [(0, c.ghost, c.inner, c.remote), (1 + t.kparoi, w.ghost, w.inner, w.inner),
                     (3, f.ghost, f.inner, f.inner), (4, y.ghost, y.inner, y.inner)]
            for kind, gg, ii, rr in lists:
                if len(gg):
                    gg, ii, rr = (np.ascontiguousarray(np.asarray(a, dtype=np.int64)) for a in (gg, ii, rr))
                    O.lib.orc_met_bord(kind, L_(len(gg)), ptr(gg), ptr(ii), ptr(rr), D(t.rokinf), D(t.roeinf),
                                       ptr(self.v), L_(n))
        if t.kparoi == 1:
            ln = t.wall_lines
            rc = O.lib.orc_met_parko(L_(len(ln)), ln.ctypes.data_as(ctypes.c_void_p),
                                     ptr(np.ascontiguousarray(t.wall_ghost)), ptr(t.mnpar), ptr(self.mu),
                                     ptr(t.dist), D(0.075), ptr(self.v), L_(n))
            assert rc == 0, "met_parko: a wall line shorter than 10 cells (the reference stops)"
        g9s, tps = [], []
        for l, b in enumerate(lay.blocks):
            g9 = O.teq_gradv(l, cs.sn, cs.vol, self.v)
            tp = O.met_prod(l, self.mut, self.v, g9)
            O.met_gradtr(l, cs.sn, cs.vol, self.v, self.fd)
            if sa:
                O.met_smsa(l, self.tc, cs.sn, cs.vol, self.v, self.mu, t.dist, g9, self.fd, self.cfke, self.ts6,
                           self.ts7)
            else:
                O.met_smko(l, self.tc, self.v, tp, self.cfke, self.ts6, self.ts7, self.cson)
            O.met_flud(l, t.kfludis, self.tc, self.v, self.mu, self.mut, self.fd)
            g9s.append(g9); tps.append(tp)
        if self.RF is not None:  # met_num.f90:575-600
            if self.other_fr:
                self.RF.met_rbse(self.other_fr, self.fd)
            if self.coin_fr:
                self.RF.met_rbsc(self.coin_fr, self.fd)
        else:
            gg = np.ascontiguousarray(np.concatenate([w.ghost, f.ghost, y.ghost]).astype(np.int64))
            ii = np.ascontiguousarray(np.concatenate([w.inner, f.inner, y.inner]).astype(np.int64))
            if len(gg):
                O.lib.orc_met_rbse(L_(len(gg)), ptr(gg), ptr(ii), ptr(self.fd), L_(n))
            if len(c.ghost):
                O.lib.orc_met_rbsc(L_(len(c.ghost)), ptr(c.ghost), ptr(c.inner), ptr(c.remote), ptr(self.fd),
                                   L_(n))
        for l, b in enumerate(lay.blocks):
            if t.klroe == 2:
                O.met_roe2o(l, self.v, self.d, cs.sn)
            else:
                O.met_roe(l, self.v, self.d, cs.sn)
            O.sch_turb(l, self.u, self.v, self.d, self.fd, self.ts6, self.ts7, cs.sn, cs.vol)
            ob = C.orc_block(b)
            snb = np.ascontiguousarray(cs.sn[9 * b.npc: 9 * (b.npc + b.nnn)])
            O.lib.orc_impli2_eqt_3d(ctypes.byref(ob), ptr(self.u), L_(n), ptr(self.dt), ptr(self.v), ptr(self.mu),
                                    ptr(self.mut), ptr(self.cfke), ptr(snb), ptr(cs.vol), D(t.cmt))
            if sa:
                O.lib.orc_met_cutsa(ctypes.byref(ob), D(t.epsk), ptr(self.v), L_(n))
            else:
                O.lib.orc_met_cut(ctypes.byref(ob), D(t.epsk), D(t.epse), ptr(self.v), L_(n))
        for l, b in enumerate(lay.blocks):
            ob = C.orc_block(b)
            if sa:
                O.lib.orc_met_samut(ctypes.byref(ob), ctypes.byref(self.tc), ptr(self.v), L_(n), ptr(self.mu),
                                    ptr(self.mut))
            else:
                g9 = g9s[l]
                if t.kwsst:
                    g9 = O.teq_gradv(l, cs.sn, cs.vol, self.v)
                O.lib.orc_met_kocmut(ctypes.byref(ob), ctypes.byref(self.tc), t.kwsst, ptr(self.v), L_(n),
                                     ptr(self.mu), ptr(self.mut), ptr(t.dist), ptr(np.ascontiguousarray(g9)),
                                     L_(b.nnn))
