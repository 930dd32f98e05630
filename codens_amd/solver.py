"""Device-resident stage driver: the host-side mirror of atsch_num.f90 for the
mean-flow paths of the synthetic cases.

Every array lives in HBM for the whole run, in the reference storage; one
stage is a fixed sequence of C-ABI launches on one stream with no host
round trip (the global time step of kdtl=0 aside, which is a reduction the
reference also synchronises on, chrono.f90:132).

Stage order (atsch_num.f90): zpres (:245) -> zvismo (:424, NS) -> chronos
(:436) -> rbord ghosts (:527; rbvc, walls, far fields, symmetry planes,
rbvc) -> zvisqc (:563, NS; then rbte/rbtc) -> sch_roe / sch_roe_euler
(:720-751) -> sch_expli (kmf=0, :1021) or implimf_3d (kmf=1, :1040).
"""
from __future__ import annotations

import numpy as np
import torch

from . import abi
from .configs import Case


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


BATCH_PHASES = ("residual", "implicit", "turbulence")


class Solver:
    def __init__(self, case: Case, stream=None, implicit="fp64", streams=1, batched=False):
        """implicit: "fp64" (implimf_3d as the reference) or "fp32" (its FP32
        defect-correction form: FP64 residual and update, FP32 sub-iterations).
        streams > 1: the per-block work of each phase runs on that many side
        streams (up to 16), each block with the workspace set of its stream.
        batched: phases run for all blocks in one launch per pass instead of
        block by block: "residual" (the Navier-Stokes residual,
        codens_sch_roe_batch), "implicit" (the FP64 implicit phase,
        codens_implimf_3d_batch), "turbulence" (the turbulence line-implicit
        phase, codens_impli2_eqt_3d_batch); True for all three, or a
        collection (or comma-separated string) of their names."""
        if implicit not in ("fp64", "fp32"):
            raise ValueError(f"implicit must be 'fp64' or 'fp32', not {implicit!r}")
        self.implicit = implicit
        self.case = case
        lay = case.layout
        self.lay = lay
        n = lay.ip11
        self.n = n
        self.blocks = [abi.block(b) for b in lay.blocks]
        self.ph = abi.phys(**case.phys)
        self.sc = abi.scheme(**case.scheme)
        self.stream = stream
        f = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")
        self.v = _dev(case.v0)
        nv = case.v0.shape[0]  # 7 with a turbulence model (planes 6-7)
        self.u = f(nv, n)
        self.d = f(nv, n)
        self.ff = f(nv, n)
        self.ps = f(n); self.temp = f(n); self.cson = f(n)
        self.mu = f(n); self.mut = f(n); self.dt = f(n)
        self.tq = f(9, n)
        self.zero = f(n)
        self.vol = _dev(case.vol)
        self.sn = _dev(case.sn)
        self.sn_blk = [self.sn[9 * b.npc: 9 * (b.npc + b.nnn)] for b in lay.blocks]
        c, w, fa = case.coincident, case.wall, case.farfield
        i64 = lambda a: _dev(a, torch.int64)
        self.coin = (i64(c.ghost), i64(c.inner), i64(c.remote if c.remote is not None else c.inner))
        self.wall = (i64(w.ghost), i64(w.inner))
        self.far = (i64(fa.ghost), i64(fa.inner), _dev(fa.normals if fa.normals is not None else np.zeros((3, 0))))
        sy = case.sym
        self.sym = (i64(sy.ghost), i64(sy.inner), _dev(sy.normals if sy.normals is not None else np.zeros((3, 0))))
        # rbte / met_rbse walk every frontier that is not coincident
        self.tq_copy = (torch.cat([self.wall[0], self.far[0], self.sym[0]]),
                        torch.cat([self.wall[1], self.far[1], self.sym[1]]))
        self.nbad = torch.zeros(1, dtype=torch.int32, device="cuda")
        t = case.turb
        if t is not None:
            self.dist = _dev(t.dist)
            if t.kparoi == 1:
                if t.wall_lines is None:
                    raise ValueError("a k-omega case needs its wall points (met_parko, kparoi = 1)")
                self.parko = (np.asarray(t.wall_start, dtype=np.int64),
                              torch.as_tensor(t.wall_lines.view(np.uint8).copy(), device="cuda"),
                              _dev(t.wall_ghost, torch.int64), _dev(t.mnpar, torch.int32))
            self.fd = f(6, n)
            self.cfke = f(n); self.ts6 = f(n); self.ts7 = f(n)
            self.g9 = [f(9, b.nnn) for b in lay.blocks]
            self.tprod = [f(b.nnn) for b in lay.blocks]
        self.cycle = 0
        self.first_ghosts = True  # rfve before the first rbord only
        self.v0 = None  # the step's initial state (4-stage explicit scheme)
        self.res0 = None
        self.graph = None
        if batched is True:
            phases = set(BATCH_PHASES)
        elif not batched:
            phases = set()
        else:
            phases = set(batched.split(",") if isinstance(batched, str) else batched)
        if not phases <= set(BATCH_PHASES):
            raise ValueError(f"batched phases must be among {BATCH_PHASES}, not {sorted(phases)}")
        if not case.ns:
            phases.discard("residual")
        if case.kmf != 1 or implicit != "fp64":
            phases.discard("implicit")
        if case.turb is None:
            phases.discard("turbulence")
        self.batch = phases
        self.blocks_arr = abi.block_array(self.blocks)
        if not 1 <= streams <= 16:
            raise ValueError("streams must be 1..16")
        self.side = [torch.cuda.Stream() for _ in range(streams)] if streams > 1 and len(self.blocks) > 1 else []
        # blocks to streams: largest first, each to the least-loaded stream
        # (longest-processing-time order), so that the streams finish together
        self.lane = [0] * len(self.blocks)
        if self.side:
            load = [0] * len(self.side)
            for l in sorted(range(len(self.blocks)), key=lambda x: -lay.blocks[x].ncells):
                k = min(range(len(load)), key=lambda x: load[x])
                self.lane[l] = k
                load[k] += lay.blocks[l].ncells

    # -- pieces of one stage ------------------------------------------------
    # -- per-block work, optionally on several streams -----------------------
    def _per_block(self, fn):
        """fn(l, b, stream) for every block.  With concurrent streams each
        block goes to its side stream (self.lane), with that stream's
        workspace set; the side streams fork from and join back to the
        caller's stream: blocks are independent between two ghost exchanges,
        so small blocks run side by side instead of leaving most of the chip
        idle."""
        if not self.side:
            for l, b in enumerate(self.blocks):
                fn(l, b, self.stream)
            return
        main = self.stream if self.stream is not None else torch.cuda.current_stream()
        for q in self.side:
            q.wait_stream(main)
        try:
            for l, b in enumerate(self.blocks):
                k = self.lane[l]
                abi.workspace_set(k)
                fn(l, b, self.side[k])
        finally:
            abi.workspace_set(0)
        for q in self.side:
            main.wait_stream(q)

    # -- pieces of one stage ------------------------------------------------
    def thermo(self):
        ns = self.case.ns

        def one(l, b, s):
            abi.zpres(b, self.ph, self.v, self.ps, self.temp, self.cson, stream=s)
            if ns:
                abi.zvismo(b, self.ph, self.temp, self.mu, self.nbad, stream=s)
        self._per_block(one)

    def time_step(self):
        cs = self.case
        self._per_block(lambda l, b, s: abi.chronos(b, self.ph, cs.eta, cs.ns, self.v, self.mu, self.mut,
                                                    self.sn_blk[l], self.vol, self.cson, self.dt, stream=s))
        if cs.kdtl == 0:
            return abi.chrono_global(self.blocks, cs.dt1min, self.dt, stream=self.stream)
        return None

    def ghosts(self):
        s = self.stream
        if self.first_ghosts and self.tq_copy[0].numel():
            # first cycle: rfve on the non-coincident frontiers
            # (atsch_num.f90:450-458); rbord then rewrites every ghost but
            # the p, T and c of symmetry ghosts
            abi.rfve(self.tq_copy[0], self.tq_copy[1], self.v, self.ps, self.temp, self.cson, stream=s)
        self.first_ghosts = False
        g, i, r = self.coin
        if g.numel():
            abi.bc_coincident(g, i, r, self.v, self.ps, self.temp, self.cson, stream=s)
        if self.wall[0].numel():
            abi.bc_wall(self.wall[0], self.wall[1], self.v, self.ps, self.temp, self.cson, stream=s)
        if self.far[0].numel():
            abi.bc_farfield(self.far[0], self.far[1], self.far[2], self.case.winf, self.ph, self.v, self.ps,
                            self.temp, self.cson, stream=s)
        if self.sym[0].numel():
            abi.bc_sym(self.sym[0], self.sym[1], self.sym[2], self.v, stream=s)
        if g.numel():
            abi.bc_coincident(g, i, r, self.v, self.ps, self.temp, self.cson, stream=s)

    def turb_stage(self):
        """The turbulence stage of met_num.f90:162-930 for SA or Wilcox
        k-omega without wall laws: per block the velocity gradients
        (teq_gradv), the production (met_prod), the gradients of the
        transported variables (met_gradtr), the model's sources and
        diffusive flux densities; their ghost values (met_rbse, met_rbsc);
        per block the upwind dissipation (met_roe2o or met_roe), the flux
        balance u(6:7) (sch_turb), the line-implicit update (impli2_eqt_3d)
        and the limiters (at_cutke), then the eddy viscosity (a block's eddy
        viscosity reads that block's state only)."""
        t = self.case.turb
        sa = t.model == "1SA"
        s = self.stream
        # met_num.f90:145-159: the turbulence ghosts of every frontier
        # (met_bord: coincident averages first), then the omega wall
        # condition of the k-omega models (met_parko)
        g, i, r = self.coin
        if g.numel():
            abi.met_bord(0, g, i, r, 0.0, 0.0, self.v, stream=s)
        if self.wall[0].numel():
            abi.met_bord(1 + t.kparoi, self.wall[0], self.wall[1], None, 0.0, 0.0, self.v, stream=s)
        if self.far[0].numel():
            abi.met_bord(3, self.far[0], self.far[1], None, t.rokinf, t.roeinf, self.v, stream=s)
        if self.sym[0].numel():
            abi.met_bord(4, self.sym[0], self.sym[1], None, 0.0, 0.0, self.v, stream=s)
        if t.kparoi == 1:
            fs, lines, ghost, mnpar = self.parko
            abi.met_parko(fs, lines, ghost, mnpar, self.mu, self.dist, 0.075, self.v, stream=s)

        def sources(l, b, s):
            abi.teq_gradv(b, self.sn_blk[l], self.vol, self.v, self.g9[l], stream=s)
            abi.met_prod(b, self.mut, self.v, self.g9[l], self.tprod[l], stream=s)
            abi.met_gradtr(b, self.sn_blk[l], self.vol, self.v, self.fd, stream=s)
            if sa:
                abi.met_smsa(b, t.tc, self.sn_blk[l], self.vol, self.v, self.mu, self.dist, self.g9[l], self.fd,
                             self.cfke, self.ts6, self.ts7, stream=s)
            else:
                abi.met_smko(b, t.tc, self.v, self.tprod[l], self.cfke, self.ts6, self.ts7, self.cson, stream=s)
            abi.met_flud(b, t.kfludis, t.tc, self.v, self.mu, self.mut, self.fd, stream=s)
        self._per_block(sources)
        s = self.stream
        if self.tq_copy[0].numel():
            abi.met_rbse(self.tq_copy[0], self.tq_copy[1], self.fd, stream=s)
        g, i, r = self.coin
        if g.numel():
            abi.met_rbsc(g, i, r, self.fd, stream=s)

        def balance(l, b, s):
            if t.klroe == 2:
                abi.met_roe2o(b, self.v, self.d, self.sn_blk[l], stream=s)
            else:
                abi.met_roe(b, self.v, self.d, self.sn_blk[l], stream=s)
            abi.sch_turb(b, self.u, self.v, self.d, self.fd, self.ts6, self.ts7, self.sn_blk[l], self.vol, stream=s)

        def close(l, b, s):
            abi.met_cut(b, 2 if sa else 1, t.epsk, t.epse, self.v, stream=s)
            if sa:
                abi.met_samut(b, t.tc, self.v, self.mu, self.mut, stream=s)
            else:
                abi.met_kocmut(b, t.tc, t.kwsst, self.sn_blk[l], self.vol, self.v, self.mu, self.mut, self.dist,
                               self.g9[l], stream=s)

        def update(l, b, s):
            balance(l, b, s)
            abi.impli2_eqt_3d(b, self.u, self.dt, self.v, self.mu, self.mut, self.cfke, self.sn_blk[l], self.vol,
                              t.cmt, stream=s)
            close(l, b, s)
        if "turbulence" in self.batch:
            # the line-implicit phase of all blocks in one launch per pass
            self._per_block(balance)
            abi.impli2_eqt_3d_batch(self.blocks_arr, len(self.blocks), self.u, self.dt, self.v, self.mu, self.mut,
                                    self.cfke, self.sn, self.vol, t.cmt, stream=self.stream)
            self._per_block(close)
        else:
            self._per_block(update)

    def stresses(self):
        self._per_block(lambda l, b, s: abi.zvisqc(b, self.ph, self.sn_blk[l], self.vol, self.v, self.temp, self.mu,
                                                   self.mut, self.tq, stream=s))
        s = self.stream
        if self.tq_copy[0].numel():
            abi.bc_tq_copy(self.tq_copy[0], self.tq_copy[1], self.tq, stream=s)
        g, i, r = self.coin
        if g.numel():
            abi.bc_tq_coincident(g, i, r, self.tq, stream=s)

    def residual(self):
        if "residual" in self.batch:
            abi.sch_roe_batch(self.blocks_arr, len(self.blocks), self.ph, self.sc, 0, self.u, self.v, self.ff, self.tq,
                              self.sn, self.ps, stream=self.stream)
            return

        def one(l, b, s):
            if self.case.ns:
                abi.sch_roe(b, self.ph, self.sc, 0, self.u, self.v, self.ff, self.tq, self.sn_blk[l], self.ps, stream=s)
            else:
                abi.sch_roe_euler(b, self.ph, self.sc, 0, self.u, self.v, self.ff, self.sn_blk[l], self.ps, stream=s)
        self._per_block(one)

    def update(self):
        cs = self.case
        if "implicit" in self.batch:
            abi.implimf_3d_batch(self.blocks_arr, len(self.blocks), self.ph, self.u, self.dt, self.v, self.d, self.ff,
                                 self.mu, self.mut, cs.lmax, self.sn, self.vol, self.cson, stream=self.stream)
            return

        def one(l, b, s):
            if cs.kmf == 0:
                abi.sch_expli(b, self.u, self.v, self.dt, self.vol, stream=s)
            else:
                fn = abi.implimf_3d if self.implicit == "fp64" else abi.implimf_3d_f32
                fn(b, self.ph, self.u, self.dt, self.v, self.d, self.ff, self.mu, self.mut, cs.lmax, self.sn_blk[l],
                   self.vol, self.ps, self.cson, stream=s)
        self._per_block(one)

    RK_ALPHA = (0.25, 1.0 / 3.0, 0.5, 1.0)

    def rk4_step(self):
        """One step of the 4-stage explicit scheme composed from the
        single-stage residual u (net outgoing flux; the reference's explicit
        update is forward Euler only, sch_expli.f90:53-66):
        v_s = v_0 - alpha_s dt/vol u(v_{s-1}), alpha = 1/4, 1/3, 1/2, 1, with
        the physical sign and the step's dt from the initial state (chronos;
        the global step when kdtl = 0).  Each stage recomputes the
        thermodynamics, ghosts, stresses and residual of its state.  Mean
        flow only."""
        cs = self.case
        if cs.turb is not None:
            raise ValueError("the 4-stage explicit scheme covers the mean flow only")
        if self.v0 is None:
            self.v0 = torch.empty_like(self.v)
        self.thermo()
        self.time_step()
        if self.stream is not None:
            with torch.cuda.stream(self.stream):
                self.v0.copy_(self.v)
        else:
            self.v0.copy_(self.v)
        for st, a in enumerate(self.RK_ALPHA):
            if st:
                self.thermo()
            self.ghosts()
            if cs.ns:
                self.stresses()
            self.residual()
            self._per_block(lambda l, b, s: abi.rk_stage(b, a, self.v0, self.u, self.v, self.dt, self.vol, nvar=5,
                                                         stream=s))
        self.cycle += 1

    def stage(self):
        """One residual + update of all blocks (atsch_num.f90:5); a replay of
        the captured graph after capture()."""
        if self.graph is not None:
            if self.stream is not None:
                with torch.cuda.stream(self.stream):
                    self.graph.replay()
            else:
                self.graph.replay()
        else:
            self._stage_body()
        self.cycle += 1

    def _stage_body(self):
        self.thermo()
        self.time_step()
        self.ghosts()
        if self.case.turb is not None:
            self.turb_stage()  # before the mean-flow fluxes (atsch_num.f90:543 < 621)
        if self.case.ns:
            self.stresses()
        self.residual()
        self.update()

    def capture(self):
        """Record one stage into a HIP graph (torch.cuda.CUDAGraph); later
        calls of stage() replay it.  The first stage runs eagerly, so that
        every workspace the stage uses is allocated before the capture (no
        allocation may happen while a stream is captured); the graph then
        holds the identical launch sequence with the identical pointers.
        Only with local time steps (kdtl=1): the global step (kdtl=0) reads a
        reduction on the host inside the stage.  No launch in the stage
        records a memset, so replays repeat the eager results exactly."""
        if self.case.kdtl == 0:
            raise ValueError("capture needs local time steps (kdtl=1): the global step syncs the host")
        if self.graph is not None:
            return
        self.stage()
        saved, self.stream = self.stream, None  # launch on the capturing stream
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        try:
            with torch.cuda.graph(g, stream=side):
                self._stage_body()
        finally:
            self.stream = saved
        torch.cuda.current_stream().wait_stream(side)
        self.graph = g

    def check(self):
        """Raises CodensError(CODENS_ENONPHYS) if zvismo met T <= 0 in any
        stage so far (where the reference aborts, zvismo.f90:122-127) or the
        state now holds a real cell with rho <= 0, p <= 0 or a non-finite
        value.  A host synchronisation; the solver calls it wherever it reads
        results anyway (density_residual, norms, host_state)."""
        nv = 7 if self.case.turb is not None else 5
        abi.check_state(self.blocks_arr, len(self.blocks), nv, self.ph, self.v, self.nbad, stream=self.stream)
        bad = int(self.nbad.item())
        if bad:
            self.nbad.zero_()
            raise abi.CodensError(3, f"non-physical state after stage {self.cycle}: {bad} cell(s) with T <= 0 in "
                                     f"zvismo, rho <= 0, p <= 0 or a non-finite value")

    def density_residual(self):
        """||(rho^{n+1}-rho^n)/dt||_2 of the last implicit stage: the increment
        d holds rho^{n+1}-rho^n on the real cells (dua_resro.f90:63-69)."""
        self.check()
        return float(np.sqrt(abi.sumsq_rate(self.blocks, self.d[0], self.zero, self.dt, stream=self.stream)))

    def norms(self):
        """The steadiness norms of the last implicit stage (residu.f90 per
        block, rscpsv.f90 over all blocks) of (v^{n+1} - v^n)/dt: the
        increment d holds v^{n+1} - v^n on the real cells.  Returns an array
        (nblocks + 1, neq, 7): L1, L2, Linf, block, i, j, k of Linf; the last
        row is the aggregate."""
        self.check()
        nv = 7 if self.case.turb is not None else 5
        if getattr(self, "_norm_out", None) is None:
            self._norm_out = torch.zeros((len(self.blocks) + 1, nv, 7), dtype=torch.float64, device="cuda")
            self._zero_v = torch.zeros_like(self.d)
        abi.residu(self.blocks_arr, len(self.blocks), nv, self.d, self._zero_v, self.dt, self._norm_out,
                   stream=self.stream)
        torch.cuda.synchronize()
        return self._norm_out.cpu().numpy()

    def run_to(self, drop=1e-8, max_iter=10000, check_every=10):
        """Implicit stages until the density residual falls to `drop` times its
        value at the first stage (resite, dua_resro.f90:63-69; the reference's
        tol = 1e-8, inimem.f90:132), or max_iter stages.  The norm is read
        every check_every stages (a host sync).  Returns (stages, history of
        (stage, drop))."""
        self.stage()
        r1 = self.density_residual()
        hist = [(1, 1.0)]
        it = 1
        while it < max_iter:
            for _ in range(min(check_every, max_iter - it)):
                self.stage()
                it += 1
            r = self.density_residual() / r1
            hist.append((it, r))
            if not np.isfinite(r) or r <= drop:
                break
        return it, hist

    def host_state(self, check=True):
        torch.cuda.synchronize()
        if check:
            self.check()
        return self.v.cpu().numpy()

    def _restart_args(self):
        ns = self.case.ns
        nturb = 2 if self.case.turb is not None else 0  # equat(6:7) = 'ke' or a 7-variable RANS model
        return ns, nturb

    def save_restart(self, path):
        """Writes the restart file of writda.f90 (as svfw.f90:58 writes it):
        per block rho, rho u, rho v, rho w, rho E, then mut (Navier-Stokes)
        and the two turbulence variables (with a turbulence model), over the
        real cells.  A copy of the state to the host, then host file I/O."""
        ns, nturb = self._restart_args()
        v = np.ascontiguousarray(self.host_state())
        mut = np.ascontiguousarray(self.mut.cpu().numpy()) if ns else None
        abi.restart_write(path, self.blocks_arr, len(self.blocks), v, mut, nturb)

    def load_restart(self, path):
        """Reads a writda.f90 restart (readda.f90:4) into the real cells of
        the state (and mut); ghost entries keep their values and the next
        stage's boundary routines set them.  Must not be called while a
        captured graph is in use with other pointers (the arrays are updated
        in place, so a captured graph stays valid)."""
        ns, nturb = self._restart_args()
        torch.cuda.synchronize()
        v = np.ascontiguousarray(self.v.cpu().numpy())
        mut = np.ascontiguousarray(self.mut.cpu().numpy()) if ns else None
        abi.restart_read(path, self.blocks_arr, len(self.blocks), v, mut, nturb)
        self.v.copy_(torch.as_tensor(v))
        if ns:
            self.mut.copy_(torch.as_tensor(mut))
        torch.cuda.synchronize()
