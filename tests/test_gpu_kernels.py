"""Kernel-level parity of every C-ABI entry point against the oracle, on
seeded random states of a two-block curvilinear layout (8x6x5 and 5x4x3
cells).  Bar: max|gpu - oracle| / max|oracle| <= 1e-12 per field."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import cpu_ref as C
from randcase import GAM, all_thermo, random_mesh, random_state, relerr

TOL = 1e-12
DIMS = [(8, 6, 5), (5, 4, 3)]


@pytest.fixture(scope="module")
def ctx():
    from codens_amd import abi
    assert torch.cuda.is_available()
    lay, sn, vol = random_mesh(DIMS, seed=11)
    return abi, lay, sn, vol


def T(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype or torch.float64)


def H(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("pinfl", [0.0, 0.3])
def test_thermo_dt(ctx, pinfl):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(1)
    n = lay.ip11
    v = random_state(n, rng, pinfl)
    ph_o = C.orc_phys(pinfl=pinfl)
    O = C.Orc(lay, ph_o)
    ph = abi.phys(pinfl=pinfl)
    ps, t, c = all_thermo(v, pinfl)
    ps_o, t_o, c_o = ps.copy(), t.copy(), c.copy()
    mu_o = np.zeros(n); mut = np.zeros(n); dt_o = np.zeros(n)
    gv = T(v); gps, gt, gc = T(ps), T(t), T(c); gmu = T(np.zeros(n)); gmut = T(mut); gdt = T(np.zeros(n))
    for l, b in enumerate(lay.blocks):
        O.zpres(l, v, ps_o, t_o, c_o)
        O.zvismo(l, mu_o, v, t_o)
        cb = abi.block(b)
        abi.zpres(cb, ph, gv, gps, gt, gc)
        abi.zvismo(cb, ph, gt, gmu)
    for name, a, b_ in (("ps", gps, ps_o), ("temp", gt, t_o), ("cson", gc, c_o), ("mu", gmu, mu_o)):
        assert relerr(H(a), b_) <= TOL, name
    for ns in (0, 1):
        for l, b in enumerate(lay.blocks):
            O.ns = bool(ns)
            O.chronos(l, 0.8, mu_o, mut, v, dt_o, sn, vol, c_o)
            abi.chronos(abi.block(b), ph, 0.8, ns, gv, gmu, gmut, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]), T(vol), gc, gdt)
        assert relerr(H(gdt), dt_o) <= TOL, f"chronos ns={ns}"


@pytest.fixture(params=[1, 0], ids=["fused", "split"])
def path(request):
    from codens_amd import abi
    prev = abi.residual_path(request.param)
    yield request.param
    abi.residual_path(prev)


@pytest.mark.parametrize("ns,ilim,pinfl", [(0, 0, 0.0), (0, 0, 0.3), (1, 0, 0.0), (1, 1, 0.0), (1, 1, 0.2)])
def test_roe_residual(ctx, path, ns, ilim, pinfl):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(2 + ilim)
    n = lay.ip11
    v = random_state(n, rng, pinfl)
    ps, _, _ = all_thermo(v, pinfl)
    tq = 0.01 * rng.standard_normal((9, n)) if ns else np.zeros((9, n))
    O = C.Orc(lay, C.orc_phys(pinfl=pinfl), muscl=1, xk=1 / 3, ilim=ilim, ns=bool(ns))
    ph, sc = abi.phys(pinfl=pinfl), abi.scheme(1, ilim, 1 / 3)
    u_o = np.zeros((5, n)); ff_o = np.ones((5, n))
    gu = T(np.zeros((5, n))); gff = T(np.ones((5, n))); gv = T(v); gps = T(ps); gtq = T(tq)
    for l, b in enumerate(lay.blocks):
        snb = T(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        if ns:
            O.sch_roe(l, 1, u_o, v, ff_o, tq, sn, ps)
            abi.sch_roe(abi.block(b), ph, sc, 1, gu, gv, gff, gtq, snb, gps)
        else:
            O.sch_roe_euler(l, 1, u_o, v, ff_o, sn, ps)
            abi.sch_roe_euler(abi.block(b), ph, sc, 1, gu, gv, gff, snb, gps)
    gu_h, gff_h = H(gu), H(gff)
    for e in range(5):
        assert relerr(gu_h[e], u_o[e]) <= TOL, f"u[{e}]"
        assert relerr(gff_h[e], ff_o[e]) <= TOL, f"ff[{e}]"


@pytest.mark.parametrize("ns,ilim", [(0, 0), (1, 0), (1, 1)])
def test_roe_residual_multitile(ns, ilim):
    """Blocks spanning several 16x16 tiles and 8-plane k slabs of the fused
    residual, with partial tiles and slabs on every high edge (37x21x19,
    16x16x8, 17x33x9): the fused kernel against the oracle within 1e-12, and
    against the face passes + gather within 1e-14 (the same face values,
    summed in the same order); with viscous terms, the all-blocks launch
    (one grid, and the persistent tile queue) against the per-block fused
    kernel bit for bit."""
    from codens_amd import abi
    lay, sn, vol = random_mesh([(37, 21, 19), (16, 16, 8), (17, 33, 9)], seed=23)
    rng = np.random.default_rng(31 + ilim)
    n = lay.ip11
    v = random_state(n, rng)
    ps, _, _ = all_thermo(v)
    tq = 0.01 * rng.standard_normal((9, n)) if ns else np.zeros((9, n))
    O = C.Orc(lay, C.orc_phys(), muscl=1, xk=1 / 3, ilim=ilim, ns=bool(ns))
    ph, sc = abi.phys(), abi.scheme(1, ilim, 1 / 3)
    u_o = np.zeros((5, n)); ff_o = np.ones((5, n))
    for l in range(len(lay.blocks)):
        if ns:
            O.sch_roe(l, 1, u_o, v, ff_o, tq, sn, ps)
        else:
            O.sch_roe_euler(l, 1, u_o, v, ff_o, sn, ps)
    res = {}
    for fused in (1, 0):
        prev = abi.residual_path(fused)
        try:
            gu = T(np.full((5, n), 3.0)); gff = T(np.ones((5, n))); gv = T(v); gps = T(ps); gtq = T(tq)
            for l, b in enumerate(lay.blocks):
                snb = T(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
                if ns:
                    abi.sch_roe(abi.block(b), ph, sc, 1, gu, gv, gff, gtq, snb, gps)
                else:
                    abi.sch_roe_euler(abi.block(b), ph, sc, 1, gu, gv, gff, snb, gps)
            res[fused] = (H(gu), H(gff))
        finally:
            abi.residual_path(prev)
    w = np.zeros(n, dtype=bool)  # the entries sch_roe writes
    for b in lay.blocks:
        w[b.idx(b.ii1 - 1, b.jj1 - 1, b.kk1 - 1):b.idx(b.ii2, b.jj2, b.kk2) + 1] = True
    for e in range(5):
        assert relerr(res[1][0][e, w], u_o[e, w]) <= TOL, f"u[{e}]"
        assert relerr(res[1][1][e], ff_o[e]) <= TOL, f"ff[{e}]"
        assert relerr(res[1][0][e, w], res[0][0][e, w]) <= 1e-14, f"fused vs split u[{e}]"
    if ns:
        # all blocks in one launch, per tile grid and as the persistent
        # tile-queue launch (twice: the second finds the queue the first left):
        # the per-block fused residual bit for bit
        arr = abi.block_array([abi.block(b) for b in lay.blocks])
        for path in (1, 2):
            prev = abi.residual_path(path)
            try:
                for rep in range(2):
                    gu = T(np.full((5, n), 3.0)); gff = T(np.ones((5, n)))
                    abi.sch_roe_batch(arr, len(lay.blocks), ph, sc, 1, gu, T(v), gff, T(tq), T(sn), T(ps))
                    np.testing.assert_array_equal(H(gu), res[1][0])
                    np.testing.assert_array_equal(H(gff), res[1][1])
            finally:
                abi.residual_path(prev)
    # both write exactly indc(i1-1,j1-1,k1-1)..indc(i2,j2,k2), ghosts zero
    for b in lay.blocks:
        n0, n1 = b.idx(b.ii1 - 1, b.jj1 - 1, b.kk1 - 1), b.idx(b.ii2, b.jj2, b.kk2)
        for f in (1, 0):
            u = res[f][0]
            assert np.all(u[:, b.npc:n0] == 3.0) and np.all(u[:, n1 + 1:b.npc + b.nnn] == 3.0), f
            assert np.array_equal(u[:, n0:n1 + 1] == 0.0, u_o[:, n0:n1 + 1] == 0.0), f


@pytest.mark.parametrize("ki2,ki4", [(0.5, 1.0 / 32), (1.0, 0.05)])
def test_dissip_jameson(ctx, ki2, ki4):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(17)
    n = lay.ip11
    v = random_state(n, rng)
    p, _, c = all_thermo(v)
    O = C.Orc(lay, C.orc_phys())
    d_o = np.zeros((5, n)); gd = T(np.zeros((5, n)))
    for l, b in enumerate(lay.blocks):
        O.dissip_jameson(l, v, d_o, sn, p, c, ki2, ki4)
        abi.dissip_jameson(abi.block(b), ki2, ki4, T(v), gd, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]), T(p), T(c))
    h = H(gd)
    for e in range(5):
        assert relerr(h[e], d_o[e]) <= TOL, e


@pytest.mark.parametrize("which", ["met_roe2o", "met_roe"])
def test_turbulence_upwind(ctx, which):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(19)
    n = lay.ip11
    t = np.zeros((7, n))
    t[:5] = random_state(n, rng)
    t[5] = 1e-3 * (1 + rng.random(n)); t[6] = 10 * (1 + rng.random(n))
    O = C.Orc(lay, C.orc_phys())
    d_o = np.zeros((7, n)); gd = T(np.zeros((7, n))); gt = T(t)
    for l, b in enumerate(lay.blocks):
        getattr(O, which)(l, t, d_o, sn)
        getattr(abi, which)(abi.block(b), gt, gd, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]))
    h = H(gd)
    for e in (5, 6):
        a, r = h[e], d_o[e]
        assert np.array_equal(np.isnan(a), np.isnan(r)), e
        ok = ~np.isnan(r)
        assert relerr(a[ok], r[ok]) <= TOL, e


def test_zvisqc(ctx):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(5)
    n = lay.ip11
    v = random_state(n, rng)
    _, t, _ = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-3 * rng.random(n)
    O = C.Orc(lay, C.orc_phys())
    ph = abi.phys()
    tq_o = np.zeros((9, n)); gtq = T(np.zeros((9, n)))
    for l, b in enumerate(lay.blocks):
        g = O.zgrad2(l, sn, vol, v, t)
        O.zfluto(l, mu, mut, tq_o, g)
        abi.zvisqc(abi.block(b), ph, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]), T(vol), T(v), T(t), T(mu), T(mut), gtq)
    h = H(gtq)
    for q in range(9):
        assert relerr(h[q], tq_o[q]) <= TOL, f"tq[{q}]"


@pytest.mark.parametrize("lmx,ityprk", [(1, 0), (3, 0), (10, 0), (2, 1)])
def test_implimf_3d(ctx, lmx, ityprk):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(7 + lmx)
    n = lay.ip11
    v = random_state(n, rng)
    ps, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((5, n)); ff = 1e-4 * rng.standard_normal((5, n))
    dt = 0.01 * (1 + rng.random(n))
    O = C.Orc(lay, C.orc_phys())
    ph = abi.phys()
    v_o = v.copy(); d_o = np.zeros((5, n))
    gv = T(v); gd = T(np.zeros((5, n)))
    for l, b in enumerate(lay.blocks):
        O.implimf_3d(l, u, dt, v_o, d_o, ff, mu, mut, lmx, sn, vol, ps, c, ityprk=ityprk)
        abi.implimf_3d(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), T(mu), T(mut), lmx,
                       T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]), T(vol), T(ps), T(c), ityprk=ityprk)
    hv, hd = H(gv), H(gd)
    for e in range(5):
        assert relerr(hd[e], d_o[e]) <= TOL, f"d[{e}]"
        assert relerr(hv[e], v_o[e]) <= TOL, f"v[{e}]"


@pytest.mark.parametrize("lmx,ityprk", [(1, 0), (4, 0), (10, 1)])
def test_implimf_3d_f32(ctx, lmx, ityprk):
    """FP32 defect-correction form against the FP64 oracle: the increment d
    within 2e-5 relative (FP32 sub-iterations: about 1e-7 per operation,
    grown over lmx sweeps) and v += d in FP64, so v - v0 carries that same
    relative error and nothing more."""
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(27 + lmx)
    n = lay.ip11
    v = random_state(n, rng)
    ps, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((5, n)); ff = 1e-4 * rng.standard_normal((5, n))
    dt = 0.01 * (1 + rng.random(n))
    O = C.Orc(lay, C.orc_phys())
    ph = abi.phys()
    v_o = v.copy(); d_o = np.zeros((5, n))
    gv = T(v); gd = T(np.full((5, n), 7.0))
    for l, b in enumerate(lay.blocks):
        O.implimf_3d(l, u, dt, v_o, d_o, ff, mu, mut, lmx, sn, vol, ps, c, ityprk=ityprk)
        abi.implimf_3d_f32(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), T(mu), T(mut), lmx,
                           T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]), T(vol), T(ps), T(c), ityprk=ityprk)
    hv, hd = H(gv), H(gd)
    inner = lay.interior()
    for e in range(5):
        assert relerr(hd[e], d_o[e]) <= 2e-5, f"d[{e}]"
        assert relerr(hv[e] - v[e], v_o[e] - v[e]) <= 2e-5, f"v[{e}]"
        assert relerr(hd[e], d_o[e]) > 0  # it is the FP32 path
    # the block ranges outside the real cells are zeroed, as implimf_3d does
    owned = np.concatenate([np.arange(b.npc, b.npc + b.nnn) for b in lay.blocks])
    ghost = np.setdiff1d(owned, inner)
    assert (hd[:, ghost] == 0).all()


@pytest.mark.parametrize("lmx,ityprk,kfmg", [(1, 0, 0), (3, 0, 0), (10, 0, 3), (2, 1, 0)])
def test_implimf_eu(ctx, lmx, ityprk, kfmg):
    """Euler implicit phase (implimf_eu.f90) against the oracle, 7-plane v and
    d as the reference stores them (planes 6-7 of d are added to v too)."""
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(17 + lmx)
    n = lay.ip11
    v = np.zeros((7, n)); v[:5] = random_state(n, rng); v[5:] = rng.random((2, n))
    ps, t, c = all_thermo(v[:5])
    u = 1e-3 * rng.standard_normal((5, n)); ff = 1e-4 * rng.standard_normal((5, n))
    dt = 0.01 * (1 + rng.random(n))
    O = C.Orc(lay, C.orc_phys(), kfmg=kfmg, ns=False)
    ph = abi.phys()
    d0 = np.zeros((7, n)); d0[5:] = 1e-3 * rng.standard_normal((2, n))
    v_o, d_o = v.copy(), d0.copy()
    gv, gd = T(v), T(d0)
    for l, b in enumerate(lay.blocks):
        O.implimf_eu(l, u, dt, v_o, d_o, ff, lmx, sn, vol, ps, c, dt1min=0.004, ityprk=ityprk)
        abi.implimf_eu(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), lmx, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]),
                       T(vol), T(ps), T(c), dt1min=0.004, kfmg=kfmg, ityprk=ityprk)
    hv, hd = H(gv), H(gd)
    for e in range(7):
        assert relerr(hd[e], d_o[e]) <= TOL, f"d[{e}]"
        assert relerr(hv[e], v_o[e]) <= TOL, f"v[{e}]"


def test_sch_expli_and_reductions(ctx):
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(9)
    n = lay.ip11
    v = random_state(n, rng); u = rng.standard_normal((5, n)); dt = 0.01 * (1 + rng.random(n))
    O = C.Orc(lay, C.orc_phys())
    v_o = v.copy(); gv = T(v)
    for l, b in enumerate(lay.blocks):
        O.sch_expli(l, u, v_o, dt, vol)
        abi.sch_expli(abi.block(b), T(u), gv, T(dt), T(vol))
    assert relerr(H(gv), v_o) <= TOL
    blocks = [abi.block(b) for b in lay.blocks]
    inner = lay.interior()
    a, bb = rng.standard_normal(n), rng.standard_normal(n)
    s = abi.sumsq_rate(blocks, T(a), T(bb), T(dt))
    ref = float(np.sum(((a[inner] - bb[inner]) / dt[inner]) ** 2))
    assert abs(s - ref) <= 1e-12 * ref
    gdt = T(dt)
    dtmin = abi.chrono_global(blocks, 0.5 * dt[inner].min(), gdt)
    assert dtmin == dt[inner].min()
    assert np.all(H(gdt) == 0.5 * dt[inner].min())
    from codens_amd.abi import CodensError
    with pytest.raises(CodensError):
        abi.chrono_global(blocks, 2.0, T(dt))


def test_rk_stage(ctx):
    """codens_rk_stage against orc_rk_stage bit for bit (both round the
    product and the difference separately), on 5 and 7 planes; ghosts of v
    untouched."""
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(19)
    n = lay.ip11
    O = C.Orc(lay, C.orc_phys())
    for nv in (5, 7):
        v0 = rng.standard_normal((nv, n)); u = rng.standard_normal((nv, n)); dt = 0.01 * (1 + rng.random(n))
        v_o = np.full((nv, n), 3.0); gv = T(v_o)
        for l, b in enumerate(lay.blocks):
            O.rk_stage(l, 1.0 / 3.0, v0, u, v_o, dt, vol)
            abi.rk_stage(abi.block(b), 1.0 / 3.0, T(v0), T(u), gv, T(dt), T(vol), nvar=nv)
        np.testing.assert_array_equal(H(gv), v_o)


def test_mesh_metrics_match_host():
    """codens_mesh_metrics (ghost nodes, sn, vol on the device) against the
    host restatement layout.metrics, which is bit-exact against the
    reference's extmhg/norm/svol (test_oracle_ref.py): equal bit for bit on
    two jittered blocks, the second mirrored so that every orientation
    factor is -0.5; the ghost nodes start as garbage and both sides
    overwrite them."""
    from codens_amd import abi
    from codens_amd import layout as L
    from randcase import random_mesh_xyz
    lay, _, _, (x, y, z) = random_mesh_xyz([(13, 9, 7), (6, 5, 4)], seed=31, jitter=0.02)
    b1 = lay.blocks[1]
    b1.view(x)[:] *= -1.0
    rng = np.random.default_rng(8)
    xs = []
    for a in (x, y, z):
        a = a.copy()
        for b in lay.blocks:
            A = b.view(a)
            ghost = np.ones(A.shape, bool)
            ghost[b.kk1 - b.kd1:b.kk2 - b.kd1 + 1, b.jj1 - b.jd1:b.jj2 - b.jd1 + 1, b.ii1 - b.id1:b.ii2 - b.id1 + 1] = False
            A[ghost] = rng.standard_normal(int(ghost.sum()))
        xs.append(a)
    hx, hy, hz = (a.copy() for a in xs)
    sn_h, vol_h = L.metrics(lay, hx, hy, hz)
    gx, gy, gz = (T(a) for a in xs)
    sn_d = torch.zeros(9 * lay.ip11, dtype=torch.float64, device="cuda")
    vol_d = torch.zeros(lay.ip11, dtype=torch.float64, device="cuda")
    for b in lay.blocks:
        abi.mesh_metrics(abi.block(b), gx, gy, gz, sn_d[9 * b.npc: 9 * (b.npc + b.nnn)], vol_d)
    for g, h in ((gx, hx), (gy, hy), (gz, hz), (sn_d, sn_h), (vol_d, vol_h)):
        np.testing.assert_array_equal(H(g), h)
    assert (vol_h[lay.interior()] > 0).all()


def test_boundary_ghosts(ctx):
    abi, lay, sn, vol = ctx
    from codens_amd import layout as L
    rng = np.random.default_rng(13)
    n = lay.ip11
    v = random_state(n, rng)
    ps, t, c = all_thermo(v)
    b = lay.blocks[0]
    O = C.Orc(lay, C.orc_phys())
    ph = abi.phys()
    g1, i1, _ = L.face_cells(b, "i1"); g2, i2, _ = L.face_cells(b, "i2")
    gw, iw, _ = L.face_cells(b, "j1")
    gf, inf, fi = L.face_cells(b, "k2")
    nrm = np.stack(L.inward_normals(b, sn, "k2", fi))
    w0 = np.array([1.0, 0.5, 0.05, -0.02, 1 / 0.4 / 1.4 + 0.5 * (0.25 + 0.0025 + 0.0004)])
    arrs_o = [v.copy(), ps.copy(), t.copy(), c.copy()]
    vg, psg, tg, cg = T(v), T(ps), T(t), T(c)
    O.bc_coincident(g1, i1, i2, *arrs_o)
    abi.bc_coincident(T(g1, torch.int64), T(i1, torch.int64), T(i2, torch.int64), vg, psg, tg, cg)
    O.bc_wall(gw, iw, *arrs_o)
    abi.bc_wall(T(gw, torch.int64), T(iw, torch.int64), vg, psg, tg, cg)
    O.bc_farfield(gf, inf, nrm[0].copy(), nrm[1].copy(), nrm[2].copy(), w0, *arrs_o)
    abi.bc_farfield(T(gf, torch.int64), T(inf, torch.int64), T(nrm), w0, ph, vg, psg, tg, cg)
    for name, a, r in zip(("v", "ps", "temp", "cson"), (vg, psg, tg, cg), arrs_o):
        assert relerr(H(a), r) <= TOL, name
    tq = rng.standard_normal((9, n)); tq_o = tq.copy(); gtq = T(tq)
    O.bc_tq_copy(gw, iw, tq_o); abi.bc_tq_copy(T(gw, torch.int64), T(iw, torch.int64), gtq)
    O.bc_tq_coincident(g2, i2, i1, tq_o); abi.bc_tq_coincident(T(g2, torch.int64), T(i2, torch.int64), T(i1, torch.int64), gtq)
    assert relerr(H(gtq), tq_o) <= TOL


def test_symmetry_ghosts_and_farfield_state(ctx):
    """codens_bc_sym (clsym.f90) against the oracle on both k faces, bit for
    bit (contraction off in the ghost kernels): v written, p/T/c untouched;
    codens_utnrd (utnrd.f90) against the oracle's restatement, bit for bit,
    and the far field it feeds."""
    abi, lay, sn, vol = ctx
    from codens_amd import layout as L
    rng = np.random.default_rng(17)
    n = lay.ip11
    v = random_state(n, rng)
    v[1:4] *= 3.0 * (1 + rng.random(n))  # strong normal components
    ps, t, c = all_thermo(v)
    O = C.Orc(lay, C.orc_phys())
    vo, vg, psg = v.copy(), T(v), T(ps)
    for l, b in enumerate(lay.blocks):
        for face in ("k1", "k2"):
            g, i, fi = L.face_cells(b, face)
            nrm = np.stack(L.inward_normals(b, sn, face, fi))
            C.orc_bc_sym(O.lib, g, i, nrm[0].copy(), nrm[1].copy(), nrm[2].copy(), vo)
            abi.bc_sym(T(g, torch.int64), T(i, torch.int64), T(nrm), vg)
    assert np.array_equal(H(vg), vo)
    assert np.array_equal(H(psg), ps)
    assert not np.array_equal(vo, v)
    ph = abi.phys()
    for mach, a, bb in [(0.5, 0.0, 0.0), (0.8, 10.0, 5.0), (2.2, -30.0, 60.0)]:
        w = np.array(abi.utnrd(mach, a, bb, 1.3, 1.1, 0.95, ph))
        wo = C.orc_utnrd(O.lib, C.orc_phys(), mach, a, bb, 1.3, 1.1, 0.95)
        assert np.array_equal(w, wo)


@pytest.mark.parametrize("neq", [5, 7])
def test_residu_norms(ctx, neq):
    """codens_residu (residu.f90 per block, rscpsv.f90 aggregate) against the
    oracle: L1 and L2 within 1e-13 (the sums run in another order), Linf and
    its location exact; exact ties locate the last real cell; two calls in a
    row (the second on other data) give the second's norms."""
    abi, lay, sn, vol = ctx
    import ctypes
    from oracle import ptr
    rng = np.random.default_rng(50 + neq)
    n = lay.ip11
    nb = len(lay.blocks)
    blocks = abi.block_array([abi.block(b) for b in lay.blocks])
    oblocks = (C.OrcBlock * nb)(*[C.orc_block(b) for b in lay.blocks])
    for trial in range(3):
        u = rng.standard_normal((7, n))
        v = u + 1e-3 * rng.standard_normal((7, n))
        dt = 0.01 * (1 + rng.random(n))
        if trial == 2:  # exact ties: r = 0.5 on every cell
            u = np.zeros((7, n)); v = np.repeat((0.5 * dt)[None, :], 7, axis=0)
        ref = np.zeros((nb + 1, neq, 7))
        C.oracle.oracle_lib().orc_residu(oblocks, nb, neq, ptr(v), ptr(u), C.L_(n), ptr(dt), ptr(ref))
        out = torch.zeros((nb + 1, neq, 7), dtype=torch.float64, device="cuda")
        abi.residu(blocks, nb, neq, T(v), T(u), T(dt), out)
        got = H(out)
        assert np.allclose(got[..., :2], ref[..., :2], rtol=1e-13, atol=0), trial
        assert np.array_equal(got[..., 2:], ref[..., 2:]), trial


def test_check_state_counts_nonphysical(ctx):
    """codens_check_state counts real cells with rho <= 0, p <= 0 or NaN,
    injected from the host; ghost cells are not counted."""
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(61)
    n = lay.ip11
    v = random_state(n, rng)
    b = lay.blocks[0]
    bad = [b.idx(b.ii1, b.jj1, b.kk1), b.idx(b.ii1 + 1, b.jj1 + 2, b.kk1 + 1), b.idx(b.ii2 - 1, b.jj2 - 1, b.kk2 - 1)]
    v[0, bad[0]] = -1.0                 # rho < 0
    # p < 0: total energy set to half the kinetic energy
    ke = 0.5 * (v[1, bad[1]] ** 2 + v[2, bad[1]] ** 2 + v[3, bad[1]] ** 2) / v[0, bad[1]]
    v[4, bad[1]] = 0.5 * ke
    assert (GAM - 1) * (v[4, bad[1]] - ke) < 0
    v[2, bad[2]] = np.nan
    v[0, b.idx(b.ii1 - 1, b.jj1, b.kk1)] = -5.0  # a ghost: not counted
    blocks = abi.block_array([abi.block(x) for x in lay.blocks])
    nbad = torch.zeros(1, dtype=torch.int32, device="cuda")
    abi.check_state(blocks, len(lay.blocks), 5, abi.phys(), T(v), nbad)
    assert int(nbad.item()) == 3


def test_bad_extents_rejected(ctx):
    abi, lay, sn, vol = ctx
    from codens_amd.abi import CodensError
    cb = abi.block(lay.blocks[0])
    cb.nnn = cb.nnn + 1
    x = T(np.zeros((5, lay.ip11)))
    with pytest.raises(CodensError):
        abi.zpres(cb, abi.phys(), x, x[0], x[1], x[2])


@pytest.mark.parametrize("form", ["fp64", "eu", "fp32"])
def test_sweep_tile_shapes_identical(form):
    """The three tile shapes of the Jacobi sub-iteration (64 x 4, 32 x 8 and
    16 x 16) and the timed choice give bit-identical increments, on a block
    whose i and j extents are multiples of no tile (70 x 37 x 9 cells)."""
    from codens_amd import abi
    lay, sn, vol = random_mesh([(70, 37, 9)], seed=23)
    b = lay.blocks[0]
    rng = np.random.default_rng(31)
    n = lay.ip11
    v = random_state(n, rng)
    ps, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((5, n)); ff = 1e-4 * rng.standard_normal((5, n))
    dt = 0.01 * (1 + rng.random(n))
    ph = abi.phys()
    snb = T(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
    out = {}
    prev = abi.sweep_tile(0)
    try:
        for mode in (0, 1, 2, -1):
            abi.sweep_tile(mode)
            gv, gd = T(v), T(np.zeros((5, n)))
            if form == "eu":
                abi.implimf_eu(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), 6, snb, T(vol), T(ps), T(c))
            elif form == "fp32":
                abi.implimf_3d_f32(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), T(mu), T(mut), 6, snb, T(vol),
                                   T(ps), T(c), ityprk=1)
            else:
                abi.implimf_3d(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), T(mu), T(mut), 6, snb, T(vol), T(ps),
                               T(c), ityprk=1)
            out[mode] = (H(gv), H(gd))
        chosen = abi.sweep_tile_of(abi.block(b), eu=form == "eu", fp32=form == "fp32")
    finally:
        abi.sweep_tile(prev)
    assert chosen in (0, 1, 2)
    for mode in (1, 2, -1):
        assert np.array_equal(out[mode][0], out[0][0]) and np.array_equal(out[mode][1], out[0][1]), mode
    assert np.abs(out[0][1]).max() > 0


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_met_bord(ctx, kind):
    """codens_met_bord against the oracle, bit for bit, for each frontier
    kind: coincident average, wall (kparoi 0 and 1), far field, symmetry."""
    from codens_amd import layout as L
    abi, lay, sn, vol = ctx
    rng = np.random.default_rng(71 + kind)
    n = lay.ip11
    v = rng.standard_normal((7, n))
    g, i, _ = L.face_cells(lay.blocks[0], "j1")
    g2, i2, _ = L.face_cells(lay.blocks[0], "j2")
    r = i2  # any real cells of the same count stand for the partner's
    O = C.Orc(lay, C.orc_phys())
    vo = v.copy()
    ga, ia, ra = (np.ascontiguousarray(a.astype(np.int64)) for a in (g, i, r))
    O.lib.orc_met_bord(kind, C.L_(len(ga)), C.ptr(ga), C.ptr(ia), C.ptr(ra), C.D(2.5e-4), C.D(7.0), C.ptr(vo),
                       C.L_(n))
    gv = T(v)
    abi.met_bord(kind, T(ga, torch.int64), T(ia, torch.int64), T(ra, torch.int64), 2.5e-4, 7.0, gv)
    assert np.array_equal(H(gv), vo)
    assert not np.array_equal(vo, v)


def test_met_parko():
    """codens_met_parko against the oracle on the j1 wall of a 9x13x5 block
    (13 cells from the wall: the 10-cell reach stays inside), bit for bit."""
    from codens_amd import abi, configs
    case = configs.wall_rans(9, 13, 5, model="2KO")
    t = case.turb
    n = case.layout.ip11
    rng = np.random.default_rng(77)
    v = rng.standard_normal((7, n)); v[6] = 40 * np.abs(v[6])
    mu = 1e-3 * (1 + rng.random(n))
    O = C.Orc(case.layout, C.orc_phys())
    vo = v.copy()
    import ctypes
    rc = O.lib.orc_met_parko(C.L_(len(t.wall_lines)), t.wall_lines.ctypes.data_as(ctypes.c_void_p),
                             C.ptr(np.ascontiguousarray(t.wall_ghost)), C.ptr(t.mnpar), C.ptr(mu), C.ptr(t.dist),
                             C.D(0.075), C.ptr(vo), C.L_(n))
    assert rc == 0
    gv = T(v)
    abi.met_parko(t.wall_start, torch.as_tensor(t.wall_lines.view(np.uint8).copy(), device="cuda"),
                  T(t.wall_ghost, torch.int64), T(t.mnpar, torch.int32), T(mu), T(t.dist), 0.075, gv)
    assert np.array_equal(H(gv), vo)
    assert not np.array_equal(vo[6], v[6])
