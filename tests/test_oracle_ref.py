"""The C restatement (oracle) against the flang build of the reference's own
routines, bit for bit, on seeded random states of a two-block curvilinear
layout, and over several whole stages of both synthetic cases."""
import numpy as np
import pytest

import oracle
import cpu_ref as C
from randcase import all_thermo, random_mesh, random_state

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (5, 4, 3)]


@pytest.fixture(scope="module")
def mesh():
    return random_mesh(DIMS, seed=3)


@pytest.mark.parametrize("ns,ilim,pinfl", [(0, 0, 0.0), (0, 1, 0.3), (1, 0, 0.0), (1, 1, 0.0), (1, 1, 0.3)])
def test_residual_chain_bit_exact(mesh, ns, ilim, pinfl):
    lay, sn, vol = mesh
    rng = np.random.default_rng(100 + 10 * ns + ilim)
    n = lay.ip11
    ph = C.orc_phys(pinfl=pinfl)
    R = C.Ref(lay, ph, ilim=ilim, ns=bool(ns))
    O = C.Orc(lay, ph, ilim=ilim, ns=bool(ns))
    v = random_state(n, rng, pinfl)
    p, t, c = all_thermo(v, pinfl)
    a = [p.copy(), t.copy(), c.copy()]
    b = [p.copy(), t.copy(), c.copy()]
    mu_r, mu_o, mut = np.zeros(n), np.zeros(n), np.zeros(n)
    dt_r, dt_o = np.zeros(n), np.zeros(n)
    for l in range(len(lay.blocks)):
        R.zpres(l, v, *a); O.zpres(l, v, *b)
        R.zvismo(l, mu_r, v, a[1]); O.zvismo(l, mu_o, v, b[1])
        R.chronos(l, 0.8, mu_r, mut, v, dt_r, sn, vol, a[2]); O.chronos(l, 0.8, mu_o, mut, v, dt_o, sn, vol, b[2])
    for x, y in zip(a + [mu_r, dt_r], b + [mu_o, dt_o]):
        assert np.array_equal(x, y)
    tq = 0.01 * rng.standard_normal((9, n)) if ns else np.zeros((9, n))
    u_r, u_o = np.zeros((7, n)), np.zeros((7, n))
    ff_r, ff_o = np.ones((7, n)), np.ones((7, n))
    for l in range(len(lay.blocks)):
        if ns:
            R.sch_roe(l, 1, u_r, v, ff_r, tq, sn, a[0]); O.sch_roe(l, 1, u_o, v, ff_o, tq, sn, b[0])
        else:
            R.sch_roe_euler(l, 1, u_r, v, ff_r, sn, a[0]); O.sch_roe_euler(l, 1, u_o, v, ff_o, sn, b[0])
    inner = lay.interior()
    assert np.array_equal(u_r[:5, inner], u_o[:5, inner])
    assert np.array_equal(ff_r[:5], ff_o[:5])


def test_viscous_and_implicit_bit_exact(mesh):
    lay, sn, vol = mesh
    rng = np.random.default_rng(7)
    n = lay.ip11
    ph = C.orc_phys()
    R, O = C.Ref(lay, ph), C.Orc(lay, ph)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    for l, b in enumerate(lay.blocks):
        gr = R.zgrad2(l, sn, vol, v, t)
        go = O.zgrad2(l, sn, vol, v, t)
        m = b.interior() - b.npc
        assert np.array_equal(gr[:, m], go[:, m])
        tr, to = np.zeros((9, n)), np.zeros((9, n))
        R.zfluto(l, mu, mut, tr, gr); O.zfluto(l, mu, mut, to, go)
        assert np.array_equal(tr, to)
    u = 1e-3 * rng.standard_normal((7, n)); ff = 1e-4 * rng.standard_normal((7, n))
    dt = 0.01 * (1 + rng.random(n))
    for lmx, ityprk in ((1, 0), (4, 0), (3, 1)):
        vr, vo = np.zeros((7, n)), np.zeros((7, n))
        vr[:5] = v; vo[:5] = v
        dr, do = np.zeros((7, n)), np.zeros((7, n))
        for l in range(len(lay.blocks)):
            R.implimf_3d(l, u, dt, vr, dr, ff, mu, mut, lmx, sn, vol, p, c, ityprk=ityprk)
            O.implimf_3d(l, u, dt, vo, do, ff, mu, mut, lmx, sn, vol, p, c, ityprk=ityprk)
        assert np.array_equal(vr[:5], vo[:5]) and np.array_equal(dr[:5], do[:5])
    vr, vo = v.copy(), v.copy()
    for l in range(len(lay.blocks)):
        R.sch_expli(l, u, vr, dt, vol); O.sch_expli(l, u, vo, dt, vol)
    assert np.array_equal(vr, vo)


def _stage_case(which):
    from codens_amd import configs
    return {"periodic_euler": lambda: configs.periodic_euler(),
            "wall_ns": lambda: configs.wall_ns(24, 16, 8, lmax=4),
            "wall_ns_sym": lambda: configs.wall_ns(24, 16, 8, lmax=4, span="sym"),
            "inclined": lambda: configs.adjoint_wall(16, 8, 8, eta=10.0, amp=0.01),
            "wall_rans_sa": lambda: configs.wall_rans(16, 12, 8, model="1SA", ilim=0, lmax=3),
            "wall_rans_ko": lambda: configs.wall_rans(16, 12, 8, model="2KO", ilim=0, lmax=3),
            "multiblock_sa": lambda: configs.multiblock_sa(6, 4, 9, None, seed=7, ilim=1, lmax=3)}[which]()


@pytest.mark.parametrize("which", ["periodic_euler", "wall_ns", "wall_ns_sym", "inclined", "wall_rans_sa",
                                   "wall_rans_ko", "multiblock_sa"])
def test_stages_bit_exact(which):
    """Whole stages of the oracle (C restatement of every kernel and ghost
    rule) against the reference build (its own kernels and its own boundary
    routines rfve in the first cycle, rbvc, clpara, utnrd + rbve + clnrd,
    clsym, rbte, rbtc, met_bord, met_rbse, met_rbsc, in rbord's and
    met_num's order), bit for bit, state and ghosts."""
    import cpu_stage
    case = _stage_case(which)
    a, b = cpu_stage.CpuStage(case, use_ref=False), cpu_stage.CpuStage(case, use_ref=True)
    for _ in range(4):
        a.stage(); b.stage()
    nv = 7 if case.turb is not None else 5
    assert np.array_equal(a.v[:nv], b.v[:nv], equal_nan=True)
    assert np.all(np.isfinite(a.v[:nv, case.layout.interior()]))
    for x, y in ((a.ps, b.ps), (a.temp, b.temp), (a.cson, b.cson), (a.tq, b.tq), (a.dt, b.dt)):
        assert np.array_equal(x, y, equal_nan=True)
    if case.turb is not None:
        assert np.array_equal(a.fd, b.fd, equal_nan=True) and np.array_equal(a.mut, b.mut)


def test_metrics_bit_exact():
    """layout.metrics (numpy) against the reference's extmhg, norm and svol."""
    import ctypes
    from codens_amd import layout as L
    from oracle import ptr
    lay = L.Layout([(6, 5, 4), (4, 3, 3)])
    rng = np.random.default_rng(1)
    x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
    for b in lay.blocks:
        K, J, I = np.meshgrid(np.arange(b.nkd), np.arange(b.njd), np.arange(b.nid), indexing="ij")
        b.view(x)[:] = I * 0.1 + 0.01 * rng.standard_normal(I.shape)
        b.view(y)[:] = J * 0.13 + 0.01 * rng.standard_normal(I.shape)
        b.view(z)[:] = K * 0.07 + 0.01 * rng.standard_normal(I.shape)
    xr, yr, zr = x.copy(), y.copy(), z.copy()
    sn, vol = L.metrics(lay, x, y, z)
    R = C.Ref(lay, C.orc_phys())
    snr = np.zeros(9 * lay.ip11); volr = np.zeros(lay.ip11)
    t = [np.zeros(lay.ip00) for _ in range(3)]
    D = ctypes.c_double
    for l, b in enumerate(lay.blocks, 1):
        R.lib.ref_extmhg(l, ptr(xr), ptr(yr), ptr(zr), D(1.0), D(0.0))
        sp = ctypes.c_void_p(snr.ctypes.data + 8 * 9 * b.npc)
        R.lib.ref_snorm(l, ptr(xr), ptr(yr), ptr(zr), sp, b.nnn)
        R.lib.ref_svol(l, ptr(xr), ptr(yr), ptr(zr), sp, b.nnn, ptr(volr), *[ptr(a) for a in t])
    assert np.array_equal(x, xr) and np.array_equal(sn, snr) and np.array_equal(vol, volr)


@pytest.mark.parametrize("ki2,ki4", [(0.5, 1.0 / 32), (1.0, 0.05), (0.0, 0.0)])
def test_dissip_jameson_bit_exact(mesh, ki2, ki4):
    lay, sn, vol = mesh
    rng = np.random.default_rng(21)
    n = lay.ip11
    ph = C.orc_phys()
    R, O = C.Ref(lay, ph), C.Orc(lay, ph)
    v = np.zeros((7, n))
    v[:5] = random_state(n, rng)
    p, t, c = all_thermo(v[:5])
    dr, do = np.zeros((7, n)), np.zeros((7, n))
    for l in range(len(lay.blocks)):
        R.dissip_jameson(l, v, dr, sn, p, c, ki2, ki4)
        O.dissip_jameson(l, v, do, sn, p, c, ki2, ki4)
    assert np.array_equal(dr[:5], do[:5])


@pytest.mark.parametrize("which", ["met_roe2o", "met_roe"])
def test_turbulence_upwind_bit_exact(mesh, which):
    lay, sn, vol = mesh
    rng = np.random.default_rng(31)
    n = lay.ip11
    ph = C.orc_phys()
    R, O = C.Ref(lay, ph), C.Orc(lay, ph)
    t = np.zeros((7, n))
    t[:5] = random_state(n, rng)
    t[5] = 1e-3 * (1 + rng.random(n)); t[6] = 10 * (1 + rng.random(n))
    dr, do = np.full((7, n), 2.0), np.full((7, n), 2.0)
    for l in range(len(lay.blocks)):
        getattr(R, which)(l, t, dr, sn)
        getattr(O, which)(l, t, do, sn)
    if which == "met_roe":  # zeroed by the reference outside every block
        for e, m in C.met_roe_outside_entries(lay):
            assert dr[e, m] == 0.0 and do[e, m] == 2.0
            dr[e, m] = do[e, m]
    # met_roe's shifted index reaches the zero-area faces of the zero-thickness
    # ghost cells (0/0): the reference itself produces NaN there
    assert np.array_equal(dr, do, equal_nan=True)
    assert np.nanmax(np.abs(do[5:7])) > 0


@pytest.mark.parametrize("lmx,ityprk,kfmg", [(1, 0, 0), (4, 0, 0), (3, 1, 0), (3, 0, 3)])
def test_implimf_eu_bit_exact(mesh, lmx, ityprk, kfmg):
    """implimf_eu.f90 (Euler implicit phase) against the oracle, bit for bit."""
    lay, sn, vol = mesh
    rng = np.random.default_rng(71 + lmx)
    n = lay.ip11
    ph = C.orc_phys()
    R, O = C.Ref(lay, ph, kfmg=kfmg, ns=False), C.Orc(lay, ph, kfmg=kfmg, ns=False)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    u = 1e-3 * rng.standard_normal((7, n)); ff = 1e-4 * rng.standard_normal((7, n))
    dt = 0.01 * (1 + rng.random(n))
    vr, vo = np.zeros((7, n)), np.zeros((7, n))
    vr[:5] = v; vo[:5] = v
    dr, do = np.full((7, n), 5.0), np.full((7, n), 5.0)
    for l in range(len(lay.blocks)):
        R.implimf_eu(l, u, dt, vr, dr, ff, lmx, sn, vol, p, c, dt1min=0.003, ityprk=ityprk)
        O.implimf_eu(l, u, dt, vo, do, ff, lmx, sn, vol, p, c, dt1min=0.003, ityprk=ityprk)
    assert np.array_equal(vr, vo) and np.array_equal(dr, do)
    assert np.abs(dr[:5]).max() > 0
