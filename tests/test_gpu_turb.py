"""Turbulence chain kernels (teq_gradv, met_prod, met_gradtr, met_flud*,
met_smsa, met_smko, sch_turb) against the oracle on seeded states of a
two-block curvilinear layout.  Bar: 1e-12 relative per field."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import cpu_ref as C
from randcase import random_mesh, random_state, relerr

TOL = 1e-12


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)


def H(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def setup():
    from codens_amd import abi
    lay, sn, vol = random_mesh([(8, 6, 5), (5, 4, 3)], seed=41)
    rng = np.random.default_rng(43)
    n = lay.ip11
    v = np.zeros((7, n))
    v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    v[6] = 50 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    mut = 1e-3 * rng.random(n)
    dist = 0.05 + rng.random(n)
    cson = np.sqrt(1.4 * (0.4 * (v[4] - 0.5 * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2) / v[0])) / v[0])
    tc = C.turb_constants()
    tcl = [getattr(tc, f) for f, _ in C.OrcTurb._fields_]
    O = C.Orc(lay, C.orc_phys())
    return abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O


def snb(sn, b):
    return T(sn[9 * b.npc: 9 * (b.npc + b.nnn)])


def test_chain(setup):
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    gv, gvol, gmu, gmut, gdist, gc = T(v), T(vol), T(mu), T(mut), T(dist), T(cson)
    vo = v.copy()
    for l, b in enumerate(lay.blocks):
        cb = abi.block(b)
        inner = b.interior()
        m = inner - b.npc
        # gradients of u, v, w (with the rho clamp)
        g9o = O.teq_gradv(l, sn, vol, vo)
        g9 = torch.zeros(9, b.nnn, dtype=torch.float64, device="cuda")
        abi.teq_gradv(cb, snb(sn, b), gvol, gv, g9)
        hg = H(g9)
        for q in range(9):
            assert relerr(hg[q, m], g9o[q, m]) <= TOL, ("teq_gradv", q)
        # production
        tpo = O.met_prod(l, mut, vo, g9o)
        tp = torch.zeros(b.nnn, dtype=torch.float64, device="cuda")
        abi.met_prod(cb, gmut, gv, g9, tp)
        assert relerr(H(tp)[m], tpo[m]) <= TOL
        # gradients of the transported variables, then the flux densities
        fdo = np.zeros((6, n)); fd = T(np.zeros((6, n)))
        O.met_gradtr(l, sn, vol, vo, fdo)
        abi.met_gradtr(cb, snb(sn, b), gvol, gv, fd)
        assert relerr(H(fd), fdo) <= TOL
        for kfl in (1, 2, 5):
            a = fdo.copy(); ga = fd.clone()
            O.met_flud(l, kfl, tc, vo, mu, mut, a)
            abi.met_flud(cb, kfl, tcl, gv, gmu, gmut, ga)
            assert relerr(H(ga), a) <= TOL, kfl
        # sources
        for which in ("smsa", "smko"):
            cf_o, t6o, t7o = np.zeros(n), np.zeros(n), np.zeros(n)
            cf, t6, t7 = T(np.zeros(n)), T(np.zeros(n)), T(np.zeros(n))
            if which == "smsa":
                O.met_smsa(l, tc, sn, vol, vo, mu, dist, g9o, fdo, cf_o, t6o, t7o)
                abi.met_smsa(cb, tcl, snb(sn, b), gvol, gv, gmu, gdist, g9, fd, cf, t6, t7)
            else:
                O.met_smko(l, tc, vo, tpo, cf_o, t6o, t7o, cson)
                abi.met_smko(cb, tcl, gv, tp, cf, t6, t7, gc)
            for name, a, r in (("cfke", cf, cf_o), ("ts6", t6, t6o), ("ts7", t7, t7o)):
                assert relerr(H(a)[inner], r[inner]) <= TOL, (which, name)
        # flux balance
        rng = np.random.default_rng(7 + l)
        d = 1e-4 * rng.standard_normal((7, n))
        ts6 = 1e-3 * rng.standard_normal(n); ts7 = 1e-3 * rng.standard_normal(n)
        uo = np.zeros((7, n)); gu = T(np.zeros((7, n)))
        O.sch_turb(l, uo, vo, d, fdo, ts6, ts7, sn, vol)
        abi.sch_turb(cb, gu, gv, T(d), fd, T(ts6), T(ts7), snb(sn, b), gvol)
        hu = H(gu)
        for e in (5, 6):
            assert relerr(hu[e], uo[e]) <= TOL, ("sch_turb", e)
    assert relerr(H(gv), vo) == 0.0


@pytest.mark.parametrize("cmt", [1.0, 2.0, 1.17])
def test_impli2_eqt_3d(setup, cmt):
    """Line-implicit phase of the transport equations against the oracle."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(91)
    u = np.zeros((7, n)); u[5:] = 1e-3 * rng.standard_normal((2, n))
    dt = 0.01 * (1 + rng.random(n)); cfke = rng.random(n)
    vo = v.copy(); gv = T(v)
    for l, b in enumerate(lay.blocks):
        O.lib.orc_impli2_eqt_3d(ctypes.byref(C.orc_block(b)), ptr(u), C.L_(n), ptr(dt), ptr(vo), ptr(mu), ptr(mut),
                                ptr(cfke), ptr(np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])), ptr(vol),
                                C.D(cmt))
        abi.impli2_eqt_3d(abi.block(b), T(u), T(dt), gv, T(mu), T(mut), T(cfke), snb(sn, b), T(vol), cmt)
    hv = H(gv)
    inner = lay.interior()
    for e in (5, 6):
        assert relerr(hv[e, inner] - v[e, inner], vo[e, inner] - v[e, inner]) <= TOL, e
    assert np.array_equal(hv[:5], v[:5])


@pytest.mark.parametrize("kwsst", [0, 1])
def test_eddy_viscosity(setup, kwsst):
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        mo = np.zeros(n); gm = T(np.zeros(n))
        O.lib.orc_met_samut(ctypes.byref(ob), ctypes.byref(tc), ptr(v), C.L_(n), ptr(mu), ptr(mo))
        abi.met_samut(abi.block(b), tcl, T(v), T(mu), gm)
        assert relerr(H(gm)[inner], mo[inner]) <= TOL
        vo = v.copy(); gv = T(v)
        mo = np.zeros(n); gm = T(np.zeros(n))
        go = np.zeros((9, b.nnn))
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        if kwsst:
            O.lib.orc_teq_gradv(ctypes.byref(ob), ptr(sb), ptr(vol), ptr(vo), C.L_(n), ptr(go))
        O.lib.orc_met_kocmut(ctypes.byref(ob), ctypes.byref(tc), kwsst, ptr(vo), C.L_(n), ptr(mu), ptr(mo), ptr(dist),
                             ptr(go), C.L_(b.nnn))
        abi.met_kocmut(abi.block(b), tcl, kwsst, T(sb), T(vol), gv, T(mu), gm, T(dist), T(np.zeros((9, b.nnn))))
        assert relerr(H(gm)[inner], mo[inner]) <= TOL
        assert np.array_equal(H(gv), vo)


@pytest.mark.parametrize("which", [1, 2])
def test_limiters(setup, which):
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(93)
    vv = v.copy(); vv[5] = 1e-3 * rng.standard_normal(n); vv[6] = rng.standard_normal(n); vv[5, ::7] = 1e-6
    vo = vv.copy(); gv = T(vv)
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        if which == 2:
            O.lib.orc_met_cutsa(ctypes.byref(ob), C.D(1e-6), ptr(vo), C.L_(n))
        else:
            O.lib.orc_met_cut(ctypes.byref(ob), C.D(1e-6), C.D(1e-4), ptr(vo), C.L_(n))
        abi.met_cut(abi.block(b), which, 1e-6, 1e-4, gv)
    assert np.array_equal(H(gv), vo)


def test_flux_density_ghosts(setup):
    """met_rbse / met_rbsc on the 6 flux-density planes, against the oracle's
    plane-wise ghost copy and face average (rbte/rbtc restatement)."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(95)
    b0, b1 = lay.blocks
    # ghosts of the i-low face of block 0 (copy) and of block 1 (average with block 0 cells)
    g0 = np.array([b0.idx(b0.ii1 - 1, j, k) for k in range(b0.kk1, b0.kk2) for j in range(b0.jj1, b0.jj2)],
                  np.int64)
    i0 = g0 + 1
    g1 = np.array([b1.idx(b1.ii1 - 1, j, k) for k in range(b1.kk1, b1.kk2) for j in range(b1.jj1, b1.jj2)],
                  np.int64)
    i1 = g1 + 1
    r1 = np.array([b0.idx(b0.ii2 - 1, j, k) for k in range(b1.kk1, b1.kk2) for j in range(b1.jj1, b1.jj2)],
                  np.int64)
    fd = rng.standard_normal((6, n))
    t9 = np.zeros((9, n)); t9[:6] = fd
    O.lib.orc_bc_tq_copy(C.L_(len(g0)), ptr(g0), ptr(i0), ptr(t9), C.L_(n))
    O.lib.orc_bc_tq_coincident(C.L_(len(g1)), ptr(g1), ptr(i1), ptr(r1), ptr(t9), C.L_(n))
    gfd = T(fd)
    G = lambda a: torch.as_tensor(a, device="cuda")
    abi.met_rbse(G(g0), G(i0), gfd)
    abi.met_rbsc(G(g1), G(i1), G(r1), gfd)
    assert np.array_equal(H(gfd), t9[:6])


def test_sources_kor_chien(setup):
    """met_smkor and met_smch against the oracle."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(97)
    fd = 1e-2 * rng.standard_normal((6, n))
    mnpar = rng.integers(0, 5, n).astype(np.int32)
    utau = np.array([1e-5, -3e-3, 2e-2, 4e-4])
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g9 = rng.standard_normal((9, b.nnn)); tp = 1e-3 * rng.random(b.nnn)
        oo = [np.zeros(n) for _ in range(3)]
        O.lib.orc_met_smkor(ctypes.byref(ob), ctypes.byref(tc), C.D(0.65), ptr(v), C.L_(n), ptr(fd), ptr(g9),
                            C.L_(b.nnn), ptr(tp), *[ptr(a) for a in oo])
        go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smkor(abi.block(b), tcl, 0.65, T(v), T(fd), T(g9), T(tp), *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL
        oo = [np.zeros(n) for _ in range(2)]
        O.lib.orc_met_smch(ctypes.byref(ob), C.D(1.35), C.D(1.8), C.D(2e-4), ptr(v), C.L_(n), ptr(mu), ptr(dist),
                           ptr(mnpar), ptr(utau), ptr(tp), *[ptr(a) for a in oo])
        go = [T(np.zeros(n)) for _ in range(2)]
        abi.met_smch(abi.block(b), 1.35, 1.8, 2e-4, T(v), T(mu), T(dist), torch.as_tensor(mnpar, device="cuda"),
                     T(utau), T(tp), *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL


@pytest.mark.parametrize("kfracom", [0, 1])
def test_menter(setup, kfracom):
    """Menter model kernels (met_smmt with the wall correction of
    met_mtcorf1 on three walls, two of them crossing, met_fludmt, met_komut
    both forms) against the oracle."""
    import ctypes
    from oracle import ptr
    from menter_case import MENTER, walls, walls_of_block
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    imx = 1
    rng = np.random.default_rng(301 + kfracom)
    desc, ncin, lines, starts, owner = walls(lay, [(0, "j1"), (1, "i1"), (0, "k1")])
    mnpar = rng.integers(0, len(ncin) + 3, n).astype(np.int32)
    gmn = torch.as_tensor(mnpar, device="cuda")
    dw = 0.002 + 0.2 * rng.random(n)
    mt = 1e-3 * (0.5 + rng.random(n))
    fd = np.zeros((6, n))
    fd[:3] = 1e-3 * rng.standard_normal((3, n))
    fd[3:] = 10 * rng.standard_normal((3, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        ln, st = walls_of_block(lines, starts, owner, l)
        o = dict(cfke=np.full(n, 7.0), gkgo=np.full(lay.ip00, 7.0), frac=np.full(n, 7.0), ts6=np.full(n, 7.0),
                 ts7=np.full(n, 7.0))
        g = {k: T(a) for k, a in o.items()}
        O.lib.orc_met_smmt(ctypes.byref(ob), ptr(MENTER), kfracom, imx, ptr(v), C.L_(n), ptr(mu), ptr(mt), ptr(dw),
                           ptr(fd), ptr(tp), ptr(o["cfke"]), ptr(o["gkgo"]), ptr(o["frac"]), ptr(o["ts6"]),
                           ptr(o["ts7"]), ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), ptr(mnpar))
        gl = torch.as_tensor(ln.view(np.uint8), device="cuda") if len(ln) else None
        abi.met_smmt(abi.block(b), MENTER, kfracom, imx, T(v), T(mu), T(mt), T(dw), T(fd), T(tp), g["cfke"],
                     g["gkgo"], g["frac"], g["ts6"], g["ts7"], face_start=st, lines=gl, mnpar=gmn)
        for k in o:
            assert relerr(H(g[k]), o[k]) <= TOL, (l, k)
        if kfracom:
            assert (o["frac"][inner] < 1).any()
        # the realizability-limited sources (blending function always on)
        g9 = rng.standard_normal((9, b.nnn)) * 10 ** rng.uniform(0.5, 3.5, b.nnn)
        o2 = dict(cfke=np.full(n, 7.0), frac=np.full(n, 7.0), ts6=np.full(n, 7.0), ts7=np.full(n, 7.0))
        g2 = {k: T(a) for k, a in o2.items()}
        O.lib.orc_met_smmtr(ctypes.byref(ob), ptr(MENTER), imx, ptr(v), C.L_(n), ptr(mu), ptr(mt), ptr(dw), ptr(fd),
                            ptr(g9), C.L_(b.nnn), ptr(tp), ptr(o2["cfke"]), ptr(o2["frac"]), ptr(o2["ts6"]),
                            ptr(o2["ts7"]), ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), ptr(mnpar))
        abi.met_smmtr(abi.block(b), MENTER, imx, T(v), T(mu), T(mt), T(dw), T(fd), T(g9), T(tp), g2["cfke"],
                      g2["frac"], g2["ts6"], g2["ts7"], face_start=st, lines=gl, mnpar=gmn)
        for k in o2:
            assert relerr(H(g2[k]), o2[k]) <= TOL, (l, "smmtr", k)
        # flux densities with the blending function just computed
        fo = fd.copy()
        O.lib.orc_met_fludmt(ctypes.byref(ob), ptr(MENTER), kfracom, ptr(mu), ptr(mt), ptr(o["frac"]), ptr(fo),
                             C.L_(n))
        gf = T(fd)
        abi.met_fludmt(abi.block(b), MENTER, kfracom, T(mu), T(mt), T(o["frac"]), gf)
        assert relerr(H(gf), fo) <= TOL
        # eddy viscosity, both forms (the SST form clamps rho and forms the
        # velocity gradients first)
        for komsst in (0, 1):
            vo = v.copy()
            g9o = O.teq_gradv(l, sn, vol, vo) if komsst else np.zeros((9, b.nnn))
            mo = np.full(n, 7.0)
            O.lib.orc_met_komut(ctypes.byref(ob), C.D(MENTER[0]), komsst, ptr(vo), C.L_(n), ptr(mu), ptr(mo),
                                ptr(dw), ptr(np.ascontiguousarray(g9o)), C.L_(b.nnn))
            gv, gm, gg = T(v), T(np.full(n, 7.0)), T(np.zeros((9, b.nnn)))
            abi.met_komut(abi.block(b), MENTER[0], komsst, snb(sn, b), T(vol), gv, T(mu), gm, T(dw), gg)
            assert relerr(H(gm)[inner], mo[inner]) <= TOL, komsst
    if kfracom:
        # the wall correction alone over all walls, in list order
        fr = rng.random(n); fr[rng.random(n) < 0.3] = 0.995
        fo = fr.copy()
        O.lib.orc_met_mtcorf1(lines.ctypes.data_as(ctypes.c_void_p), C.L_(len(lines)), imx, ptr(mnpar), ptr(fo))
        gfr = T(fr)
        abi.met_mtcorf1(starts, torch.as_tensor(lines.view(np.uint8), device="cuda"), imx, gmn, gfr)
        assert np.array_equal(H(gfr), fo)


def test_smdes():
    """met_smdes (SA with the DES length from met_dist) against the oracle,
    on a mesh whose node coordinates the test holds."""
    import ctypes
    from oracle import ptr
    from codens_amd import abi
    from randcase import random_mesh_xyz
    lay, sn, vol, (x, y, z) = random_mesh_xyz([(8, 6, 5), (5, 4, 3)], seed=41)
    rng = np.random.default_rng(383)
    n = lay.ip11
    v = np.zeros((7, n)); v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    dist = 0.01 + 0.2 * rng.random(n)
    tc = C.turb_constants()
    tcl = [getattr(tc, f) for f, _ in C.OrcTurb._fields_]
    O = C.Orc(lay, C.orc_phys())
    fd = 1e-2 * rng.standard_normal((6, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g9 = rng.standard_normal((9, b.nnn))
        do = np.zeros(b.nnn); oo = [np.zeros(n) for _ in range(3)]
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        O.lib.orc_met_smdes(ctypes.byref(ob), ctypes.byref(tc), ptr(x), ptr(y), ptr(z), ptr(do), ptr(sb), ptr(vol),
                            ptr(v), C.L_(n), ptr(mu), ptr(dist), ptr(g9), ptr(fd), *[ptr(a) for a in oo])
        gd = T(np.zeros(b.nnn)); go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smdes(abi.block(b), tcl, T(x), T(y), T(z), gd, snb(sn, b), T(vol), T(v), T(mu), T(dist), T(g9), T(fd),
                      *go)
        assert relerr(H(gd)[inner - b.npc], do[inner - b.npc]) <= TOL
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL


@pytest.mark.parametrize("ischema", [1, 7])
def test_smsasas(setup, ischema):
    """met_smsasas (SA with the SAS destruction, both met_laplaciens forms)
    against the oracle."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(489 + ischema)
    fd = 1e-2 * rng.standard_normal((6, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g9 = 3 * rng.standard_normal((9, b.nnn))
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        oo = [np.zeros(n) for _ in range(3)]
        O.lib.orc_met_smsasas(ctypes.byref(ob), ctypes.byref(tc), ischema, ptr(sb), ptr(vol), ptr(v), C.L_(n),
                              ptr(mu), ptr(dist), ptr(g9), ptr(fd), *[ptr(a) for a in oo])
        go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smsasas(abi.block(b), tcl, ischema, snb(sn, b), T(vol), T(v), T(mu), T(dist), T(g9), T(fd), *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL


def test_kl_model(setup):
    """Smith k-l model kernels (met_smkl with met_bark, met_klmut) against
    the oracle."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    kl = np.array([18.0, 1.2, 1.43, 0.41])
    rng = np.random.default_rng(497)
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g9 = rng.standard_normal((9, b.nnn)); tp = 1e-3 * rng.random(b.nnn)
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        oo = [np.zeros(n) for _ in range(3)]
        O.lib.orc_met_smkl(ctypes.byref(ob), ptr(kl), ptr(sb), ptr(vol), ptr(vk), C.L_(n), ptr(mu), ptr(mut), ptr(dist),
                           ptr(fd), ptr(g9), ptr(tp), *[ptr(a) for a in oo])
        go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smkl(abi.block(b), kl, snb(sn, b), T(vol), T(vk), T(mu), T(mut), T(dist), T(fd), T(g9), T(tp), *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL
        mo = np.zeros(n)
        O.lib.orc_met_klmut(ctypes.byref(ob), ptr(kl), ptr(vk), C.L_(n), ptr(mu), ptr(mo), ptr(dist))
        gm = T(np.zeros(n))
        abi.met_klmut(abi.block(b), kl, T(vk), T(mu), gm, T(dist))
        assert relerr(H(gm)[inner], mo[inner]) <= TOL


KE = np.array([1.55, 2.0, 0.09])  # cke1, cke2, cmu (atlecdon.f90:124-128)


@pytest.mark.parametrize("variant,lparoi,isch", [(0, 0, 7), (1, 0, 7), (2, 1, 7), (3, 0, 7), (3, 0, 1)])
def test_ke_sources(setup, variant, lparoi, isch):
    """k-epsilon sources (met_smke: met_exgr, met_bark, met_bare, then
    met_smkes / met_smker / met_smkec / met_smkesas) against the oracle:
    cfke, ts6, ts7, bark and bare within 1e-12, the ghost layers that
    met_exgr writes into the velocity gradients exactly."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(511 + variant)
    vk = v.copy(); vk[6] = 5e-2 * (1 + rng.random(n)) * v[0]
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        m = inner - b.npc
        g9 = rng.standard_normal((9, b.nnn)); tp = 1e-2 * rng.random(b.nnn)
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        go9 = g9.copy()
        ob2 = [np.zeros(b.nnn) for _ in range(2)]
        oo = [np.zeros(n) for _ in range(3)]
        O.lib.orc_met_smke(ctypes.byref(ob), ptr(KE), variant, lparoi, isch, ptr(sb), ptr(vol), ptr(vk), C.L_(n),
                           ptr(mu), ptr(mut), ptr(tp), ptr(go9), ptr(cson), ptr(ob2[0]), ptr(ob2[1]),
                           *[ptr(a) for a in oo])
        gg9 = T(g9)
        gb2 = [T(np.zeros(b.nnn)) for _ in range(2)]
        go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smke(abi.block(b), KE, variant, lparoi, isch, snb(sn, b), T(vol), T(vk), T(mu), T(mut), T(tp), gg9,
                     T(cson), *gb2, *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL
        for a, o in zip(gb2, ob2):
            assert relerr(H(a)[m], o[m]) <= TOL
        assert np.array_equal(H(gg9), go9)
        assert np.abs(oo[2][inner]).max() > 0


@pytest.mark.parametrize("form,lparoi", [(0, 0), (1, 0), (2, 0), (2, 1)])
def test_ke_eddy_viscosity(setup, form, lparoi):
    """met_kemut (Jones-Launder, with the SST limiter, Launder-Sharma)
    against the oracle within 1e-12."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    rng = np.random.default_rng(531)
    vk = v.copy(); vk[6] = 5e-2 * (1 + rng.random(n)) * v[0]
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        vo = vk.copy(); mo = np.zeros(n); g9o = np.zeros((9, b.nnn))
        O.lib.orc_met_kemut(ctypes.byref(ob), C.D(KE[2]), form, lparoi, ptr(sb), ptr(vol), ptr(vo), C.L_(n), ptr(mu),
                            ptr(mo), ptr(dist), ptr(g9o))
        gv = T(vk); gm = T(np.zeros(n)); gg9 = T(np.zeros((9, b.nnn)))
        abi.met_kemut(abi.block(b), KE[2], form, lparoi, snb(sn, b), T(vol), gv, T(mu), gm, T(dist), gg9)
        assert relerr(H(gm)[inner], mo[inner]) <= TOL
        if form == 1:
            assert relerr(H(gg9), g9o) <= TOL


@pytest.mark.parametrize("isch", [7, 1])
def test_kl_sas(setup, isch):
    """met_smklsas (Smith k-l with the scale-adaptive term) against the
    oracle within 1e-12."""
    import ctypes
    from oracle import ptr
    abi, lay, sn, vol, v, mu, mut, dist, cson, tc, tcl, O = setup
    n = lay.ip11
    kl = np.array([18.0, 1.2, 1.43, 0.41])
    rng = np.random.default_rng(557 + isch)
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g9 = rng.standard_normal((9, b.nnn)); tp = 1e-3 * rng.random(b.nnn)
        sb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        oo = [np.zeros(n) for _ in range(3)]
        O.lib.orc_met_smklsas(ctypes.byref(ob), ptr(kl), isch, ptr(sb), ptr(vol), ptr(vk), C.L_(n), ptr(mu), ptr(mut),
                              ptr(dist), ptr(fd), ptr(g9), ptr(tp), *[ptr(a) for a in oo])
        go = [T(np.zeros(n)) for _ in range(3)]
        abi.met_smklsas(abi.block(b), kl, isch, snb(sn, b), T(vol), T(vk), T(mu), T(mut), T(dist), T(fd), T(g9),
                        T(tp), *go)
        for a, o in zip(go, oo):
            assert relerr(H(a)[inner], o[inner]) <= TOL
