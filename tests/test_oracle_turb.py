"""Turbulence-transport chain of met_num (Spalart-Allmaras and Wilcox k-omega
sources, gradients, production, diffusive flux densities, flux balance):
the oracle against the flang build of the reference routines, bit for bit."""
import ctypes

import numpy as np
import pytest

import cpu_ref as C
import oracle
from oracle import ptr
from randcase import random_mesh, random_state

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (5, 4, 3)]


@pytest.fixture(scope="module")
def setup():
    lay, sn, vol = random_mesh(DIMS, seed=41)
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
    R = C.Ref(lay, C.orc_phys())
    R.set_turb(tc)
    return lay, sn, vol, v, mu, mut, dist, cson, tc, R, oracle.oracle_lib()


# teq_gradv, met_gradtr and met_smsa take the global metric array and find the
# block's slice themselves (npsn = ndir*npfb(l)+1); sch_turb takes the slice.
def snb(sn, b):
    return np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])


def test_gradients_production_fluxdensity(setup):
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        # teq_gradv (with its rho clamp)
        vr, vo = v.copy(), v.copy()
        vr[0, b.npc + 5] = -1.0; vo[0, b.npc + 5] = -1.0
        gr = np.zeros((9, lay.ip00)); s = np.zeros(lay.ip00)
        R.lib.ref_teq_gradv(l + 1, ptr(sn), ptr(vol), ptr(vr), ptr(s), ptr(gr), ptr(cmu))
        go = np.zeros((9, b.nnn))
        O.orc_teq_gradv(ctypes.byref(ob), ptr(snb(sn, b)), ptr(vol), ptr(vo), C.L_(n), ptr(go))
        m = b.interior() - b.npc
        assert np.array_equal(vr, vo)
        assert np.array_equal(gr[:, m], go[:, m])
        # met_prod
        tr, to = np.zeros(lay.ip00), np.zeros(b.nnn)
        R.lib.ref_met_prod(l + 1, ptr(mut), ptr(v), ptr(np.ascontiguousarray(gr)), ptr(tr))
        O.orc_met_prod(ctypes.byref(ob), ptr(mut), ptr(v), C.L_(n), ptr(go), ptr(to))
        assert np.array_equal(tr[m], to[m])
        # met_gradtr then the three flux-density forms
        fr, fo = np.zeros((6, n)), np.zeros((6, n))
        w = [np.zeros(lay.ip00) for _ in range(4)]
        R.lib.ref_met_gradtr(l + 1, ptr(sn), ptr(vol), ptr(v), ptr(mu), ptr(mut), *[ptr(a) for a in w],
                             ptr(fr), ptr(cmu))
        O.orc_met_gradtr(ctypes.byref(ob), ptr(snb(sn, b)), ptr(vol), ptr(v), C.L_(n), ptr(fo))
        assert np.array_equal(fr, fo)
        for kfl in (1, 2, 5):
            a, c = fr.copy(), fo.copy()
            R.lib.ref_met_flud(l + 1, kfl, ptr(v), ptr(mu), ptr(mut), ptr(a), *[ptr(x) for x in w])
            O.orc_met_flud(ctypes.byref(ob), kfl, ctypes.byref(tc), ptr(v), C.L_(n), ptr(mu), ptr(mut), ptr(c))
            assert np.array_equal(a, c), kfl


def test_sources_and_balance(setup):
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    cmu = np.ones(n)
    b = lay.blocks[0]   # met_smsa reads its grad(nu~) array by the block-local index
    ob = C.orc_block(b)
    m = b.interior() - b.npc
    gr = np.zeros((9, lay.ip00)); s = np.zeros(lay.ip00)
    vv = v.copy()
    R.lib.ref_teq_gradv(1, ptr(sn), ptr(vol), ptr(vv), ptr(s), ptr(gr), ptr(cmu))
    fd = np.zeros((6, n)); w = [np.zeros(lay.ip00) for _ in range(4)]
    R.lib.ref_met_gradtr(1, ptr(sn), ptr(vol), ptr(vv), ptr(mu), ptr(mut), *[ptr(a) for a in w], ptr(fd),
                         ptr(cmu))
    # Spalart-Allmaras
    cr, co = np.zeros(n), np.zeros(n)
    t6r, t7r, t6o, t7o = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    w5 = [np.zeros(lay.ip00) for _ in range(5)]
    R.lib.ref_met_smsa(1, ptr(sn), ptr(vol), ptr(vv), ptr(mu), ptr(dist), ptr(gr), ptr(fd), ptr(cr),
                       *[ptr(a) for a in w5], ptr(t6r), ptr(t7r), ptr(cmu))
    go = np.ascontiguousarray(gr[:, :b.nnn])
    O.orc_met_smsa(ctypes.byref(ob), ctypes.byref(tc), ptr(snb(sn, b)), ptr(vol), ptr(vv), C.L_(n), ptr(mu),
                   ptr(dist), ptr(go), ptr(fd), ptr(co), ptr(t6o), ptr(t7o))
    inner = b.interior()
    for x, y, name in ((t6r, t6o, "ts6"), (t7r, t7o, "ts7"), (cr, co, "cfke")):
        assert np.array_equal(x[inner], y[inner]), name
    # k-omega
    tprod = np.zeros(lay.ip00)
    R.lib.ref_met_prod(1, ptr(mut), ptr(vv), ptr(gr), ptr(tprod))
    cr, co = np.zeros(n), np.zeros(n)
    t6r, t7r, t6o, t7o = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    R.lib.ref_met_smko(1, ptr(vv), ptr(fd), ptr(tprod), ptr(cr), ptr(t6r), ptr(t7r), ptr(cson))
    O.orc_met_smko(ctypes.byref(ob), ctypes.byref(tc), ptr(vv), C.L_(n), ptr(np.ascontiguousarray(tprod[:b.nnn])),
                   ptr(co), ptr(t6o), ptr(t7o), ptr(cson))
    for x, y, name in ((t6r, t6o, "ts6"), (t7r, t7o, "ts7"), (cr, co, "cfke")):
        assert np.array_equal(x[inner], y[inner]), name
    # flux balance
    rng = np.random.default_rng(5)
    d = 1e-4 * rng.standard_normal((7, n))
    ur, uo = np.zeros((7, n)), np.zeros((7, n))
    f = np.zeros((6, lay.ip00)); ncin = np.zeros(1, dtype=np.int32)
    R.lib.ref_sch_turb(1, ptr(ur), ptr(vv), ptr(d), ptr(fd), ptr(t6r), ptr(t7r), ptr(ncin), ptr(snb(sn, b)), b.nnn,
                       ptr(vol), ptr(f))
    O.orc_sch_turb(ctypes.byref(ob), ptr(uo), ptr(vv), C.L_(n), ptr(d), ptr(fd), ptr(t6r), ptr(t7r),
                   ptr(snb(sn, b)), ptr(vol))
    assert np.array_equal(ur[5:7], uo[5:7])


@pytest.mark.parametrize("model,cmt", [(0, 1.0), (1, 2.0), (4, 1.17)])
def test_impli2_eqt_3d(setup, model, cmt):
    """Line-implicit phase of the transport equations (impli2_eqt_3d.f90,
    lparoi = 0) for the models' cmt constants."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(47 + model)
    u = np.zeros((7, n)); u[5:] = 1e-3 * rng.standard_normal((2, n))
    dt = 0.01 * (1 + rng.random(n)); cfke = rng.random(n)
    R.lib.ref_set_equatt(model)
    t = [np.zeros(lay.ip00) for _ in range(10)]
    vr, vo = v.copy(), v.copy()
    for l, b in enumerate(lay.blocks):
        R.lib.ref_impli2_eqt_3d(l + 1, ptr(u), ptr(dt), ptr(vr), ptr(mu), ptr(mut), ptr(cfke), ptr(snb(sn, b)),
                                b.nnn, ptr(vol), *[ptr(a) for a in t])
        O.orc_impli2_eqt_3d(ctypes.byref(C.orc_block(b)), ptr(u), C.L_(n), ptr(dt), ptr(vo), ptr(mu), ptr(mut),
                            ptr(cfke), ptr(snb(sn, b)), ptr(vol), C.D(cmt))
    assert np.array_equal(vr, vo)
    assert np.abs(vr[5:] - v[5:]).max() > 0


@pytest.mark.parametrize("kwsst", [0, 1])
def test_eddy_viscosity(setup, kwsst):
    """met_samut (SA) and met_kocmut (k-omega, with the SST limiter for
    kwsst = 1, whose gradients come from teq_gradv)."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    cmu = np.ones(n)
    otc = tc
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        mr, mo = np.full(n, -3.0), np.full(n, -3.0)
        R.lib.ref_met_samut(l + 1, ptr(v), ptr(mu), ptr(mr))
        O.orc_met_samut(ctypes.byref(ob), ctypes.byref(otc), ptr(v), C.L_(n), ptr(mu), ptr(mo))
        assert np.array_equal(mr, mo)
        vr, vo = v.copy(), v.copy()
        mr, mo = np.full(n, -3.0), np.full(n, -3.0)
        g = np.zeros((9, lay.ip00)); s = np.zeros(lay.ip00)
        R.lib.ref_met_kocmut(l + 1, kwsst, ptr(sn), ptr(vol), ptr(s), ptr(g), ptr(dist), ptr(vr), ptr(mu), ptr(mr),
                             ptr(cmu))
        go = np.zeros((9, b.nnn))
        if kwsst:
            O.orc_teq_gradv(ctypes.byref(ob), ptr(snb(sn, b)), ptr(vol), ptr(vo), C.L_(n), ptr(go))
        O.orc_met_kocmut(ctypes.byref(ob), ctypes.byref(otc), kwsst, ptr(vo), C.L_(n), ptr(mu), ptr(mo), ptr(dist),
                         ptr(go), C.L_(b.nnn))
        assert np.array_equal(vr, vo)
        assert np.array_equal(mr, mo)


@pytest.mark.parametrize("which", [1, 2])
def test_limiters(setup, which):
    """met_cut (which = 1) and met_cutsa (which = 2) of at_cutke."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(53)
    vv = v.copy()
    vv[5] = rng.standard_normal(n) * 1e-3
    vv[6] = rng.standard_normal(n)
    vv[5, ::7] = 1e-6  # exactly at the threshold
    vr, vo = vv.copy(), vv.copy()
    for l, b in enumerate(lay.blocks):
        R.lib.ref_met_cut(l + 1, which, C.D(1e-6), C.D(1e-4), ptr(vr))
        ob = C.orc_block(b)
        if which == 2:
            O.orc_met_cutsa(ctypes.byref(ob), C.D(1e-6), ptr(vo), C.L_(n))
        else:
            O.orc_met_cut(ctypes.byref(ob), C.D(1e-6), C.D(1e-4), ptr(vo), C.L_(n))
    assert np.array_equal(vr, vo)
    assert not np.array_equal(vr, vv)


def test_sources_kor_chien(setup):
    """met_smkor (k-omega with cross diffusion and strain limiter) and
    met_smch (Chien k-epsilon, y+ from the wall point's friction velocity)."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(57)
    fd = 1e-2 * rng.standard_normal((6, n))
    mnpar = rng.integers(0, 5, n).astype(np.int32)
    utau = np.array([1e-5, -3e-3, 2e-2, 4e-4])
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        out_r = [np.full(n, 7.0) for _ in range(3)]
        out_o = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smkor(l + 1, C.D(0.65), ptr(v), ptr(fd), ptr(g), ptr(tp), *[ptr(a) for a in out_r])
        O.orc_met_smkor(ctypes.byref(ob), ctypes.byref(tc), C.D(0.65), ptr(v), C.L_(n), ptr(fd), ptr(gb),
                        C.L_(b.nnn), ptr(tp), *[ptr(a) for a in out_o])
        for a, o in zip(out_r, out_o):
            assert np.array_equal(a, o)
        out_r = [np.full(n, 7.0) for _ in range(2)]
        out_o = [np.full(n, 7.0) for _ in range(2)]
        R.lib.ref_met_smch(l + 1, C.D(1.35), C.D(1.8), C.D(2e-4), ptr(v), ptr(mu), ptr(dist), ptr(mnpar), ptr(utau),
                           ptr(tp), *[ptr(a) for a in out_r])
        O.orc_met_smch(ctypes.byref(ob), C.D(1.35), C.D(1.8), C.D(2e-4), ptr(v), C.L_(n), ptr(mu), ptr(dist),
                       ptr(mnpar), ptr(utau), ptr(tp), *[ptr(a) for a in out_o])
        for a, o in zip(out_r, out_o):
            assert np.array_equal(a, o)
            assert np.isfinite(a[b.interior()]).all()


def test_smdes():
    """met_smdes (SA with the DES length from met_dist's longest cell edge)
    on a mesh whose node coordinates the test holds: sources, cfke and the
    delta the reference returns; the DES length is active in some cells and
    not in others."""
    from randcase import random_mesh_xyz
    lay, sn, vol, (x, y, z) = random_mesh_xyz(DIMS, seed=41)
    rng = np.random.default_rng(83)
    n = lay.ip11
    v = np.zeros((7, n)); v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    dist = 0.01 + 0.2 * rng.random(n)
    tc = C.turb_constants()
    R = C.Ref(lay, C.orc_phys()); R.set_turb(tc)
    O = oracle.oracle_lib()
    fd = 1e-2 * rng.standard_normal((6, n))
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        w = [np.zeros(lay.ip00) for _ in range(5)]
        dr = np.full(lay.ip00, 7.0); r = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smdes(l + 1, ptr(x), ptr(y), ptr(z), ptr(dr), ptr(sn), ptr(vol), ptr(v), ptr(mu), ptr(dist),
                            ptr(g), ptr(fd), ptr(r[0]), *[ptr(a) for a in w], ptr(r[1]), ptr(r[2]), ptr(cmu))
        do = np.full(lay.ip00, 7.0); o = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smdes(ctypes.byref(ob), ctypes.byref(tc), ptr(x), ptr(y), ptr(z), ptr(do), ptr(snb(sn, b)),
                        ptr(vol), ptr(v), C.L_(n), ptr(mu), ptr(dist), ptr(gb), ptr(fd), *[ptr(a) for a in o])
        inner = b.interior()
        m = inner - b.npc
        assert np.array_equal(dr[m], do[m])
        for a, c in zip(r, o):
            assert np.array_equal(a[inner], c[inner])
        des = 0.65 * do[m] < dist[inner]
        assert des.any() and not des.all()


@pytest.mark.parametrize("ischema", [1, 7])
def test_smsasas(setup, ischema):
    """met_smsasas (SA with the SAS destruction, met_laplaciens in its plain
    (ischema 1) and weighted forms): sources and cfke, bit for bit; the SAS
    branch is taken in some cells and not in others."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(89 + ischema)
    fd = 1e-2 * rng.standard_normal((6, n))
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = 3 * rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        w = [np.zeros(lay.ip00) for _ in range(5)]
        tn = np.zeros((4, lay.ip00))
        r = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smsasas(l + 1, ischema, ptr(sn), ptr(vol), ptr(v), ptr(mu), ptr(dist), ptr(g), ptr(fd),
                              ptr(r[0]), *[ptr(a) for a in w], ptr(r[1]), ptr(r[2]), ptr(tn), ptr(cmu))
        o = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smsasas(ctypes.byref(ob), ctypes.byref(tc), ischema, ptr(snb(sn, b)), ptr(vol), ptr(v), C.L_(n),
                          ptr(mu), ptr(dist), ptr(gb), ptr(fd), *[ptr(a) for a in o])
        inner = b.interior()
        for a, c in zip(r, o):
            assert np.array_equal(a[inner], c[inner])
        # where the SAS branch is taken the source differs from met_smsa's
        s = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smsa(ctypes.byref(ob), ctypes.byref(tc), ptr(snb(sn, b)), ptr(vol), ptr(v), C.L_(n), ptr(mu),
                       ptr(dist), ptr(gb), ptr(fd), *[ptr(a) for a in s])
        differs = r[1][inner] != s[1][inner]
        assert differs.any()


KL = np.array([18.0, 1.2, 1.43, 0.41])  # cklb1, ckle2, sigmal, xkappa (atlecdon.f90:226-231)


def test_kl_model(setup):
    """Smith k-l model: met_smkl (with met_bark) and met_klmut, bit for bit."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(97)
    R.lib.ref_set_kl(ptr(KL))
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]  # rho l
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    mnpar = np.zeros(n, dtype=np.int32)
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        w = [np.zeros(lay.ip00) for _ in range(5)]
        r = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smkl(l + 1, ptr(sn), ptr(vol), ptr(vk), ptr(mu), ptr(mut), ptr(dist), ptr(mnpar), ptr(g),
                           ptr(fd), ptr(tp), ptr(r[0]), *[ptr(a) for a in w], ptr(r[1]), ptr(r[2]), ptr(cmu))
        o = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smkl(ctypes.byref(ob), ptr(KL), ptr(snb(sn, b)), ptr(vol), ptr(vk), C.L_(n), ptr(mu), ptr(mut),
                       ptr(dist), ptr(fd), ptr(gb), ptr(tp), *[ptr(a) for a in o])
        for a, c in zip(r, o):
            assert np.array_equal(a[inner], c[inner])
            assert np.isfinite(a[inner]).all()
        mr, mo = np.full(n, 7.0), np.full(n, 7.0)
        R.lib.ref_met_klmut(l + 1, ptr(vk), ptr(mu), ptr(mr), ptr(dist))
        O.orc_met_klmut(ctypes.byref(ob), ptr(KL), ptr(vk), C.L_(n), ptr(mu), ptr(mo), ptr(dist))
        assert np.array_equal(mr, mo)


@pytest.mark.parametrize("isch", [7, 1])
def test_kl_sas(setup, isch):
    """Smith k-l scale-adaptive sources (met_smklsas with met_bark and
    met_laplaciens), bit for bit; the shielding function selects the SAS
    term on part of the cells only."""
    lay, sn, vol, v, mu, mut, dist, cson, tc, R, O = setup
    n = lay.ip11
    rng = np.random.default_rng(99 + isch)
    R.lib.ref_set_kl(ptr(KL))
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    mnpar = np.zeros(n, dtype=np.int32)
    cmu = np.ones(n)
    differs = 0
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        w = [np.zeros(lay.ip00) for _ in range(5)]
        r = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smklsas(l + 1, isch, ptr(sn), ptr(vol), ptr(vk), ptr(mu), ptr(mut), ptr(dist), ptr(mnpar),
                              ptr(g), ptr(fd), ptr(tp), ptr(r[0]), *[ptr(a) for a in w], ptr(r[1]), ptr(r[2]),
                              ptr(np.zeros((4, lay.ip00))), ptr(cmu))
        o = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smklsas(ctypes.byref(ob), ptr(KL), isch, ptr(snb(sn, b)), ptr(vol), ptr(vk), C.L_(n), ptr(mu),
                          ptr(mut), ptr(dist), ptr(fd), ptr(gb), ptr(tp), *[ptr(a) for a in o])
        for a, c in zip(r, o):
            assert np.array_equal(a[inner], c[inner])
            assert np.isfinite(a[inner]).all()
        p = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smkl(ctypes.byref(ob), ptr(KL), ptr(snb(sn, b)), ptr(vol), ptr(vk), C.L_(n), ptr(mu), ptr(mut),
                       ptr(dist), ptr(fd), ptr(gb), ptr(tp), *[ptr(a) for a in p])
        d = o[2][inner] != p[2][inner]
        differs += int(d.sum())
        assert d.sum() < len(inner)  # the SAS term is off where fd <= 0.8
    assert differs > 0
