"""The oracle's boundary-ghost restatements against the reference's own
boundary routines (flang build), bit for bit.

Frontiers of a seeded two-block curvilinear layout, in the reference's
listing order: coincident pairs (rbvc.f90, rbtc.f90, met_rbsc.f90), an
adiabatic wall (clpara.f90), far fields (utnrd.f90 + rbve.f90 + clnrd.f90),
symmetry planes (clsym.f90), and the copies of rbte.f90 / met_rbse.f90.  The
far-field states are drawn so that every characteristic branch of clnrd
(supersonic and subsonic inflow and outflow) occurs, and the full rbord.f90
sequence (rbvc, the frontiers in order, rbvc again) is compared as one.
"""
import numpy as np
import pytest

import oracle
import cpu_ref as C
from codens_amd import layout as L
from randcase import GAM, all_thermo, random_mesh_xyz

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (6, 6, 4)]
# (block, face, kind); coincident pairs are consecutive and of equal shape
FRONTIERS = [
    (0, "i1", "rc"), (0, "i2", "rc"),
    (1, "j1", "rc"), (1, "j2", "rc"),
    (0, "j1", "para"), (0, "j2", "nrd"), (0, "k1", "sym"), (0, "k2", "nrd"),
    (1, "i1", "para"), (1, "i2", "nrd"), (1, "k1", "nrd"), (1, "k2", "sym"),
]
STAG = dict(roa1=1.0 * (1 + 0.2 * 0.64) ** 2.5, aa1=np.sqrt(1 + 0.2 * 0.64), pa1=(1 + 0.2 * 0.64) ** 3.5 / GAM)


@pytest.fixture(scope="module")
def setup():
    lay, sn, vol, xyz = random_mesh_xyz(DIMS, seed=11)
    fr = []
    for l, f, kind in FRONTIERS:
        g, i, fi = L.face_cells(lay.blocks[l], f)
        nrm = np.stack(L.inward_normals(lay.blocks[l], sn, f, fi))
        fr.append(dict(l=l, face=f, kind=kind, g=g, i=i, nrm=nrm))
    partner = [-1] * len(fr)
    for a in range(0, 4, 2):
        partner[a], partner[a + 1] = a + 1, a
    for a, b in ((0, 1), (2, 3)):
        fr[a]["r"], fr[b]["r"] = fr[b]["i"], fr[a]["i"]
    R = C.Ref(lay, C.orc_phys())
    RF = C.RefFrontiers(R, [f["g"] for f in fr], [f["i"] for f in fr], partner)
    O = C.Orc(lay, C.orc_phys())
    return lay, sn, xyz, fr, RF, O


def state(n, rng, amp=1.6):
    """Random states whose velocities reach |q| ~ 2c, so the far-field normal
    Mach numbers cover both signs and both sides of 1."""
    rho = 1 + 0.1 * rng.standard_normal(n)
    p = 1 / GAM * (1 + 0.1 * rng.standard_normal(n))
    c = np.sqrt(GAM * p / rho)
    u, vv, w = (amp * c * rng.standard_normal(n) for _ in range(3))
    return np.stack([rho, rho * u, rho * vv, rho * w, p / (GAM - 1) + 0.5 * rho * (u * u + vv * vv + w * w)])


def arrays(v):
    p, t, c = all_thermo(v)
    return [v.copy(), p.copy(), t.copy(), c.copy()]


def same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True), float(np.nanmax(np.abs(x - y)))


def test_utnrd_state():
    ph = C.orc_phys()
    lay, sn, vol, xyz = random_mesh_xyz(DIMS, seed=11)
    R = C.Ref(lay, ph)
    for mach, alpha, beta in [(0.5, 0.0, 0.0), (0.8, 10.0, 5.0), (2.0, -7.5, 30.0), (0.3, 90.0, -45.0)]:
        g, i, fi = L.face_cells(lay.blocks[0], "j2")
        nrm = np.stack(L.inward_normals(lay.blocks[0], sn, "j2", fi))
        RF = C.RefFrontiers(R, [g], [i], [-1])
        v = state(lay.ip11, np.random.default_rng(1))
        w_ref = RF.farfield(0, 0, mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"], nrm, *arrays(v))
        w_orc = C.orc_utnrd(oracle.oracle_lib(), ph, mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"])
        assert np.array_equal(w_ref, w_orc), (mach, alpha, beta, w_ref - w_orc)
    # closed form: at rest the state is the stagnation state
    w = C.orc_utnrd(oracle.oracle_lib(), ph, 0.0, 3.0, 4.0, 1.3, 1.1, 0.9)
    assert np.allclose(w, [1.3, 0, 0, 0, 0.9 / 0.4])
    # static state and velocity direction
    m = 0.8
    w = C.orc_utnrd(oracle.oracle_lib(), ph, m, 10.0, 5.0, STAG["roa1"], STAG["aa1"], STAG["pa1"])
    q = w[1:4] / w[0]
    qq = np.sqrt(q @ q)
    p = 0.4 * (w[4] - 0.5 * w[0] * qq * qq)
    assert abs(qq / np.sqrt(GAM * p / w[0]) - m) < 1e-13
    a, b = np.radians(10.0), np.radians(5.0)
    assert np.allclose(q / qq, [np.cos(a) * np.cos(b), -np.sin(b), np.sin(a) * np.cos(b)], atol=1e-14)


def test_farfield_all_branches(setup):
    lay, sn, xyz, fr, RF, O = setup
    rng = np.random.default_rng(5)
    regimes = set()
    for trial in range(4):
        v = state(lay.ip11, rng)
        a, b = arrays(v), arrays(v)
        mach, alpha, beta = [(0.5, 0, 0), (1.7, 20, -10), (0.9, -30, 40), (2.5, 170, 5)][trial]
        for k, f in enumerate(fr):
            if f["kind"] != "nrd":
                continue
            # branch of every face, from the adjacent cell the ghost copies (rbve)
            ni = f["i"]
            rho = v[0, ni]
            qn = (v[1, ni] * f["nrm"][0] + v[2, ni] * f["nrm"][1] + v[3, ni] * f["nrm"][2]) / rho
            c = np.sqrt(GAM * 0.4 * (v[4, ni] - 0.5 * (v[1, ni] ** 2 + v[2, ni] ** 2 + v[3, ni] ** 2) / rho) / rho)
            regimes |= set(np.digitize(qn / c, [-1.0, 0.0, 1.0]).tolist())
            w_ref = RF.farfield(k, f["l"], mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"], f["nrm"], *a)
            w0 = C.orc_utnrd(O.lib, O.ph, mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"])
            assert np.array_equal(w_ref, w0)
            O.bc_farfield(f["g"], f["i"], f["nrm"][0].copy(), f["nrm"][1].copy(), f["nrm"][2].copy(), w0, *b)
        same(a, b)
    # supersonic outflow, subsonic outflow, subsonic inflow, supersonic inflow
    assert regimes == {0, 1, 2, 3}, regimes


def test_wall_sym_coincident(setup):
    lay, sn, xyz, fr, RF, O = setup
    rng = np.random.default_rng(6)
    v = state(lay.ip11, rng)
    a, b = arrays(v), arrays(v)
    RF.rbvc([0, 1, 2, 3], *a)
    g = np.concatenate([fr[k]["g"] for k in range(4)])
    i = np.concatenate([fr[k]["i"] for k in range(4)])
    r = np.concatenate([fr[k]["r"] for k in range(4)])
    O.bc_coincident(g, i, r, *b)
    same(a, b)
    for k, f in enumerate(fr):
        if f["kind"] == "para":
            RF.clpara(k, *a)
            O.bc_wall(f["g"], f["i"], *b)
        elif f["kind"] == "sym":
            RF.clsym(k, f["nrm"], a[0])
            C.orc_bc_sym(O.lib, f["g"], f["i"], f["nrm"][0].copy(), f["nrm"][1].copy(), f["nrm"][2].copy(), b[0])
    same(a, b)


def test_rbord_sequence(setup):
    """rbord.f90:137-429: rbvc, every frontier in order, rbvc again."""
    lay, sn, xyz, fr, RF, O = setup
    rng = np.random.default_rng(7)
    v = state(lay.ip11, rng)
    a, b = arrays(v), arrays(v)
    mach, alpha, beta = 0.8, 10.0, 5.0
    w0 = C.orc_utnrd(O.lib, O.ph, mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"])
    coin = [0, 1, 2, 3]
    g = np.concatenate([fr[k]["g"] for k in coin])
    i = np.concatenate([fr[k]["i"] for k in coin])
    r = np.concatenate([fr[k]["r"] for k in coin])
    RF.rbvc(coin, *a)
    O.bc_coincident(g, i, r, *b)
    for k, f in enumerate(fr):
        if f["kind"] == "para":
            RF.clpara(k, *a)
            O.bc_wall(f["g"], f["i"], *b)
        elif f["kind"] == "nrd":
            RF.farfield(k, f["l"], mach, alpha, beta, STAG["roa1"], STAG["aa1"], STAG["pa1"], f["nrm"], *a)
            O.bc_farfield(f["g"], f["i"], f["nrm"][0].copy(), f["nrm"][1].copy(), f["nrm"][2].copy(), w0, *b)
        elif f["kind"] == "sym":
            RF.clsym(k, f["nrm"], a[0])
            C.orc_bc_sym(O.lib, f["g"], f["i"], f["nrm"][0].copy(), f["nrm"][1].copy(), f["nrm"][2].copy(), b[0])
    RF.rbvc(coin, *a)
    O.bc_coincident(g, i, r, *b)
    same(a, b)
    # the corner ghosts a coincident frontier reads are not interior: every
    # ghost of the second rbvc pass is the average of two interior cells
    inner = set(lay.interior().tolist())
    assert all(int(x) in inner for x in np.concatenate([i, r]))


def test_stress_and_turbulence_ghosts(setup):
    lay, sn, xyz, fr, RF, O = setup
    rng = np.random.default_rng(8)
    n = lay.ip11
    tq = rng.standard_normal((9, n))
    fd = rng.standard_normal((6, n))
    noncoin = [k for k, f in enumerate(fr) if f["kind"] != "rc"]
    coin = [0, 1, 2, 3]
    gg = np.concatenate([fr[k]["g"] for k in noncoin]); ii = np.concatenate([fr[k]["i"] for k in noncoin])
    g = np.concatenate([fr[k]["g"] for k in coin])
    i = np.concatenate([fr[k]["i"] for k in coin])
    r = np.concatenate([fr[k]["r"] for k in coin])
    ta, tb = tq.copy(), tq.copy()
    RF.rbte(noncoin, ta)
    RF.rbtc(coin, ta)
    O.lib.orc_bc_tq_copy(C.L_(len(gg)), C.ptr(gg), C.ptr(ii), C.ptr(tb), C.L_(n))
    O.lib.orc_bc_tq_coincident(C.L_(len(g)), C.ptr(g), C.ptr(i), C.ptr(r), C.ptr(tb), C.L_(n))
    assert np.array_equal(ta, tb)
    fa, fb = fd.copy(), fd.copy()
    RF.met_rbse(noncoin, fa)
    RF.met_rbsc(coin, fa)
    O.lib.orc_met_rbse(C.L_(len(gg)), C.ptr(gg), C.ptr(ii), C.ptr(fb), C.L_(n))
    O.lib.orc_met_rbsc(C.L_(len(g)), C.ptr(g), C.ptr(i), C.ptr(r), C.ptr(fb), C.L_(n))
    assert np.array_equal(fa, fb)


def test_zvisqc_is_gradients_closure_then_ghosts(setup):
    """zvisqc.f90:4 as a whole (the reference's block loop of zgrad2 and
    zfluto, then rbte over the non-coincident and rbtc over the coincident
    frontiers, zvisqc.f90:134-205) against the oracle's zgrad2 and zfluto per
    block followed by its two stress-ghost restatements, bit for bit."""
    lay, sn, xyz, fr, RF, O = setup
    _, _, vol, _ = random_mesh_xyz(DIMS, seed=11)
    rng = np.random.default_rng(9)
    n = lay.ip11
    v = state(n, rng, amp=0.3)
    _, t, _ = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-3 * rng.random(n)
    noncoin = [k for k, f in enumerate(fr) if f["kind"] != "rc"]
    coin = [0, 1, 2, 3]
    ta = np.full((9, n), -7.0)
    RF.zvisqc(noncoin, coin, v, mu, mut, ta, sn, vol, t)
    tb = np.full((9, n), -7.0)
    for l in range(len(lay.blocks)):
        O.zfluto(l, mu, mut, tb, O.zgrad2(l, sn, vol, v, t))
    gg = np.concatenate([fr[k]["g"] for k in noncoin]); ii = np.concatenate([fr[k]["i"] for k in noncoin])
    g = np.concatenate([fr[k]["g"] for k in coin])
    i = np.concatenate([fr[k]["i"] for k in coin])
    r = np.concatenate([fr[k]["r"] for k in coin])
    O.lib.orc_bc_tq_copy(C.L_(len(gg)), C.ptr(gg), C.ptr(ii), C.ptr(tb), C.L_(n))
    O.lib.orc_bc_tq_coincident(C.L_(len(g)), C.ptr(g), C.ptr(i), C.ptr(r), C.ptr(tb), C.L_(n))
    assert np.array_equal(ta, tb)
    assert np.all(ta[:, np.concatenate([gg, g])] != -7.0)  # every listed ghost written


def test_inward_normals_match_initns(setup):
    lay, sn, xyz, fr, RF, O = setup
    x, y, z = xyz
    for k, f in enumerate(fr):
        b = lay.blocks[f["l"]]
        got = RF.initns(k, f["l"], f["face"], b, x, y, z)
        for c in range(3):
            assert np.array_equal(got[c], f["nrm"][c]), (f["face"], c)
