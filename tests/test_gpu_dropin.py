"""The Fortran drop-in modules (codens_amd/fortran) end to end: the reference's
own call signatures, module globals and host arrays, run on the GPU through
the C-ABI, against the reference routines on the CPU, on a two-block
curvilinear layout (so every host<->device window starts inside the arrays).

The drop-in library is oracle/_ref/libcodens_dropin.so: the reference harness
and reference objects with the drop-in modules linked in place of the
reference's zpres, zvismo, zvisqc, rbord, chronos, sch_roe_euler, sch_roe, sch_expli,
implimf_3d, implimf_eu, dissip_jameson, met_roe2o and met_roe."""
import numpy as np
import pytest

import oracle
import cpu_ref as C
from randcase import all_thermo, random_mesh, random_state, relerr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(oracle.ref_lib() is None or oracle.dropin_lib() is None,
                                 reason="reference build or drop-in build (oracle/_ref) not present")]

DIMS = [(8, 6, 5), (5, 4, 3)]
TOL = 1e-12


@pytest.fixture(scope="module")
def mesh():
    return random_mesh(DIMS, seed=5)


def pair(lay, ph, **kw):
    return C.Ref(lay, ph, **kw), C.Ref(lay, ph, lib=oracle.dropin_lib(), **kw)


def close(a, b, tol=TOL):
    """Relative max error per row (variable) within tol."""
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    for e in range(a.shape[0]):
        assert relerr(b[e], a[e]) <= tol, (e, relerr(b[e], a[e]))


def test_thermo_and_time_step(mesh):
    lay, sn, vol = mesh
    rng = np.random.default_rng(41)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    v = random_state(n, rng)
    out = {k: [np.full(n, -7.0) for _ in range(5)] for k in "rg"}
    mut = 1e-4 * rng.random(n)
    for l in range(len(lay.blocks)):
        for X, (ps, t, c, mu, dt) in ((R, out["r"]), (G, out["g"])):
            X.zpres(l, v, ps, t, c)
            X.zvismo(l, mu, v, t)
            X.chronos(l, 0.8, mu, mut, v, dt, sn, vol, c)
    for a, b in zip(out["r"], out["g"]):
        close(a, b)
        assert np.array_equal(a == -7.0, b == -7.0)  # same entries left alone


@pytest.mark.parametrize("ns,ilim", [(0, 0), (1, 0), (1, 1)])
def test_roe_residual(mesh, ns, ilim):
    lay, sn, vol = mesh
    rng = np.random.default_rng(43 + ilim)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph, ilim=ilim, ns=bool(ns))
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    tq = 0.01 * rng.standard_normal((9, n))
    res = {}
    for X, k in ((R, "r"), (G, "g")):
        u = np.full((7, n), 3.0)
        ff = np.ones((7, n))
        for l in range(len(lay.blocks)):
            if ns:
                X.sch_roe(l, 1, u, v, ff, tq, sn, p)
            else:
                X.sch_roe_euler(l, 1, u, v, ff, sn, p)
        res[k] = (u, ff)
    inner = lay.interior()
    close(res["r"][0][:5, inner], res["g"][0][:5, inner])
    close(res["r"][1][:5], res["g"][1][:5])
    assert np.array_equal(res["r"][0][5:], res["g"][0][5:])


def test_zvisqc(mesh):
    """zvisqc over both blocks with its ghosts: rbte over the non-coincident
    frontiers, rbtc over two coincident pairs (block 0's i faces, block 1's j
    faces); the drop-in computes the real cells on the GPU and calls the
    reference's rbte and rbtc for the ghosts."""
    from codens_amd import layout as L
    lay, sn, vol = mesh
    rng = np.random.default_rng(53)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    v = random_state(n, rng)
    _, t, _ = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-3 * rng.random(n)
    faces = [(0, "i1"), (0, "i2"), (1, "j1"), (1, "j2"), (0, "j1"), (0, "j2"), (0, "k1"), (0, "k2"),
             (1, "i1"), (1, "i2"), (1, "k1"), (1, "k2")]
    gh, inn = [], []
    for l, f in faces:
        g, i, _ = L.face_cells(lay.blocks[l], f)
        gh.append(g); inn.append(i)
    partner = [1, 0, 3, 2] + [-1] * 8
    out = {}
    for X, k in ((R, "r"), (G, "g")):
        RF = C.RefFrontiers(X, gh, inn, partner)
        tq = np.full((9, n), -7.0)
        RF.zvisqc(list(range(4, 12)), [0, 1, 2, 3], v, mu, mut, tq, sn, vol, t)
        out[k] = tq
    close(out["r"], out["g"])
    assert np.array_equal(out["r"] == -7.0, out["g"] == -7.0)  # the same entries written


def test_rbord():
    """rbord through the drop-in (the ghost kernels on the GPU, grouped by
    kind) against the reference's own sequence -- rbvc, each frontier's
    routine in listing order (clpara, utnrd + rbve + clnrd, clsym), rbvc again
    -- on the two-block layout of test_oracle_bc.py: two coincident pairs,
    walls, far fields whose normal Mach numbers take both signs on both sides
    of 1, and symmetry planes."""
    import test_oracle_bc as B
    from codens_amd import layout as L
    from randcase import random_mesh_xyz
    lay, sn, vol, xyz = random_mesh_xyz(B.DIMS, seed=11)
    fr = []
    for l, f, kind in B.FRONTIERS:
        g, i, fi = L.face_cells(lay.blocks[l], f)
        fr.append(dict(l=l, kind=kind, g=g, i=i, nrm=np.stack(L.inward_normals(lay.blocks[l], sn, f, fi))))
    partner = [1, 0, 3, 2] + [-1] * (len(fr) - 4)
    gh, inn = [f["g"] for f in fr], [f["i"] for f in fr]
    R, G = C.Ref(lay, C.orc_phys()), C.Ref(lay, C.orc_phys(), lib=oracle.dropin_lib())
    RR, RG = C.RefFrontiers(R, gh, inn, partner), C.RefFrontiers(G, gh, inn, partner)
    v = B.state(lay.ip11, np.random.default_rng(7))
    a, b = B.arrays(v), B.arrays(v)
    mach, alpha, beta = 0.8, 10.0, 5.0
    st = (B.STAG["roa1"], B.STAG["aa1"], B.STAG["pa1"])
    coin = [0, 1, 2, 3]
    RR.rbvc(coin, *a)
    for k, f in enumerate(fr):
        if f["kind"] == "para":
            RR.clpara(k, *a)
        elif f["kind"] == "nrd":
            RR.farfield(k, f["l"], mach, alpha, beta, *st, f["nrm"], *a)
        elif f["kind"] == "sym":
            RR.clsym(k, f["nrm"], a[0])
    RR.rbvc(coin, *a)
    x, y, z = xyz
    RG.rbord([f["kind"] for f in fr], [f["l"] for f in fr], coin,
             RG.normals([f["nrm"] if f["kind"] in ("nrd", "sym") else None for f in fr]), *b, mach, alpha, beta,
             *st, y, z)
    for x_, y_ in zip(a, b):
        assert np.array_equal(np.isnan(x_), np.isnan(y_))
        close(np.nan_to_num(x_), np.nan_to_num(y_))
    assert not np.array_equal(a[0], v)  # the ghosts were written


def _dropin_stage(G, lay, A, sn, vol):
    """zpres, zvismo, chronos, sch_roe (minmod, viscous) and implimf_3d
    (lmax = 3) of every block, through the drop-ins, on the host arrays of A."""
    for l in range(len(lay.blocks)):
        G.zpres(l, A["v"], A["p"], A["t"], A["c"])
        G.zvismo(l, A["mu"], A["v"], A["t"])
        G.chronos(l, 0.8, A["mu"], A["mut"], A["v"], A["dt"], sn, vol, A["c"])
        G.sch_roe(l, 0, A["u"], A["v"], A["ff"], A["tq"], sn, A["p"])
        G.implimf_3d(l, A["u"], A["dt"], A["v"], A["d"], A["ff"], A["mu"], A["mut"], 3, sn, vol, A["p"], A["c"])


def test_residency(mesh):
    """The mirrors' residency protocol under the drop-ins: a resident stage
    leaves the host arrays as they were until codens_host_sync_all, which
    brings back exactly the copy-mode results; a second stage with no host
    access in between moves only zvismo's 4-byte counter each way per block;
    codens_host_modified refuses an array with device-newer windows; turning
    residency off brings everything back."""
    from codens_amd import abi
    lay, sn, vol = mesh
    rng = np.random.default_rng(59)
    n, nb = lay.ip11, len(lay.blocks)
    G = C.Ref(lay, C.orc_phys(), lib=oracle.dropin_lib(), ilim=1, ns=True)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    base = dict(v=np.vstack([v, np.zeros((2, n))]), p=p, t=t, c=c, mu=np.zeros(n), mut=1e-4 * rng.random(n),
                dt=np.zeros(n), u=np.zeros((7, n)), ff=np.zeros((7, n)), d=np.zeros((7, n)),
                tq=0.01 * rng.standard_normal((9, n)))

    def fresh():
        return {k: np.ascontiguousarray(a.copy()) for k, a in base.items()}
    assert abi.host_residency(0) == 0
    A = fresh()
    _dropin_stage(G, lay, A, sn, vol)
    S1 = {k: a.copy() for k, a in A.items()}
    _dropin_stage(G, lay, A, sn, vol)
    S2 = A
    B = fresh()
    abi.host_residency(1)
    try:
        _dropin_stage(G, lay, B, sn, vol)
        for k in ("v", "p", "u", "d", "dt"):
            assert np.array_equal(B[k], base[k]), k  # nothing came back yet
        abi.host_sync_all()
        for k in S1:
            assert np.array_equal(B[k], S1[k]), k
        h0, d0 = abi.host_traffic()
        _dropin_stage(G, lay, B, sn, vol)
        h1, d1 = abi.host_traffic()
        assert (h1 - h0, d1 - d0) == (4 * nb, 4 * nb)
        with pytest.raises(abi.CodensError):
            abi.host_modified(B["v"])
        abi.host_sync(B["v"])
        abi.host_modified(B["v"])  # now current on the host: allowed
    finally:
        abi.host_residency(0)
    for k in S2:
        assert np.array_equal(B[k], S2[k]), k


@pytest.mark.parametrize("lmx,ityprk", [(1, 0), (4, 0), (3, 1)])
def test_implicit_and_explicit_update(mesh, lmx, ityprk):
    lay, sn, vol = mesh
    rng = np.random.default_rng(47)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((7, n)); ff = 1e-4 * rng.standard_normal((7, n))
    dt = 0.01 * (1 + rng.random(n))
    res = {}
    for X, k in ((R, "r"), (G, "g")):
        vv = np.zeros((7, n)); vv[:5] = v
        d = np.zeros((7, n))
        for l in range(len(lay.blocks)):
            X.implimf_3d(l, u, dt, vv, d, ff, mu, mut, lmx, sn, vol, p, c, ityprk=ityprk)
        ve = v.copy()
        for l in range(len(lay.blocks)):
            X.sch_expli(l, u, ve, dt, vol)
        res[k] = (vv, d, ve)
    for a, b in zip(res["r"], res["g"]):
        close(a[:5], b[:5])


def test_implimf_3d_ps_contract(mesh):
    """The drop-in implimf_3d takes the base-flux pressure from v (zpres's
    formula) instead of reading ps: with ps = zpres(v) it matches the
    reference routine within 1e-12; with ps perturbed by 1e-3 on the real
    cells its result does not move (bit for bit), while the reference's
    does, by far more than 1e-12."""
    lay, sn, vol = mesh
    rng = np.random.default_rng(53)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((7, n)); ff = 1e-4 * rng.standard_normal((7, n))
    dt = 0.01 * (1 + rng.random(n))
    pp = p * (1 + 1e-3 * rng.standard_normal(n))

    def run(X, pres):
        vv = np.zeros((7, n)); vv[:5] = v
        d = np.zeros((7, n))
        for l in range(len(lay.blocks)):
            X.implimf_3d(l, u, dt, vv, d, ff, mu, mut, 3, sn, vol, pres, c, ityprk=0)
        return d[:5]
    dr, dg = run(R, p), run(G, p)
    close(dr, dg)
    dr2, dg2 = run(R, pp), run(G, pp)
    assert np.array_equal(dg, dg2)
    assert max(relerr(dr2[e], dr[e]) for e in range(5)) > 1e-8


@pytest.mark.parametrize("lmx,kfmg", [(3, 0), (2, 3)])
def test_implimf_eu(mesh, lmx, kfmg):
    lay, sn, vol = mesh
    rng = np.random.default_rng(61 + lmx)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph, kfmg=kfmg, ns=False)
    v = random_state(n, rng)
    p, t, c = all_thermo(v)
    u = 1e-3 * rng.standard_normal((7, n)); ff = 1e-4 * rng.standard_normal((7, n))
    dt = 0.01 * (1 + rng.random(n))
    res = {}
    for X, k in ((R, "r"), (G, "g")):
        vv = np.zeros((7, n)); vv[:5] = v; vv[5:] = 0.5
        d = np.zeros((7, n)); d[5:] = 0.25
        for l in range(len(lay.blocks)):
            X.implimf_eu(l, u, dt, vv, d, ff, lmx, sn, vol, p, c, dt1min=0.004)
        res[k] = (vv, d)
    for a, b in zip(res["r"], res["g"]):
        close(a, b)


def test_dissip_jameson(mesh):
    lay, sn, vol = mesh
    rng = np.random.default_rng(53)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    v = np.zeros((7, n)); v[:5] = random_state(n, rng)
    p, t, c = all_thermo(v[:5])
    dr, dg = np.zeros((7, n)), np.zeros((7, n))
    for l in range(len(lay.blocks)):
        R.dissip_jameson(l, v, dr, sn, p, c, 0.5, 1.0 / 32)
        G.dissip_jameson(l, v, dg, sn, p, c, 0.5, 1.0 / 32)
    close(dr[:5], dg[:5])


@pytest.mark.parametrize("which", ["met_roe2o", "met_roe"])
def test_turbulence_upwind(mesh, which):
    lay, sn, vol = mesh
    rng = np.random.default_rng(59)
    n = lay.ip11
    ph = C.orc_phys()
    R, G = pair(lay, ph)
    t = np.zeros((7, n)); t[:5] = random_state(n, rng)
    t[5] = 1e-3 * (1 + rng.random(n)); t[6] = 10 * (1 + rng.random(n))
    dr, dg = np.full((7, n), 2.0), np.full((7, n), 2.0)
    for l in range(len(lay.blocks)):
        getattr(R, which)(l, t, dr, sn)
        getattr(G, which)(l, t, dg, sn)
    if which == "met_roe":  # entries outside every block, see met_roe_outside_entries
        for e, m in C.met_roe_outside_entries(lay):
            assert dr[e, m] == 0.0 and dg[e, m] == 2.0
            dr[e, m] = dg[e, m]
    # met_roe's shifted index gives NaN on the zero-thickness ghost faces in
    # the reference itself; the same entries must be NaN on both sides
    assert np.array_equal(np.isnan(dr), np.isnan(dg))
    ok = ~np.isnan(dr[5:7])
    assert np.abs(dr[5:7][ok] - dg[5:7][ok]).max() <= TOL * np.abs(dr[5:7][ok]).max()
    assert np.array_equal(dr[:5], dg[:5])


@pytest.fixture(scope="module")
def tsetup(mesh):
    lay, sn, vol = mesh
    rng = np.random.default_rng(201)
    n = lay.ip11
    v = np.zeros((7, n)); v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    v[6] = 50 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-3 * rng.random(n)
    dist = 0.05 + rng.random(n)
    p, t, c = all_thermo(v[:5])
    R, G = pair(lay, C.orc_phys())
    tc = C.turb_constants()
    R.set_turb(tc); G.set_turb(tc)
    return lay, sn, vol, v, mu, mut, dist, c, R, G


_RESIDENT = [False]


def _both(R, G, name, *arrays_and_scalars, outs=()):
    """Call wrapper `name` on the reference and on the drop-in library with
    copies of the listed numpy arrays; returns the two lists of arrays.  With
    the mirrors resident, the drop-in's results are brought back and its
    arrays forgotten, rows of 2-D arrays included (they are freed, and another
    array may take an address the library keeps)."""
    from oracle import ptr
    res = []
    for X in (R, G):
        args = [a.copy() if isinstance(a, np.ndarray) else a for a in arrays_and_scalars]
        getattr(X.lib, name)(*[ptr(a) if isinstance(a, np.ndarray) else a for a in args])
        if _RESIDENT[0] and X is G:
            from codens_amd import abi
            abi.host_sync_all()
            for a in args:  # every array the drop-in may have keyed: a 2-D array's rows are the
                if isinstance(a, np.ndarray):  # fd5x..fd6z, toxx..qcz arrays of a bundle
                    abi.host_modified(a)
                    for row in (a if a.ndim == 2 else ()):
                        abi.host_modified(row)
        res.append([a for a in args if isinstance(a, np.ndarray)])
    return res


def test_turbulence_routines(tsetup):
    """teq_gradv, met_prod, met_gradtr, met_flud* (1, 2, 5), met_smsa,
    met_smko, met_smkor, met_smch, sch_turb, met_samut, met_kocmut (0, 1),
    met_cut, met_cutsa through the reference's own signatures."""
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    cmu = np.ones(n)
    rng = np.random.default_rng(203)
    D = C.D
    for l, b in enumerate(lay.blocks):
        inner = b.interior()
        g0 = np.zeros((9, lay.ip00)); s0 = np.zeros(lay.ip00)
        r_, g_ = _both(R, G, "ref_teq_gradv", l + 1, sn, vol, v, s0, g0, cmu)
        vr, gr, vg, gg = r_[2], r_[4], g_[2], g_[4]
        assert np.array_equal(vr, vg)
        m = inner - b.npc
        close(gr[:, m], gg[:, m])
        g = gr
        (_, _, _, tr), (_, _, _, tg) = _both(R, G, "ref_met_prod", l + 1, mut, v, g, np.zeros(lay.ip00))
        close(tr[m], tg[m])
        tp = tr
        w = [np.zeros(lay.ip00) for _ in range(4)]
        r_, g_ = _both(R, G, "ref_met_gradtr", l + 1, sn, vol, v, mu, mut, *w, np.zeros((6, n)), cmu)
        fdr, fdg = r_[-2], g_[-2]
        close(fdr[:, inner], fdg[:, inner])
        fd = fdr
        for kfl in (1, 2, 5):
            r_, g_ = _both(R, G, "ref_met_flud", l + 1, kfl, v, mu, mut, fd, *w)
            close(r_[3][:, inner], g_[3][:, inner])
        w5 = [np.zeros(lay.ip00) for _ in range(5)]
        outs = lambda: [np.zeros(n) for _ in range(3)]
        r_, g_ = _both(R, G, "ref_met_smsa", l + 1, sn, vol, v, mu, dist, g, fd, np.zeros(n), *w5, np.zeros(n),
                       np.zeros(n), cmu)
        for x, y in zip([r_[7], r_[13], r_[14]], [g_[7], g_[13], g_[14]]):  # cfke, ts6, ts7
            close(x[inner], y[inner])
        r_, g_ = _both(R, G, "ref_met_smko", l + 1, v, fd, tp, *outs(), c)
        for x, y in zip(r_[3:6], g_[3:6]):
            close(x[inner], y[inner])
        r_, g_ = _both(R, G, "ref_met_smkor", l + 1, D(0.65), v, fd, g, tp, *outs())
        for x, y in zip(r_[4:7], g_[4:7]):
            close(x[inner], y[inner])
        mnpar = rng.integers(0, 5, n).astype(np.int32)
        utau = np.zeros(lay.ip00); utau[:4] = [1e-5, -3e-3, 2e-2, 4e-4]  # utau(ip42), ip42 = ip00
        r_, g_ = _both(R, G, "ref_met_smch", l + 1, D(1.35), D(1.8), D(2e-4), v, mu, dist, mnpar, utau, tp,
                       np.zeros(n), np.zeros(n))
        for x, y in zip(r_[6:8], g_[6:8]):
            close(x[inner], y[inner])
        d = 1e-4 * rng.standard_normal((7, n)); ts = 1e-3 * rng.standard_normal((2, n))
        ncin = np.zeros(1, dtype=np.int32)
        snb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
        r_, g_ = _both(R, G, "ref_sch_turb", l + 1, np.zeros((7, n)), v, d, fd, ts[0].copy(), ts[1].copy(), ncin,
                       snb, b.nnn, vol, np.zeros((6, lay.ip00)))
        close(r_[0][5:7, inner], g_[0][5:7, inner])
        r_, g_ = _both(R, G, "ref_met_samut", l + 1, v, mu, np.zeros(n))
        close(r_[2][inner], g_[2][inner])
        for kwsst in (0, 1):
            r_, g_ = _both(R, G, "ref_met_kocmut", l + 1, kwsst, sn, vol, np.zeros(lay.ip00),
                           np.zeros((9, lay.ip00)), dist, v, mu, np.zeros(n), cmu)
            close(r_[7][inner], g_[7][inner])
            assert np.array_equal(r_[5], g_[5])
        vv = v.copy(); vv[5] = 1e-3 * rng.standard_normal(n); vv[6] = rng.standard_normal(n)
        for which in (1, 2):
            r_, g_ = _both(R, G, "ref_met_cut", l + 1, which, D(1e-6), D(1e-4), vv)
            assert np.array_equal(r_[0], g_[0])


def test_turbulence_routines_resident(tsetup):
    """The same routines with the mirrors resident: the turbulence drop-ins'
    bundles (one bundle key named with 6 and with 3 arrays), plain mirrors and
    whole-array windows give the copy mode's results."""
    from codens_amd import abi
    _RESIDENT[0] = True
    abi.host_residency(1)
    try:
        test_turbulence_routines(tsetup)
    finally:
        _RESIDENT[0] = False
        abi.host_residency(0)


@pytest.mark.parametrize("model", [0, 1])
def test_impli2_eqt_3d(tsetup, model):
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    rng = np.random.default_rng(207)
    u = np.zeros((7, n)); u[5:] = 1e-3 * rng.standard_normal((2, n))
    dt = 0.01 * (1 + rng.random(n)); cfke = rng.random(n)
    R.lib.ref_set_equatt(model); G.lib.ref_set_equatt(model)
    t = [np.zeros(lay.ip00) for _ in range(10)]
    vr, vg = v.copy(), v.copy()
    from oracle import ptr
    for X, vx in ((R, vr), (G, vg)):
        for l, b in enumerate(lay.blocks):
            snb = np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])
            X.lib.ref_impli2_eqt_3d(l + 1, ptr(u), ptr(dt), ptr(vx), ptr(mu), ptr(mut), ptr(cfke), ptr(snb), b.nnn,
                                    ptr(vol), *[ptr(a) for a in t])
    inner = lay.interior()
    close((vr - v)[5:7, inner], (vg - v)[5:7, inner])
    assert np.array_equal(vr[:5], vg[:5])


@pytest.mark.parametrize("kfracom", [0, 1])
def test_menter_routines(tsetup, kfracom):
    """met_smmt (with met_mtcorf1 on the walls of the block), met_mtcorf1 on
    all walls, met_fludmt and met_komut (both forms) through the reference's
    signatures, the walls set in both builds' module `boundary`."""
    import ctypes
    from oracle import ptr
    from menter_case import MENTER, walls
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    imx = 1
    rng = np.random.default_rng(401 + kfracom)
    desc, ncin, lines, starts, owner = walls(lay, [(0, "j1"), (1, "i1"), (0, "k1")])
    for X in (R, G):
        X.lib.ref_set_walls(desc.shape[1], desc.ctypes.data_as(ctypes.c_void_p))
    mnpar = rng.integers(0, len(ncin) + 3, n).astype(np.int32)
    dw = 0.002 + 0.2 * rng.random(n)
    fd = np.zeros((6, n))
    fd[:3] = 1e-3 * rng.standard_normal((3, n))
    fd[3:] = 10 * rng.standard_normal((3, n))
    cmu = np.ones(n)
    try:
        for komsst in (0, 1):
            for X in (R, G):
                X.lib.ref_set_menter(ptr(MENTER), kfracom, komsst, imx)
            for l, b in enumerate(lay.blocks):
                inner = b.interior()
                tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
                if komsst == 0:
                    r_, g_ = _both(R, G, "ref_met_smmt", l + 1, v, mu, mut, dw, mnpar, ncin, fd, tp,
                                   np.full(n, 7.0), np.full(lay.ip00, 7.0), np.full(n, 7.0), np.full(n, 7.0),
                                   np.full(n, 7.0))
                    # cfke, gkgo (block-local), frac, ts6, ts7
                    close(r_[8][inner], g_[8][inner])
                    close(r_[9][:b.nnn], g_[9][:b.nnn])
                    close(r_[10], g_[10])
                    close(r_[11][inner], g_[11][inner])
                    close(r_[12][inner], g_[12][inner])
                    frac = r_[10]
                    g9 = np.zeros((9, lay.ip00))
                    g9[:, :b.nnn] = rng.standard_normal((9, b.nnn)) * 10 ** rng.uniform(0.5, 3.5, b.nnn)
                    r2, g2 = _both(R, G, "ref_met_smmtr", l + 1, v, mu, mut, dw, mnpar, ncin, fd, g9, tp,
                                   np.full(n, 7.0), np.full(n, 7.0), np.full(n, 7.0), np.full(n, 7.0))
                    # cfke, frac, ts6, ts7
                    for q in (9, 11, 12):
                        close(r2[q][inner], g2[q][inner])
                    close(r2[10], g2[10])
                    r_, g_ = _both(R, G, "ref_met_fludmt", l + 1, v, mu, mut, frac, fd)
                    close(r_[4][:, inner], g_[4][:, inner])
                g0 = np.zeros((9, lay.ip00))
                r_, g_ = _both(R, G, "ref_met_komut", l + 1, sn, vol, np.zeros(lay.ip00), g0, dw, v, mu,
                               np.full(n, 7.0), cmu)
                close(r_[7][inner], g_[7][inner])
        if kfracom:
            fr = rng.random(n); fr[rng.random(n) < 0.3] = 0.995
            r_, g_ = _both(R, G, "ref_met_mtcorf1", 0, ncin, dw, mnpar, fr)
            assert np.array_equal(r_[3], g_[3])
    finally:
        for X in (R, G):
            X.lib.ref_set_walls(0, None)


def test_smdes(mesh):
    """met_smdes through the reference's signature (node coordinates of the
    mesh, delta returned)."""
    from randcase import random_mesh_xyz
    lay, sn, vol, (x, y, z) = random_mesh_xyz([(8, 6, 5), (5, 4, 3)], seed=41)
    rng = np.random.default_rng(409)
    n = lay.ip11
    v = np.zeros((7, n)); v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    dist = 0.01 + 0.2 * rng.random(n)
    R, G = pair(lay, C.orc_phys())
    tc = C.turb_constants()
    R.set_turb(tc); G.set_turb(tc)
    fd = 1e-2 * rng.standard_normal((6, n))
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        w5 = [np.zeros(lay.ip00) for _ in range(5)]
        r_, g_ = _both(R, G, "ref_met_smdes", l + 1, x, y, z, np.zeros(lay.ip00), sn, vol, v, mu, dist, g, fd,
                       np.zeros(n), *w5, np.zeros(n), np.zeros(n), cmu)
        close(r_[3][inner - b.npc], g_[3][inner - b.npc])  # delta
        for q in (11, 17, 18):  # cfke, ts6, ts7
            close(r_[q][inner], g_[q][inner])


@pytest.mark.parametrize("ischema", [1, 7])
def test_smsasas(tsetup, ischema):
    """met_smsasas through the reference's signature (the module's ischema
    selects met_laplaciens's form)."""
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    rng = np.random.default_rng(419 + ischema)
    fd = 1e-2 * rng.standard_normal((6, n))
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = 3 * rng.standard_normal((9, b.nnn))
        w5 = [np.zeros(lay.ip00) for _ in range(5)]
        r_, g_ = _both(R, G, "ref_met_smsasas", l + 1, ischema, sn, vol, v, mu, dist, g, fd, np.zeros(n), *w5,
                       np.zeros(n), np.zeros(n), np.zeros((4, lay.ip00)), cmu)
        for q in (7, 13, 14):  # cfke, ts6, ts7
            close(r_[q][inner], g_[q][inner])


def test_kl_model(tsetup):
    """met_smkl and met_klmut through the reference's signatures."""
    from oracle import ptr
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    kl = np.array([18.0, 1.2, 1.43, 0.41])
    for X in (R, G):
        X.lib.ref_set_kl(ptr(kl))
    rng = np.random.default_rng(431)
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        w5 = [np.zeros(lay.ip00) for _ in range(5)]
        r_, g_ = _both(R, G, "ref_met_smkl", l + 1, sn, vol, vk, mu, mut, dist, np.zeros(n, dtype=np.int32), g, fd, tp,
                       np.zeros(n), *w5, np.zeros(n), np.zeros(n), cmu)
        for q in (10, 16, 17):  # cfke, ts6, ts7
            close(r_[q][inner], g_[q][inner])
        r_, g_ = _both(R, G, "ref_met_klmut", l + 1, vk, mu, np.zeros(n), dist)
        close(r_[2][inner], g_[2][inner])


@pytest.mark.parametrize("code,lparoi,isch", [(0, 0, 7), (2, 0, 7), (3, 1, 7), (4, 0, 7), (4, 0, 1)])
def test_ke_model(tsetup, code, lparoi, isch):
    """met_smke (each variant, by equatt(4:4)) and met_kemut (2JL, 2JL with
    kesst = 1, 2LS) through the reference's signatures."""
    from oracle import ptr
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    ke = np.array([1.55, 2.0, 0.09])
    for X in (R, G):
        X.lib.ref_set_ke(ptr(ke), lparoi, 0, code)
    rng = np.random.default_rng(611 + code)
    vk = v.copy(); vk[6] = 5e-2 * (1 + rng.random(n)) * v[0]
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        inner = b.interior()
        m = inner - b.npc
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-2 * rng.random(b.nnn)
        w4 = [np.zeros(lay.ip00) for _ in range(4)]
        r_, g_ = _both(R, G, "ref_met_smke", l + 1, isch, sn, vol, vk, mu, mut, tp, np.zeros(n), g, *w4,
                       np.zeros(lay.ip00), np.zeros(lay.ip00), np.zeros(n), np.zeros(n), np.zeros((3, lay.ip00)), c,
                       cmu)
        for q in (6, 14, 15):  # cfke, ts6, ts7
            close(r_[q][inner], g_[q][inner])
        for q in (12, 13):  # bark, bare
            close(r_[q][m], g_[q][m])
        close(r_[7][:, :b.nnn], g_[7][:, :b.nnn])  # velocity gradients with met_exgr's ghost layers
    for form, kes, ecode in ((0, 0, 0), (1, 1, 1), (2, 0, 5)):
        for X in (R, G):
            X.lib.ref_set_ke(ptr(ke), lparoi, kes, ecode)
        for l, b in enumerate(lay.blocks):
            inner = b.interior()
            r_, g_ = _both(R, G, "ref_met_kemut", l + 1, sn, vol, vk, mu, np.zeros(n), dist, np.zeros((9, lay.ip00)),
                           np.zeros(lay.ip00), cmu)
            close(r_[4][inner], g_[4][inner])
            if form == 1:
                close(r_[2][:, inner], g_[2][:, inner])


def test_kl_sas(tsetup):
    """met_smklsas through the reference's signature."""
    from oracle import ptr
    lay, sn, vol, v, mu, mut, dist, c, R, G = tsetup
    n = lay.ip11
    kl = np.array([18.0, 1.2, 1.43, 0.41])
    for X in (R, G):
        X.lib.ref_set_kl(ptr(kl))
    rng = np.random.default_rng(433)
    vk = v.copy(); vk[6] = 0.05 * (1 + rng.random(n)) * v[0]
    fd = np.zeros((6, n)); fd[:3] = 1e-3 * rng.standard_normal((3, n)); fd[3:] = 1e-2 * rng.standard_normal((3, n))
    cmu = np.ones(n)
    for isch in (7, 1):
        for l, b in enumerate(lay.blocks):
            inner = b.interior()
            g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
            tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
            w5 = [np.zeros(lay.ip00) for _ in range(5)]
            r_, g_ = _both(R, G, "ref_met_smklsas", l + 1, isch, sn, vol, vk, mu, mut, dist,
                           np.zeros(n, dtype=np.int32), g, fd, tp, np.zeros(n), *w5, np.zeros(n), np.zeros(n),
                           np.zeros((4, lay.ip00)), cmu)
            for q in (10, 16, 17):  # cfke, ts6, ts7
                close(r_[q][inner], g_[q][inner])


def test_dropin_stage_rate():
    """PCIe-inclusive rate of the drop-in path on the 256x128x64 wall block
    (2,097,152 cells): the routines of an implicit stage that the drop-ins
    replace -- zpres, zvismo, chronos, zvisqc (with its stress ghosts),
    sch_roe (minmod), implimf_3d (lmax = 10) -- called with the reference's
    host arrays, so that every call moves its block windows host -> device
    and its outputs back.  The rest of the stage (the state's ghost cells)
    runs on the oracle between them and is not timed.  Prints the rate over
    three timed stages (after one untimed) and checks that the state stays
    physical; then the same with the mirrors resident (codens_host_residency),
    first with the oracle's rbord on the host between the drop-ins, then with
    the ghost step on the device, with the bytes each stage moves."""
    import time
    import cpu_stage
    from codens_amd import configs
    case = configs.wall_ns()
    st = cpu_stage.CpuStage(case)
    sc = case.scheme
    G = C.Ref(case.layout, C.orc_phys(**case.phys), lib=oracle.dropin_lib(), muscl=sc["muscl"], xk=sc["xk"],
              ilim=sc["ilim"], ns=True)
    spent = [0.0]

    class Timed:
        """The stage's kernel calls: the drop-in ones timed, the others the oracle's."""
        def __getattr__(self, name):
            if name in ("zpres", "zvismo", "chronos", "sch_roe", "implimf_3d"):
                fn = getattr(G, name)

                def call(*a, **k):
                    t0 = time.perf_counter()
                    fn(*a, **k)
                    spent[0] += time.perf_counter() - t0
                return call
            return getattr(st.O, name)
    st.K = Timed()
    RFG, coin, other = st.frontiers(G)

    def zvisqc():
        t0 = time.perf_counter()
        RFG.zvisqc(other, coin, st.v, st.mu, st.mut, st.tq, case.sn, case.vol, st.temp)
        spent[0] += time.perf_counter() - t0
    st.zvisqc = zvisqc
    st.stage()  # first call: device buffers made, metrics and volumes uploaded
    spent[0] = 0.0
    for _ in range(3):
        st.stage()
    sec = spent[0] / 3
    rate = case.layout.ncells / sec / 1e6
    print("drop-in path, 256x128x64: %.2f ms per stage in the drop-in routines, %.1f M cell-updates/s "
          "(PCIe-inclusive)" % (1e3 * sec, rate))
    inner = case.layout.interior()
    assert np.isfinite(st.v[:, inner]).all() and (st.v[0, inner] > 0).all()

    # The same with the mirrors resident: the stage's only host code, the
    # oracle's rbord, gets v, p, T, c synchronised before it and marked
    # written after it (timed with the drop-ins: it is PCIe traffic).
    from codens_amd import abi
    host_arrays = (st.v, st.ps, st.temp, st.cson)
    orig_rbord = st.rbord

    def rbord():
        t0 = time.perf_counter()
        for a in host_arrays:
            abi.host_sync(a)
        spent[0] += time.perf_counter() - t0
        orig_rbord()
        t0 = time.perf_counter()
        for a in host_arrays:
            abi.host_modified(a)
        spent[0] += time.perf_counter() - t0
    st.rbord = rbord
    abi.host_residency(1)
    try:
        st.stage()
        spent[0] = 0.0
        h0, d0 = abi.host_traffic()
        for _ in range(3):
            st.stage()
        h1, d1 = abi.host_traffic()
    finally:
        abi.host_residency(0)
    sec_r = spent[0] / 3
    cells = case.layout.ncells
    print("drop-in path, resident mirrors: %.2f ms per stage, %.1f M cell-updates/s; %.0f B/cell host->device, "
          "%.0f B/cell device->host per stage (rbord's v, p, T, c on the host)"
          % (1e3 * sec_r, cells / sec_r / 1e6, (h1 - h0) / 3 / cells, (d1 - d0) / 3 / cells))
    assert np.isfinite(st.v[:, inner]).all() and (st.v[0, inner] > 0).all()

    # A driver whose ghost step runs on the device as well: the C-ABI boundary
    # routines (rbord's sequence) on the drop-ins' own device copies of v, p,
    # T, c, whose windows it then declares device-newer.  No host code touches
    # the stage's arrays, so a resident stage moves only zvismo's counter.
    import ctypes
    import torch
    from codens_amd._lib import lib
    lay = case.layout
    n = lay.ip11
    i64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device="cuda")
    f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")
    cg, w, fa, sy = case.coincident, case.wall, case.farfield, case.sym
    L = dict(cg=i64(cg.ghost), ci=i64(cg.inner), cr=i64(cg.remote if cg.remote is not None else cg.inner),
             wg=i64(w.ghost), wi=i64(w.inner), fg=i64(fa.ghost), fi=i64(fa.inner),
             fn=f64(fa.normals if fa.normals is not None else np.zeros((3, 0))),
             yg=i64(sy.ghost), yi=i64(sy.inner), yn=f64(sy.normals if sy.normals is not None else np.zeros((3, 0))))
    P = lambda k: ctypes.c_void_p(L[k].data_ptr())
    ph = abi.phys(**case.phys)
    w0 = (ctypes.c_double * 5)(*[float(x) for x in case.winf])
    torch.cuda.synchronize()
    arrays = ((st.v, 5), (st.ps, 1), (st.temp, 1), (st.cson, 1))

    def rbord_dev():
        t0 = time.perf_counter()
        d = []
        for a, rows in arrays:
            for b in lay.blocks:
                base = abi.host_mirror(a, 8 * b.npc, 8 * b.nnn, 8 * n, rows)
            d.append(ctypes.c_void_p(base))
        V, Pp, T, Cc = d
        chk = lambda rc: rc == 0 or pytest.fail(lib.codens_last_error().decode())
        if L["cg"].numel():
            chk(lib.codens_bc_coincident(L["cg"].numel(), P("cg"), P("ci"), P("cr"), V, n, Pp, T, Cc, None))
        if L["wg"].numel():
            chk(lib.codens_bc_wall(L["wg"].numel(), P("wg"), P("wi"), V, n, Pp, T, Cc, None))
        if L["fg"].numel():
            chk(lib.codens_bc_farfield(L["fg"].numel(), P("fg"), P("fi"), P("fn"), w0, ctypes.byref(ph), V, n, Pp,
                                       T, Cc, None))
        if L["yg"].numel():
            chk(lib.codens_bc_sym(L["yg"].numel(), P("yg"), P("yi"), P("yn"), V, n, None))
        if L["cg"].numel():
            chk(lib.codens_bc_coincident(L["cg"].numel(), P("cg"), P("ci"), P("cr"), V, n, Pp, T, Cc, None))
        for a, rows in arrays:
            for b in lay.blocks:
                abi.host_fetch(a, 8 * b.npc, 8 * b.nnn, 8 * n, rows)
        spent[0] += time.perf_counter() - t0
    st.rbord = rbord_dev
    abi.host_residency(1)
    try:
        st.stage()
        spent[0] = 0.0
        h0, d0 = abi.host_traffic()
        for _ in range(3):
            st.stage()
        h1, d1 = abi.host_traffic()
    finally:
        abi.host_residency(0)
    sec_d = spent[0] / 3
    print("drop-in path, resident mirrors and device ghosts: %.2f ms per stage, %.1f M cell-updates/s; "
          "%d B host->device, %d B device->host per stage" % (1e3 * sec_d, cells / sec_d / 1e6, (h1 - h0) // 3,
                                                             (d1 - d0) // 3))
    assert (h1 - h0, d1 - d0) == (3 * 4 * len(lay.blocks), 3 * 4 * len(lay.blocks))
    assert np.isfinite(st.v[:, inner]).all() and (st.v[0, inner] > 0).all()

    # The same ghost step through the rbord drop-in: a stage made only of
    # drop-in calls, resident.
    fs = case.farspec
    kinds = [x[0] for x in st.fr]
    nrm = RFG.normals([x[1].normals[:, x[2]] if x[0] in ("nrd", "sym") else None for x in st.fr])
    yz = np.zeros(n)

    def rbord_dropin():
        t0 = time.perf_counter()
        RFG.rbord(kinds, [x[3] for x in st.fr], coin, nrm, st.v, st.ps, st.temp, st.cson, fs.mach, fs.alpha,
                  fs.beta, fs.roa1, fs.aa1, fs.pa1, yz, yz)
        spent[0] += time.perf_counter() - t0
    st.rbord = rbord_dropin
    abi.host_residency(1)
    try:
        st.stage()
        spent[0] = 0.0
        h0, d0 = abi.host_traffic()
        for _ in range(3):
            st.stage()
        h1, d1 = abi.host_traffic()
    finally:
        abi.host_residency(0)
    sec_e = spent[0] / 3
    print("drop-in path, resident mirrors and the rbord drop-in: %.2f ms per stage, %.1f M cell-updates/s; "
          "%d B host->device, %d B device->host per stage" % (1e3 * sec_e, cells / sec_e / 1e6, (h1 - h0) // 3,
                                                             (d1 - d0) // 3))
    assert (h1 - h0, d1 - d0) == (3 * 4 * len(lay.blocks), 3 * 4 * len(lay.blocks))
    assert np.isfinite(st.v[:, inner]).all() and (st.v[0, inner] > 0).all()
