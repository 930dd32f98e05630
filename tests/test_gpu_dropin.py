"""The Fortran drop-in modules (codens_amd/fortran) end to end: the reference's
own call signatures, module globals and host arrays, run on the GPU through
the C-ABI, against the reference routines on the CPU, on a two-block
curvilinear layout (so every host<->device window starts inside the arrays).

The drop-in library is oracle/_ref/libcodens_dropin.so: the reference harness
and reference objects with the drop-in modules linked in place of the
reference's zpres, zvismo, chronos, sch_roe_euler, sch_roe, sch_expli,
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
