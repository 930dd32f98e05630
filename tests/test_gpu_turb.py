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
