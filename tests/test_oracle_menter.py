"""Menter k-omega model (met_smmt with its blending function, met_mtcorf1,
met_fludmt, met_komut): the oracle against the flang build of the reference
routines, bit for bit."""
import ctypes

import numpy as np
import pytest

import cpu_ref as C
import oracle
from menter_case import MENTER, walls, walls_of_block
from oracle import ptr
from randcase import random_mesh, random_state

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (5, 4, 3)]
IMX = 1  # imxclko: the first imxclko+2 cells of a line (set to 1) stay inside these blocks
WALLS = [(0, "j1"), (1, "i1"), (0, "k1")]  # the k1 lines of block 0 cross its j1 lines


@pytest.fixture(scope="module")
def setup():
    lay, sn, vol = random_mesh(DIMS, seed=61)
    rng = np.random.default_rng(63)
    n = lay.ip11
    v = np.zeros((7, n))
    v[:5] = random_state(n, rng)
    v[5] = 1e-3 * (1 + rng.random(n)) * v[0]
    v[6] = 50 * (1 + rng.random(n)) * v[0]
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    mut = 1e-3 * (0.5 + rng.random(n))
    dist = 0.002 + 0.2 * rng.random(n)
    R = C.Ref(lay, C.orc_phys())
    R.set_turb(C.turb_constants())
    desc, ncin, lines, starts, owner = walls(lay, WALLS)
    R.lib.ref_set_walls(len(WALLS), desc.ctypes.data_as(ctypes.c_void_p))
    npts = len(ncin)
    mnpar = rng.integers(0, npts + 3, n).astype(np.int32)
    return lay, sn, vol, v, mu, mut, dist, R, oracle.oracle_lib(), (ncin, lines, starts, owner, mnpar)


def _set(R, kfracom, komsst=0):
    R.lib.ref_set_menter(ptr(MENTER), kfracom, komsst, IMX)


def test_mtcorf1(setup):
    """Wall correction of the blending function: per block and for all
    blocks at once (l = 0), on a frac with values above and below 0.99."""
    lay, sn, vol, v, mu, mut, dist, R, O, (ncin, lines, starts, owner, mnpar) = setup
    rng = np.random.default_rng(65)
    _set(R, 1)
    for l in (1, 2, 0):
        frac = rng.random(lay.ip11)
        frac[rng.random(lay.ip11) < 0.3] = 0.995
        fr, fo = frac.copy(), frac.copy()
        R.lib.ref_met_mtcorf1(l, ptr(ncin), ptr(dist), ptr(mnpar), ptr(fr))
        if l == 0:
            ln = lines
        else:
            ln, _ = walls_of_block(lines, starts, owner, l - 1)
        O.orc_met_mtcorf1(ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), IMX, ptr(mnpar), ptr(fo))
        assert np.array_equal(fr, fo)
        assert not np.array_equal(fr, frac)


@pytest.mark.parametrize("kfracom", [0, 1])
def test_smmt(setup, kfracom):
    """Sources, blending function and gkgo (kfracom = 1 with the walls)."""
    lay, sn, vol, v, mu, mut, dist, R, O, (ncin, lines, starts, owner, mnpar) = setup
    n = lay.ip11
    rng = np.random.default_rng(67 + kfracom)
    _set(R, kfracom)
    fd = np.zeros((6, n))
    fd[:3] = 1e-3 * rng.standard_normal((3, n))
    fd[3:] = 10 * rng.standard_normal((3, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        r = dict(cfke=np.full(n, 7.0), gkgo=np.full(lay.ip00, 7.0), frac=np.full(n, 7.0), ts6=np.full(n, 7.0),
                 ts7=np.full(n, 7.0))
        o = {k: a.copy() for k, a in r.items()}
        R.lib.ref_met_smmt(l + 1, ptr(v), ptr(mu), ptr(mut), ptr(dist), ptr(mnpar), ptr(ncin), ptr(fd), ptr(tp),
                           ptr(r["cfke"]), ptr(r["gkgo"]), ptr(r["frac"]), ptr(r["ts6"]), ptr(r["ts7"]))
        ln, st = walls_of_block(lines, starts, owner, l)
        O.orc_met_smmt(ctypes.byref(ob), ptr(MENTER), kfracom, IMX, ptr(v), C.L_(n), ptr(mu), ptr(mut), ptr(dist),
                       ptr(fd), ptr(tp), ptr(o["cfke"]), ptr(o["gkgo"]), ptr(o["frac"]), ptr(o["ts6"]), ptr(o["ts7"]),
                       ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), ptr(mnpar))
        for k in r:
            assert np.array_equal(r[k], o[k]), (l, k)
        inner = b.interior()
        assert np.isfinite(r["ts7"][inner]).all()
        if kfracom:
            f = r["frac"][inner]
            assert (f > 0).all() and (f <= 1).all() and (f < 1).any()


@pytest.mark.parametrize("kfracom", [0, 1])
def test_fludmt(setup, kfracom):
    lay, sn, vol, v, mu, mut, dist, R, O, _ = setup
    n = lay.ip11
    rng = np.random.default_rng(71)
    _set(R, kfracom)
    frac = rng.random(n)
    fd0 = rng.standard_normal((6, n))
    for l, b in enumerate(lay.blocks):
        fr, fo = fd0.copy(), fd0.copy()
        R.lib.ref_met_fludmt(l + 1, ptr(v), ptr(mu), ptr(mut), ptr(frac), ptr(fr))
        O.orc_met_fludmt(ctypes.byref(C.orc_block(b)), ptr(MENTER), kfracom, ptr(mu), ptr(mut), ptr(frac), ptr(fo),
                         C.L_(n))
        assert np.array_equal(fr, fo)
        assert not np.array_equal(fr, fd0)


@pytest.mark.parametrize("komsst", [0, 1])
def test_komut(setup, komsst):
    """Eddy viscosity; the SST form runs teq_gradv on the reference side, so
    the oracle gets the gradients the reference computed."""
    lay, sn, vol, v, mu, mut, dist, R, O, _ = setup
    n = lay.ip11
    _set(R, 1, komsst)
    cmu = np.ones(n)
    for l, b in enumerate(lay.blocks):
        g = np.zeros((9, lay.ip00)); t = np.zeros(lay.ip00)
        vr, mr = v.copy(), np.full(n, 7.0)
        R.lib.ref_met_komut(l + 1, ptr(sn), ptr(vol), ptr(t), ptr(g), ptr(dist), ptr(vr), ptr(mu), ptr(mr), ptr(cmu))
        mo = np.full(n, 7.0)
        gb = np.ascontiguousarray(g[:, :b.nnn])
        O.orc_met_komut(ctypes.byref(C.orc_block(b)), C.D(MENTER[0]), komsst, ptr(vr), C.L_(n), ptr(mu), ptr(mo),
                        ptr(dist), ptr(gb), C.L_(b.nnn))
        assert np.array_equal(mr, mo)
        assert np.isfinite(mr[b.interior()]).all()


def test_smmtr(setup):
    """Sources with Durbin's realizability limit (blending function and
    wall correction always on)."""
    lay, sn, vol, v, mu, mut, dist, R, O, (ncin, lines, starts, owner, mnpar) = setup
    n = lay.ip11
    rng = np.random.default_rng(73)
    _set(R, 1)
    fd = np.zeros((6, n))
    fd[:3] = 1e-3 * rng.standard_normal((3, n))
    fd[3:] = 10 * rng.standard_normal((3, n))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        g = np.zeros((9, lay.ip00))
        g[:, :b.nnn] = rng.standard_normal((9, b.nnn)) * 10 ** rng.uniform(0.5, 3.5, b.nnn)
        gb = np.ascontiguousarray(g[:, :b.nnn])
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-3 * rng.random(b.nnn)
        r = dict(cfke=np.full(n, 7.0), frac=np.full(n, 7.0), ts6=np.full(n, 7.0), ts7=np.full(n, 7.0))
        o = {k: a.copy() for k, a in r.items()}
        R.lib.ref_met_smmtr(l + 1, ptr(v), ptr(mu), ptr(mut), ptr(dist), ptr(mnpar), ptr(ncin), ptr(fd), ptr(g),
                            ptr(tp), ptr(r["cfke"]), ptr(r["frac"]), ptr(r["ts6"]), ptr(r["ts7"]))
        ln, st = walls_of_block(lines, starts, owner, l)
        O.orc_met_smmtr(ctypes.byref(ob), ptr(MENTER), IMX, ptr(v), C.L_(n), ptr(mu), ptr(mut), ptr(dist), ptr(fd),
                        ptr(gb), C.L_(b.nnn), ptr(tp), ptr(o["cfke"]), ptr(o["frac"]), ptr(o["ts6"]), ptr(o["ts7"]),
                        ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), ptr(mnpar))
        for k in r:
            assert np.array_equal(r[k], o[k]), (l, k)
        # the limit is active somewhere and not everywhere
        o2 = {k: np.full(n, 7.0) for k in r}
        O.orc_met_smmt(ctypes.byref(ob), ptr(MENTER), 1, IMX, ptr(v), C.L_(n), ptr(mu), ptr(mut), ptr(dist), ptr(fd),
                       ptr(tp), ptr(o2["cfke"]), ptr(np.zeros(lay.ip00)), ptr(o2["frac"]), ptr(o2["ts6"]),
                       ptr(o2["ts7"]), ln.ctypes.data_as(ctypes.c_void_p), C.L_(len(ln)), ptr(mnpar))
        inner = b.interior()
        same = r["ts7"][inner] == o2["ts7"][inner]
        assert same.any() and not same.all()
