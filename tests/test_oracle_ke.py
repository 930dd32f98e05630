"""k-epsilon family of the turbulence stage, oracle against the reference
build, bit for bit: met_smke (met_exgr, met_bark, met_bare, then the variant
met_smkes, met_smker, met_smkec or met_smkesas) and met_kemut (Jones-Launder,
with the SST limiter, Launder-Sharma)."""
import ctypes

import numpy as np
import pytest

import cpu_ref as C
import oracle
from oracle import ptr
from randcase import random_mesh, random_state

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build unavailable")

DIMS = [(9, 7, 5), (6, 8, 7)]
KE = np.array([1.55, 2.0, 0.09])  # cke1, cke2, cmu (atlecdon.f90:124-128)
# variant -> letter code of ref_set_ke (0 ' ', 2 'R', 3 'C', 4 'L')
CODES = {0: 0, 1: 2, 2: 3, 3: 4}


@pytest.fixture(scope="module")
def ke():
    lay, sn, vol = random_mesh(DIMS, seed=61)
    rng = np.random.default_rng(63)
    n = lay.ip11
    v = np.zeros((7, n))
    v[:5] = random_state(n, rng)
    v[5] = 1e-2 * (1 + rng.random(n)) * v[0]       # rho k
    v[6] = 5e-2 * (1 + rng.random(n)) * v[0]       # rho epsilon
    mu = 1e-3 * (1 + 0.1 * rng.random(n))
    mut = 1e-3 * rng.random(n)
    dist = 0.05 + rng.random(n)
    cson = np.sqrt(1.4 * (0.4 * (v[4] - 0.5 * (v[1] ** 2 + v[2] ** 2 + v[3] ** 2) / v[0])) / v[0])
    R = C.Ref(lay, C.orc_phys())
    R.set_turb(C.turb_constants())
    return lay, sn, vol, v, mu, mut, dist, cson, R, oracle.oracle_lib(), rng


def snb(sn, b):
    return np.ascontiguousarray(sn[9 * b.npc: 9 * (b.npc + b.nnn)])


@pytest.mark.parametrize("variant,lparoi,isch", [(0, 0, 7), (0, 1, 7), (1, 0, 7), (2, 0, 7), (2, 1, 7),
                                                 (3, 0, 7), (3, 0, 1), (3, 1, 7)])
def test_met_smke(ke, variant, lparoi, isch):
    lay, sn, vol, v, mu, mut, dist, cson, R, O, rng = ke
    n = lay.ip11
    R.lib.ref_set_ke(ptr(KE), lparoi, 0, CODES[variant])
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        inner = b.interior()
        g = np.zeros((9, lay.ip00)); g[:, :b.nnn] = rng.standard_normal((9, b.nnn))
        gb = np.ascontiguousarray(g[:, :b.nnn])
        tp = np.zeros(lay.ip00); tp[:b.nnn] = 1e-2 * rng.random(b.nnn)
        w = [np.zeros(lay.ip00) for _ in range(4)]
        tn = np.zeros((3, lay.ip00))
        rb = [np.full(lay.ip00, 7.0) for _ in range(2)]
        r = [np.full(n, 7.0) for _ in range(3)]
        R.lib.ref_met_smke(l + 1, isch, ptr(sn), ptr(vol), ptr(v), ptr(mu), ptr(mut), ptr(tp), ptr(r[0]), ptr(g),
                           *[ptr(a) for a in w], ptr(rb[0]), ptr(rb[1]), ptr(r[1]), ptr(r[2]), ptr(tn), ptr(cson),
                           ptr(np.ones(n)))
        ob_ = [np.full(b.nnn, 7.0) for _ in range(2)]
        o = [np.full(n, 7.0) for _ in range(3)]
        O.orc_met_smke(ctypes.byref(ob), ptr(KE), variant, lparoi, isch, ptr(snb(sn, b)), ptr(vol), ptr(v), C.L_(n),
                       ptr(mu), ptr(mut), ptr(np.ascontiguousarray(tp[:b.nnn])), ptr(gb), ptr(cson), ptr(ob_[0]),
                       ptr(ob_[1]), *[ptr(a) for a in o])
        m = inner - b.npc
        for a, c in zip(r, o):  # cfke, ts6, ts7
            assert np.array_equal(a[inner], c[inner])
            assert np.isfinite(a[inner]).all()
        if variant != 1:  # bark and bare are computed for every variant
            for a, c in zip(rb, ob_):
                assert np.array_equal(a[m], c[m])
        assert np.array_equal(g[:, :b.nnn], gb)  # met_exgr's ghost layers
        assert np.abs(o[2][inner]).max() > 0


@pytest.mark.parametrize("form,lparoi", [(0, 0), (0, 1), (1, 0), (2, 0), (2, 2)])
def test_met_kemut(ke, form, lparoi):
    lay, sn, vol, v, mu, mut, dist, cson, R, O, rng = ke
    n = lay.ip11
    R.lib.ref_set_ke(ptr(KE), lparoi, 1 if form == 1 else 0, 5 if form == 2 else (1 if form == 1 else 0))
    for l, b in enumerate(lay.blocks):
        ob = C.orc_block(b)
        vr, vo = v.copy(), v.copy()
        g = np.zeros((9, lay.ip00)); gb = np.zeros((9, b.nnn))
        mr, mo = np.full(n, 7.0), np.full(n, 7.0)
        R.lib.ref_met_kemut(l + 1, ptr(sn), ptr(vol), ptr(vr), ptr(mu), ptr(mr), ptr(dist), ptr(g),
                            ptr(np.zeros(lay.ip00)), ptr(np.ones(n)))
        O.orc_met_kemut(ctypes.byref(ob), C.D(KE[2]), form, lparoi, ptr(snb(sn, b)), ptr(vol), ptr(vo), C.L_(n),
                        ptr(mu), ptr(mo), ptr(dist), ptr(gb))
        assert np.array_equal(mr, mo)
        assert np.array_equal(vr, vo)
        assert np.isfinite(mo[b.interior()]).all()
