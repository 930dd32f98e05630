"""The oracle's steadiness norms (residu.f90 per block, rscpsv.f90 over all
blocks) against the reference's own residu, bit for bit per block, and the
aggregate against its closed forms over all real cells."""
import ctypes

import numpy as np
import pytest

import oracle
import cpu_ref as C
from oracle import ptr
from randcase import random_mesh

pytestmark = pytest.mark.skipif(oracle.ref_lib() is None, reason="reference build (oracle/_ref) not present")

DIMS = [(8, 6, 5), (5, 4, 3), (7, 3, 4)]


def orc_residu(lay, neq, v, u, dt):
    blocks = (C.OrcBlock * len(lay.blocks))(*[C.orc_block(b) for b in lay.blocks])
    out = np.zeros((len(lay.blocks) + 1, neq, 7))
    oracle.oracle_lib().orc_residu(blocks, len(lay.blocks), neq, ptr(v), ptr(u), C.L_(lay.ip11), ptr(dt), ptr(out))
    return out


def inputs(lay, neq, rng):
    n = lay.ip11
    u = rng.standard_normal((7, n))
    v = u + 1e-3 * rng.standard_normal((7, n)) * (1 + 10 * rng.random((7, n)))
    dt = 0.01 * (1 + rng.random(n))
    return v, u, dt


@pytest.mark.parametrize("neq", [5, 7])
def test_residu_bit_exact(neq):
    lay, sn, vol = random_mesh(DIMS, seed=4)
    rng = np.random.default_rng(40 + neq)
    v, u, dt = inputs(lay, neq, rng)
    R = C.Ref(lay, C.orc_phys())
    out = orc_residu(lay, neq, v, u, dt)
    for l in range(len(lay.blocks)):
        d1, d2, dm = np.zeros(7), np.zeros(7), np.zeros(7)
        im, jm, km = (np.zeros(7, dtype=np.int32) for _ in range(3))
        R.lib.ref_residu(l + 1, neq, ptr(u), ptr(v), ptr(dt), ptr(d1), ptr(d2), ptr(dm), ptr(im), ptr(jm), ptr(km))
        got = out[l]
        assert np.array_equal(got[:, 0], d1[:neq]) and np.array_equal(got[:, 1], d2[:neq])
        assert np.array_equal(got[:, 2], dm[:neq])
        assert np.array_equal(got[:, 4:7], np.stack([im, jm, km], 1)[:neq].astype(float))


def test_residu_aggregate_and_ties():
    lay, sn, vol = random_mesh(DIMS, seed=4)
    rng = np.random.default_rng(9)
    v, u, dt = inputs(lay, 5, rng)
    out = orc_residu(lay, 5, v, u, dt)
    inner = lay.interior()
    r = (v[:5, inner] - u[:5, inner]) / dt[inner]
    g = out[-1]
    assert np.allclose(g[:, 0], np.abs(r).mean(1), rtol=1e-14, atol=0)
    assert np.allclose(g[:, 1], np.sqrt((r * r).mean(1)), rtol=1e-14, atol=0)
    assert np.array_equal(g[:, 2], np.abs(r).max(1))
    # the aggregate's location is that of the block holding the maximum
    for e in range(5):
        l = int(g[e, 3])
        assert out[l, e, 2] == g[e, 2] and np.array_equal(out[l, e, 4:], g[e, 4:])
    # ties: a constant |r| puts the location on the last real cell of the block
    # (storage order) and the aggregate on the last block
    z = np.zeros_like(u)
    v2 = np.repeat((0.5 * dt)[None, :], 7, axis=0)  # r = 0.5 exactly
    out = orc_residu(lay, 5, v2, z, dt)
    for l, b in enumerate(lay.blocks):
        assert np.all(out[l, :, 4:] == [b.ii2 - 1, b.jj2 - 1, b.kk2 - 1])
    assert np.all(out[-1, :, 3] == len(lay.blocks) - 1)
    # a planted maximum is found where it is
    b = lay.blocks[1]
    n = b.idx(b.ii1 + 2, b.jj1 + 1, b.kk1 + 1)
    v3 = v.copy()
    v3[3, n] = u[3, n] + 1e3 * dt[n]
    out = orc_residu(lay, 5, v3, u, dt)
    assert abs(out[-1, 3, 2] - 1e3) < 1e-9
    assert np.array_equal(out[-1, 3, 3:], [1, b.ii1 + 2, b.jj1 + 1, b.kk1 + 1])
