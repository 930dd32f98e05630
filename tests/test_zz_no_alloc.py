"""codens_implimf_3d_noalloc (the implicit phase with the library's static
scratch, no allocation) against codens_implimf_3d on the same inputs: two
calls in a row, the second on a smaller block, each within the 1e-12 bar
(the same kernels on the same operands: bit for bit in practice), and a
block over the scratch's maximum is refused with CODENS_EINVAL."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from randcase import all_thermo, random_mesh, random_state, relerr


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=torch.float64)


def _run(fn, lay, sn, vol, seed, lmx):
    from codens_amd import abi
    rng = np.random.default_rng(seed)
    n = lay.ip11
    v = random_state(n, rng)
    ps, t, c = all_thermo(v)
    mu = 1e-3 * (1 + 0.1 * rng.random(n)); mut = 1e-4 * rng.random(n)
    u = 1e-3 * rng.standard_normal((5, n)); ff = 1e-4 * rng.standard_normal((5, n))
    dt = 0.01 * (1 + rng.random(n))
    gv = T(v); gd = T(np.full((5, n), 5.0))
    ph = abi.phys()
    for b in lay.blocks:
        fn(abi.block(b), ph, T(u), T(dt), gv, gd, T(ff), T(mu), T(mut), lmx, T(sn[9 * b.npc: 9 * (b.npc + b.nnn)]),
           T(vol), T(ps), T(c))
    torch.cuda.synchronize()
    return gv.cpu().numpy(), gd.cpu().numpy()


def test_noalloc_matches_implimf_3d():
    from codens_amd import abi
    big = random_mesh([(70, 37, 19)], seed=3)
    small = random_mesh([(9, 7, 5), (6, 5, 4)], seed=4)
    for (lay, sn, vol), seed, lmx in ((big, 11, 10), (small, 12, 3)):
        v0, d0 = _run(abi.implimf_3d, lay, sn, vol, seed, lmx)
        v1, d1 = _run(abi.implimf_3d_noalloc, lay, sn, vol, seed, lmx)
        for e in range(5):
            assert relerr(d1[e], d0[e]) <= 1e-12, e
            assert relerr(v1[e], v0[e]) <= 1e-12, e
        assert np.abs(d0).max() > 0


def test_noalloc_rejects_oversized_block():
    from codens_amd import abi
    from codens_amd.abi import CodensError
    lay, sn, vol = random_mesh([(9, 7, 5)], seed=5)
    cb = abi.block(lay.blocks[0])
    big = abi.CodensBlock(0, 0, 301, 0, 101, 0, 81, 1, 300, 1, 100, 1, 80, 302 * 102 * 82, 0, 0)
    n = 302 * 102 * 82
    z = torch.zeros((5, n), dtype=torch.float64, device="cuda")
    o = torch.ones(n, dtype=torch.float64, device="cuda")
    with pytest.raises(CodensError) as e:
        abi.implimf_3d_noalloc(big, abi.phys(), z, o, z, z, z, o, o, 3, torch.zeros(9 * n, dtype=torch.float64,
                               device="cuda"), o, o, o)
    assert e.value.code == 1 and "maximum" in str(e.value)
    assert cb.nnn < 2400000
