"""The uneven multi-block RANS case (1-to-1 coincident patches along a seeded
adjacency tree): its topology, on the CPU.

* the full 48-block draw: every axis in 16..192 cells, 38M..42M cells in all;
* every ghost cell is filled by exactly one face rule;
* every coincident ghost pairs its own adjacent cell with the cell one step
  across the face in the neighbouring block, found from the blocks' lattice
  origins (an independent restatement of the placement), and each patch is
  listed from both sides;
* the node coordinates of the two blocks agree on every shared face node.
"""
import numpy as np

from codens_amd import configs


def _where(case, n):
    """(block, lattice cell) of flat storage index n."""
    for l, b in enumerate(case.layout.blocks):
        if b.npc <= n < b.npc + b.nnn:
            m = n - b.npc
            i = b.id1 + m % b.nid
            j = b.jd1 + (m // b.nid) % b.njd
            k = b.kd1 + m // (b.nid * b.njd)
            return l, case.origin[l] + np.array([i - b.ii1, j - b.jj1, k - b.kk1])
    raise AssertionError(n)


def test_full_draw():
    dims, _ = configs.multiblock_dims()
    d = np.array(dims)
    assert d.shape == (48, 3) and d.min() >= 16 and d.max() <= 192
    assert 38_000_000 <= int(np.prod(d, axis=1).sum()) <= 42_000_000


def test_topology_reduced():
    case = configs.multiblock_sa(9, 4, 11, None, seed=7)
    c, w, f = case.coincident, case.wall, case.farfield
    allg = np.concatenate([c.ghost, w.ghost, f.ghost])
    assert len(allg) == len(np.unique(allg))
    nface = sum(2 * ((b.ni - 1) * (b.nj - 1) + (b.nj - 1) * (b.nk - 1) + (b.ni - 1) * (b.nk - 1))
                for b in case.layout.blocks)
    assert len(allg) == nface
    assert len(case.links) == 8 and len(c.ghost) > 0 and len(w.ghost) > 0
    pairs = set()
    for g, i, r in zip(c.ghost, c.inner, c.remote):
        lg, pg = _where(case, g)
        li, pi = _where(case, i)
        lr, pr = _where(case, r)
        assert lg == li != lr
        step = pr - pi
        assert np.abs(step).sum() == 1          # one lattice step across the face
        assert np.array_equal(pg, pi + step)     # the ghost sits on that side
        pairs.add((int(i), int(r)))
    assert all((r, i) in pairs for i, r in pairs)
    # node coordinates on the shared faces agree (uniform lattice, spacing h)
    h = 1.0 / 128.0
    xs = case.sn  # metrics are built from the same nodes: face areas h^2
    for l, b in enumerate(case.layout.blocks):
        nnn = b.nnn
        base = 9 * b.npc
        for d in range(3):
            sd = xs[base + (d * 3 + d) * nnn: base + (d * 3 + d + 1) * nnn]
            inner = b.interior()
            assert np.allclose(sd[inner - b.npc], h * h, rtol=1e-12)


def test_wall_distance_positive():
    case = configs.multiblock_sa(9, 4, 11, None, seed=7)
    inner = case.layout.interior()
    d = case.turb.dist[inner]
    assert d.min() >= 0.5 / 128.0 - 1e-15


def test_ghost_node_fold_is_the_transpose_of_extrapolation():
    """The adjoint harness folds adjoints of ghost nodes back onto the nodes
    they copy; that must be the exact transpose of the extrapolation the
    mesh metrics use (layout.extrapolate_ghost_nodes): <E a, w> = <a, F w>
    for random real-node values a and random weights w on all nodes."""
    import numpy as np
    from codens_amd import layout as L
    from codens_amd.adjoint import _fold_ghost_nodes
    lay = L.Layout([(5, 4, 3)])
    b = lay.blocks[0]
    rng = np.random.default_rng(3)
    sl = (slice(b.kk1 - b.kd1, b.kk2 - b.kd1 + 1), slice(b.jj1 - b.jd1, b.jj2 - b.jd1 + 1),
          slice(b.ii1 - b.id1, b.ii2 - b.id1 + 1))
    for _ in range(3):
        real = rng.standard_normal(b.view(np.zeros(lay.ip11))[sl].shape)
        x = np.zeros(lay.ip11); y = np.zeros(lay.ip11); z = np.zeros(lay.ip11)
        b.view(x)[sl] = real
        L.extrapolate_ghost_nodes(b, x, y, z)
        w = rng.standard_normal((b.nkd, b.njd, b.nid))
        lhs = float((b.view(x) * w).sum())
        fw = _fold_ghost_nodes(b, w)
        rhs = float((real * fw[sl]).sum())
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
        # the fold leaves nothing on ghost nodes that the extrapolation fills
        mask = np.ones_like(w, dtype=bool); mask[sl] = False
        e = np.zeros(lay.ip11); b.view(e)[sl] = 1.0
        L.extrapolate_ghost_nodes(b, e, np.zeros(lay.ip11), np.zeros(lay.ip11))
        filled = (b.view(e) != 0) & mask
        assert np.all(fw[filled] == 0)
