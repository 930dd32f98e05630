"""Discrete adjoint of the residual and of one 4-stage RK step (the wall
force of a viscous wall block against the state and the wall shape).

* the residual restated for differentiation equals the production kernels'
  residual within 1e-12 (the value part of the tangent-linear launch);
* dot-product tests: <J_R xdot, ubar> = <xdot, J_R^T ubar> for the residual
  (state, metrics, volumes) and for the mesh metrics (wall displacements);
* the tangent-linear sweep (forward mode) and the reverse sweep agree on the
  objective's directional derivatives within 1e-10;
* central finite differences (eps 1e-7) on 64 seeded unit directions of
  (state, wall displacement) match the reverse-sweep gradient within 1e-6.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from randcase import relerr


def _setup(ci, cj, ck, nwall):
    from codens_amd import configs
    from codens_amd.adjoint import AdjointProblem
    case = configs.adjoint_wall(ci, cj, ck, nwall_nodes=nwall)
    P = AdjointProblem(case)
    v0 = torch.as_tensor(case.v0, device="cuda")
    delta = torch.zeros(len(case.wall_nodes), dtype=torch.float64, device="cuda")
    return P, case, v0, delta


@pytest.fixture(scope="module")
def small():
    return _setup(32, 16, 8, 256)


def _dirs(P, case, rng, count):
    """Seeded unit directions of (state on the real cells, displacement)."""
    n = case.layout.ip11
    real = P.real.cpu().numpy()
    out = []
    for _ in range(count):
        dv = rng.standard_normal((5, n)) * real
        dd = rng.standard_normal(len(case.wall_nodes))
        s = np.sqrt((dv * dv).sum() + (dd * dd).sum())
        out.append((torch.as_tensor(dv / s, device="cuda"), torch.as_tensor(dd / s, device="cuda")))
    return out


def test_generic_residual_matches_production(small):
    P, case, v0, _ = small
    b = case.layout.blocks[0]
    sn = torch.as_tensor(case.sn[9 * b.npc: 9 * (b.npc + b.nnn)], device="cuda")
    vol = torch.as_tensor(case.vol, device="cuda")
    up = P.residual(v0, sn, vol)
    ug, _ = P.jvp(v0, sn, vol, None, None, None)
    inner = case.layout.interior()
    upn, ugn = up.cpu().numpy(), ug.cpu().numpy()
    for e in range(5):
        assert relerr(ugn[e, inner], upn[e, inner]) <= 1e-12, e
    assert np.abs(upn[:, inner]).max() > 0


def test_residual_dot_product(small):
    P, case, v0, _ = small
    b = case.layout.blocks[0]
    rng = np.random.default_rng(5)
    sn = torch.as_tensor(case.sn[9 * b.npc: 9 * (b.npc + b.nnn)], device="cuda")
    vol = torch.as_tensor(case.vol, device="cuda")
    real = P.real
    for trial in range(3):
        T = lambda *s: torch.as_tensor(rng.standard_normal(s), device="cuda")
        vd = T(5, case.layout.ip11) * 1e-3
        snd = T(9 * b.nnn) * 1e-6
        vold = T(case.layout.ip11) * 1e-9
        ub = T(5, case.layout.ip11) * real
        _, ud = P.jvp(v0, sn, vol, vd, snd, vold)
        vb, snb, volb = P.vjp(v0, sn, vol, ub)
        lhs = float((ud * ub).sum())
        parts = [float((vd * vb).sum()), float((snd * snb).sum()), float((vold * volb).sum())]
        rhs = sum(parts)
        assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), 1e-300), (trial, lhs, rhs, parts)
        assert all(abs(x) > 0 for x in parts)


def test_metric_dot_product(small):
    P, case, v0, _ = small
    rng = np.random.default_rng(6)
    d0 = torch.as_tensor(rng.standard_normal(len(case.wall_nodes)) * 1e-3, device="cuda")
    _, _, xyz = P.metrics(d0)
    b = case.layout.blocks[0]
    for trial in range(3):
        dd = torch.as_tensor(rng.standard_normal(len(case.wall_nodes)) * 1e-4, device="cuda")
        snbar = torch.as_tensor(rng.standard_normal(9 * b.nnn), device="cuda")
        volbar = torch.as_tensor(rng.standard_normal(case.layout.ip11), device="cuda")
        snd, vold = P.metric_jvp(d0, dd)
        lhs = float((snd * snbar).sum() + (vold * volbar).sum())
        rhs = float((dd * P.metric_vjp(snbar, volbar, xyz)).sum())
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs), (trial, lhs, rhs)


def test_device_metrics_match_host(small):
    """The adjoint's mesh (wall nodes moved by a displacement) through the
    device metrics equals the host restatement bit for bit."""
    from codens_amd import layout as L
    P, case, v0, _ = small
    rng = np.random.default_rng(12)
    d = rng.standard_normal(len(case.wall_nodes)) * 1e-3
    sn, vol, (x, y, z) = P.metrics(d)
    hx, hy, hz = P.x0.copy(), P.y0.copy(), P.z0.copy()
    hy[P.wall_nodes] += d
    sn_h, vol_h = L.metrics(case.layout, hx, hy, hz)
    b = case.layout.blocks[0]
    np.testing.assert_array_equal(sn.cpu().numpy(), sn_h[9 * b.npc: 9 * (b.npc + b.nnn)])
    np.testing.assert_array_equal(vol.cpu().numpy(), vol_h)
    np.testing.assert_array_equal(y.cpu().numpy(), hy)


def test_tangent_matches_reverse(small):
    from codens_amd.adjoint import gradient, directional
    P, case, v0, delta = small
    J, gv, gd = gradient(P, v0, delta)
    assert abs(J) > 0 and float(gv.abs().max()) > 0 and float(gd.abs().max()) > 0
    assert float(gv[:, P.real == 0].abs().max()) == 0.0
    for vdir, ddir in _dirs(P, case, np.random.default_rng(7), 4):
        t = directional(P, v0, delta, vdir, ddir)
        r = float((gv * vdir).sum() + (gd * ddir).sum())
        assert abs(t - r) <= 1e-10 * abs(r), (t, r)


def test_finite_differences_64_directions(small):
    from codens_amd.adjoint import gradient
    P, case, v0, delta = small
    _, gv, gd = gradient(P, v0, delta)
    eps = 1e-7
    errs = []
    d0 = delta.cpu().numpy()
    for vdir, ddir in _dirs(P, case, np.random.default_rng(20261205), 64):
        dd = ddir.cpu().numpy()
        jp = P.J_value(v0 + eps * vdir, d0 + eps * dd)
        jm = P.J_value(v0 - eps * vdir, d0 - eps * dd)
        fd = (jp - jm) / (2 * eps)
        ad = float((gv * vdir).sum() + (gd * ddir).sum())
        errs.append(abs(fd - ad) / abs(ad))
    print("FD relative errors: max %.3e median %.3e" % (max(errs), float(np.median(errs))))
    assert max(errs) <= 1e-6, errs


def test_full_size_adjoint():
    """The specified size: 128x64x32 block, 4096 seeded wall nodes.  On 64
    seeded unit directions of (state, wall displacement): the tangent-linear
    sweep against the reverse-sweep gradient within 1e-10, and central finite
    differences (eps 1e-7) against it within 1e-6."""
    from codens_amd.adjoint import gradient, directional
    P, case, v0, delta = _setup(128, 64, 32, 4096)
    _, gv, gd = gradient(P, v0, delta)
    d0 = delta.cpu().numpy()
    eps = 1e-7
    e_fd, e_tl = [], []
    for vdir, ddir in _dirs(P, case, np.random.default_rng(11), 64):
        t = directional(P, v0, delta, vdir, ddir)
        r = float((gv * vdir).sum() + (gd * ddir).sum())
        dd = ddir.cpu().numpy()
        fd = (P.J_value(v0 + eps * vdir, d0 + eps * dd) - P.J_value(v0 - eps * vdir, d0 - eps * dd)) / (2 * eps)
        e_tl.append(abs(t - r) / abs(r))
        e_fd.append(abs(fd - r) / abs(r))
    print("128x64x32, 64 directions: FD max %.3e median %.3e; tangent max %.3e"
          % (max(e_fd), float(np.median(e_fd)), max(e_tl)))
    assert max(e_tl) <= 1e-10, e_tl
    assert max(e_fd) <= 1e-6, e_fd
