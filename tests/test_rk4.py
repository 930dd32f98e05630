"""The 4-stage explicit scheme of the checker (CpuStage.rk4_step over the
oracle's orc_rk_stage), pinned by its closed form: on a linear residual
u = mu (vol/dt) v the stages v_s = v_0 - alpha_s (dt/vol) u(v_{s-1}),
alpha = 1/4, 1/3, 1/2, 1, give v_4 = v_0 (1 - mu + mu^2/2 - mu^3/6 + mu^4/24),
the degree-4 Taylor polynomial of exp(-mu).  A wrong alpha order, a dropped
stage or a flipped sign changes the polynomial."""
import numpy as np
import pytest

import cpu_stage
from codens_amd import configs


@pytest.mark.parametrize("mu", [0.3, -0.7, 1.9])
def test_rk4_step_closed_form(mu):
    case = configs.periodic_euler(16, 8, 4)
    st = cpu_stage.CpuStage(case)
    vol = case.vol

    def time_step():
        st.dt[:] = 0.01 * (1.0 + np.arange(st.dt.size) % 3)

    def residual():
        with np.errstate(divide="ignore", invalid="ignore"):
            st.u[:] = np.where(vol > 0, mu * (vol / st.dt), 0.0) * st.v

    st.thermo = lambda: None
    st.time_step = time_step
    st.residual = residual
    v0 = st.v.copy()
    st.rk4_step()
    inner = case.layout.interior()
    p = 1.0 - mu + mu ** 2 / 2 - mu ** 3 / 6 + mu ** 4 / 24
    np.testing.assert_allclose(st.v[:, inner], v0[:, inner] * p, rtol=4e-15, atol=0)
    ghost = np.setdiff1d(np.arange(case.layout.ip11), inner)
    np.testing.assert_array_equal(st.v[:, ghost], v0[:, ghost])


def test_rk4_stage_matches_numpy():
    """orc_rk_stage on random operands: v0 - (alpha (dt/vol)) u with the
    product and difference rounded separately, real cells only."""
    case = configs.periodic_euler(16, 8, 4)
    st = cpu_stage.CpuStage(case)
    rng = np.random.default_rng(7)
    n = case.layout.ip11
    v0 = rng.standard_normal((5, n)); u = rng.standard_normal((5, n))
    dt = 0.5 + rng.random(n); vol = 0.5 + rng.random(n)
    v = np.zeros((5, n))
    st.O.rk_stage(0, 1.0 / 3.0, v0, u, v, dt, vol)
    inner = case.layout.interior()
    c = (1.0 / 3.0) * (dt[inner] / vol[inner])
    np.testing.assert_array_equal(v[:, inner], v0[:, inner] - c * u[:, inner])
    ghost = np.setdiff1d(np.arange(n), inner)
    assert not v[:, ghost].any()
