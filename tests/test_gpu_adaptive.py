"""Adaptive stepping on the GPU: the device CFL extrema against the oracle
(max speed of the velocity, max |theta_hat|, NaN propagation), and an
adaptive run whose dt sequence, replayed by the oracle's RK4, gives the same
state to 1e-12."""

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return pg


@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_cfl_extrema_vs_oracle(pg, alpha):
    from pygsquig_amd.plan import Plan

    N, L = 64, 2 * np.pi
    g = pg.make_grid(N, L)
    X = np.fft.fft2(ics.kolmogorov_53(N, seed=3, alpha=alpha))
    plan = Plan(g, alpha, 0.0, 8)
    vmax, amax = plan.cfl_extrema(torch.from_numpy(X).cuda())
    u, v = R.compute_velocity_from_theta(X, R.make_grid(N, L), alpha)
    assert abs(vmax - np.sqrt(u * u + v * v).max()) <= 1e-12 * vmax
    assert abs(amax - np.abs(X).max()) <= 1e-13 * amax
    Xn = X.copy()
    Xn[3, 5] = np.nan
    _, an = plan.cfl_extrema(torch.from_numpy(Xn).cuda())
    assert np.isnan(an)


def test_adaptive_evolve_replays_on_oracle(pg):
    from pygsquig_amd.adaptive import AdaptivegSQGSolver, CFLConfig

    N, L, alpha, nu, p = 64, 2 * np.pi, 1.0, 1e-12, 8
    g = pg.make_grid(N, L)
    cfg = CFLConfig(cfl_safety=0.8, target_cfl=0.5, dt_min=1e-6, dt_max=0.02, growth_factor=1.2)
    s = AdaptivegSQGSolver(g, alpha, nu, p, cfl_config=cfg)
    theta0 = 0.5 * ics.kolmogorov_53(N, seed=5)
    st = s.initialize(theta0=torch.from_numpy(theta0).cuda())
    res = s.evolve(st, 0.05, save_interval=0.025)
    fin = res["final_state"]
    assert float(fin["time"]) == pytest.approx(0.05, abs=1e-14)
    dts = s.timestepper.dt_history
    assert res["statistics"]["total_steps"] == len(dts)
    assert all(1e-6 <= d <= 0.02 for d in dts)
    # the dt actually used per step, capped at the save and final times
    used, t = [], 0.0
    saves = [0.025, 0.05]
    for d in dts:
        nxt = next(x for x in saves if x > t + 1e-15)
        step = min(d, 0.05 - t, nxt - t)
        used.append(step)
        t += step
    so = R.Solver(R.make_grid(N, L), alpha, nu, p)
    ost = so.initialize(theta0)
    for d in used:
        ost = so.step(ost, d)
    got = fin["theta_hat"].cpu().numpy()
    err = np.linalg.norm(got - ost["theta_hat"]) / np.linalg.norm(ost["theta_hat"])
    print("adaptive steps", len(used), "rel", err)
    assert err < 1e-12
