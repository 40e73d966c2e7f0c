"""CFL controller arithmetic (core/adaptive_timestep.py): configuration
checks with the reference's messages and the timestep rules, against values
computed by hand from their definitions."""

import math

import pytest

from pygsquig_amd.adaptive import (AdaptiveTimestepper, CFLConfig, compute_diffusion_cfl,
                                   compute_timestep, estimate_initial_timestep)


class _G:
    def __init__(self, N, L):
        self.N, self.L = N, L


@pytest.mark.parametrize("kw,msg", [({"cfl_safety": 0}, "cfl_safety must be in (0, 1]"),
                                    ({"target_cfl": 1.5}, "target_cfl must be in (0, 1]"),
                                    ({"dt_min": 1.0, "dt_max": 0.5}, "must be less than dt_max"),
                                    ({"growth_factor": 1.0}, "growth_factor must be > 1"),
                                    ({"shrink_factor": 1.0}, "shrink_factor must be in (0, 1)")])
def test_config_validation(kw, msg):
    with pytest.raises(ValueError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        CFLConfig(**kw)


def test_timestep_rules():
    g = _G(128, 2 * math.pi)
    dx = g.L / g.N
    cfg = CFLConfig(cfl_safety=0.8, dt_min=1e-8, dt_max=1.0, growth_factor=1.1,
                    shrink_factor=0.5)
    # advection-limited, no history
    dt, d = compute_timestep(2.0, g, cfg, 0.0, 8)
    assert dt == pytest.approx(0.8 * dx / 2.0, rel=1e-15)
    assert d["limited_by"] == "diffusion" or d["dt_diff"] == math.inf
    assert d["cfl_adv"] == pytest.approx(0.8)
    # hyperviscous limit dx^(2p) / (nu k_max^(2p-2)), k_max = pi N / L
    nu, p = 1e-10, 4
    kmax = math.pi * g.N / g.L
    assert compute_diffusion_cfl(g, nu, p) == pytest.approx(dx ** 8 / (nu * kmax ** 6), rel=1e-14)
    assert compute_diffusion_cfl(g, 1e-3, 1) == pytest.approx(0.25 * dx ** 2 / 1e-3, rel=1e-14)
    # growth limit and CFL-violation shrink
    dt_g, _ = compute_timestep(0.1, g, cfg, 0.0, 8, current_dt=1e-3)
    assert dt_g == pytest.approx(1.1e-3, rel=1e-14)
    dt_s, _ = compute_timestep(10.0, g, cfg, 0.0, 8, current_dt=dx)  # current CFL = 10 > 1
    assert dt_s == pytest.approx(0.5 * dx, rel=1e-14)
    # clipping to [dt_min, dt_max]
    assert compute_timestep(0.0, g, cfg)[0] == 1.0
    assert estimate_initial_timestep(g, 0.0, 8, 0.5) == pytest.approx(0.5 * dx)


def test_stability_check_and_statistics():
    ts = AdaptiveTimestepper(_G(64, 2 * math.pi))
    assert ts.check_stability(1.0, 5.0) == (True, None)
    assert ts.check_stability(1.0, 20.0)[0] is False
    assert ts.check_stability(1.0, float("nan"))[1] == "NaN or Inf detected"
    ts.compute_timestep({"time": 0.0}, 1.0)
    ts.compute_timestep({"time": 0.1}, 1.0)
    assert ts.suggest_retry_timestep(0.2) == pytest.approx(0.1)
    st = ts.get_statistics()
    assert st["n_steps"] == 2 and st["n_rejected"] == 1
    assert st["efficiency"] == pytest.approx(1 - 1 / 3)
