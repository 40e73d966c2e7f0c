"""The YAML run configuration (io/config.py) and the runner's host logic
(scripts/run.py): schema, defaults, validation messages, YAML round trip,
the fixed-step chunking against a step-by-step replay of the reference's
loop, and the CFL formula against its closed form."""

import math

import numpy as np
import pytest
import torch
import yaml

from pygsquig_amd.config import RunConfig, load_config
from pygsquig_amd.run import _steps_to_event, adaptive_timestep, main

FULL = {
    "grid": {"N": 128, "L": 6.283185307179586},
    "solver": {"alpha": 1.0,
               "dissipation": {"type": "hyperviscosity", "nu_p": 1e-12, "p": 8},
               "damping": {"type": "linear_drag", "mu": 0.1, "k_cutoff_factor": 0.5},
               "time_integration": {"method": "RK4", "dt": 5e-4, "adaptive_cfl": False,
                                    "scheme": "if"}},
    "forcing": {"type": "ring", "kf": 8.0, "dk": 2.0, "epsilon": 0.1, "seed": 7},
    "scalars": {"enabled": True,
                "species": [{"name": "dye", "kappa": 1e-3, "initial_condition": "gaussian",
                             "initial_params": {"center": [3.0, 3.0], "width": 0.4}},
                            {"name": "tracer", "initial_condition": "uniform",
                             "initial_params": {"value": 2.0}}]},
    "output": {"fields": ["theta", "velocity", "scalars"], "diagnostics": ["energy_spectrum"]},
    "simulation": {"t_end": 0.01, "output_interval": 0.004, "checkpoint_interval": 0.005,
                   "log_interval": 0.002},
    "initial_condition": {"type": "random", "seed": 3, "amplitude": 0.5},
}


def test_full_config_and_defaults():
    c = RunConfig.from_dict(FULL)
    assert c.grid.N == 128 and c.solver.dissipation.nu_p == 1e-12
    assert c.solver.damping.mu == 0.1 and c.solver.time_integration.scheme == "if"
    assert c.forcing.kf == 8.0 and c.forcing.tau_f == 0.0
    assert [s.name for s in c.scalars.species] == ["dye", "tracer"]
    assert c.scalars.species[1].kappa == 0.0
    d = RunConfig.from_dict({"grid": {}, "solver": {"alpha": 0.5}})
    assert d.grid.N == 256 and abs(d.grid.L - 2 * np.pi) < 1e-15
    assert d.solver.dissipation.p == 8 and d.solver.dissipation.nu_p == 1e-16
    ti = d.solver.time_integration
    assert (ti.method, ti.dt, ti.adaptive_cfl, ti.cfl_safety, ti.scheme) == ("RK4", 0.001, True,
                                                                             0.8, "explicit")
    assert d.forcing is None and d.scalars is None and d.simulation.t_end == 100.0
    assert d.output.fields == ["theta"] and d.initial_condition.type == "random"
    # disabled scalars are dropped, as in the reference
    e = RunConfig.from_dict({"solver": {"alpha": 1}, "scalars": {"enabled": False,
                                                                 "species": [{"name": "a"}]}})
    assert e.scalars is None


@pytest.mark.parametrize("patch,msg", [
    ({"grid": {"N": 63}}, "N must be a positive even integer, got 63"),
    ({"grid": {"L": -1}}, "L must be positive, got -1"),
    ({"solver": {"alpha": 2.5}}, "alpha must be in [-2, 2], got 2.5"),
    ({"solver": {"alpha": 1, "dissipation": {"p": 6}}}, "p must be 2, 4, or 8, got 6"),
    ({"solver": {"alpha": 1, "dissipation": {"type": "spectral"}}}, "Unknown dissipation type"),
    ({"solver": {"alpha": 1, "damping": {"mu": -1}}}, "mu must be non-negative, got -1"),
    ({"solver": {"alpha": 1, "time_integration": {"cfl_safety": 1.0}}},
     "cfl_safety must be in (0, 1), got 1.0"),
    ({"solver": {"alpha": 1, "time_integration": {"method": "Euler"}}},
     "Unknown time integration method: Euler"),
    ({"solver": {"alpha": 1, "time_integration": {"scheme": "etd"}}}, "scheme must be"),
    ({"forcing": {"kf": 0}}, "kf must be positive, got 0"),
    ({"output": {"fields": ["pressure"]}}, "Unknown fields"),
    ({"output": {"diagnostics": ["helicity"]}}, "Unknown diagnostics"),
    ({"simulation": {"t_end": 1, "log_interval": 0}}, "log_interval must be positive, got 0"),
    ({"initial_condition": {"type": "checkpoint"}}, "checkpoint_path required when type='checkpoint'"),
    ({"scalars": {"enabled": True, "species": [{"name": "a"}, {"name": "a"}]}},
     "Duplicate scalar names found"),
    ({"scalars": {"enabled": True, "species": [{"name": "a", "kappa": -1}]}},
     "kappa must be non-negative, got -1"),
])
def test_validation_messages(patch, msg):
    d = {"grid": {}, "solver": {"alpha": 1.0}}
    d.update(patch)
    with pytest.raises(ValueError) as e:
        RunConfig.from_dict(d)
    assert msg in str(e.value)


def test_yaml_round_trip(tmp_path):
    c = RunConfig.from_dict(FULL)
    p = tmp_path / "c.yml"
    c.to_yaml(p)
    assert load_config(p) == c
    raw = yaml.safe_load(p.read_text())
    assert "checkpoint_path" not in raw["initial_condition"]  # None values dropped
    assert raw["scalars"]["species"][0]["initial_params"]["width"] == 0.4


def _reference_loop(t, dt, t_end, every):
    """Steps after which scripts/run.py:332-504 logs / saves / checkpoints,
    replayed one step at a time (fixed dt; the last step shortened)."""
    last = {k: t for k in every}
    events = []
    n = 0
    while t < t_end:
        h = min(dt, t_end - t)
        if h <= 1e-12 * dt:
            break
        t += h
        n += 1
        for k in every:
            if t - last[k] >= every[k]:
                events.append((n, k))
                last[k] = t
    return n, events


def _chunked_loop(t, dt, t_end, every):
    last = {k: t for k in every}
    events, n = [], 0
    while t < t_end:
        k = _steps_to_event(t, dt, t_end, [(last[q], every[q]) for q in every])
        h = dt
        if k == 0:
            h, k = t_end - t, 1
        if h <= 1e-12 * dt:
            break
        for _ in range(k):
            t += h
        n += k
        for q in every:
            if t - last[q] >= every[q]:
                events.append((n, q))
                last[q] = t
    return n, events


@pytest.mark.parametrize("dt,t_end,every", [
    (1e-3, 0.05, {"log": 0.004, "out": 0.01, "ckpt": 0.025}),
    (3e-3, 0.1, {"log": 0.007, "out": 0.013, "ckpt": 1.0}),
    (0.1, 1.0, {"log": 0.1, "out": 0.3, "ckpt": 0.5}),
    (0.07, 1.0, {"log": 0.2, "out": 2.0, "ckpt": 10.0}),  # shortened last step
    (1e-4, 0.02, {"log": 1e-4, "out": 5e-3, "ckpt": 1e-2}),
])
def test_chunking_matches_step_by_step_loop(dt, t_end, every):
    assert _chunked_loop(0.0, dt, t_end, every) == _reference_loop(0.0, dt, t_end, every)


def test_adaptive_timestep_closed_form():
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal((2, 32, 32))
    dx = 2 * np.pi / 32
    want = 0.8 * min(dx / (np.pi * np.abs(u).max()), dx / (np.pi * np.abs(v).max()))
    got = adaptive_timestep(torch.from_numpy(u), torch.from_numpy(v), grid_dx=dx, dt_max=None,
                            cfl_safety=0.8)
    assert abs(got - want) <= 1e-15 * want
    assert adaptive_timestep(torch.from_numpy(u), torch.from_numpy(v), grid_dx=dx,
                             dt_max=1e-3) == 1e-3
    # explicit-diffusion limit 2 safety / (nu k2_max)
    d = adaptive_timestep(torch.from_numpy(u), torch.from_numpy(v), grid_dx=dx, nu=1.0,
                          k2_max=1e4, dt_max=None, cfl_safety=0.5)
    assert d == min(0.5 * 2.0 / 1e4, 0.5 / 0.8 * want)
    z = torch.zeros(4, 4, dtype=torch.float64)
    assert math.isinf(adaptive_timestep(z, z, grid_dx=dx, dt_max=None))


def test_dry_run_validates_without_running(tmp_path):
    p = tmp_path / "c.yml"
    RunConfig.from_dict(FULL).to_yaml(p)
    out = tmp_path / "out"
    assert main([str(p), "--dry-run", "--output-dir", str(out)]) == 0
    assert not out.exists()
    bad = tmp_path / "bad.yml"
    bad.write_text(yaml.safe_dump({"grid": {"N": 7}, "solver": {"alpha": 1}}))
    with pytest.raises(ValueError, match="N must be a positive even integer"):
        main([str(bad), "--dry-run"])
