"""Adaptive (CFL-controlled) time stepping: core/adaptive_timestep.py and
core/adaptive_solver.py of pygSQuiG on the GPU solver.

The controller arithmetic is the reference's (same configuration checks,
advection and hyperviscous limits, growth and shrink rules, statistics).
The quantities it needs from the state -- max |u| of the velocity and max
|theta_hat| for the growth check -- come from one device reduction per step
(``gsqg_cfl_extrema``); the step itself is the fused device step of
``gSQGSolver``.  Difference: ``compute_timestep`` takes the maximum speed
rather than the (u, v) arrays, since the velocity never leaves the device.
"""

from __future__ import annotations

import math
import time as pytime
from dataclasses import dataclass
from typing import Any, Callable, Optional

import torch

from .grid import Grid
from .solver import gSQGSolver


@dataclass
class CFLConfig:
    """core/adaptive_timestep.py:19-54."""

    cfl_safety: float = 0.8
    dt_min: float = 1e-8
    dt_max: float = 1.0
    growth_factor: float = 1.1
    shrink_factor: float = 0.5
    target_cfl: float = 0.5
    advection_weight: float = 1.0

    def __post_init__(self):
        if not 0 < self.cfl_safety <= 1:
            raise ValueError(f"cfl_safety must be in (0, 1], got {self.cfl_safety}")
        if not 0 < self.target_cfl <= 1:
            raise ValueError(f"target_cfl must be in (0, 1], got {self.target_cfl}")
        if self.dt_min >= self.dt_max:
            raise ValueError(f"dt_min ({self.dt_min}) must be less than dt_max ({self.dt_max})")
        if self.growth_factor <= 1:
            raise ValueError(f"growth_factor must be > 1, got {self.growth_factor}")
        if not 0 < self.shrink_factor < 1:
            raise ValueError(f"shrink_factor must be in (0, 1), got {self.shrink_factor}")


def compute_advection_dt(max_velocity: float, dx: float) -> float:
    """dx / max|u| with the reference's floor of 1e-10 on the speed
    (core/adaptive_timestep.py:73-110)."""
    return dx / max(max_velocity, 1e-10)


def compute_diffusion_cfl(grid: Grid, nu: float, p: int = 2) -> float:
    """core/adaptive_timestep.py:113-148."""
    if nu == 0:
        return math.inf
    dx = grid.L / grid.N
    k_max = math.pi * grid.N / grid.L
    if p == 1:
        return 0.25 * dx**2 / nu
    return dx ** (2 * p) / (nu * k_max ** (2 * p - 2))


def compute_timestep(max_velocity: float, grid: Grid, config: CFLConfig, nu_p: float = 0.0,
                     p: int = 8, current_dt: Optional[float] = None):
    """core/adaptive_timestep.py:151-215 (new dt, diagnostics)."""
    dx = grid.L / grid.N
    dt_adv = compute_advection_dt(max_velocity, dx)
    dt_diff = compute_diffusion_cfl(grid, nu_p, p)
    dt_safe = config.cfl_safety * min(dt_adv, dt_diff)
    if current_dt is not None:
        dt_safe = min(dt_safe, current_dt * config.growth_factor)
        if current_dt / dt_adv > 1.0:  # CFL violation: shrink immediately
            dt_safe = current_dt * config.shrink_factor
    dt_new = min(max(dt_safe, config.dt_min), config.dt_max)
    diagnostics = {
        "dt": dt_new,
        "dt_adv": dt_adv,
        "dt_diff": dt_diff,
        "cfl_adv": dt_new / dt_adv,
        "cfl_diff": dt_new / dt_diff if dt_diff < math.inf else 0.0,
        "max_velocity": max_velocity,
        "limited_by": "advection" if dt_adv < dt_diff else "diffusion",
    }
    return dt_new, diagnostics


def estimate_initial_timestep(grid: Grid, nu_p: float = 0.0, p: int = 8,
                              target_cfl: float = 0.5) -> float:
    """core/adaptive_timestep.py:346-377."""
    dx = grid.L / grid.N
    dt_adv = target_cfl * dx / 1.0
    if nu_p > 0:
        return min(dt_adv, compute_diffusion_cfl(grid, nu_p, p) * target_cfl)
    return dt_adv


class AdaptiveTimestepper:
    """core/adaptive_timestep.py:218-343."""

    def __init__(self, grid: Grid, config: Optional[CFLConfig] = None, verbose: bool = False):
        self.grid = grid
        self.config = config or CFLConfig()
        self.verbose = verbose
        self.dt_history: list = []
        self.cfl_history: list = []
        self.time_history: list = []
        self.n_steps = 0
        self.n_rejected = 0
        self.n_dt_changes = 0

    def compute_timestep(self, state: dict, max_velocity: float, nu_p: float = 0.0, p: int = 8):
        current_dt = self.dt_history[-1] if self.dt_history else None
        dt_new, diags = compute_timestep(max_velocity, self.grid, self.config, nu_p, p,
                                         current_dt)
        if current_dt is not None:
            ratio = dt_new / current_dt
            if abs(ratio - 1.0) > 0.1:
                self.n_dt_changes += 1
                if self.verbose:
                    print(f"Timestep change: {current_dt:.3e} -> {dt_new:.3e} (ratio {ratio:.2f})")
        self.dt_history.append(dt_new)
        self.cfl_history.append(diags["cfl_adv"])
        self.time_history.append(float(state.get("time", 0.0)))
        self.n_steps += 1
        return dt_new, diags

    @staticmethod
    def check_stability(max_abs_before: float, max_abs_after: float):
        """core/adaptive_timestep.py:245-275 on the two states' max |theta_hat|
        (NaN propagates into the maximum)."""
        if not math.isfinite(max_abs_after):
            return False, "NaN or Inf detected"
        growth = max_abs_after / (max_abs_before + 1e-10)
        if growth > 10.0:
            return False, f"Excessive growth: {growth:.2f}x"
        return True, None

    def suggest_retry_timestep(self, dt_failed: float) -> float:
        self.n_rejected += 1
        dt_retry = dt_failed * self.config.shrink_factor
        if self.verbose:
            print(f"Timestep rejected, retrying with dt={dt_retry:.3e}")
        return max(dt_retry, self.config.dt_min)

    def get_statistics(self) -> dict:
        if not self.dt_history:
            return {}
        dts, cfls = self.dt_history, self.cfl_history
        return {
            "n_steps": self.n_steps,
            "n_rejected": self.n_rejected,
            "n_dt_changes": self.n_dt_changes,
            "dt_min_used": min(dts),
            "dt_max_used": max(dts),
            "dt_mean": sum(dts) / len(dts),
            "cfl_mean": sum(cfls) / len(cfls),
            "cfl_max": max(cfls),
            "efficiency": 1.0 - self.n_rejected / (self.n_steps + self.n_rejected),
        }

    def reset_history(self):
        self.dt_history.clear()
        self.cfl_history.clear()
        self.time_history.clear()
        self.n_steps = self.n_rejected = self.n_dt_changes = 0


class AdaptivegSQGSolver:
    """core/adaptive_solver.py:22-360 on the device solver.  ``scheme`` is
    passed to gSQGSolver ("explicit" as in the reference, or "if")."""

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8,
                 cfl_config: Optional[CFLConfig] = None, verbose: bool = False,
                 scheme: str = "explicit"):
        self.base_solver = gSQGSolver(grid, alpha, nu_p, p, scheme=scheme)
        self.grid, self.alpha, self.nu_p, self.p = grid, alpha, nu_p, p
        self.verbose = verbose
        self.timestepper = AdaptiveTimestepper(grid, cfl_config, verbose=verbose)
        self.current_dt = estimate_initial_timestep(
            grid, nu_p, p, target_cfl=cfl_config.target_cfl if cfl_config else 0.5)

    def initialize(self, **kwargs) -> dict:
        return self.base_solver.initialize(**kwargs)

    def _extrema(self, theta_hat):
        from .solver import _to_device_c128

        th = _to_device_c128(theta_hat, self.grid)
        return self.base_solver._plan(None, None).cfl_extrema(th)

    def step(self, state: dict, forcing: Optional[Callable] = None,
             damping: Optional[Callable] = None, dt_max: Optional[float] = None, **kwargs):
        """core/adaptive_solver.py:81-165: (new state, step info)."""
        vmax, amax = self._extrema(state["theta_hat"])
        dt, diags = self.timestepper.compute_timestep(state, vmax, self.nu_p, self.p)
        if dt_max is not None:
            dt = min(dt, dt_max)
        try:
            new = self.base_solver.step(state, dt, forcing=forcing, damping=damping, **kwargs)
            ok, reason = self.timestepper.check_stability(amax, self._extrema(new["theta_hat"])[1])
            if not ok:
                if self.verbose:
                    print(f"Step rejected: {reason}")
                dt = self.timestepper.suggest_retry_timestep(dt)
                new = self.base_solver.step(state, dt, forcing=forcing, damping=damping, **kwargs)
                diags["retry"] = True
        except (RuntimeError, ValueError) as e:
            if self.verbose:
                print(f"Step failed with error: {e}")
            dt = self.timestepper.suggest_retry_timestep(dt)
            new = self.base_solver.step(state, dt, forcing=forcing, damping=damping, **kwargs)
            diags["retry"] = True
            diags["error"] = str(e)
        self.current_dt = dt
        return new, {"dt_used": dt, "dt_diagnostics": diags, "time": new["time"],
                     "step": new["step"]}

    def evolve(self, state: dict, t_final: float, forcing=None, damping=None,
               save_interval: Optional[float] = None, callback: Optional[Callable] = None,
               max_steps: int = 100000, min_progress_rate: float = 1e-6, **kwargs) -> dict:
        """core/adaptive_solver.py:167-318."""
        results: dict[str, Any] = {"times": [], "states": [], "dt_history": [], "diagnostics": []}
        wall0 = pytime.perf_counter()
        t0 = float(state["time"])
        last_check = wall0
        if save_interval is not None:
            results["times"].append(float(state["time"]))
            results["states"].append(dict(state))
            next_save = float(state["time"]) + save_interval
        else:
            next_save = t_final + 1
        cur, steps, last_dt = state, 0, None
        while float(cur["time"]) < t_final:
            if steps >= max_steps:
                raise RuntimeError(f"Maximum steps ({max_steps}) reached. Current time: "
                                   f"{float(cur['time']):.6f}, Target: {t_final:.6f}, "
                                   f"Last dt: {last_dt if last_dt is not None else 'N/A'}")
            now = pytime.perf_counter()
            if now - last_check > 1.0:
                rate = (float(cur["time"]) - t0) / (now - wall0)
                if now > wall0 and rate < min_progress_rate:
                    raise RuntimeError(f"Progress too slow. Rate: {rate:.2e} < "
                                       f"{min_progress_rate:.2e}.")
                last_check = now
            dt_max = t_final - float(cur["time"])
            if save_interval is not None:
                dt_max = min(dt_max, next_save - float(cur["time"]))
            new, info = self.step(cur, forcing=forcing, damping=damping, dt_max=dt_max, **kwargs)
            if float(new["time"]) <= float(cur["time"]):
                raise RuntimeError(f"Time not advancing: dt={info['dt_used']}, "
                                   f"time={float(cur['time'])}")
            cur, steps, last_dt = new, steps + 1, info["dt_used"]
            if save_interval is not None and float(cur["time"]) >= next_save:
                results["times"].append(float(cur["time"]))
                results["states"].append(dict(cur))
                results["dt_history"].append(info["dt_used"])
                results["diagnostics"].append(info["dt_diagnostics"])
                next_save += save_interval
            if callback is not None:
                callback(cur, info)
        if save_interval is not None and float(cur["time"]) > results["times"][-1]:
            results["times"].append(float(cur["time"]))
            results["states"].append(dict(cur))
        wall = pytime.perf_counter() - wall0
        stats = self.timestepper.get_statistics()
        stats.update({"total_steps": steps, "wall_time": wall,
                      "simulated_time": float(cur["time"]) - t0})
        stats["time_ratio"] = stats["simulated_time"] / wall if wall > 0 else math.inf
        results["statistics"] = stats
        results["final_state"] = cur
        return results

    def get_diagnostics(self, state: dict) -> dict:
        """core/adaptive_solver.py:320-345."""
        diags = self.base_solver.get_diagnostics(state)
        vmax, _ = self._extrema(state["theta_hat"])
        current = self.timestepper.dt_history[-1] if self.timestepper.dt_history else None
        _, dd = compute_timestep(vmax, self.grid, self.timestepper.config, self.nu_p, self.p,
                                 current)
        diags.update({f"cfl_{k}": v for k, v in dd.items() if k.startswith("cfl_")})
        diags["current_dt"] = self.current_dt
        diags["max_velocity"] = vmax
        return diags
