"""Run a simulation from a YAML configuration (scripts/run.py).

``python -m pygsquig_amd.run config.yml --output-dir out`` reads the
reference's configuration schema (``pygsquig_amd.config``) and drives this
package's solvers, with the reference runner's outputs, each written as a
NumPy archive because h5py is not part of this stack:

- ``config.yml``: the configuration as loaded;
- ``diagnostics/timeseries.jsonl``: one JSON object per log interval
  (scripts/run.py:206-231: theta_rms, theta_max, energy, enstrophy and the
  solver's get_diagnostics; ``timeseries.npz`` holds the same columns);
- ``fields/fields_<step>.npz``: the fields named in ``output.fields`` every
  ``output_interval`` (scripts/run.py:410-483);
- ``checkpoints/step_<step>.npz`` every ``checkpoint_interval`` and
  ``final_step_<step>.npz`` at the end (``pygsquig_amd.checkpoint``).

Differences from scripts/run.py, all deliberate:

- Damping: the reference builds ``CombinedDamping`` with the solver's own
  hyperviscosity (nu_p, p) (:322-330), so small scales are damped twice.
  Here the damping carries only the linear drag, for |k| < kf *
  k_cutoff_factor (the reference hard-codes kf / 2, i.e. factor 0.5).
- Adaptive steps: the reference passes (dx, dy, cfl_safety, dt_max)
  positionally into ``adaptive_timestep(u, v, grid_dx, nu, k2_max, dt_max,
  cfl_safety)`` (:340-347), so dy lands in nu and cfl_safety in k2_max.
  Here each goes to its own keyword.
- Fixed steps run as chunks of ``multistep`` that end where the reference's
  per-step loop would log, save or checkpoint, so they replay as HIP graphs.
- Forcing phases come from the solver's per-step keys (``RingForcing``,
  seed ``forcing.seed`` or 42), not from a split JAX key.
"""

from __future__ import annotations

import argparse
import json
import logging
import math
import signal
import sys
import time
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from .config import RunConfig, load_config

log = logging.getLogger("pygsquig_amd.run")


def adaptive_timestep(u, v, grid_dx: float, nu: float = 0.0, k2_max: float = 0.0,
                      dt_max: Optional[float] = 0.1, cfl_safety: float = 0.8) -> float:
    """core/time_integrator.py:68-165: safety * min(dx / (pi max|u|),
    dx / (pi max|v|)), the explicit-diffusion limit 2 safety / (nu k2_max)
    when nu and k2_max are positive, and dt_max."""
    u_max = float(torch.amax(torch.abs(torch.as_tensor(u))))
    v_max = float(torch.amax(torch.abs(torch.as_tensor(v))))
    dt_x = grid_dx / (math.pi * u_max) if u_max > 0 else math.inf
    dt_y = grid_dx / (math.pi * v_max) if v_max > 0 else math.inf
    dt = cfl_safety * min(dt_x, dt_y)
    if nu > 0 and k2_max > 0:
        dt = min(dt, cfl_safety * 2.0 / (nu * k2_max))
    if dt_max is not None:
        dt = min(dt, dt_max)
    return float(dt)


def _scalar_init(species, grid, seed: int) -> dict:
    """Initial scalar fields (scripts/run.py:140-190); random fields draw
    from numpy.random.default_rng."""
    out = {}
    N, L = grid.N, grid.L
    for sp in species:
        ic, prm = sp.initial_condition, sp.initial_params
        if ic == "random":
            rs = np.random.default_rng(prm.get("seed", seed + 1))
            f = prm.get("amplitude", 1.0) * rs.standard_normal((N, N))
        elif ic == "uniform":
            f = prm.get("value", 1.0) * np.ones((N, N))
        elif ic == "gaussian":
            x0, y0 = prm.get("center", [L / 2, L / 2])
            w = prm.get("width", 1.0)
            x, y = grid.x.cpu().numpy(), grid.y.cpu().numpy()
            f = np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * w ** 2))
        else:
            f = np.zeros((N, N))
        out[sp.name] = torch.from_numpy(np.ascontiguousarray(f))
    return out


def build_simulation(cfg: RunConfig, checkpoint: Optional[str] = None,
                     device: Optional[torch.device] = None):
    """(grid, solver, state, forcing, damping) of a configuration
    (scripts/run.py:92-202, 301-330)."""
    from .checkpoint import load_checkpoint
    from .forcing import CombinedDamping, RingForcing
    from .grid import make_grid
    from .scalars import gSQGSolverWithScalars
    from .solver import gSQGSolver

    ti = cfg.solver.time_integration
    if ti.method != "RK4":
        raise NotImplementedError(f"time integration method {ti.method!r}: this build runs RK4")
    grid = make_grid(cfg.grid.N, cfg.grid.L, device)
    diss = cfg.solver.dissipation
    species = cfg.scalars.species if cfg.scalars is not None and cfg.scalars.enabled else []
    if species:
        from .sources import source_from_config

        solver = gSQGSolverWithScalars(grid, cfg.solver.alpha, diss.nu_p, diss.p, scheme=ti.scheme,
                                       passive_scalars={sp.name: {
                                           "kappa": sp.kappa,
                                           "source": source_from_config(sp.source)}
                                           for sp in species})
    else:
        solver = gSQGSolver(grid, cfg.solver.alpha, diss.nu_p, diss.p, scheme=ti.scheme)

    forcing = None
    fc = cfg.forcing
    if fc is not None and fc.type == "ring":
        forcing = RingForcing(kf=fc.kf, dk=fc.dk, epsilon=fc.epsilon, tau_f=fc.tau_f,
                              seed=fc.seed if fc.seed else 42)
    damping = None
    dc = cfg.solver.damping
    if dc is not None and dc.type != "none":
        kf = fc.kf if fc is not None else 20.0
        # CombinedDamping drags |k| < kf' / 2; kf' = 2 k_cutoff_factor kf
        damping = CombinedDamping(mu=dc.mu, kf=2.0 * dc.k_cutoff_factor * kf, nu_p=0.0, p=diss.p)

    ic = cfg.initial_condition
    path = checkpoint or (ic.checkpoint_path if ic.type == "checkpoint" else None)
    if path is not None:
        state = load_checkpoint(path, device=grid.device)
        state.pop("config", None)
    elif ic.type in ("random", "zero"):
        seed = ic.seed if ic.seed is not None else 0
        theta0 = None if ic.type == "random" else torch.zeros((grid.N, grid.N), dtype=torch.float64)
        if species:
            state = solver.initialize(theta0, seed=seed if theta0 is None else None,
                                      scalar_init=_scalar_init(species, grid, seed))
        else:
            state = solver.initialize(theta0, seed=seed if theta0 is None else None)
        if ic.amplitude != 1.0:
            state["theta_hat"] = state["theta_hat"] * ic.amplitude
    else:
        raise NotImplementedError(f"Initial condition type '{ic.type}' not implemented")
    return grid, solver, state, forcing, damping


def _steps_to_event(t: float, dt: float, t_end: float, marks) -> int:
    """Number of full steps of dt from t until the first step after which
    the reference's per-step loop acts: an interval (t_last, every) has
    elapsed, or no further full step fits before t_end."""
    k, tk = 0, t
    while True:
        nxt = tk + dt
        if nxt > t_end:
            return k
        k, tk = k + 1, nxt
        if any(tk - last >= every for last, every in marks) or t_end - tk < dt:
            return k


class Simulation:
    """One configured run: ``Simulation(cfg, out_dir).run()`` returns the
    final state."""

    def __init__(self, cfg: RunConfig, output_dir, checkpoint: Optional[str] = None,
                 device: Optional[torch.device] = None):
        self.cfg = cfg
        self.out = Path(output_dir)
        self.grid, self.solver, self.state, self.forcing, self.damping = build_simulation(
            cfg, checkpoint, device)
        self.dirs = {k: self.out / k for k in ("checkpoints", "fields", "diagnostics")}
        self.series: list = []
        self.stop = False

    # ------------------------------------------------------------ outputs
    def _t(self) -> float:
        return float(self.state["time"])

    def _step_no(self) -> int:
        return int(self.state["step"])

    def diagnostics(self, dt: float) -> dict:
        """scripts/run.py:206-231 (theta_rms = sqrt(2 * 0.5 <theta^2>))."""
        from .diagnostics import compute_enstrophy, compute_total_energy

        th, a = self.state["theta_hat"], self.cfg.solver.alpha
        sd = self.solver.get_diagnostics(self.state)
        d = {"time": self._t(), "step": self._step_no(), "dt": dt,
             "theta_rms": math.sqrt(2.0 * sd["enstrophy"]), "theta_max": sd["max_theta"],
             "energy": compute_total_energy(th, self.grid, a),
             "enstrophy": compute_enstrophy(th, self.grid, a)}
        d.update({k: v for k, v in sd.items() if k not in ("time", "step")})
        return d

    def _scalars(self) -> dict:
        ss = self.state.get("scalar_state")
        return dict(ss.scalars) if ss is not None else {}

    def save_fields(self, name: str, fields) -> Path:
        from .grid import ifft2
        from .operators import compute_streamfunction, compute_velocity_from_theta

        g, a = self.grid, self.cfg.solver.alpha
        th = self.state["theta_hat"]
        data = {}
        for f in fields:
            if f == "theta":
                data["theta"] = ifft2(th, g)
            elif f == "streamfunction":
                data["streamfunction"] = ifft2(compute_streamfunction(th, g, a), g)
            elif f == "velocity":
                data["u"], data["v"] = compute_velocity_from_theta(th, g, a)
            elif f == "scalars":
                for n, s in self._scalars().items():
                    data[f"scalar_{n}"] = ifft2(s, g)
            elif f == "vorticity":
                log.warning("Vorticity output not yet implemented")
        arrays = {k: torch.as_tensor(v).detach().cpu().numpy() for k, v in data.items()}
        d = self.cfg.solver.dissipation
        arrays.update(time=np.array(self._t()), step=np.array(self._step_no()),
                      alpha=np.array(float(a)), nu_p=np.array(float(d.nu_p)), p=np.array(int(d.p)))
        path = self.dirs["fields"] / name
        np.savez(path, **arrays)
        return path

    def save_checkpoint(self, name: str) -> Path:
        from .checkpoint import save_checkpoint

        path = self.dirs["checkpoints"] / name
        save_checkpoint(self.state, path, self.cfg.to_dict())
        return path

    # --------------------------------------------------------------- loop
    def _advance(self, n: int, dt: float) -> None:
        st = self.state
        if n == 1:
            self.state = self.solver.step(st, dt, forcing=self.forcing, damping=self.damping)
        else:
            self.state = self.solver.multistep(st, n, dt, forcing=self.forcing, damping=self.damping)

    def run(self) -> dict:
        from .operators import compute_velocity_from_theta

        cfg, sim, ti = self.cfg, self.cfg.simulation, self.cfg.solver.time_integration
        for d in self.dirs.values():
            d.mkdir(parents=True, exist_ok=True)
        cfg.to_yaml(self.out / "config.yml")
        t0_wall = time.time()
        t = self._t()
        last = {"log": t, "out": t, "ckpt": t}
        every = {"log": sim.log_interval, "out": sim.output_interval,
                 "ckpt": sim.checkpoint_interval}
        dx = self.grid.L / self.grid.N
        log.info("run from t=%.4g to t=%.4g (N=%d, alpha=%g, scheme=%s)", t, sim.t_end,
                 self.grid.N, cfg.solver.alpha, ti.scheme)
        with open(self.dirs["diagnostics"] / "timeseries.jsonl", "a") as series:
            while t < sim.t_end and not self.stop:
                if ti.adaptive_cfl:
                    u, v = compute_velocity_from_theta(self.state["theta_hat"], self.grid,
                                                       cfg.solver.alpha)
                    dt = min(adaptive_timestep(u, v, grid_dx=dx, dt_max=ti.dt_max,
                                               cfl_safety=ti.cfl_safety), sim.t_end - t)
                    n = 1
                else:
                    dt = ti.dt
                    n = _steps_to_event(t, dt, sim.t_end, [(last[k], every[k]) for k in last])
                    if n == 0:  # the last, shortened step
                        dt, n = sim.t_end - t, 1
                if dt <= 1e-12 * ti.dt:
                    break
                self._advance(n, dt)
                t = self._t()
                if t - last["log"] >= every["log"]:
                    d = self.diagnostics(dt)
                    self.series.append(d)
                    series.write(json.dumps(d) + "\n")
                    series.flush()
                    log.info("t=%.5g step=%d dt=%.3g E=%.6g theta_max=%.4g", t, d["step"], dt,
                             d["energy"], d["theta_max"])
                    last["log"] = t
                if t - last["out"] >= every["out"]:
                    p = self.save_fields(f"fields_{self._step_no():08d}.npz", cfg.output.fields)
                    log.info("fields saved: %s", p)
                    last["out"] = t
                if t - last["ckpt"] >= every["ckpt"]:
                    p = self.save_checkpoint(f"step_{self._step_no():08d}.npz")
                    log.info("checkpoint saved: %s", p)
                    last["ckpt"] = t
                if sim.wall_time_limit and time.time() - t0_wall > sim.wall_time_limit:
                    log.warning("Wall time limit reached!")
                    self.stop = True
        self.save_checkpoint(f"final_step_{self._step_no():08d}.npz")
        if t - last["out"] > 0:
            self.save_fields(f"fields_final_{self._step_no():08d}.npz", ["theta", "scalars"])
        if self.series:
            keys = sorted({k for d in self.series for k in d})
            np.savez(self.dirs["diagnostics"] / "timeseries.npz",
                     **{k: np.array([d.get(k, np.nan) for d in self.series]) for k in keys})
        log.info("done: t=%.5g step=%d in %.1f s; output in %s", t, self._step_no(),
                 time.time() - t0_wall, self.out)
        return self.state


def main(argv=None) -> int:
    """Command line of scripts/run.py:234-253 (--device selects nothing
    here: the solver runs on the current GPU; "cpu" is refused)."""
    ap = argparse.ArgumentParser(prog="python -m pygsquig_amd.run",
                                 description="Run a gSQG simulation from a YAML configuration.")
    ap.add_argument("config")
    ap.add_argument("--device", choices=["cpu", "gpu", "tpu"], default="gpu")
    ap.add_argument("--checkpoint", default=None, help="checkpoint (.npz) to resume from")
    ap.add_argument("--output-dir", default="./output")
    ap.add_argument("--dry-run", action="store_true", help="validate the configuration only")
    ap.add_argument("--log-level", choices=["DEBUG", "INFO", "WARNING", "ERROR"], default="INFO")
    a = ap.parse_args(argv)
    logging.basicConfig(level=getattr(logging, a.log_level),
                        format="%(asctime)s %(levelname)s %(message)s")
    cfg = load_config(a.config)
    if a.dry_run:
        log.info("Configuration validated successfully: grid %dx%d, L=%g, alpha=%g, t_end=%g",
                 cfg.grid.N, cfg.grid.N, cfg.grid.L, cfg.solver.alpha, cfg.simulation.t_end)
        return 0
    if a.device != "gpu":
        log.error("this build runs on the GPU only (got --device %s)", a.device)
        return 2
    out = Path(a.output_dir)
    out.mkdir(parents=True, exist_ok=True)
    logging.getLogger().addHandler(logging.FileHandler(out / "simulation.log"))
    sim = Simulation(cfg, out, checkpoint=a.checkpoint)
    prev = signal.signal(signal.SIGINT, lambda *_: setattr(sim, "stop", True))
    try:
        sim.run()
    finally:
        signal.signal(signal.SIGINT, prev)
    return 0


if __name__ == "__main__":
    sys.exit(main())
