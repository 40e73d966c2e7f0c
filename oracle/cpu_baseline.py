"""Timed CPU baseline: the oracle's reference-structured step on the host.

TEST/BENCH INFRASTRUCTURE ONLY (see ``oracle/__init__``).  ``bench.py``
calls this as its ``cpu_baseline`` leg.  The path has the reference's
structure -- full complex (N, N) arrays, 9 c2c FFTs per RHS plus 2 per
forcing call, 7 per scalar RHS (pygsquig/core/solver.py:28-58,
forcing/ring_forcing.py:153-155, scalars/passive_scalar.py:64-103) -- with
``scipy.fft`` c2c transforms on ``workers`` threads.  JAX (the reference's
engine) is not installed here, so this NumPy/SciPy restatement is the
stand-in ("port").
"""

from __future__ import annotations

import os
import time

import numpy as np

from . import ics
from . import reference_sqg as R
from . import rng


class ScipyFFT:
    def __init__(self, workers: int):
        import scipy.fft as sf

        self.sf = sf
        self.workers = workers

    def fft2(self, a):
        return self.sf.fft2(a, workers=self.workers)

    def ifft2(self, a):
        return self.sf.ifft2(a, workers=self.workers)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_forced_scalars(N: int = 1024, n_steps: int = 3, workers: int | None = None,
                        n_scalars: int = 2, dt: float = 5e-4, scheme: str = "if",
                        budget_s: float = 30.0) -> dict:
    """Seconds per RK4 step of the forced, damped configuration with passive
    scalars (theta + scalars), after one untimed warm-up step.  Stops early
    once ``budget_s`` is spent."""
    workers = workers or host_cores()
    R.set_fft_engine(ScipyFFT(workers))
    try:
        L = 2 * np.pi
        g = R.make_grid(N, L)
        s = R.Solver(g, 1.0, 0.0, 8, scheme=scheme)
        f = R.RingForcing(kf=40, dk=2, epsilon=0.1)
        d = R.CombinedDamping(mu=0.1, kf=40, nu_p=1e-8 * (512 / N) ** 8, p=8)
        st = s.initialize(ics.kolmogorov_53(N, seed=2))
        sc = {f"s{i}": R.fft2(ics.gaussian_blob(N, c, L / 16))
              for i, c in enumerate(ics.SCALAR_CENTRES[:n_scalars])}
        kap = {k: 0.0 for k in sc}

        def one(n):
            nonlocal st, sc
            key = rng.step_key(42, n)
            st = s.step(st, dt, forcing=lambda y: f(y, key=key, grid=g), damping=d)
            if sc:
                sc = s.step_scalars(sc, st["theta_hat"], dt, kap)

        one(0)
        t0 = time.perf_counter()
        done = 0
        for n in range(1, n_steps + 1):
            one(n)
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        sec = (time.perf_counter() - t0) / done
        return {"sec_per_step": sec, "steps": done, "workers": workers, "N": N,
                "n_scalars": n_scalars, "scheme": scheme, "cpu": cpu_model()}
    finally:
        R.set_fft_engine(None)


def time_plain(N: int = 256, n_steps: int = 100, workers: int | None = None,
               nu: float = 1e-4, dt: float = 1e-3, budget_s: float = 20.0) -> dict:
    """Seconds per RK4 step of the plain unforced configuration (theta0 =
    sin(4x) cos(4y), IF scheme as the GPU run), after one warm-up step."""
    workers = workers or host_cores()
    R.set_fft_engine(ScipyFFT(workers))
    try:
        g = R.make_grid(N, 2 * np.pi)
        s = R.Solver(g, 1.0, nu, 8, scheme="if")
        st = s.step(s.initialize(ics.sin4(N)), dt)
        t0 = time.perf_counter()
        done = 0
        for _ in range(n_steps):
            st = s.step(st, dt)
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        return {"sec_per_step": (time.perf_counter() - t0) / done, "steps": done,
                "workers": workers, "N": N, "cpu": cpu_model()}
    finally:
        R.set_fft_engine(None)


def time_large_rhs(N: int = 8192, workers: int | None = None, n_scalars: int = 4,
                   alpha: float = 0.5) -> dict:
    """One RK4 step of the large configuration is 4 stages of (theta RHS with
    ring forcing and damping) plus the scalar step (4 stages of n_scalars
    scalar RHS) plus one velocity.  A whole step takes minutes on the host, so
    this times one evaluation of each kind and reports their per-step sum:
    4 (t_theta + n_scalars t_scalar) + t_velocity."""
    workers = workers or host_cores()
    R.set_fft_engine(ScipyFFT(workers))
    try:
        L = 2 * np.pi
        g = R.make_grid(N, L)
        s = R.Solver(g, alpha, 0.0, 8, scheme="if")
        f = R.RingForcing(kf=100, dk=2, epsilon=0.1)
        d = R.CombinedDamping(mu=0.1, kf=100, nu_p=1e-8 * (512 / N) ** 8, p=8)
        th = s.initialize(ics.kolmogorov_53(N, seed=2, alpha=alpha))["theta_hat"]
        sh = R.fft2(ics.gaussian_blob(N, ics.SCALAR_CENTRES[0], L / 16))
        key = rng.step_key(42, 1)
        t0 = time.perf_counter()
        rhs = R.core_rhs(th, g, alpha, 0.0, 8) + f(th, key=key, grid=g) + d(th, g)
        t_theta = time.perf_counter() - t0
        t0 = time.perf_counter()
        u, v = s.compute_velocity(th)
        t_vel = time.perf_counter() - t0
        t0 = time.perf_counter()
        rs = R.scalar_rhs(sh, u, v, g, 0.0)
        t_scalar = time.perf_counter() - t0
        del rhs, rs
        sec = 4 * (t_theta + n_scalars * t_scalar) + t_vel
        return {"sec_per_step": sec, "t_theta_rhs": t_theta, "t_scalar_rhs": t_scalar,
                "t_velocity": t_vel, "workers": workers, "N": N, "cpu": cpu_model()}
    finally:
        R.set_fft_engine(None)


def time_batch(members, n_steps: int = 2, workers: int | None = None, nu: float = 1e-4,
               dt: float = 1e-3, budget_s: float = 25.0) -> dict:
    """Seconds per RK4 step of looping the oracle solver over the members
    (N, alpha) of the uneven ensemble (IF, as the GPU run), after one
    warm-up step of every member."""
    workers = workers or host_cores()
    R.set_fft_engine(ScipyFFT(workers))
    try:
        grids, solvers, states = {}, [], []
        for m, (N, a) in enumerate(members):
            g = grids.setdefault(N, R.make_grid(N, 2 * np.pi))
            s = R.Solver(g, a, nu, 8, scheme="if")
            solvers.append(s)
            states.append(s.step(s.initialize(ics.kolmogorov_53(N, seed=100 + m, alpha=a)), dt))
        t0 = time.perf_counter()
        done = 0
        for _ in range(n_steps):
            states = [s.step(st, dt) for s, st in zip(solvers, states)]
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        return {"sec_per_step": (time.perf_counter() - t0) / done, "steps": done,
                "workers": workers, "members": len(members), "cpu": cpu_model()}
    finally:
        R.set_fft_engine(None)
