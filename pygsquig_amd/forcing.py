"""Ring forcing and damping (pygsquig/forcing/ring_forcing.py, damping.py).

Passing a ``RingForcing`` or ``CombinedDamping`` to ``gSQGSolver.step``
lowers it into the fused kernels (plan flags and epilogue terms).  Calling
the objects directly returns the same spectra the reference returns, computed
on the device.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .rng import key_to_u64


class RingForcing:
    """forcing/ring_forcing.py:16-170.

    Phases: the reference draws ``jax.random.uniform(key, (N, N), 0, 2*pi)``
    (:90); this implementation draws them from the SplitMix64 counter hash of
    ``pygsquig_amd.rng`` (same key -> same phases on host and device).  When no
    key is passed to ``step``, step ``n`` uses ``step_key(seed, n)``.

    ``tau_f > 0`` makes the forcing an Ornstein-Uhlenbeck process (:94-117):
    every call updates ``forcing_state`` in place, exactly as the reference's
    calls do, so a solver step (four RK4 stages, each calling the forcing)
    advances it four times, all with that step's key and dt.  The state is a
    device tensor of the ring modes only (2 complex values per entry of the
    ring list: the mode and its mirror); the reference's full (N, N) array is
    zero everywhere else.  Such a forcing is not lowered into the fused
    kernels: the solver calls it per stage.
    """

    def __init__(self, kf: float, dk: float = 1.0, epsilon: float = 0.1, tau_f: float = 0.0,
                 amplitude: Optional[float] = None, seed: int = 42):
        self.kf = kf
        self.dk = dk
        self.epsilon = epsilon
        self.tau_f = tau_f
        self.amplitude = float(np.sqrt(epsilon)) if amplitude is None else float(amplitude)
        self.seed = int(seed)
        if tau_f < 0:
            raise ValueError(f"tau_f must be non-negative, got {tau_f}")
        self.forcing_state = None
        self._state_grid = None

    def signature(self):
        return ("ring", float(self.kf), float(self.dk), float(self.epsilon), float(self.amplitude))

    def _get_forcing_mask(self, grid) -> torch.Tensor:
        """:64-76."""
        kmag = torch.sqrt(grid.k2)
        return torch.abs(kmag - self.kf) <= self.dk / 2

    def __call__(self, theta_hat, key=None, dt=None, grid=None):
        """F scaled so that <theta Re ifft2 F> = epsilon (:119-170)."""
        if grid is None:
            raise TypeError("RingForcing.__call__ needs grid=")
        from .plan import forcing_plan

        k = key_to_u64(key) if key is not None else self.seed
        th = torch.as_tensor(theta_hat).to(device=grid.device, dtype=torch.complex128).contiguous()
        out = torch.zeros_like(th)
        plan = forcing_plan(grid, self)
        if self.tau_f == 0:
            plan.ring_forcing(th, k, out)
            return out
        if dt is None:
            raise TypeError("RingForcing with tau_f > 0 needs dt=")
        gkey = (grid.N, float(grid.L), str(grid.device))
        first = self.forcing_state is None
        if first:
            self.forcing_state = torch.zeros((max(plan.ring_count(), 1), 2),
                                             dtype=torch.complex128, device=grid.device)
            self._state_grid = gkey
        elif self._state_grid != gkey:
            raise ValueError("forcing_state was built on another grid; reset it to None first")
        decay = float(np.exp(-float(dt) / self.tau_f))
        vfac = float(np.sqrt(1.0 - decay ** 2))
        plan.ring_forcing_ou(th, k, decay, vfac, self.forcing_state, first, out)
        return out


def large_scale_damping(theta_hat, grid, mu: float, kf: float):
    """forcing/damping.py:15-42."""
    if mu == 0:
        return torch.zeros_like(theta_hat)
    mask = torch.sqrt(grid.k2) < kf / 2
    return -mu * theta_hat * mask


class CombinedDamping:
    """forcing/damping.py:49-92: -mu theta [|k| < kf/2] - nu_p |k|^p theta."""

    def __init__(self, mu: float = 0.0, kf: float = 20.0, nu_p: float = 0.0, p: int = 8):
        self.mu = mu
        self.kf = kf
        self.nu_p = nu_p
        self.p = p
        if p not in [2, 4, 8]:
            raise ValueError(f"p must be 2, 4, or 8, got {p}")

    def signature(self):
        return ("damp", float(self.mu), float(self.kf), float(self.nu_p), int(self.p))

    def __call__(self, theta_hat, grid=None, **_unused):
        if grid is None:
            raise TypeError("CombinedDamping.__call__ needs grid=")
        from .operators import hyperviscosity

        return large_scale_damping(theta_hat, grid, self.mu, self.kf) + hyperviscosity(
            theta_hat, grid, self.nu_p, self.p)
