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
    """

    def __init__(self, kf: float, dk: float = 1.0, epsilon: float = 0.1, tau_f: float = 0.0,
                 amplitude: Optional[float] = None, seed: int = 42):
        self.kf = kf
        self.dk = dk
        self.epsilon = epsilon
        self.tau_f = tau_f
        self.amplitude = float(np.sqrt(epsilon)) if amplitude is None else float(amplitude)
        self.seed = int(seed)
        self.forcing_state = None

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
        if self.tau_f != 0:
            raise NotImplementedError("tau_f > 0 (stateful OU forcing) is not supported")
        from .plan import forcing_plan

        k = key_to_u64(key) if key is not None else self.seed
        th = torch.as_tensor(theta_hat).to(device=grid.device, dtype=torch.complex128).contiguous()
        out = torch.zeros_like(th)
        forcing_plan(grid, self).ring_forcing(th, k, out)
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
