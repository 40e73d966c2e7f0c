"""Seeded synthetic inputs for the benchmark configurations (host, NumPy).

- ``sin4``: sin(4x) cos(4y) on the ij grid (the unforced configuration).
- ``kolmogorov_53``: isotropic random-phase field with shell energy
  0.5*|k|^(alpha-2)*|theta_hat|^2 ~ k^(-5/3) on kmin <= |k| <= kmax
  (pygsquig/utils/diagnostics.py:21-28 weighting), phases from
  ``numpy.random.default_rng(seed)``, rescaled to the requested rms.
- ``gaussian_blob``: exp(-r^2 / (2 sigma^2)), r measured without periodic
  wrap, as pygsquig/scripts/run.py:152-156 builds scalar fields.
"""

from __future__ import annotations

import numpy as np

SCALAR_CENTRES = ((np.pi, np.pi), (np.pi / 2, 3 * np.pi / 2),
                  (3 * np.pi / 2, np.pi / 2), (3 * np.pi / 2, 3 * np.pi / 2))


def coords(N: int, L: float):
    x1d = np.arange(N) * (L / N)
    return np.meshgrid(x1d, x1d, indexing="ij")


def sin4(N: int, L: float = 2 * np.pi) -> np.ndarray:
    x, y = coords(N, L)
    return np.sin(4 * x) * np.cos(4 * y)


def kolmogorov_53(N: int, seed: int, alpha: float = 1.0, L: float = 2 * np.pi, rms: float = 1.0,
                  kmin: float = 1.0, kmax: float | None = None) -> np.ndarray:
    kmax = N / 3 if kmax is None else kmax
    unit = 2 * np.pi / L
    ki = np.fft.fftfreq(N, d=1.0 / N)
    kk = np.sqrt(ki[:, None] ** 2 + ki[None, :] ** 2)
    inband = (kk >= kmin) & (kk <= kmax)
    kphys = np.where(inband, kk * unit, 1.0)
    power = np.where(inband, kphys ** (-5.0 / 3.0) * kphys ** (2.0 - alpha) / (2 * np.pi * kphys), 0.0)
    phase = np.random.default_rng(seed).uniform(0.0, 2 * np.pi, size=(N, N))
    field = np.fft.ifft2(np.sqrt(power) * np.exp(1j * phase)).real
    r = np.sqrt(np.mean(field**2))
    return field if r == 0 else field * (rms / r)


def gaussian_blob(N: int, center, sigma: float, L: float = 2 * np.pi) -> np.ndarray:
    x, y = coords(N, L)
    return np.exp(-((x - center[0]) ** 2 + (y - center[1]) ** 2) / (2 * sigma**2))
