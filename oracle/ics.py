"""Seeded initial conditions shared by the oracle and the GPU runs.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).  The same generators are
re-implemented for the product in ``pygsquig_amd/ics.py`` (they are input
data, generated once on the host and injected into both paths); the tests
check the two agree bit for bit.

- ``sin4``: theta0 = sin(4x) cos(4y) on the ij grid x_i = i*L/N.
- ``kolmogorov_53``: random-phase isotropic field whose shell energy under the
  reference's weight 0.5*|k|^(alpha-2)*|theta_hat|^2
  (``pygsquig/utils/diagnostics.py:21-28``) follows k^(-5/3) on
  1 <= |k| <= N/3; phases uniform from ``numpy.random.default_rng(seed)``
  (PCG64); Hermitian by construction (real field); rescaled to rms theta = 1.
- ``gaussian_blob``: exp(-r^2 / (2 sigma^2)) with a non-periodic r, as
  ``pygsquig/scripts/run.py:152-156`` builds scalar initial conditions.
"""

from __future__ import annotations

import numpy as np


def grid_coords(N: int, L: float):
    x1d = np.arange(N) * (L / N)
    return np.meshgrid(x1d, x1d, indexing="ij")


def sin4(N: int, L: float = 2 * np.pi):
    x, y = grid_coords(N, L)
    return np.sin(4 * x) * np.cos(4 * y)


def kolmogorov_53(N: int, seed: int, alpha: float = 1.0, L: float = 2 * np.pi,
                  rms: float = 1.0, kmin: float = 1.0, kmax: float | None = None):
    """Physical (N, N) float64 field."""
    if kmax is None:
        kmax = N / 3
    dk = 2 * np.pi / L
    kint = np.fft.fftfreq(N, d=1.0 / N)
    kxi, kyi = np.meshgrid(kint, kint, indexing="ij")
    kmag = np.sqrt(kxi**2 + kyi**2)  # in units of dk
    band = (kmag >= kmin) & (kmag <= kmax)
    ks = np.where(band, kmag * dk, 1.0)
    # |theta_hat|^2 ~ E(k) * k^(2-alpha) / (2 pi k) with E ~ k^-5/3
    amp2 = np.where(band, ks ** (-5.0 / 3.0) * ks ** (2.0 - alpha) / (2 * np.pi * ks), 0.0)
    gen = np.random.default_rng(seed)
    phases = gen.uniform(0.0, 2 * np.pi, size=(N, N))
    spec = np.sqrt(amp2) * np.exp(1j * phases)
    theta = np.fft.ifft2(spec).real  # real part = Hermitian projection
    r = np.sqrt(np.mean(theta**2))
    if r == 0:
        return theta
    return theta * (rms / r)


def gaussian_blob(N: int, center, sigma: float, L: float = 2 * np.pi):
    x, y = grid_coords(N, L)
    x0, y0 = center
    r2 = (x - x0) ** 2 + (y - y0) ** 2
    return np.exp(-r2 / (2 * sigma**2))


SCALAR_CENTRES = [(np.pi, np.pi), (np.pi / 2, 3 * np.pi / 2),
                  (3 * np.pi / 2, np.pi / 2), (3 * np.pi / 2, 3 * np.pi / 2)]
