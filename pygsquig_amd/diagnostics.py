"""Diagnostics: pygsquig/utils/diagnostics.py counterparts on the GPU.

The reductions run in the library's diagnostics kernels
(``gsqg_spectral_diagnostics``): one pass over the full spectrum gives the
binned energy spectrum, the integer-shell spectrum and the total-energy,
enstrophy and palinstrophy sums.  This module adds the reference's bin edges
and final normalisations, in the reference's own operation order, so the
results follow its definitions (including its conventions: no 1/N^2 in the
binned spectrum, enstrophy = 0.5 * sum |k^alpha theta_hat|^2).
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .grid import Grid


def _edges(grid: Grid):
    """utils/diagnostics.py:56-63: n_bins = min(N // 4, 64) linear bins over
    [0, N * dk / 2]."""
    dk = 2 * np.pi / grid.L
    k_max = grid.N * dk / 2
    n_bins = min(grid.N // 4, 64)
    return np.linspace(0, k_max, n_bins + 1), dk


def _sums(theta_hat, grid: Grid, alpha: float):
    from .plan import transform_plan

    X = torch.as_tensor(theta_hat).to(device=grid.device, dtype=torch.complex128).contiguous()
    if X.shape != (grid.N, grid.N):
        raise ValueError(f"theta_hat must have shape ({grid.N}, {grid.N}), got {tuple(X.shape)}")
    edges, dk = _edges(grid)
    bins, shells, sums = transform_plan(grid).spectral_diagnostics(X, float(alpha), edges, dk)
    return bins, shells, sums, edges


def compute_energy_spectrum(theta_hat, grid: Grid, alpha: float):
    """E(k) of utils/diagnostics.py:32-84: (k_bins, E_k) as NumPy arrays."""
    bins, _, _, edges = _sums(theta_hat, grid, alpha)
    k_bins = 0.5 * (edges[:-1] + edges[1:])
    E_k = bins / (edges[1:] - edges[:-1])
    return k_bins, E_k


def compute_total_energy(theta_hat, grid: Grid, alpha: float) -> float:
    """utils/diagnostics.py:170-190: sum 0.5 (k^2)^((alpha-2)/2) |theta_hat|^2 / N^2."""
    return float(_sums(theta_hat, grid, alpha)[2][0] / grid.N**2)


def compute_enstrophy(theta_hat, grid: Grid, alpha: float) -> float:
    """utils/diagnostics.py:119-141: 0.5 * sum |k|^(2 alpha) |theta_hat|^2."""
    return float(0.5 * _sums(theta_hat, grid, alpha)[2][1])


def compute_palinstrophy(theta_hat, grid: Grid, alpha: float) -> float:
    """utils/diagnostics.py:193-216: sum 0.5 (k^2)^(1+alpha) |theta_hat|^2 / N^2."""
    return float(_sums(theta_hat, grid, alpha)[2][2] / grid.N**2)


def compute_shell_spectrum(theta_hat, grid: Grid, alpha: float) -> np.ndarray:
    """Integer-shell spectrum E[s] = sum over rint(|k|/dk) == s of
    0.5 |k|^(alpha-2) |theta_hat|^2 / N^4, s = 0 .. N/2 (no reference
    counterpart; the closure metric compares shells k <= N/3)."""
    return _sums(theta_hat, grid, alpha)[1]


def compute_energy_flux(theta_hat, forcing_hat: Optional[torch.Tensor], grid: Grid) -> float:
    """utils/diagnostics.py:144-167: Re sum conj(theta_hat) F_hat / N^2."""
    if forcing_hat is None:
        return 0.0
    t = torch.as_tensor(theta_hat).to(device=grid.device, dtype=torch.complex128)
    f = torch.as_tensor(forcing_hat).to(device=grid.device, dtype=torch.complex128)
    return float(torch.real(torch.sum(torch.conj(t) * f)) / grid.N**2)
