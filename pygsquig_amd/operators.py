"""Spectral operators (pygsquig/core/operators.py counterpart).

These standalone functions keep the reference's signatures for callers that
use them directly.  The transforms are the library's R2C/C2R kernels; the
velocity inversion runs the fused multiplier + C2R kernels.  The solver's
step never calls these: its RHS is fused into the column/row kernels.
"""

from __future__ import annotations

import torch

from .grid import Grid, fft2, ifft2


def gradient(theta_hat, grid: Grid):
    """core/operators.py:14-34."""
    return ifft2(1j * grid.kx * theta_hat, grid), ifft2(1j * grid.ky * theta_hat, grid)


def laplacian(theta_hat, grid: Grid):
    """core/operators.py:37-49."""
    return -grid.k2 * theta_hat


def _frac_op(grid: Grid, power: float) -> torch.Tensor:
    kmag = torch.sqrt(grid.k2)
    pos = grid.k2 > 0
    safe = torch.where(pos, kmag, torch.ones_like(kmag))
    return torch.where(pos, safe ** power, torch.zeros_like(kmag))


def fractional_laplacian(theta_hat, grid: Grid, alpha: float):
    """core/operators.py:52-72: |k|^alpha, zero at k=0."""
    return _frac_op(grid, alpha) * theta_hat


def perpendicular_gradient(psi_hat, grid: Grid):
    """core/operators.py:75-97: (u, v) = (-d_y psi, d_x psi)."""
    return ifft2(-1j * grid.ky * psi_hat, grid), ifft2(1j * grid.kx * psi_hat, grid)


def jacobian(theta, u, v, grid: Grid):
    """core/operators.py:100-130: dealiased u.grad(theta) in physical space."""
    th_hat = fft2(theta, grid)
    tx, ty = gradient(th_hat, grid)
    j_hat = fft2(u * tx + v * ty, grid) * grid.dealias_mask
    return ifft2(j_hat, grid)


def compute_streamfunction(theta_hat, grid: Grid, alpha: float):
    """core/operators.py:133-155 (no minus sign)."""
    return _frac_op(grid, -alpha) * theta_hat


def velocity_plan(grid: Grid, alpha: float):
    """Shared plan whose multipliers carry |k|^-alpha (velocity, diagnostics)."""
    from .plan import Plan, _PLAN_CACHE_VEL

    key = (grid.N, grid.L, str(grid.device), float(alpha))
    plan = _PLAN_CACHE_VEL.get(key)
    if plan is None:
        plan = Plan(grid, alpha, 0.0, 8)
        _PLAN_CACHE_VEL[key] = plan
    return plan


def compute_velocity_from_theta(theta_hat, grid: Grid, alpha: float):
    """core/operators.py:158-181, fused: |k|^-alpha and the perpendicular
    gradient are applied in the column kernel, the C2R in the row kernel."""
    plan = velocity_plan(grid, alpha)
    th = theta_hat.to(device=grid.device, dtype=torch.complex128).contiguous()
    u = torch.empty((grid.N, grid.N), dtype=torch.float64, device=grid.device)
    v = torch.empty_like(u)
    plan.velocity(th, u, v)
    return u, v


def hyperviscosity(theta_hat, grid: Grid, nu_p: float, p: int):
    """core/operators.py:184-207: -nu_p * k2**(p/2) * theta_hat, [0,0] = 0."""
    d = -nu_p * (grid.k2 ** (p / 2)) * theta_hat
    d = d.clone()
    d[0, 0] = 0.0
    return d
