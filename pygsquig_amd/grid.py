"""Grid and transforms (pygsquig/core/grid.py counterpart).

``make_grid`` computes the wavenumbers and masks on the host with the
reference's operation order (core/grid.py:72-107) and keeps both the host
(NumPy) and device (torch) copies: the host copies feed the plan's per-mode
flag table, the device tensors are the ``Grid`` attributes users read.
``fft2``/``ifft2`` run the library's R2C/C2R kernels and take the
reference's arguments, ``fft2(field)`` and ``ifft2(field_hat)``.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np
import torch

from . import _lib


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.GsqgLibraryError("pygsquig_amd needs a ROCm GPU (torch.cuda is unavailable)")
    return torch.device("cuda", torch.cuda.current_device())


def wavenumbers_1d(N: int, L: float) -> np.ndarray:
    """``jnp.fft.fftfreq(N, d=L/N) * 2 * jnp.pi`` (core/grid.py:94) using
    JAX's fftfreq formula, integer frequencies divided by ``d * n``."""
    dx = L / N
    k = np.empty(N, dtype=np.float64)
    k[: N // 2] = np.arange(0, N // 2, dtype=np.float64)
    k[N // 2:] = np.arange(-(N // 2), 0, dtype=np.float64)
    return k / (dx * N) * 2 * np.pi


def dealias_cutoff(N: int, L: float) -> float:
    """(2/3) * pi * N / L (core/grid.py:102-104)."""
    k_max = np.pi * N / L
    return (2.0 / 3.0) * k_max


@dataclasses.dataclass
class Grid:
    """Same attributes as the reference's Grid (core/grid.py:15-50), as
    device tensors of shape (N, N), axis 0 = x."""

    N: int
    L: float
    x: torch.Tensor
    y: torch.Tensor
    kx: torch.Tensor
    ky: torch.Tensor
    k2: torch.Tensor
    dealias_mask: torch.Tensor
    k1d: np.ndarray = dataclasses.field(repr=False, default=None)

    @property
    def device(self) -> torch.device:
        return self.kx.device

    # Host-side per-mode quantities of the half plane ky in [0, N/2] with the
    # reference's operation order.
    def half_plane_kmag(self) -> np.ndarray:
        k = self.k1d
        kx = k[:, None]
        ky = np.concatenate([k[: self.N // 2], k[self.N // 2: self.N // 2 + 1]])[None, :]
        k2 = kx**2 + ky**2
        return np.sqrt(k2)  # (N, N/2 + 1): [kx index][ky index]


def make_grid(N: int, L: float, device: Optional[torch.device] = None) -> Grid:
    """core/grid.py:72-107.  Odd N raises ValueError (:84-85)."""
    if N % 2 != 0:
        raise ValueError(f"N must be even, got {N}")
    dev = device if device is not None else default_device()
    dx = L / N
    x1d = np.arange(N) * dx
    x, y = np.meshgrid(x1d, x1d, indexing="ij")
    k1d = wavenumbers_1d(N, L)
    kx, ky = np.meshgrid(k1d, k1d, indexing="ij")
    k2 = kx**2 + ky**2
    mask = np.sqrt(k2) < dealias_cutoff(N, L)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return Grid(N, L, t(x), t(y), t(kx), t(ky), t(k2), t(mask), k1d)


def _square_n(shape) -> int:
    if len(shape) != 2 or shape[0] != shape[1]:
        raise ValueError(f"expected a square (N, N) field, got shape {tuple(shape)}")
    return int(shape[0])


def _as_device(a, grid: Optional[Grid], dtype) -> torch.Tensor:
    """``a`` as a contiguous device tensor of ``dtype``: on ``grid``'s
    device when a grid is given, else on the tensor's own GPU (or the
    current GPU for host data)."""
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
    if grid is not None:
        dev = grid.device
        if t.shape != (grid.N, grid.N):
            raise ValueError(f"expected shape {(grid.N, grid.N)}, got {tuple(t.shape)}")
    else:
        _square_n(t.shape)
        dev = t.device if t.is_cuda else default_device()
    return t.to(device=dev, dtype=dtype).contiguous()


def _plan_for(N: int, device: torch.device, grid: Optional[Grid]):
    from .plan import transform_plan, transform_plan_n

    return transform_plan(grid) if grid is not None else transform_plan_n(N, device)


def fft2(field, grid: Optional[Grid] = None) -> torch.Tensor:
    """Unnormalised 2D FFT -> full (N, N) complex128 (core/grid.py:110-124).

    Called as the reference calls it, ``fft2(field)``: N comes from the
    field's shape (a transform does not depend on L) and the transform plan
    is cached per (N, device).  ``grid`` may be passed to pin the device.
    A real field runs one R2C transform; a complex field runs the full c2c
    transform the reference's ``jnp.fft.fft2`` does, as
    F(Re f) + i F(Im f) from two R2C transforms."""
    if torch.is_tensor(field) and field.is_complex() or (
            not torch.is_tensor(field) and np.iscomplexobj(field)):
        z = _as_device(field, grid, torch.complex128)
        return fft2(z.real.contiguous(), grid) + 1j * fft2(z.imag.contiguous(), grid)
    f = _as_device(field, grid, torch.float64)
    N = f.shape[0]
    out = torch.empty((N, N), dtype=torch.complex128, device=f.device)
    _plan_for(N, f.device, grid).fft2(f, out)
    return out


def ifft2(field_hat, grid: Optional[Grid] = None) -> torch.Tensor:
    """``ifft2(field_hat).real`` (core/grid.py:127-138), also for a
    non-Hermitian ``field_hat`` (the C2R runs on its Hermitian part, which is
    exactly what ``.real`` of the c2c inverse keeps).  Called as
    ``ifft2(field_hat)``; ``grid`` is optional as for ``fft2``."""
    X = _as_device(field_hat, grid, torch.complex128)
    N = X.shape[0]
    out = torch.empty((N, N), dtype=torch.float64, device=X.device)
    _plan_for(N, X.device, grid).ifft2(X, out)
    return out
