"""pygsquig_amd: MI355X-native pseudo-spectral gSQG time step.

Drop-in for pygSQuiG's hot path (``make_grid``, ``gSQGSolver.step``,
``gSQGSolverOptimized.multistep``, ``gSQGSolverWithScalars``, ``RingForcing``,
``CombinedDamping``) on PyTorch-ROCm device tensors, with the arithmetic in
hand-written HIP kernels for gfx950 behind the C ABI of include/gsqg.h.
"""

from .grid import Grid, fft2, ifft2, make_grid
from .forcing import CombinedDamping, RingForcing, large_scale_damping
from .operators import (compute_streamfunction, compute_velocity_from_theta,
                        fractional_laplacian, gradient, hyperviscosity, jacobian, laplacian,
                        perpendicular_gradient)
from .solver import gSQGSolver, gSQGSolverOptimized
from .diagnostics import (compute_energy_flux, compute_energy_spectrum, compute_enstrophy,
                          compute_palinstrophy, compute_shell_spectrum, compute_total_energy)
from .ensemble import EnsembleSolver
from .closure import ClosedSolver, ClosureNet
from ._lib import GsqgLibraryError

__all__ = [
    "Grid", "make_grid", "fft2", "ifft2", "RingForcing", "CombinedDamping",
    "large_scale_damping", "gradient", "laplacian", "fractional_laplacian",
    "perpendicular_gradient", "jacobian", "compute_streamfunction",
    "compute_velocity_from_theta", "hyperviscosity", "gSQGSolver", "gSQGSolverOptimized",
    "compute_energy_spectrum", "compute_total_energy", "compute_enstrophy",
    "compute_palinstrophy", "compute_shell_spectrum", "compute_energy_flux", "EnsembleSolver",
    "ClosureNet", "ClosedSolver", "GsqgLibraryError",
]
__version__ = "0.1.0"
