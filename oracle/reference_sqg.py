"""NumPy FP64 restatement of pygSQuiG's pseudo-spectral gSQG step.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

Data layout follows the reference exactly: every spectrum is a full complex
(N, N) array, C-order, axis 0 = x / kx and axis 1 = y / ky
(``pygsquig/core/grid.py:90,95``); forward transforms are unnormalised c2c
``fft2`` of real fields and inverse transforms are ``ifft2(...).real``
(``core/grid.py:110-138``).  The FFT engine is pocketfft through ``numpy.fft``
(or ``scipy.fft`` with ``workers`` for the timed CPU baseline), the same
algorithm family as jaxlib's CPU FFT.

Pinning (what fixes each function without importing the reference; the tests
named in brackets hold the checks):
- make_grid: closed forms of ``tests/test_grid.py`` (wavenumber order, Nyquist,
  dealias fraction, N=4 mask) [test_oracle_reference.py::test_grid_*].  The last
  ulp of k follows JAX's published ``fftfreq`` (``k / (d*n)``); numpy's
  ``fftfreq`` (``k * (1/(n*d))``) differs in the last bit for ~8 % of entries at
  N >= 256, which can only move ring/damping masks at exactly integer radii.
  That ulp is *parity unpinned* against jaxlib (no jaxlib here).
- fft2/ifft2: brute-force DFT at N <= 16, round trip, Parseval.
- operators: closed forms (sin/cos derivatives, SQG velocity of a single
  shell, alpha=0 and alpha=2 reductions, zero-mean Jacobian, divergence-free
  velocity) and the inviscid invariant <theta * RHS> = 0.
- rk4_step: exact RK4 amplification polynomial on linear modes and 4th-order
  convergence.
- RingForcing: <theta * Re ifft2(F)> = epsilon after rescaling, mask counts,
  |F| = amplitude on the ring, realified DC/Nyquist.
- damping: mode selection and values per ``tests/test_damping.py``.
- scalars: pure diffusion closed form, mass conservation; with source terms,
  the RK4 amplification polynomial of a linear source, a pointwise RK4 of a
  quadratic one, a source held at the step's start time, and the coupled
  sources' explicit update [test_oracle_sources.py].
- diagnostics: single-mode closed forms.
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Optional

import numpy as np

from . import rng as _rng

# ---------------------------------------------------------------------------
# FFT engine (numpy by default; the CPU baseline swaps in scipy.fft + workers)
# ---------------------------------------------------------------------------


class _NumpyFFT:
    @staticmethod
    def fft2(a):
        return np.fft.fft2(a)

    @staticmethod
    def ifft2(a):
        return np.fft.ifft2(a)


_ENGINE = _NumpyFFT()


def set_fft_engine(engine) -> None:
    """Replace the c2c engine (an object with ``fft2``/``ifft2``)."""
    global _ENGINE
    _ENGINE = engine if engine is not None else _NumpyFFT()


# ---------------------------------------------------------------------------
# Grid  (pygsquig/core/grid.py:15-107)
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class Grid:
    N: int
    L: float
    x: np.ndarray
    y: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    k2: np.ndarray
    dealias_mask: np.ndarray


def wavenumbers_1d(N: int, L: float) -> np.ndarray:
    """k in FFT order, ``fftfreq(N, d=L/N) * 2 * pi`` (core/grid.py:94) with
    JAX's fftfreq formula ``k_int / (d * n)``."""
    dx = L / N
    kint = np.empty(N, dtype=np.float64)
    kint[: N // 2] = np.arange(0, N // 2, dtype=np.float64)
    kint[N // 2 :] = np.arange(-(N // 2), 0, dtype=np.float64)
    return kint / (dx * N) * 2 * np.pi


def make_grid(N: int, L: float) -> Grid:
    """core/grid.py:72-107.  Odd N raises ValueError (:84-85)."""
    if N % 2 != 0:
        raise ValueError(f"N must be even, got {N}")
    dx = L / N
    x1d = np.arange(N) * dx
    x, y = np.meshgrid(x1d, x1d, indexing="ij")
    k1d = wavenumbers_1d(N, L)
    kx, ky = np.meshgrid(k1d, k1d, indexing="ij")
    k2 = kx**2 + ky**2
    k_max = np.pi * N / L
    cutoff = (2.0 / 3.0) * k_max
    dealias_mask = np.sqrt(k2) < cutoff
    return Grid(N, L, x, y, kx, ky, k2, dealias_mask)


def fft2(field):
    """core/grid.py:110-124: unnormalised c2c of a real field."""
    return _ENGINE.fft2(np.asarray(field, dtype=np.complex128))


def ifft2(field_hat):
    """core/grid.py:127-138: ``.real`` of the normalised c2c inverse."""
    return _ENGINE.ifft2(field_hat).real


# ---------------------------------------------------------------------------
# Operators  (pygsquig/core/operators.py)
# ---------------------------------------------------------------------------


def gradient(theta_hat, grid: Grid):
    """operators.py:14-34."""
    dx_hat = 1j * grid.kx * theta_hat
    dy_hat = 1j * grid.ky * theta_hat
    return ifft2(dx_hat), ifft2(dy_hat)


def laplacian(theta_hat, grid: Grid):
    """operators.py:37-49."""
    return -grid.k2 * theta_hat


def fractional_laplacian(theta_hat, grid: Grid, alpha: float):
    """operators.py:52-72: |k|^alpha, zero at k=0."""
    kmag = np.sqrt(grid.k2)
    safe = np.where(grid.k2 > 0, kmag, 1.0)
    op = np.where(grid.k2 > 0, safe**alpha, 0.0)
    return op * theta_hat


def perpendicular_gradient(psi_hat, grid: Grid):
    """operators.py:75-97: (u, v) = (-d_y psi, d_x psi)."""
    u_hat = -1j * grid.ky * psi_hat
    v_hat = 1j * grid.kx * psi_hat
    return ifft2(u_hat), ifft2(v_hat)


def jacobian(theta, u, v, grid: Grid):
    """operators.py:100-130: dealiased u.grad(theta), returned in physical
    space.  Note that it re-transforms the physical theta (:117)."""
    th_hat = fft2(theta)
    tx, ty = gradient(th_hat, grid)
    jphys = u * tx + v * ty
    j_hat = fft2(jphys) * grid.dealias_mask
    return ifft2(j_hat)


def inverse_fractional_op(grid: Grid, alpha: float):
    """The |k|^-alpha multiplier of compute_streamfunction (operators.py:149-153)."""
    kmag = np.sqrt(grid.k2)
    safe = np.where(grid.k2 > 0, kmag, 1.0)
    return np.where(grid.k2 > 0, safe ** (-alpha), 0.0)


def compute_streamfunction(theta_hat, grid: Grid, alpha: float):
    """operators.py:133-155 (no minus sign)."""
    return inverse_fractional_op(grid, alpha) * theta_hat


def compute_velocity_from_theta(theta_hat, grid: Grid, alpha: float):
    """operators.py:158-181."""
    return perpendicular_gradient(compute_streamfunction(theta_hat, grid, alpha), grid)


def hyperviscosity(theta_hat, grid: Grid, nu_p: float, p: int):
    """operators.py:184-207: -nu |k|^p theta_hat (k2**(p/2)), [0,0] = 0."""
    d = -nu_p * (grid.k2 ** (p / 2)) * theta_hat
    d = np.array(d, dtype=np.complex128, copy=True)
    d[0, 0] = 0.0
    return d


# ---------------------------------------------------------------------------
# RHS and RK4  (core/solver.py:28-58, core/time_integrator.py:14-36,
#               core/solver_optimized.py:26-62)
# ---------------------------------------------------------------------------


def advection_hat(theta_hat, grid: Grid, alpha: float):
    """The advective part of the RHS, fft2(-J(theta, u, v)) with (u, v) from
    theta_hat itself (solver.py:46-53).  Shared by core_rhs, the optimized
    RHS and the integrating-factor scheme's nonlinear term."""
    theta = ifft2(theta_hat)
    u, v = compute_velocity_from_theta(theta_hat, grid, alpha)
    advection = -jacobian(theta, u, v, grid)
    return fft2(advection)


def core_rhs(theta_hat, grid: Grid, alpha: float, nu_p: float, p: int):
    """solver.py:28-58."""
    return advection_hat(theta_hat, grid, alpha) + hyperviscosity(theta_hat, grid, nu_p, p)


def core_rhs_optimized(theta_hat, grid: Grid, alpha: float, nu_p: float, p: int):
    """solver_optimized.py:26-45 (dissipation without the [0,0] zeroing)."""
    return advection_hat(theta_hat, grid, alpha) + (-nu_p * (grid.k2 ** (p / 2)) * theta_hat)


def rk4_step(state, rhs_fn: Callable, dt: float):
    """time_integrator.py:14-36, same operation order."""
    k1 = rhs_fn(state)
    k2 = rhs_fn(state + 0.5 * dt * k1)
    k3 = rhs_fn(state + 0.5 * dt * k2)
    k4 = rhs_fn(state + dt * k3)
    return state + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6


def rk4_step_multistep_order(state, rhs_fn: Callable, dt: float):
    """solver_optimized.py:48-62: final combine ``(dt / 6.0) * (...)``."""
    k1 = rhs_fn(state)
    k2 = rhs_fn(state + 0.5 * dt * k1)
    k3 = rhs_fn(state + 0.5 * dt * k2)
    k4 = rhs_fn(state + dt * k3)
    return state + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)


def if_rk4_step(state, nonlin_fn: Callable, lin: np.ndarray, dt: float):
    """Integrating-factor RK4 (Lawson) for d/dt y = lin*y + N(y) with a
    diagonal ``lin``.  Not in the reference: the reference integrates the
    hyperviscous term explicitly (operators.py:202 inside every RK stage),
    which is unstable at the production parameters.  Exact for linear modes:
    N == 0 gives y * exp(lin*dt)."""
    e1 = np.exp(lin * (0.5 * dt))
    e2 = np.exp(lin * dt)
    k1 = nonlin_fn(state)
    y2 = e1 * (state + (0.5 * dt) * k1)
    k2 = nonlin_fn(y2)
    y3 = e1 * state + (0.5 * dt) * k2
    k3 = nonlin_fn(y3)
    y4 = e2 * state + dt * (e1 * k3)
    k4 = nonlin_fn(y4)
    acc = e2 * k1
    acc = acc + 2.0 * (e1 * k2)
    acc = acc + 2.0 * (e1 * k3)
    return e2 * state + (dt / 6.0) * (acc + k4)


# ---------------------------------------------------------------------------
# Ring forcing  (pygsquig/forcing/ring_forcing.py)
# ---------------------------------------------------------------------------


def ensure_hermitian_symmetry(field_hat):
    """ring_forcing.py:204-237: only DC and the three Nyquist corners are
    made real; nothing else is symmetrised."""
    f = np.array(field_hat, dtype=np.complex128, copy=True)
    N = f.shape[0]
    f[0, 0] = f[0, 0].real
    if N % 2 == 0:
        h = N // 2
        f[h, 0] = f[h, 0].real
        f[0, h] = f[0, h].real
        f[h, h] = f[h, h].real
    return f


class RingForcing:
    """ring_forcing.py:16-170 with injected phases.

    ``phases`` may be given explicitly to ``__call__``; otherwise they are
    drawn from ``key`` by :func:`oracle.rng.forcing_phases` (the reference
    uses ``jax.random.uniform``, :90, which cannot be reproduced here)."""

    def __init__(self, kf, dk=1.0, epsilon=0.1, tau_f=0.0, amplitude=None):
        self.kf = kf
        self.dk = dk
        self.epsilon = epsilon
        self.tau_f = tau_f
        self.amplitude = np.sqrt(epsilon) if amplitude is None else amplitude
        self.forcing_state = None

    def forcing_mask(self, grid: Grid):
        """:64-76."""
        kmag = np.sqrt(grid.k2)
        return np.abs(kmag - self.kf) <= self.dk / 2

    def raw_forcing(self, grid: Grid, key=None, phases=None, dt=None):
        """F before rescaling: amplitude * mask * exp(i phi), OU-filtered if
        tau_f > 0 (:105-117, a stateful side effect), DC/Nyquist realified."""
        mask = self.forcing_mask(grid)
        if phases is None:
            phases = _rng.forcing_phases(_rng.key_to_u64(key), grid.N)
        f = self.amplitude * mask * np.exp(1j * phases)
        if self.tau_f > 0:
            if self.forcing_state is None:
                self.forcing_state = f
            else:
                decay = np.exp(-dt / self.tau_f)
                self.forcing_state = decay * self.forcing_state + np.sqrt(1 - decay**2) * f
            f = self.forcing_state
        return ensure_hermitian_symmetry(f)

    @staticmethod
    def rate_and_scale(theta_hat, f_hat, epsilon):
        """:152-165: rate = mean(theta * Re ifft2 F); scale = eps/rate if
        |rate| > 1e-10 else 1."""
        rate = np.mean(ifft2(theta_hat) * ifft2(f_hat))
        scale = epsilon / rate if abs(rate) > 1e-10 else 1.0
        return rate, scale

    def __call__(self, theta_hat, key=None, dt=None, grid=None, phases=None):
        f = self.raw_forcing(grid, key=key, phases=phases, dt=dt)
        _, scale = self.rate_and_scale(theta_hat, f, self.epsilon)
        return scale * f


# ---------------------------------------------------------------------------
# Damping  (pygsquig/forcing/damping.py)
# ---------------------------------------------------------------------------


def large_scale_mask(grid: Grid, kf: float):
    return np.sqrt(grid.k2) < kf / 2


def large_scale_damping(theta_hat, grid: Grid, mu: float, kf: float):
    """damping.py:15-42."""
    if mu == 0:
        return np.zeros_like(theta_hat)
    return -mu * theta_hat * large_scale_mask(grid, kf)


class CombinedDamping:
    """damping.py:49-92."""

    def __init__(self, mu=0.0, kf=20.0, nu_p=0.0, p=8):
        if p not in [2, 4, 8]:
            raise ValueError(f"p must be 2, 4, or 8, got {p}")
        self.mu, self.kf, self.nu_p, self.p = mu, kf, nu_p, p

    def __call__(self, theta_hat, grid: Grid):
        return large_scale_damping(theta_hat, grid, self.mu, self.kf) + hyperviscosity(
            theta_hat, grid, self.nu_p, self.p
        )


# ---------------------------------------------------------------------------
# Passive scalars  (pygsquig/scalars/passive_scalar.py:24-103)
# ---------------------------------------------------------------------------


def scalar_advection_hat(scalar_hat, u, v, grid: Grid):
    """fft2(-J(s, u, v)) with a given (frozen) velocity (passive_scalar.py:86-95)."""
    return fft2(-jacobian(ifft2(scalar_hat), u, v, grid))


def scalar_rhs(scalar_hat, u, v, grid: Grid, kappa: float, source_hat=None):
    """passive_scalar.py:63-103: advection + diffusion (+ source_hat)."""
    adv_hat = scalar_advection_hat(scalar_hat, u, v, grid)
    diff = -kappa * grid.k2 * scalar_hat
    rhs = adv_hat + diff
    if source_hat is not None:
        rhs = rhs + source_hat
    return rhs


def scalar_source_hat(scalar_hat, source_fn, grid: Grid, t: float):
    """PassiveScalarEvolver.compute_source (passive_scalar.py:172-190):
    fft2(S(ifft2(s_hat), grid, t))."""
    return fft2(source_fn(ifft2(scalar_hat), grid, t))


def scalar_step(scalar_hat, u, v, grid: Grid, kappa: float, dt: float, source_fn=None,
                t: float = 0.0, scheme: str = "explicit"):
    """PassiveScalarEvolver.step (passive_scalar.py:212-232): RK4 of
    compute_rhs with frozen (u, v), the source at the step's start time t in
    every stage.  scheme "if": Lawson RK4 with the diffusion exact (the
    scheme the fused kernels run for an IF solver's scalars)."""
    def src(y):
        return None if source_fn is None else scalar_source_hat(y, source_fn, grid, t)

    if scheme == "if":
        def nonlin(y):
            out = scalar_advection_hat(y, u, v, grid)
            return out if source_fn is None else out + src(y)

        return if_rk4_step(scalar_hat, nonlin, -kappa * grid.k2, dt)
    return rk4_step(scalar_hat, lambda y: scalar_rhs(y, u, v, grid, kappa, src(y)), dt)


def apply_coupled_sources(scalars: dict, coupled_sources, grid: Grid, t: float, dt: float):
    """MultiSpeciesEvolver._apply_coupled_sources (passive_scalar.py:378-403):
    s_hat += dt * fft2(S), every coupled source seeing the fields as they
    were before any was added."""
    out = dict(scalars)
    phys = {name: ifft2(sh) for name, sh in scalars.items()}
    for source_fn in coupled_sources:
        for name, source in source_fn(phys, grid, t).items():
            if name in out:
                out[name] = out[name] + dt * fft2(source)
    return out


# ---------------------------------------------------------------------------
# Solver  (core/solver.py:61-245, core/solver_with_scalars.py:131-178)
# ---------------------------------------------------------------------------


class Solver:
    """gSQGSolver restated.  ``scheme='explicit'`` is the reference's;
    ``scheme='if'`` integrates the linear terms (solver hyperviscosity and
    the damping object's terms) exactly.

    ``combine='multistep'`` restates gSQGSolverOptimized
    (solver_optimized.py:83-199): its RHS never zeroes the hyperviscosity at
    k = 0 (:42-45), its unforced step combines as ``(dt/6.0)*(...)`` (:62),
    and with forcing or damping it falls back to ``rk4_step`` (:152-164)."""

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8,
                 scheme: str = "explicit", combine: str = "step"):
        if not -2 <= alpha <= 2:
            raise ValueError(f"alpha must be in [-2, 2], got {alpha}")
        if p not in [2, 4, 8]:
            raise ValueError(f"p must be 2, 4, or 8, got {p}")
        self.grid, self.alpha, self.nu_p, self.p = grid, alpha, nu_p, p
        self.scheme = scheme
        self.combine = combine

    def initialize(self, theta0=None):
        """solver.py:96-124 (the seed path uses jax.random and is not restated)."""
        if theta0 is not None:
            th = fft2(theta0)
        else:
            th = np.zeros((self.grid.N, self.grid.N), dtype=np.complex128)
        return {"theta_hat": th, "time": 0.0, "step": 0}

    def compute_rhs(self, theta_hat, forcing=None, damping=None):
        """solver.py:150-180 with per-stage closures."""
        if self.combine == "multistep":
            rhs = core_rhs_optimized(theta_hat, self.grid, self.alpha, self.nu_p, self.p)
        else:
            rhs = core_rhs(theta_hat, self.grid, self.alpha, self.nu_p, self.p)
        if forcing is not None:
            rhs = rhs + forcing(theta_hat)
        if damping is not None:
            rhs = rhs + damping(theta_hat)
        return rhs

    def linear_operator(self, damping: Optional[CombinedDamping] = None):
        """Diagonal linear part used by the IF scheme."""
        g = self.grid
        lin = -self.nu_p * (g.k2 ** (self.p / 2))
        lin[0, 0] = 0.0
        if damping is not None:
            ld = -damping.nu_p * (g.k2 ** (damping.p / 2))
            ld[0, 0] = 0.0
            if damping.mu != 0:
                ld = ld - damping.mu * large_scale_mask(g, damping.kf)
            lin = lin + ld
        return lin

    def step(self, state, dt, forcing=None, damping=None):
        th = state["theta_hat"]
        damp_obj = damping if isinstance(damping, CombinedDamping) else None
        if damp_obj is not None:
            damping = lambda y: damp_obj(y, self.grid)  # noqa: E731
        if self.scheme == "if":
            if damping is not None and damp_obj is None:
                raise ValueError("the IF scheme needs a CombinedDamping object")
            damping = damp_obj
            lin = self.linear_operator(damping)
            g = self.grid

            def nonlin(y):
                r = advection_hat(y, g, self.alpha)
                if forcing is not None:
                    r = r + forcing(y)
                return r

            new = if_rk4_step(th, nonlin, lin, dt)
        else:
            fn = lambda y: self.compute_rhs(y, forcing, damping)  # noqa: E731
            if self.combine == "multistep" and forcing is None and damping is None:
                new = rk4_step_multistep_order(th, fn, dt)
            else:
                new = rk4_step(th, fn, dt)
        return {"theta_hat": new, "time": state["time"] + dt, "step": state["step"] + 1}

    def compute_velocity(self, theta_hat):
        return compute_velocity_from_theta(theta_hat, self.grid, self.alpha)

    def step_scalars(self, scalars: dict, theta_hat_new, dt, kappas: dict):
        """solver_with_scalars.py:163-171 + passive_scalar.py:212-258: scalars
        advance with the velocity of the *new* theta_hat, frozen over their
        four stages."""
        u, v = self.compute_velocity(theta_hat_new)
        out = {}
        g = self.grid
        for name, s_hat in scalars.items():
            kap = kappas.get(name, 0.0)
            if self.scheme == "if":
                out[name] = if_rk4_step(
                    s_hat, lambda s: scalar_advection_hat(s, u, v, g), -kap * g.k2, dt)
            else:
                out[name] = rk4_step(s_hat, lambda s: scalar_rhs(s, u, v, g, kap), dt)
        return out

    def get_diagnostics(self, state):
        """solver.py:216-245."""
        th = state["theta_hat"]
        theta = ifft2(th)
        u, v = self.compute_velocity(th)
        return {
            "kinetic_energy": float(0.5 * np.mean(u**2 + v**2)),
            "enstrophy": float(0.5 * np.mean(theta**2)),
            "max_theta": float(np.max(np.abs(theta))),
            "time": float(state["time"]),
            "step": int(state["step"]),
        }


# ---------------------------------------------------------------------------
# Diagnostics  (pygsquig/utils/diagnostics.py)
# ---------------------------------------------------------------------------


def energy_spectrum(theta_hat, grid: Grid, alpha: float):
    """diagnostics.py:21-84: linear bins over [0, pi*N/L], n_bins =
    min(N//4, 64), energy density 0.5*|k|^(alpha-2)*|theta_hat|^2 with
    [0,0]=0, no 1/N^2, divided by the bin width."""
    kmag = np.sqrt(grid.k2)
    ksafe = np.where(kmag > 0, kmag, 1.0)
    dens = 0.5 * ksafe ** (alpha - 2) * np.abs(theta_hat) ** 2
    dens[0, 0] = 0.0
    dk = 2 * np.pi / grid.L
    k_max = grid.N * dk / 2
    n_bins = min(grid.N // 4, 64)
    edges = np.linspace(0, k_max, n_bins + 1)
    centres = 0.5 * (edges[:-1] + edges[1:])
    e = np.zeros(n_bins)
    for i in range(n_bins):
        m = (kmag >= edges[i]) & (kmag < edges[i + 1])
        e[i] = np.sum(dens[m])
    e /= edges[1:] - edges[:-1]
    return centres, e


def total_energy(theta_hat, grid: Grid, alpha: float):
    """diagnostics.py:170-190."""
    k2s = np.where(grid.k2 > 0, grid.k2, 1.0)
    dens = 0.5 * k2s ** ((alpha - 2) / 2) * np.abs(theta_hat) ** 2
    dens[0, 0] = 0.0
    return float(np.sum(dens) / grid.N**2)


def enstrophy(theta_hat, grid: Grid, alpha: float):
    """diagnostics.py:119-141."""
    q = fractional_laplacian(theta_hat, grid, alpha)
    return float(0.5 * np.mean(np.abs(q) ** 2) * grid.N**2)


def palinstrophy(theta_hat, grid: Grid, alpha: float):
    """diagnostics.py:193-216: sum 0.5 (k2)^(1+alpha) |theta_hat|^2 / N^2, [0,0] = 0."""
    k2s = np.where(grid.k2 > 0, grid.k2, 1.0)
    dens = 0.5 * k2s ** (1 + alpha) * np.abs(theta_hat) ** 2
    dens[0, 0] = 0.0
    return float(np.sum(dens) / grid.N**2)


def energy_flux(theta_hat, forcing_hat, grid: Grid):
    """diagnostics.py:144-167: Re sum conj(theta_hat) F_hat / N^2."""
    if forcing_hat is None:
        return 0.0
    return float(np.real(np.sum(np.conj(theta_hat) * forcing_hat)) / grid.N**2)


def shell_spectrum(theta_hat, grid: Grid, alpha: float):
    """Integer-shell spectrum E(k) = sum over round(|k|/dk) == k of
    0.5*|k|^(alpha-2)*|theta_hat|^2 / N^4 (the closure metric)."""
    kmag = np.sqrt(grid.k2)
    ksafe = np.where(kmag > 0, kmag, 1.0)
    dens = 0.5 * ksafe ** (alpha - 2) * np.abs(theta_hat) ** 2 / grid.N**4
    dens[0, 0] = 0.0
    dk = 2 * np.pi / grid.L
    shell = np.rint(kmag / dk).astype(np.int64)
    nb = grid.N // 2 + 1
    out = np.zeros(nb)
    ok = shell < nb
    np.add.at(out, shell[ok], dens[ok])
    return out
