"""Pins of the oracle's passive-scalar source terms and the scalar step that
calls them (pygsquig/scalars/passive_scalar.py:63-103,172-258,378-403 and
scalars/source_terms.py), each by a closed form that a plausible mistake
fails, with the mistake planted beside it."""

import numpy as np
import pytest

from oracle import reference_sqg as R
from oracle import sources as S

TWO_PI = 2 * np.pi


def _field(N, seed):
    rs = np.random.default_rng(seed)
    g = R.make_grid(N, TWO_PI)
    sh = R.fft2(rs.standard_normal((N, N)))
    return g, sh * (np.sqrt(g.k2) < N / 4)


def _rk4_poly(z):
    return 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24


@pytest.mark.parametrize("kappa", [0.0, 0.05])
def test_linear_source_gives_rk4_polynomial(kappa):
    """u = v = 0 and S = lambda s: every mode is multiplied by the RK4
    polynomial R((lambda - kappa k^2) dt)."""
    N, dt, lam = 32, 0.1, -0.7
    g, sh = _field(N, 1)
    z = np.zeros((N, N))
    out = R.scalar_step(sh, z, z, g, kappa, dt, S.ExponentialGrowth(lam), t=0.3)
    want = _rk4_poly((lam - kappa * g.k2) * dt) * sh
    assert np.abs(out - want).max() < 1e-12 * np.abs(sh).max()
    # planted: the source dropped from stages 2-4 (only k1 carries it)
    bad = sh + dt * (lam * sh + 0 * sh)
    assert np.abs(bad - want).max() > 1e-3 * np.abs(sh).max()


def test_linear_source_if_scheme_is_exact_diffusion_times_rk4():
    """IF: the diffusion integrates exactly, the source by RK4:
    exp(-kappa k^2 dt) R(lambda dt) per mode."""
    N, dt, lam, kappa = 32, 0.1, 0.4, 0.05
    g, sh = _field(N, 2)
    z = np.zeros((N, N))
    out = R.scalar_step(sh, z, z, g, kappa, dt, S.ExponentialGrowth(lam), scheme="if")
    want = np.exp(-kappa * g.k2 * dt) * _rk4_poly(lam * dt) * sh
    assert np.abs(out - want).max() < 1e-12 * np.abs(sh).max()


def test_quadratic_source_is_pointwise_rk4():
    """u = v = 0, kappa = 0, S = -k s^2: the step is RK4 of ds/dt = -k s^2
    at every grid point (ifft2(fft2(f)) = f for a real f)."""
    N, dt, k = 32, 0.05, 0.8
    g, sh = _field(N, 3)
    z = np.zeros((N, N))
    out = R.ifft2(R.scalar_step(sh, z, z, g, 0.0, dt, S.ChemicalReaction(k)))
    s = R.ifft2(sh)
    f = lambda y: -k * y * y  # noqa: E731
    k1 = f(s)
    k2 = f(s + 0.5 * dt * k1)
    k3 = f(s + 0.5 * dt * k2)
    k4 = f(s + dt * k3)
    want = s + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6
    assert np.abs(out - want).max() < 1e-12
    # planted: the sign of the reaction flipped
    bad = R.ifft2(R.scalar_step(sh, z, z, g, 0.0, dt, S.ChemicalReaction(-k)))
    assert np.abs(bad - want).max() > 1e-3


def test_threshold_switches_the_reaction_off():
    N = 16
    g, sh = _field(N, 4)
    s = R.ifft2(sh)
    src = S.ChemicalReaction(1.3, threshold=0.1)(s, g, 0.0)
    assert np.all(src[s <= 0.1] == 0.0)
    np.testing.assert_array_equal(src[s > 0.1], -1.3 * s[s > 0.1] ** 2)


def test_source_is_held_at_the_step_start_time():
    """u = v = 0, kappa = 0, S(t) = A sin(w t + phi): the right-hand side is
    the same in all four stages, so the step adds exactly dt S(t0)."""
    N, dt, t0 = 16, 0.2, 0.7
    g, sh = _field(N, 5)
    z = np.zeros((N, N))
    src = S.TimePeriodicSource(1.5, 3.0, 0.25, spatial_pattern=lambda gr: np.cos(gr.x))
    out = R.scalar_step(sh, z, z, g, 0.0, dt, src, t=t0)
    want = sh + dt * R.fft2(1.5 * np.sin(3.0 * t0 + 0.25) * np.cos(g.x))
    assert np.abs(out - want).max() < 1e-10
    # planted: the source taken at the stage times t0, t0 + dt/2, t0 + dt
    def staged(t):
        return R.fft2(1.5 * np.sin(3.0 * t + 0.25) * np.cos(g.x))
    bad = sh + dt * (staged(t0) + 4 * staged(t0 + dt / 2) + staged(t0 + dt)) / 6
    assert np.abs(bad - want).max() > 1e-3


def test_localized_source_periodic_distance():
    """A Gaussian at x0 near the box edge wraps: its value at x = 2 pi - d
    equals that at x = d around x0 = 0."""
    N = 64
    g = R.make_grid(N, TWO_PI)
    src = S.LocalizedSource(2.0, 0.0, np.pi, 0.5)(np.zeros((N, N)), g, 0.0)
    j = N // 2  # y = pi
    for i in range(1, 8):
        assert abs(src[i, j] - src[N - i, j]) < 1e-14
    assert abs(src[0, j] - 2.0) < 1e-15
    r2 = (g.x[3, j]) ** 2
    assert abs(src[3, j] - 2.0 * np.exp(-r2 / 0.25)) < 1e-15


def test_linear_combination_and_constant_source():
    """LinearCombination sums weighted sources; a constant source adds
    exactly dt S with no velocity and no diffusion."""
    N, dt = 16, 0.1
    g, sh = _field(N, 6)
    z = np.zeros((N, N))
    loc = S.LocalizedSource(1.0, 1.0, 2.0, 0.7)
    comb = S.LinearCombination([(2.0, loc), (-0.5, loc)])
    out = R.scalar_step(sh, z, z, g, 0.0, dt, comb)
    want = sh + dt * R.fft2(1.5 * loc(None, g, 0.0))
    assert np.abs(out - want).max() < 1e-10


def test_coupled_sources_explicit_update():
    """A -> B conversion at rate k: A += -dt k A, B += dt k A, both from the
    fields before the update."""
    N, dt, k = 16, 0.1, 0.6
    g, a = _field(N, 7)
    _, b = _field(N, 8)

    def conv(phys, grid, t):
        return {"A": -k * phys["A"], "B": k * phys["A"]}

    out = R.apply_coupled_sources({"A": a, "B": b}, [conv], g, 1.0, dt)
    assert np.abs(out["A"] - (a - dt * k * a)).max() < 1e-12 * np.abs(a).max()
    assert np.abs(out["B"] - (b + dt * k * a)).max() < 1e-12 * np.abs(a).max()
    # planted: the second entry seeing the already-updated A
    a1 = a - dt * k * a
    assert np.abs((b + dt * k * a1) - out["B"]).max() > 1e-4 * np.abs(a).max()


def test_source_rhs_adds_to_advection_and_diffusion():
    """scalar_rhs with a source equals the unsourced rhs plus source_hat."""
    N = 32
    g, sh = _field(N, 9)
    _, th = _field(N, 10)
    u, v = R.compute_velocity_from_theta(th, g, 1.0)
    src = R.scalar_source_hat(sh, S.ChemicalReaction(0.3), g, 0.0)
    a = R.scalar_rhs(sh, u, v, g, 0.01, src)
    b = R.scalar_rhs(sh, u, v, g, 0.01) + src
    np.testing.assert_array_equal(a, b)
