"""Pins of the oracle's Ornstein-Uhlenbeck ring forcing (tau_f > 0,
forcing/ring_forcing.py:94-117, 144-145), each with planted-mistake twins
that must fail the same check.

- Same key every call: the white value w is fixed, so the state follows
  s_1 = w, s_n = d s_(n-1) + v w with d = exp(-dt / tau_f),
  v = sqrt(1 - d^2); in closed form s_n = s* + d^(n-1) (w - s*) with the
  fixed point s* = v w / (1 - d).
- Independent keys: the reference's own test (tests/test_ring_forcing.py:
  233-267) asks that the correlation of the n-th forcing with the first
  decays like exp(-n dt / tau_f) over the first five calls (rtol 0.2).
- The variance factor keeps the process stationary: with unit-modulus white
  noise, the mean |s|^2 over the ring stays at amp^2.
- tau_f = 0 keeps no state and draws fresh white noise on every call.
"""

import numpy as np
import pytest

from oracle import reference_sqg as R

TWO_PI = 2 * np.pi


def _theta_const(N):
    """theta = 1: <theta f> = 0 off DC, so the rescale leaves F unscaled
    (ring_forcing.py:161-165) and F is the raw OU state."""
    return R.fft2(np.ones((N, N)))


class _Mutant(R.RingForcing):
    """RingForcing with one planted mistake in its OU update."""

    def __init__(self, variant, **kw):
        super().__init__(**kw)
        self.variant = variant

    def raw_forcing(self, grid, key=None, phases=None, dt=None):
        mask = self.forcing_mask(grid)
        if phases is None:
            phases = R._rng.forcing_phases(R._rng.key_to_u64(key), grid.N)
        f = self.amplitude * mask * np.exp(1j * phases)
        d = np.exp(-dt / self.tau_f)
        v = np.sqrt(1 - d ** 2)
        if self.forcing_state is None:
            self.forcing_state = v * f if self.variant == "filtered_first" else f
        elif self.variant == "swapped":
            self.forcing_state = v * self.forcing_state + d * f
        elif self.variant == "no_variance_factor":
            self.forcing_state = d * self.forcing_state + (1 - d) * f
        else:
            self.forcing_state = d * self.forcing_state + v * f
        return R.ensure_hermitian_symmetry(self.forcing_state)


def _fixed_key_sequence(forcing, n_calls, N=32, dt=0.01):
    g = R.make_grid(N, TWO_PI)
    th = _theta_const(N)
    return [forcing(th, key=7, dt=dt, grid=g) for _ in range(n_calls)], g


def _fixed_key_closed_form_ok(forcing, N=32, dt=0.01, tau=0.1):
    outs, g = _fixed_key_sequence(forcing, 6, N, dt)
    mask = forcing.forcing_mask(g)
    w = forcing.amplitude * mask * np.exp(1j * R._rng.forcing_phases(R._rng.key_to_u64(7), N))
    d = np.exp(-dt / tau)
    v = np.sqrt(1 - d ** 2)
    star = v * w / (1 - d)
    ok = True
    for n, F in enumerate(outs, start=1):
        want = R.ensure_hermitian_symmetry(star + d ** (n - 1) * (w - star))
        ok &= np.allclose(F, want, rtol=0, atol=1e-14)
    return ok


def test_ou_fixed_key_closed_form():
    f = R.RingForcing(kf=10.0, dk=2.0, epsilon=0.1, tau_f=0.1)
    assert _fixed_key_closed_form_ok(f)
    # the state is the pre-realification value, full plane (ring modes only)
    assert f.forcing_state.shape == (32, 32)
    assert np.all(f.forcing_state[~f.forcing_mask(R.make_grid(32, TWO_PI))] == 0)


@pytest.mark.parametrize("variant", ["swapped", "filtered_first", "no_variance_factor"])
def test_ou_fixed_key_planted_mistakes_fail(variant):
    f = _Mutant(variant, kf=10.0, dk=2.0, epsilon=0.1, tau_f=0.1)
    assert not _fixed_key_closed_form_ok(f)


def test_ou_mutant_harness_is_faithful():
    f = _Mutant("none", kf=10.0, dk=2.0, epsilon=0.1, tau_f=0.1)
    assert _fixed_key_closed_form_ok(f)


def test_ou_state_keeps_imaginary_part_at_realified_modes():
    """The reference filters the raw forcing and realifies only its output
    (:144-150): at a ring mode on the Nyquist row the state keeps its
    imaginary part while the returned forcing is real there.  (Realifying
    the state instead gives the same forcing, since the update is linear
    with real coefficients.)  kf = 16 at N = 32 puts (16, 0) on the ring."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    th = _theta_const(N)
    f = R.RingForcing(kf=16.0, dk=1.0, tau_f=0.1)
    for k in (3, 4, 5):
        F = f(th, key=k, dt=0.01, grid=g)
    assert f.forcing_mask(g)[16, 0]
    assert F[16, 0].imag == 0 and f.forcing_state[16, 0].imag != 0
    assert F[16, 0].real == f.forcing_state[16, 0].real


def test_ou_correlation_decays_like_exp_minus_t_over_tau():
    """tests/test_ring_forcing.py:233-267 restated on the oracle."""
    N, tau, dt = 32, 0.1, 0.01
    g = R.make_grid(N, TWO_PI)
    f = R.RingForcing(kf=10.0, tau_f=tau)
    th = _theta_const(N)
    f0 = f(th, key=R._rng.step_key(42, 0), dt=dt, grid=g)
    corr = []
    for i in range(20):
        F = f(th, key=R._rng.step_key(42, i + 1), dt=dt, grid=g)
        m = (np.abs(f0) > 1e-10) & (np.abs(F) > 1e-10)
        n0 = np.sqrt(np.mean(np.abs(f0[m]) ** 2))
        n1 = np.sqrt(np.mean(np.abs(F[m]) ** 2))
        corr.append(np.mean(np.conj(f0[m]) * F[m]).real / (n0 * n1))
    # the first entry is one call after f0
    expected = np.exp(-np.arange(1, 21) * dt / tau)
    np.testing.assert_allclose(corr[:5], expected[:5], rtol=0.2)


def test_ou_variance_stationary_and_white_noise_uncorrelated():
    N, tau, dt = 64, 0.05, 0.01
    g = R.make_grid(N, TWO_PI)
    th = _theta_const(N)
    f = R.RingForcing(kf=20.0, dk=2.0, epsilon=0.1, tau_f=tau)
    m = f.forcing_mask(g)
    for i in range(200):
        f(th, key=R._rng.step_key(5, i), dt=dt, grid=g)
    ratio = np.mean(np.abs(f.forcing_state[m]) ** 2) / f.amplitude ** 2
    assert abs(ratio - 1) < 0.1
    bad = _Mutant("no_variance_factor", kf=20.0, dk=2.0, epsilon=0.1, tau_f=tau)
    for i in range(200):
        bad(th, key=R._rng.step_key(5, i), dt=dt, grid=g)
    assert np.mean(np.abs(bad.forcing_state[m]) ** 2) / bad.amplitude ** 2 < 0.5
    # tau_f = 0: no state, fresh noise (tests/test_ring_forcing.py:208-230)
    w = R.RingForcing(kf=20.0, dk=2.0, tau_f=0.0)
    a = w(th, key=1, dt=dt, grid=g)
    b = w(th, key=2, dt=dt, grid=g)
    assert w.forcing_state is None
    c = np.mean(np.conj(a[m]) * b[m]) / np.sqrt(np.mean(np.abs(a[m]) ** 2) *
                                                  np.mean(np.abs(b[m]) ** 2))
    assert abs(c) < 0.2
