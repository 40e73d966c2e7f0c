"""Pins the NumPy oracle (oracle/reference_sqg.py) against closed forms,
invariants and brute force.  These mirror the reference's own known-answer
tests (tests/test_grid.py, test_operators.py, test_solver.py,
test_time_integrator.py, test_ring_forcing.py, test_damping.py,
test_passive_scalar.py) and add checks that a dropped term, a wrong sign or a
transposed operand would fail."""

import numpy as np
import pytest

from oracle import reference_sqg as R
from oracle import ics

TWO_PI = 2 * np.pi


# --------------------------------------------------------------------- grid
def test_grid_wavenumber_order_and_nyquist():
    g = R.make_grid(4, TWO_PI)  # tests/test_grid.py:49-58
    np.testing.assert_array_equal(g.kx[:, 0], [0, 1, -2, -1])
    np.testing.assert_array_equal(g.ky[0, :], [0, 1, -2, -1])
    g8 = R.make_grid(8, TWO_PI)  # tests/test_grid.py:60-69
    assert g8.kx[4, 0] == -4 and g8.ky[0, 4] == -4
    assert g.x.shape == (4, 4) and g.k2[0, 0] == 0


def test_grid_dealias_mask():
    g = R.make_grid(4, TWO_PI)  # tests/test_grid.py:197-209
    assert g.dealias_mask[0, 0] and g.dealias_mask[0, 1] and not g.dealias_mask[0, 2]
    g32 = R.make_grid(32, TWO_PI)  # tests/test_grid.py:84-103
    assert not g32.dealias_mask[16, 0]
    frac = g32.dealias_mask.mean()
    assert 0.25 < frac < 0.45


def test_grid_odd_raises():
    with pytest.raises(ValueError):
        R.make_grid(7, 1.0)


def test_grid_coordinates_and_k2():
    N, L = 16, 3.0
    g = R.make_grid(N, L)
    np.testing.assert_allclose(g.x[:, 0], np.arange(N) * L / N, rtol=1e-14)
    np.testing.assert_allclose(g.y[0, :], np.arange(N) * L / N, rtol=1e-14)
    np.testing.assert_allclose(g.k2, g.kx**2 + g.ky**2, rtol=0)
    kint = np.fft.fftfreq(N, 1.0 / N)
    np.testing.assert_allclose(g.kx[:, 0], kint * TWO_PI / L, rtol=1e-14)


# --------------------------------------------------------------------- FFT
def _brute_dft2(f):
    N = f.shape[0]
    n = np.arange(N)
    W = np.exp(-2j * np.pi * np.outer(n, n) / N)
    return W @ f @ W.T


@pytest.mark.parametrize("N", [4, 8, 16])
def test_fft2_brute_force(N):
    rs = np.random.default_rng(N)
    f = rs.standard_normal((N, N))
    np.testing.assert_allclose(R.fft2(f), _brute_dft2(f), atol=1e-12 * N * N)
    X = rs.standard_normal((N, N)) + 1j * rs.standard_normal((N, N))
    brute_inv = np.conj(_brute_dft2(np.conj(X))) / N**2
    np.testing.assert_allclose(R.ifft2(X), brute_inv.real, atol=1e-13)


def test_fft_round_trip_and_parseval():
    rs = np.random.default_rng(1)
    f = rs.standard_normal((32, 32))
    np.testing.assert_allclose(R.ifft2(R.fft2(f)), f, rtol=1e-14, atol=1e-14)
    F = R.fft2(f)
    np.testing.assert_allclose(np.sum(f**2), np.sum(np.abs(F) ** 2) / 32**2, rtol=1e-12)


def test_ifft2_real_equals_hermitian_projection():
    """.real of the c2c inverse only sees P(X)(k) = (X(k) + conj X(-k))/2."""
    rs = np.random.default_rng(2)
    N = 16
    X = rs.standard_normal((N, N)) + 1j * rs.standard_normal((N, N))
    Xm = np.conj(np.roll(np.flip(X, (0, 1)), 1, (0, 1)))
    P = 0.5 * (X + Xm)
    np.testing.assert_allclose(R.ifft2(X), np.fft.irfft2(P[:, : N // 2 + 1], s=(N, N)), atol=1e-14)
    # the anti-Hermitian part contributes nothing
    np.testing.assert_allclose(R.ifft2(X - P), 0.0, atol=1e-14)


# ---------------------------------------------------------------- operators
def test_gradient_closed_form():
    N = 32
    g = R.make_grid(N, TWO_PI)
    f = np.sin(2 * g.x) * np.cos(3 * g.y)
    fx, fy = R.gradient(R.fft2(f), g)
    np.testing.assert_allclose(fx, 2 * np.cos(2 * g.x) * np.cos(3 * g.y), atol=1e-12)
    np.testing.assert_allclose(fy, -3 * np.sin(2 * g.x) * np.sin(3 * g.y), atol=1e-12)


@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0, 2.0])
def test_velocity_single_shell_closed_form(alpha):
    """psi_hat = |k|^-alpha theta_hat (no minus sign, operators.py:149-155),
    (u, v) = (-d_y psi, d_x psi)."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    theta = np.cos(2 * g.x) * np.sin(3 * g.y)
    k = np.sqrt(13.0)
    u, v = R.compute_velocity_from_theta(R.fft2(theta), g, alpha)
    psi = theta * k ** (-alpha)
    np.testing.assert_allclose(u, -3 * np.cos(2 * g.x) * np.cos(3 * g.y) * k ** (-alpha), atol=1e-12)
    np.testing.assert_allclose(v, -2 * np.sin(2 * g.x) * np.sin(3 * g.y) * k ** (-alpha), atol=1e-12)
    del psi


def test_streamfunction_alpha2_is_inverse_laplacian():
    N = 16
    g = R.make_grid(N, TWO_PI)
    rs = np.random.default_rng(3)
    th = R.fft2(rs.standard_normal((N, N)))
    psi = R.compute_streamfunction(th, g, 2.0)
    ref = np.where(g.k2 > 0, th / np.where(g.k2 > 0, g.k2, 1.0), 0.0)
    np.testing.assert_allclose(psi, ref, rtol=1e-13, atol=1e-13)


def test_jacobian_linear_flow():
    """tests/test_operators.py:260-278: u=1, v=0, theta=sin x -> cos x."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    J = R.jacobian(np.sin(g.x), np.ones((N, N)), np.zeros((N, N)), g)
    np.testing.assert_allclose(J, np.cos(g.x), atol=1e-12)
    J2 = R.jacobian(np.sin(g.y), np.zeros((N, N)), np.ones((N, N)), g)
    np.testing.assert_allclose(J2, np.cos(g.y), atol=1e-12)


def test_jacobian_zero_mean_and_inviscid_invariant():
    N = 32
    g = R.make_grid(N, TWO_PI)
    theta = ics.kolmogorov_53(N, seed=5, kmax=N / 3)
    th = R.fft2(theta)
    u, v = R.compute_velocity_from_theta(th, g, 1.0)
    J = R.jacobian(R.ifft2(th), u, v, g)
    assert abs(np.mean(J)) < 1e-14
    rhs = R.core_rhs(th, g, 1.0, 0.0, 8)
    # <theta, RHS> = 0 for the inviscid dealiased SQG flux (test_solver.py:182-201)
    assert abs(np.real(np.sum(np.conj(th) * rhs))) < 1e-10 * np.sum(np.abs(th) ** 2)
    # DC-only state has zero RHS (test_solver.py:168-180)
    dc = np.zeros((N, N), complex)
    dc[0, 0] = 3.0
    assert np.max(np.abs(R.core_rhs(dc, g, 1.0, 1e-3, 8))) == 0.0


def test_hyperviscosity_is_k_to_the_p_and_dc_zero():
    """operators.py:202: k2**(p/2) = |k|^p (not |k|^{2p}); [0,0] = 0."""
    N = 16
    g = R.make_grid(N, TWO_PI)
    th = np.ones((N, N), complex)
    hv = R.hyperviscosity(th, g, 0.5, 4)
    np.testing.assert_allclose(hv[1, 0], -0.5 * 1.0)
    np.testing.assert_allclose(hv[2, 1], -0.5 * 25.0)
    assert hv[0, 0] == 0


# --------------------------------------------------------------------- RK4
def test_rk4_linear_ode_exact_polynomial():
    lam, dt = -1.7, 0.1
    y = np.array([1.0 + 0.5j])
    y1 = R.rk4_step(y, lambda s: lam * s, dt)
    z = lam * dt
    np.testing.assert_allclose(y1, y * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), rtol=1e-15)


def test_rk4_order():
    errs = []
    for n in [10, 20, 40]:
        y = np.array([1.0])
        dt = 1.0 / n
        for _ in range(n):
            y = R.rk4_step(y, lambda s: np.cos(s), dt)
        errs.append(y[0])
    # Richardson: ratio of successive differences ~ 2^4
    ratio = (errs[0] - errs[1]) / (errs[1] - errs[2])
    assert 14.5 < ratio < 17.5


def test_single_shell_decays_by_rk4_polynomial():
    """sin(4x)cos(4y) is a single |k| shell: the Jacobian vanishes and only the
    explicit hyperviscous decay R(-dt nu |k|^p) remains.  p=4, nu=1e-3,
    dt=1e-2 keeps dt*nu*|k|^4 <= 2.6 at the corner, inside RK4's stability
    interval, so round-off at other modes is not amplified."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    nu, p, dt = 1e-3, 4, 1e-2
    s = R.Solver(g, alpha=1.0, nu_p=nu, p=p)
    st = s.initialize(ics.sin4(N))
    th0 = st["theta_hat"].copy()
    st = s.step(st, dt)
    z = -dt * nu * 32.0 ** (p / 2)
    fac = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    np.testing.assert_allclose(st["theta_hat"], th0 * fac, atol=1e-12 * np.abs(th0).max())
    # IF scheme decays by exactly exp(z)
    s_if = R.Solver(g, alpha=1.0, nu_p=nu, p=p, scheme="if")
    st2 = s_if.step(s_if.initialize(ics.sin4(N)), dt)
    np.testing.assert_allclose(st2["theta_hat"], th0 * np.exp(z), atol=1e-12 * np.abs(th0).max())


def test_if_scheme_converges_to_explicit():
    N = 16
    g = R.make_grid(N, TWO_PI)
    th0 = ics.kolmogorov_53(N, seed=2, kmax=5) * 0.5
    outs = {}
    for scheme in ["explicit", "if"]:
        s = R.Solver(g, 1.0, nu_p=1e-4, p=4, scheme=scheme)
        st = s.initialize(th0)
        for _ in range(20):
            st = s.step(st, 5e-3)
        outs[scheme] = st["theta_hat"]
    rel = np.linalg.norm(outs["if"] - outs["explicit"]) / np.linalg.norm(outs["explicit"])
    assert rel < 1e-6


# ----------------------------------------------------------------- forcing
def test_ring_forcing_rate_is_epsilon():
    """ring_forcing.py:152-168: after rescaling <theta * Re ifft2 F> = eps."""
    N = 64
    g = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=7))
    f = R.RingForcing(kf=8, dk=2, epsilon=0.1)
    F = f(th, key=123, grid=g)
    rate = np.mean(R.ifft2(th) * R.ifft2(F))
    np.testing.assert_allclose(rate, 0.1, rtol=1e-12)
    mask = f.forcing_mask(g)
    raw = f.raw_forcing(g, key=123)
    np.testing.assert_allclose(np.abs(raw[mask]), np.sqrt(0.1), rtol=1e-14)
    assert np.all(raw[~mask] == 0)
    # zero state: no rescale (scale = 1)
    F0 = f(np.zeros((N, N), complex), key=123, grid=g)
    np.testing.assert_allclose(F0, raw)


def test_ring_mask_selection():
    """test_ring_forcing.py:44-91: |k| in [kf-dk/2, kf+dk/2]; no k=0."""
    N = 64
    g = R.make_grid(N, TWO_PI)
    f = R.RingForcing(kf=10, dk=2)
    m = f.forcing_mask(g)
    km = np.sqrt(g.k2)
    assert np.all((km[m] >= 9 - 1e-9) & (km[m] <= 11 + 1e-9))
    assert not m[0, 0]
    # integer-lattice count in the closed annulus
    ki = np.fft.fftfreq(N, 1.0 / N)
    kk = np.sqrt(ki[:, None] ** 2 + ki[None, :] ** 2)
    assert m.sum() == np.sum((kk >= 9) & (kk <= 11))


def test_hermitian_corners_real():
    X = np.full((8, 8), 1 + 2j)
    Y = R.ensure_hermitian_symmetry(X)
    for (i, j) in [(0, 0), (4, 0), (0, 4), (4, 4)]:
        assert Y[i, j].imag == 0
    assert Y[1, 2] == 1 + 2j


# ----------------------------------------------------------------- damping
def test_damping_values():
    """damping.py:15-42, 75-92 / tests/test_damping.py:16-92."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    th = np.ones((N, N), complex)
    d = R.large_scale_damping(th, g, mu=0.3, kf=8)
    km = np.sqrt(g.k2)
    np.testing.assert_allclose(d[km < 4], -0.3)
    assert np.all(d[km >= 4] == 0)
    assert np.all(R.large_scale_damping(th, g, 0.0, 8) == 0)
    cd = R.CombinedDamping(mu=0.3, kf=8, nu_p=1e-3, p=2)
    np.testing.assert_allclose(cd(th, g), d + R.hyperviscosity(th, g, 1e-3, 2))
    with pytest.raises(ValueError):
        R.CombinedDamping(p=3)


# ----------------------------------------------------------------- scalars
def test_scalar_pure_diffusion_and_conservation():
    N = 32
    g = R.make_grid(N, TWO_PI)
    s0 = ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 8)
    sh = R.fft2(s0)
    zero = np.zeros((N, N))
    kap, dt = 0.01, 0.05
    s1 = R.rk4_step(sh, lambda s: R.scalar_rhs(s, zero, zero, g, kap), dt)
    z = -kap * g.k2 * dt
    np.testing.assert_allclose(s1, sh * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), atol=1e-12)
    # advection conserves the mean exactly (k=0 untouched)
    th = R.fft2(ics.kolmogorov_53(N, seed=3))
    u, v = R.compute_velocity_from_theta(th, g, 1.0)
    s2 = R.rk4_step(sh, lambda s: R.scalar_rhs(s, u, v, g, 0.0), 1e-2)
    np.testing.assert_allclose(s2[0, 0].real, sh[0, 0].real, rtol=1e-13)


# ------------------------------------------------------------- diagnostics
def test_diagnostics_single_mode():
    """test_solver.py:286-301 style closed forms."""
    N = 32
    g = R.make_grid(N, TWO_PI)
    s = R.Solver(g, 1.0)
    st = s.initialize(np.cos(3 * g.x))
    d = s.get_diagnostics(st)
    np.testing.assert_allclose(d["enstrophy"], 0.25, rtol=1e-13)
    # psi = cos(3x)/3 -> v = -sin(3x), KE = 0.5 * 0.5
    np.testing.assert_allclose(d["kinetic_energy"], 0.25, rtol=1e-13)
    np.testing.assert_allclose(d["max_theta"], 1.0, rtol=1e-13)
    # total energy: 0.5 * |k|^(alpha-2) * sum|th|^2 / N^2 = 0.5 * (1/3) * 0.5 N^2
    np.testing.assert_allclose(R.total_energy(st["theta_hat"], g, 1.0), 0.5 / 3 * 0.5 * N**2, rtol=1e-13)
    shells = R.shell_spectrum(st["theta_hat"], g, 1.0)
    np.testing.assert_allclose(shells[3], 0.5 / 3 * 0.5, rtol=1e-13)
    assert shells.sum() == pytest.approx(shells[3])
    # two modes (+-3, 0) of |theta_hat| = N^2 / 2, alpha = 1:
    # enstrophy 0.5 * 2 * 3^2 * N^4/4; palinstrophy 2 * 0.5 * (3^2)^2 * N^4/4 / N^2
    np.testing.assert_allclose(R.enstrophy(st["theta_hat"], g, 1.0), 9 * N**4 / 4, rtol=1e-13)
    np.testing.assert_allclose(R.palinstrophy(st["theta_hat"], g, 1.0), 81 * N**2 / 4, rtol=1e-13)
    # binned spectrum: N//4 = 8 bins of width 2 over [0, 16]; |k| = 3 in bin 1
    kb, ek = R.energy_spectrum(st["theta_hat"], g, 1.0)
    np.testing.assert_allclose(kb, np.arange(1.0, 16.0, 2.0), rtol=1e-15)
    np.testing.assert_allclose(ek[1], 2 * 0.5 / 3 * N**4 / 4 / 2, rtol=1e-13)
    assert np.max(np.abs(np.delete(ek, 1))) < 1e-25 * ek[1]  # FFT roundoff only
    # energy flux <theta F> of F = theta_hat: sum |theta_hat|^2 / N^2 = 2 (N^2/2)^2 / N^2
    np.testing.assert_allclose(R.energy_flux(st["theta_hat"], st["theta_hat"], g), N**2 / 2,
                               rtol=1e-13)


# --------------------------------------------- planted mistakes are caught
def test_planted_sign_error_is_caught():
    """A velocity with the wrong sign of the perpendicular gradient fails the
    closed-form velocity check used above."""
    N = 16
    g = R.make_grid(N, TWO_PI)
    theta = np.cos(2 * g.x) * np.sin(3 * g.y)
    th = R.fft2(theta)
    psi = R.compute_streamfunction(th, g, 1.0)
    wrong_u = R.ifft2(+1j * g.ky * psi)  # planted: lost minus sign
    exact_u = -3 * np.cos(2 * g.x) * np.cos(3 * g.y) / np.sqrt(13.0)
    assert np.max(np.abs(wrong_u - exact_u)) > 0.1


def test_planted_dropped_rk4_term_is_caught():
    lam, dt = -1.7, 0.1

    def bad_rk4(y, f, h):
        k1 = f(y)
        k2 = f(y + 0.5 * h * k1)
        k3 = f(y + 0.5 * h * k2)
        k4 = f(y + h * k3)
        return y + h * (k1 + 2 * k2 + k3 + k4) / 6  # planted: 2*k3 -> k3

    z = lam * dt
    exact = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    assert abs(bad_rk4(1.0, lambda s: lam * s, dt) - exact) > 1e-4
