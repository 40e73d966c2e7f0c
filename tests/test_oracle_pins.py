"""Closed-form pins of the oracle's advection sign, its integrating-factor
(Lawson) RK4 stage weights and the optimized solver's combine, each with
planted-mistake twins that must fail the same check.

The existing invariants (<theta * RHS> = 0, the DC-only RHS, single-shell
decay) cannot see the sign of the advection term, and "IF converges to
explicit" is too loose to see a dropped exponential factor.  These checks
are exact:

- theta = cos x + cos 2y (two shells, |k| = 1 and 2) gives psi = cos x +
  2^-a cos 2y, u = -psi_y = 2^(1-a) sin 2y, v = psi_x = -sin x, and
  J = u theta_x + v theta_y = (2 - 2^(1-a)) sin x sin 2y, so the inviscid
  RHS of core/solver.py:52 is -(2 - 2^(1-a)) sin x sin 2y.
- s = cos x advected by u = sin 2y, v = -sin x: J = u s_x = -sin x sin 2y,
  so the scalar RHS (scalars/passive_scalar.py:86-95) is +sin x sin 2y.
- Lawson RK4 with a diagonal linear part lam and N(y) = a y returns
  exp(lam dt) R(a dt) y with R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24.
- With a constant N(y) = c it returns
  exp(lam dt) y + dt/6 (exp(lam dt) + 4 exp(lam dt/2) + 1) c.
"""

import numpy as np
import pytest

from oracle import reference_sqg as R

TWO_PI = 2 * np.pi


def _two_shell(N):
    g = R.make_grid(N, TWO_PI)
    theta = np.cos(g.x) + np.cos(2 * g.y)
    return g, theta


def _exact_rhs(g, alpha):
    return -(2.0 - 2.0 ** (1.0 - alpha)) * np.sin(g.x) * np.sin(2 * g.y)


# ------------------------------------------------------------- advection sign
def _rhs_core(th, g, a):
    return R.core_rhs(th, g, a, 0.0, 8)


def _rhs_opt(th, g, a):
    return R.core_rhs_optimized(th, g, a, 0.0, 8)


def _rhs_adv(th, g, a):
    return R.advection_hat(th, g, a)


def _mutant_plus_jacobian(th, g, a):
    """Planted: +J instead of -J (core/solver.py:52)."""
    theta = R.ifft2(th)
    u, v = R.compute_velocity_from_theta(th, g, a)
    return R.fft2(R.jacobian(theta, u, v, g))


def _mutant_swapped_velocity(th, g, a):
    """Planted: (u, v) = (psi_x, -psi_y), the perpendicular gradient rotated the
    other way (core/operators.py:90-91)."""
    theta = R.ifft2(th)
    psi = R.compute_streamfunction(th, g, a)
    u, v = R.ifft2(1j * g.kx * psi), R.ifft2(-1j * g.ky * psi)
    return R.fft2(-R.jacobian(theta, u, v, g))


def _mutant_psi_power(th, g, a):
    """Planted: psi = |k|^+a theta (the exponent's sign, operators.py:149-153)."""
    theta = R.ifft2(th)
    kmag = np.sqrt(g.k2)
    op = np.where(g.k2 > 0, np.where(g.k2 > 0, kmag, 1.0) ** a, 0.0)
    u, v = R.perpendicular_gradient(op * th, g)
    return R.fft2(-R.jacobian(theta, u, v, g))


def _rhs_error(fn, alpha, N=16):
    g, theta = _two_shell(N)
    rhs = R.ifft2(fn(R.fft2(theta), g, alpha))
    return np.max(np.abs(rhs - _exact_rhs(g, alpha)))


@pytest.mark.parametrize("fn", [_rhs_core, _rhs_opt, _rhs_adv], ids=["core", "optimized", "if_nonlinear"])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0, 1.5, 2.0, -1.0])
def test_advection_two_shell_closed_form(fn, alpha):
    assert _rhs_error(fn, alpha) < 1e-13


@pytest.mark.parametrize("mutant", [_mutant_plus_jacobian, _mutant_swapped_velocity, _mutant_psi_power])
def test_advection_planted_mistakes_fail(mutant):
    # alpha = 0.5 separates every mutant: 2 - 2^0.5 != 0, and |k|^(+-0.5) differ at |k| = 2
    assert _rhs_error(mutant, 0.5) > 0.1


def test_alpha_zero_two_shell_rhs_vanishes_but_sign_still_pinned():
    """At alpha = 0 (2D Euler in vorticity form with psi = theta here) the
    two-shell RHS is exactly zero, so the sign is pinned by alpha != 0 only."""
    assert _rhs_error(_rhs_core, 0.0) < 1e-13
    assert _rhs_error(_mutant_plus_jacobian, 0.0) < 1e-13
    assert _rhs_error(_mutant_plus_jacobian, 1.0) > 0.5


def test_explicit_and_if_steps_carry_the_advection_sign():
    """Through a whole step: (step(theta) - theta) / dt -> RHS as dt -> 0,
    for both schemes (a flipped sign anywhere in the step fails by O(1))."""
    N, dt, alpha = 16, 1e-7, 1.0
    g, theta = _two_shell(N)
    exact = _exact_rhs(g, alpha)
    for scheme in ["explicit", "if"]:
        s = R.Solver(g, alpha, nu_p=1e-3, p=2, scheme=scheme)
        st = s.initialize(theta)
        st1 = s.step(st, dt)
        # hyperviscosity adds -nu |k|^2 theta: -1e-3 cos x - 4e-3 cos 2y
        lin = -1e-3 * np.cos(g.x) - 4e-3 * np.cos(2 * g.y)
        d = (R.ifft2(st1["theta_hat"]) - theta) / dt - lin
        assert np.max(np.abs(d - exact)) < 1e-5, scheme


# ---------------------------------------------------------- scalar advection
def _scalar_case(N=16):
    g = R.make_grid(N, TWO_PI)
    return g, np.cos(g.x), np.sin(2 * g.y), -np.sin(g.x)


def test_scalar_advection_closed_form():
    g, s, u, v = _scalar_case()
    exact = np.sin(g.x) * np.sin(2 * g.y)
    got = R.ifft2(R.scalar_rhs(R.fft2(s), u, v, g, 0.0))
    assert np.max(np.abs(got - exact)) < 1e-13
    got2 = R.ifft2(R.scalar_advection_hat(R.fft2(s), u, v, g))
    assert np.max(np.abs(got2 - exact)) < 1e-13
    # planted: +J
    bad = R.ifft2(R.fft2(R.jacobian(s, u, v, g)))
    assert np.max(np.abs(bad - exact)) > 1.0
    # planted: u and v exchanged
    bad2 = R.ifft2(R.fft2(-R.jacobian(s, v, u, g)))
    assert np.max(np.abs(bad2 - exact)) > 0.5


def test_scalar_steps_carry_the_advection_sign():
    """step_scalars in both schemes: (s1 - s0)/dt -> +sin x sin 2y when the
    velocity of the new theta is (sin 2y, -sin x): with alpha = 0, psi =
    theta = cos x + 0.5 cos 2y gives u = -psi_y = sin 2y, v = psi_x = -sin x."""
    N, dt = 16, 1e-7
    g = R.make_grid(N, TWO_PI)
    th = R.fft2(np.cos(g.x) + 0.5 * np.cos(2 * g.y))
    s0 = np.cos(g.x)
    exact = np.sin(g.x) * np.sin(2 * g.y)
    for scheme in ["explicit", "if"]:
        slv = R.Solver(g, 0.0, scheme=scheme)
        out = slv.step_scalars({"c": R.fft2(s0)}, th, dt, {"c": 0.0})["c"]
        d = (R.ifft2(out) - s0) / dt
        assert np.max(np.abs(d - exact)) < 1e-5, scheme


# ------------------------------------------------------- Lawson RK4 identities
def _lawson_ok(step_fn, lam, a, c, dt):
    """Both exact identities; returns (err_linear, err_constant)."""
    y = np.array([0.7 - 0.4j, -1.3 + 0.2j])
    lam = np.array(lam)
    z = a * dt
    R4 = 1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24
    got = step_fn(y, lambda s: a * s, lam, dt)
    e_lin = np.max(np.abs(got - np.exp(lam * dt) * R4 * y))
    got_c = step_fn(y, lambda s: c * np.ones_like(s), lam, dt)
    want_c = np.exp(lam * dt) * y + dt / 6 * (np.exp(lam * dt) + 4 * np.exp(lam * dt / 2) + 1) * c
    e_c = np.max(np.abs(got_c - want_c))
    return e_lin, e_c


LAWSON_CASE = dict(lam=[-3.0, -0.7], a=-1.9 + 0.8j, c=0.3 - 1.1j, dt=0.2)


def test_if_rk4_lawson_identities():
    e_lin, e_c = _lawson_ok(R.if_rk4_step, **LAWSON_CASE)
    assert e_lin < 1e-15 and e_c < 1e-15
    # lam = 0 reduces to the classical RK4 polynomial
    e_lin0, _ = _lawson_ok(R.if_rk4_step, lam=[0.0, 0.0], a=-1.9, c=0.3, dt=0.2)
    assert e_lin0 < 1e-15


def _if_mutant(variant):
    def step(state, nonlin_fn, lin, dt):
        e1 = np.exp(lin * (0.5 * dt))
        e2 = np.exp(lin * dt)
        k1 = nonlin_fn(state)
        y2 = e1 * (state + (0.5 * dt) * k1)
        if variant == "y2_no_e1_on_k1":
            y2 = e1 * state + (0.5 * dt) * k1
        k2 = nonlin_fn(y2)
        y3 = e1 * state + (0.5 * dt) * k2
        if variant == "y3_e2":
            y3 = e2 * state + (0.5 * dt) * k2
        k3 = nonlin_fn(y3)
        y4 = e2 * state + dt * (e1 * k3)
        if variant == "y4_drop_e1":
            y4 = e2 * state + dt * k3
        k4 = nonlin_fn(y4)
        acc = e2 * k1
        if variant == "acc_e1_k1":
            acc = e1 * k1
        acc = acc + 2.0 * ((e2 if variant == "k2_e2" else e1) * k2)
        acc = acc + 2.0 * (e1 * k3)
        last = e1 * k4 if variant == "k4_e1" else k4
        return e2 * state + (dt / 6.0) * (acc + last)

    return step


@pytest.mark.parametrize("variant", ["y2_no_e1_on_k1", "y3_e2", "y4_drop_e1", "acc_e1_k1",
                                     "k2_e2", "k4_e1"])
def test_if_rk4_planted_stage_mistakes_fail(variant):
    e_lin, e_c = _lawson_ok(_if_mutant(variant), **LAWSON_CASE)
    assert max(e_lin, e_c) > 1e-4, variant


def test_if_rk4_mutant_harness_is_faithful():
    """The mutant harness with no mutation reproduces the oracle exactly, so
    the failures above come from the planted change alone."""
    rs = np.random.default_rng(3)
    y = rs.standard_normal(6) + 1j * rs.standard_normal(6)
    lin = -rs.random(6)
    f = lambda s: (0.3 - 0.2j) * s + 0.1 * s**2  # noqa: E731
    np.testing.assert_array_equal(_if_mutant("none")(y, f, lin, 0.1), R.if_rk4_step(y, f, lin, 0.1))


def test_if_solver_step_single_shell_matches_lawson_with_linear_forcing():
    """A whole IF solver step whose nonlinear part is a linear "forcing"
    a*theta on a single shell (the Jacobian of a single shell vanishes):
    exp(lam dt) R(a dt) theta exactly."""
    N, dt, nu, p, a = 16, 0.05, 2e-3, 4, -0.9
    g = R.make_grid(N, TWO_PI)
    s = R.Solver(g, 1.0, nu_p=nu, p=p, scheme="if")
    th0 = R.fft2(np.sin(2 * g.x) * np.cos(2 * g.y))  # |k|^2 = 8
    st = s.step(s.initialize(R.ifft2(th0)), dt, forcing=lambda y: a * y)
    lam = -nu * 8.0 ** (p / 2)
    z = a * dt
    want = np.exp(lam * dt) * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24) * th0
    assert np.max(np.abs(st["theta_hat"] - want)) < 1e-12 * np.abs(th0).max()


# --------------------------------------------------- optimized solver (A17)
def test_multistep_order_linear_polynomial_and_combine():
    """solver_optimized.py:48-62: the stages of the classical RK4 polynomial,
    combined as theta + (dt/6.0)*(k1 + 2k2 + 2k3 + k4)."""
    lam, dt = -1.7 + 0.3j, 0.1
    y = np.array([1.0 + 0.5j])
    z = lam * dt
    y1 = R.rk4_step_multistep_order(y, lambda s: lam * s, dt)
    np.testing.assert_allclose(y1, y * (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24), rtol=1e-15)


def test_multistep_combine_rounding_differs_from_step_combine():
    """The two combines differ only in rounding.  On the constant RHS k = 1
    the sum is exactly 6, so the optimized combine gives (dt/6.0)*6 and the
    rk4_step combine dt*6/6 (numpy divides a complex array by 6 as a product
    with the reciprocal 1/6).  At dt = 0.3 these round to 0.3 and
    0.29999999999999993, so each combine is identified bit for bit (a
    planted swap of the combine fails here)."""
    dt = 0.3
    assert (dt / 6.0) * 6.0 != dt * 6.0 * (1.0 / 6)  # the premise
    y0 = np.array([0.0 + 0.0j])
    ones = lambda s: np.ones_like(s)  # noqa: E731
    a = R.rk4_step_multistep_order(y0, ones, dt)
    b = R.rk4_step(y0, ones, dt)
    assert a[0] == (dt / 6.0) * 6.0 and b[0] == dt * 6.0 * (1.0 / 6)
    assert a[0] != b[0]


def test_optimized_rhs_has_no_dc_zeroing():
    """core_rhs_optimized multiplies theta_hat[0,0] by -nu*0**(p/2) (no
    zeroing, solver_optimized.py:42-45): a non-finite DC turns the DC RHS
    into NaN, while core_rhs zeroes it (operators.py:205).  For finite input
    the two are equal."""
    N = 16
    g = R.make_grid(N, TWO_PI)
    rs = np.random.default_rng(5)
    th = R.fft2(rs.standard_normal((N, N)))
    a = R.core_rhs(th, g, 1.0, 1e-3, 4)
    b = R.core_rhs_optimized(th, g, 1.0, 1e-3, 4)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14 * np.abs(a).max())
    with np.errstate(invalid="ignore"):
        hv = R.hyperviscosity(np.full((N, N), np.inf + 0j), g, 1e-3, 4)
    assert hv[0, 0] == 0
    with np.errstate(invalid="ignore"):
        d = -1e-3 * (g.k2 ** 2.0) * np.full((N, N), np.inf + 0j)
    assert np.isnan(d[0, 0].real)


def test_optimized_solver_forced_step_uses_step_combine():
    """gSQGSolverOptimized.step with a forcing falls back to rk4_step
    (solver_optimized.py:152-164) while its unforced step uses the (dt/6.0)
    combine: with the constant forcing 1 and nu = 0 on theta = 0, the forced
    step gives dt*6/6 (rounded as in the test above) and no forcing gives
    exactly 0."""
    N, dt = 8, 0.3
    g = R.make_grid(N, TWO_PI)
    s = R.Solver(g, 1.0, combine="multistep")
    st = s.initialize(np.zeros((N, N)))
    f = lambda y: np.ones_like(y)  # noqa: E731
    out = s.step(st, dt, forcing=f)["theta_hat"]
    assert np.all(out == dt * 6.0 * (1.0 / 6))
    assert np.all(s.step(st, dt)["theta_hat"] == 0)
