"""Pins of the closure network's CPU restatement (oracle/closure_cnn.py):
bfloat16 rounding on known values, the periodic convolution against a
brute-force sum, its adjoint and weight gradient by the bilinear identity,
GELU on closed-form values, the full backward pass against central finite
differences (float64), the first Adam step's closed form, and planted
mistakes that these checks must catch."""

import math

import numpy as np
import pytest

from oracle import closure_cnn as O


def test_bf16_rounding_known_values():
    ulp = 2.0 ** -7  # bfloat16 spacing in [1, 2)
    vals = np.array([1.0, 1 + ulp / 2, 1 + 3 * ulp / 2, 1 + ulp / 2 + 2 ** -20, -2.5, 0.0,
                     3.0e38, 1e-40], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -2.5, 0.0], np.float32)
    got = O.bf16(vals)
    np.testing.assert_array_equal(got[:6], want)  # ties to even, above-tie rounds up
    assert np.isfinite(got[6]) and abs(got[6] / 3.0e38 - 1) < 2 ** -8
    assert abs(float(got[7]) - 1e-40) <= 2.0 ** -134  # subnormal: half the bf16 spacing 2^-133
    assert np.isnan(O.bf16(np.array([np.nan], np.float32)))[0]
    # every result is representable: low 16 bits clear
    r = O.bf16(np.random.default_rng(0).standard_normal(1000).astype(np.float32))
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def brute_conv(a, W):
    B, H, Wd, Ci = a.shape
    Co = W.shape[2]
    out = np.zeros((B, H, Wd, Co))
    for b in range(B):
        for i in range(H):
            for j in range(Wd):
                for t in range(9):
                    dy, dx = t // 3 - 1, t % 3 - 1
                    out[b, i, j] += a[b, (i + dy) % H, (j + dx) % Wd] @ W[t]
    return out


def test_conv_matches_brute_force_sum():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((2, 4, 5, 3))
    W = rng.standard_normal((9, 3, 2))
    np.testing.assert_allclose(O.conv(a, W, np.float64), brute_conv(a, W), rtol=1e-12, atol=1e-12)


def test_planted_tap_reversal_is_caught():
    """A convolution with the taps reversed (a plausible index mistake) fails the brute-force check."""
    rng = np.random.default_rng(2)
    a = rng.standard_normal((1, 4, 5, 2))
    W = rng.standard_normal((9, 2, 2))
    wrong = O.conv(a, W[::-1].copy(), np.float64)
    assert np.abs(wrong - brute_conv(a, W)).max() > 1e-3


def test_conv_adjoint_and_weight_gradient_identities():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 4, 6, 3))
    W = rng.standard_normal((9, 3, 4))
    d = rng.standard_normal((2, 4, 6, 4))
    lhs = np.sum(O.conv(a, W, np.float64) * d)
    assert abs(lhs - np.sum(a * O.conv_adjoint(d, W, np.float64))) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(W * O.conv_wgrad(a, d, np.float64))) < 1e-10 * abs(lhs)
    # single-output-channel forms (layer 5)
    W1 = rng.standard_normal((9, 3))
    d1 = rng.standard_normal((2, 4, 6))
    l1 = np.sum(O.conv(a, W1, np.float64) * d1)
    assert abs(l1 - np.sum(a * O.conv_adjoint(d1, W1, np.float64))) < 1e-10 * abs(l1)
    assert abs(l1 - np.sum(W1 * O.conv_wgrad(a, d1, np.float64))) < 1e-10 * abs(l1)


def test_gelu_closed_forms():
    phi = lambda x: 0.5 * (1 + math.erf(x / math.sqrt(2)))  # noqa: E731
    z = np.array([0.0, 1.0, -1.0, 3.0])
    np.testing.assert_allclose(O.gelu(z, exact=True), z * np.array([phi(x) for x in z]), rtol=1e-15)
    assert O.gelu_prime(np.array([0.0]), exact=True)[0] == 0.5
    h = 1e-6
    zz = np.linspace(-4, 4, 17)
    fd = (O.gelu(zz + h, exact=True) - O.gelu(zz - h, exact=True)) / (2 * h)
    np.testing.assert_allclose(O.gelu_prime(zz, exact=True), fd, atol=1e-8)
    # float32 forms agree with the float64 ones to float32 precision
    np.testing.assert_allclose(O.gelu(zz.astype(np.float32)), O.gelu(zz, exact=True), rtol=1e-6,
                               atol=1e-7)


def _small_problem(seed=4, H=4, W=4, B=2):
    rng = np.random.default_rng(seed)
    flat = O.init_params(seed).astype(np.float64)
    flat += 0.05 * rng.standard_normal(flat.size)  # non-zero biases too
    planes = rng.standard_normal((B, 3, H, W))
    target = rng.standard_normal((B, H, W))
    return flat, planes, target


def test_backward_matches_finite_differences():
    flat, planes, target = _small_problem()
    scale = (0.8, 1.1, 0.9)

    def loss(p):
        tau, _ = O.forward(p, planes, scale, 1.3, exact=True)
        return O.loss_grad(tau, target, exact=True)[0]

    tau, cache = O.forward(flat, planes, scale, 1.3, exact=True)
    _, dtau = O.loss_grad(tau, target, exact=True)
    g = O.backward(flat, cache, dtau * 1.3, exact=True)
    rng = np.random.default_rng(5)
    sl = O.unflatten(np.arange(O.n_params()))
    for name in ("W0", "b0", "W2", "b3", "W4", "W5", "b5"):
        d = np.zeros(O.n_params())
        idx = np.asarray(sl[name]).ravel()
        d[idx] = rng.standard_normal(idx.size)
        h = 1e-5
        fd = (loss(flat + h * d) - loss(flat - h * d)) / (2 * h)
        an = float(g @ d)
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-8), (name, fd, an)


def test_planted_dropped_gelu_derivative_is_caught():
    """A backward pass that drops the GELU derivative disagrees with finite differences."""
    flat, planes, target = _small_problem(seed=6)
    scale = (1.0, 1.0, 1.0)
    tau, cache = O.forward(flat, planes, scale, 1.0, exact=True)
    _, dtau = O.loss_grad(tau, target, exact=True)
    good = O.backward(flat, cache, dtau, exact=True)
    bad_cache = {"x": cache["x"], "a": cache["a"], "z": [np.full_like(z, 50.0) for z in cache["z"]]}
    bad = O.backward(flat, bad_cache, dtau, exact=True)  # gelu'(50) = 1: derivative dropped
    sl = O.unflatten(np.arange(O.n_params()))
    idx = np.asarray(sl["W1"]).ravel()
    assert np.linalg.norm(bad[idx] - good[idx]) > 1e-3 * np.linalg.norm(good[idx])


def test_bf16_forward_tracks_exact_forward():
    flat, planes, _ = _small_problem(seed=7, H=8, W=8, B=1)
    t32, _ = O.forward(flat.astype(np.float32), planes, (1, 1, 1), 1.0)
    t64, _ = O.forward(flat, planes, (1, 1, 1), 1.0, exact=True)
    assert np.linalg.norm(t32 - t64) / np.linalg.norm(t64) < 3e-2


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(8)
    w = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    z = np.zeros(100, np.float32)
    w1, m1, v1 = O.adam(w, g, z, z, 1, lr=1e-3)
    # m_hat = g, v_hat = g^2: the step is lr * g / (|g| + eps)
    np.testing.assert_allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-5)


def test_parameter_layout():
    assert O.n_params() == 150081
    sl = O.unflatten(np.arange(O.n_params()))
    assert sl["W0"].ravel()[0] == 0 and sl["b5"][0] == O.n_params() - 1
    assert sl["W1"].shape == (9, 64, 64) and sl["W1"].ravel()[0] == 9 * 3 * 64 + 64
