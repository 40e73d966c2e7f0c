"""CPU restatement of the learned-closure network (test infrastructure only).

The reference has no closure; the network is specified by the workload it is
benchmarked on: a 6-layer periodic-padded 3x3 CNN with 64 channels and GELU,
inputs (theta, u, v) on the coarse grid, output the subgrid Jacobian flux
tau, BF16 weights and activations with FP32 accumulation, trained by Adam on
the mean squared error.  This module fixes the exact arithmetic the HIP
kernels implement, so their outputs and gradients can be compared:

- bf16(x) rounds float32 to the nearest bfloat16, ties to even;
- inputs: x = bf16(float32(plane * in_scale)), channels (theta, u, v);
- layer 0: z0 = bf16(conv(x, bf16(W0)) + b0), a0 = bf16(gelu(z0));
- layers 1-4: z = bf16(conv(a_prev, bf16(W)) + b), a = bf16(gelu(z));
- layer 5: tau = out_scale * (conv(a4, bf16(W5)) + b5), float32;
- conv(a, W)[i, j, co] = sum over taps t = 3 (dy + 1) + (dx + 1) and input
  channels ci of a[(i + dy) mod H, (j + dx) mod W, ci] * W[t, ci, co];
- gelu(z) = z/2 (1 + erf(z / sqrt 2)) in float32;
- loss = mean((tau - target)^2); dtau = 2 (tau - target) / count;
- dZ = bf16(dA * gelu'(z)) for layers 0-4 (stored gradients are bf16).

Axis convention: images are (H, W) = (axis 0, axis 1) of the solver's
physical fields, channels last.

``exact=True`` switches every bf16 rounding off and computes in float64:
that is the form the finite-difference checks pin.
"""

from __future__ import annotations

import math

import numpy as np

N_CH = 64
N_LAYERS = 6
TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).reshape(a.shape)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z, exact=False):
    if exact:
        z = np.asarray(z, dtype=np.float64)
        return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))
    z = np.asarray(z, dtype=np.float32)
    e = _erf(z.astype(np.float64) / math.sqrt(2.0)).astype(np.float32)
    return (np.float32(0.5) * z * (np.float32(1.0) + e)).astype(np.float32)


def gelu_prime(z, exact=False):
    if exact:
        z = np.asarray(z, dtype=np.float64)
        return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)
    z = np.asarray(z, dtype=np.float32)
    z64 = z.astype(np.float64)
    e = _erf(z64 / math.sqrt(2.0)).astype(np.float32)
    g = (np.exp(-0.5 * z64 * z64) / math.sqrt(2 * math.pi)).astype(np.float32)
    return (np.float32(0.5) * (np.float32(1.0) + e) + z * g).astype(np.float32)


def shapes():
    """Parameter names and shapes in the flat order of the device vector."""
    s = [("W0", (9, 3, N_CH)), ("b0", (N_CH,))]
    for l in range(1, 5):
        s += [(f"W{l}", (9, N_CH, N_CH)), (f"b{l}", (N_CH,))]
    s += [("W5", (9, N_CH)), ("b5", (1,))]
    return s


def n_params() -> int:
    return sum(int(np.prod(sh)) for _, sh in shapes())


def unflatten(flat):
    out, o = {}, 0
    for name, sh in shapes():
        n = int(np.prod(sh))
        out[name] = np.asarray(flat[o:o + n]).reshape(sh)
        o += n
    return out


def flatten(p) -> np.ndarray:
    return np.concatenate([np.asarray(p[name]).ravel() for name, _ in shapes()])


def init_params(seed: int = 0) -> np.ndarray:
    """He-style uniform initialisation (float32 flat vector), small output layer."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, sh in shapes():
        if name.startswith("b"):
            p[name] = np.zeros(sh, np.float32)
            continue
        fan_in = 9 * sh[1] if len(sh) == 3 else 9 * sh[1]
        lim = math.sqrt(6.0 / fan_in) if name != "W5" else 0.1 * math.sqrt(3.0 / fan_in)
        p[name] = rng.uniform(-lim, lim, size=sh).astype(np.float32)
    return flatten(p).astype(np.float32)


def conv(a, W, dtype):
    """Periodic 3x3 cross-correlation: a (B, H, W, Ci), W (9, Ci, Co) or (9, Ci)."""
    B, H, Wd, Ci = a.shape
    Wm = W if W.ndim == 3 else W[:, :, None]
    out = np.zeros((B, H, Wd, Wm.shape[2]), dtype)
    for t, (dy, dx) in enumerate(TAPS):
        sh = np.roll(a, shift=(-dy, -dx), axis=(1, 2)).astype(dtype)
        out += sh @ Wm[t].astype(dtype)
    return out if W.ndim == 3 else out[..., 0]


def conv_adjoint(d, W, dtype):
    """Adjoint of conv in its input: d (B, H, W, Co) -> (B, H, W, Ci)."""
    Wm = W if W.ndim == 3 else W[:, :, None]
    dd = d if d.ndim == 4 else d[..., None]
    out = np.zeros(dd.shape[:3] + (Wm.shape[1],), dtype)
    for t, (dy, dx) in enumerate(TAPS):
        sh = np.roll(dd, shift=(dy, dx), axis=(1, 2)).astype(dtype)
        out += sh @ Wm[t].T.astype(dtype)
    return out


def conv_wgrad(a, d, dtype):
    """Gradient of sum(conv(a, W) * d) in W: a (B,H,W,Ci), d (B,H,W,Co) or (B,H,W)."""
    dd = d if d.ndim == 4 else d[..., None]
    Ci, Co = a.shape[3], dd.shape[3]
    g = np.zeros((9, Ci, Co), dtype)
    df = dd.reshape(-1, Co).astype(dtype)
    for t, (dy, dx) in enumerate(TAPS):
        sh = np.roll(a, shift=(-dy, -dx), axis=(1, 2)).reshape(-1, Ci).astype(dtype)
        g[t] = sh.T @ df
    return g if d.ndim == 4 else g[:, :, 0]


def features(planes, in_scale, exact=False):
    """planes (B, 3, H, W) float64 -> x (B, H, W, 3)."""
    x = np.moveaxis(np.asarray(planes, np.float64) * np.asarray(in_scale, np.float64)[None, :, None, None], 1, -1)
    return x if exact else bf16(x.astype(np.float32))


def forward(flat, planes, in_scale, out_scale=1.0, exact=False):
    """Forward pass; returns (tau (B, H, W), cache)."""
    p = unflatten(flat)
    dt = np.float64 if exact else np.float32
    rnd = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    x = features(planes, in_scale, exact)
    cache = {"x": x, "z": [], "a": []}
    a = x
    for l in range(5):
        W, b = p[f"W{l}"], p[f"b{l}"]
        z = rnd(conv(a, rnd(W), dt) + b.astype(dt))
        a = rnd(gelu(z, exact))
        cache["z"].append(z)
        cache["a"].append(a)
    tau = (conv(a, rnd(p["W5"]), dt) + dt(p["b5"][0])) * dt(out_scale)
    return tau.astype(dt), cache


def loss_grad(tau, target, exact=False):
    dt = np.float64 if exact else np.float32
    diff = tau.astype(dt) - np.asarray(target).astype(dt)
    n = diff.size
    return float(np.mean(diff.astype(np.float64) ** 2)), (dt(2.0) * diff / dt(n)).astype(dt)


def backward(flat, cache, dtau, exact=False):
    """Gradient of the loss in the flat parameter vector (same layout)."""
    p = unflatten(flat)
    dt = np.float64 if exact else np.float32
    rnd = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    g = {}
    a4 = cache["a"][4]
    g["W5"] = conv_wgrad(a4, dtau, dt)
    g["b5"] = np.array([np.sum(dtau, dtype=dt)], dt)
    dA = conv_adjoint(dtau, rnd(p["W5"]), dt)
    dZ = rnd(dA * gelu_prime(cache["z"][4], exact))
    for l in range(4, 0, -1):
        a_prev = cache["a"][l - 1]
        g[f"W{l}"] = conv_wgrad(a_prev, dZ, dt)
        g[f"b{l}"] = np.sum(dZ.reshape(-1, N_CH), axis=0, dtype=dt)
        dA = conv_adjoint(dZ, rnd(p[f"W{l}"]), dt)
        dZ = rnd(dA * gelu_prime(cache["z"][l - 1], exact))
    g["W0"] = conv_wgrad(cache["x"], dZ, dt)
    g["b0"] = np.sum(dZ.reshape(-1, N_CH), axis=0, dtype=dt)
    return flatten(g).astype(dt)


def adam(flat, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected, float32); returns new (flat, m, v)."""
    f32 = np.float32
    g = grad.astype(f32)
    m = (f32(b1) * m + f32(1 - b1) * g).astype(f32)
    v = (f32(b2) * v + f32(1 - b2) * g * g).astype(f32)
    c1 = f32(1.0 - b1 ** t)
    c2 = f32(1.0 - b2 ** t)
    mh = m / c1
    vh = v / c2
    w = (flat.astype(f32) - f32(lr) * mh / (np.sqrt(vh) + f32(eps))).astype(f32)
    return w, m, v


def closure_rhs(theta_hat, grid, alpha, flat, in_scale, out_scale):
    """The closure's term in the right-hand side of the coarse model:
    -mask * fft2(tau), tau = out_scale * net(theta, u, v) of theta_hat (the
    flux is added to the physical Jacobian before its dealiased transform)."""
    from oracle import reference_sqg as R

    th = R.ifft2(theta_hat)
    u, v = R.compute_velocity_from_theta(theta_hat, grid, alpha)
    planes = np.stack([th, u, v])[None]
    tau, _ = forward(flat, planes, in_scale, out_scale)
    return -(grid.dealias_mask * R.fft2(tau[0].astype(np.float64)))
