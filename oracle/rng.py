"""Counter-based generator for forcing phases (oracle side).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``).

The reference draws its ring-forcing phases with ``jax.random.uniform(key,
(N, N), minval=0, maxval=2*pi)`` (``pygsquig/forcing/ring_forcing.py:90``).
``jax.random`` (threefry) lives in jaxlib, which is not present here, so its
bit stream cannot be reproduced; instead every stochastic field is generated
from a seed by *this* generator and injected into both the oracle and the GPU
path.  The HIP kernels evaluate exactly the same integer arithmetic
(``gsqg_phase_bits`` in ``pygsquig_amd/csrc/gsqg_common.h``).

Generator: the SplitMix64 output function (Steele, Lea and Flood, "Fast
splittable pseudorandom number generators", OOPSLA 2014; the constants are
those of Vigna's public-domain ``splitmix64.c``).  The phase of full-plane mode
index ``m = i*N + j`` under a 64-bit key is

    bits = mix64(key + (m + 1) * 0x9E3779B97F4A7C15)      (mod 2**64)
    u    = (bits >> 11) * 2**-53                          in [0, 1)
    phi  = u * (2*pi)                                     (binary64 2*pi)

i.e. ``bits`` is the (m+1)-th SplitMix64 output of a generator seeded with
``key``.  Pinned by the published SplitMix64 sequence for seed 1234567 (the
known-answer test in ``tests/test_oracle_rng.py``).

The per-step key used by ``multistep`` (which has no caller-supplied keys) is
``step_key(seed, n) = mix64(seed + (n + 1) * 0xD1B54A32D192ED03)``.
"""

from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
STEP_MUL = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
TWO_PI = 2.0 * np.pi


def mix64(z):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def splitmix64_sequence(seed: int, n: int) -> np.ndarray:
    """First ``n`` outputs of SplitMix64 seeded with ``seed``."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) + idx * GOLDEN)


def key_to_u64(key) -> int:
    """Map a user key (int, or a 2-word uint32 array such as a JAX PRNGKey
    ``[hi, lo]``) to the 64-bit key the generator uses."""
    if key is None:
        raise ValueError("a forcing key is required")
    if isinstance(key, (int, np.integer)):
        return int(key) & 0xFFFFFFFFFFFFFFFF
    arr = np.asarray(key).astype(np.uint64).ravel()
    if arr.size == 1:
        return int(arr[0]) & 0xFFFFFFFFFFFFFFFF
    if arr.size == 2:
        return ((int(arr[0]) & 0xFFFFFFFF) << 32) | (int(arr[1]) & 0xFFFFFFFF)
    raise ValueError(f"unsupported key shape {np.asarray(key).shape}")


def step_key(seed: int, n: int) -> int:
    """Key of step ``n`` (0-based) of a run seeded with ``seed``."""
    with np.errstate(over="ignore"):
        v = mix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.uint64(n + 1) * STEP_MUL)
    return int(v)


def phase_bits(key: int, idx) -> np.ndarray:
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(np.uint64(key) + (idx + np.uint64(1)) * GOLDEN)


def forcing_phases(key: int, N: int) -> np.ndarray:
    """(N, N) phases in [0, 2*pi) for every full-plane mode (axis 0 = kx)."""
    bits = phase_bits(key, np.arange(N * N, dtype=np.uint64))
    u = (bits >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return (u * TWO_PI).reshape(N, N)
