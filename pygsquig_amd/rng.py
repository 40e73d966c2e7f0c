"""Host side of the forcing-phase generator.

The device kernels hash every mode index with the SplitMix64 output function
(Steele, Lea and Flood, OOPSLA 2014; constants of Vigna's splitmix64.c):
phase(key, m) = 2*pi * ((mix64(key + (m+1) * 0x9E3779B97F4A7C15) >> 11) * 2**-53).
This module maps user keys to the 64-bit key and derives per-step keys; it
replaces ``jax.random.PRNGKey``/``split`` at the reference's call sites
(scripts/run.py:380-382, forcing/ring_forcing.py:90).
"""

from __future__ import annotations

import numpy as np

_MASK = 0xFFFFFFFFFFFFFFFF


def _mix64(z: int) -> int:
    z &= _MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return z ^ (z >> 31)


def key_to_u64(key) -> int:
    """int, or a 2-word uint32 key such as a JAX PRNGKey ``[hi, lo]``."""
    if key is None:
        raise ValueError("a forcing key is required")
    if isinstance(key, (int, np.integer)):
        return int(key) & _MASK
    try:
        import torch

        if isinstance(key, torch.Tensor):
            key = key.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    arr = np.asarray(key).astype(np.uint64).ravel()
    if arr.size == 1:
        return int(arr[0]) & _MASK
    if arr.size == 2:
        return ((int(arr[0]) & 0xFFFFFFFF) << 32) | (int(arr[1]) & 0xFFFFFFFF)
    raise ValueError(f"unsupported key shape {np.asarray(key).shape}")


def step_key(seed: int, n: int) -> int:
    """Key of step n of a run seeded with ``seed`` (same formula as the kernels)."""
    return _mix64((seed & _MASK) + ((n + 1) * 0xD1B54A32D192ED03 & _MASK))


def PRNGKey(seed: int) -> int:
    """Stand-in for ``jax.random.PRNGKey``: a 64-bit integer key."""
    return _mix64(int(seed) & _MASK)


def split(key, num: int = 2):
    """Stand-in for ``jax.random.split``: ``num`` derived 64-bit keys."""
    k = key_to_u64(key)
    return [_mix64(k + (i + 1) * 0x9E3779B97F4A7C15 & _MASK) for i in range(num)]
