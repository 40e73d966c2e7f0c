"""Known-answer tests for the forcing-phase generator (oracle/rng.py)."""

import numpy as np

from oracle import rng


def test_splitmix64_published_sequence():
    # First five outputs of SplitMix64 seeded with 1234567 (Vigna's
    # splitmix64.c reference implementation).
    expected = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                4593380528125082431, 16408922859458223821]
    assert rng.splitmix64_sequence(1234567, 5).tolist() == expected


def test_phase_bits_are_splitmix_outputs():
    seq = rng.splitmix64_sequence(42, 16)
    assert rng.phase_bits(42, np.arange(16)).tolist() == seq.tolist()


def test_forcing_phases_range_and_uniformity():
    ph = rng.forcing_phases(7, 64)
    assert ph.shape == (64, 64)
    assert ph.min() >= 0.0 and ph.max() < 2 * np.pi
    # mean of a uniform on [0, 2pi): pi +- 5 sigma
    sigma = 2 * np.pi / np.sqrt(12) / 64
    assert abs(ph.mean() - np.pi) < 5 * sigma


def test_key_mapping():
    assert rng.key_to_u64(5) == 5
    assert rng.key_to_u64(np.array([1, 2], dtype=np.uint32)) == (1 << 32) | 2
    assert rng.step_key(3, 0) != rng.step_key(3, 1)
