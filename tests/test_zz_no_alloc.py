"""The allocation-free spectral diagnostics entry point
(gsqg_spectral_diagnostics_noalloc) against gsqg_spectral_diagnostics
through the C ABI, on two calls in a row with the second on a smaller input
(another plan size and fewer bins).  Both run the same kernels with the same
block count; their per-bin sums accumulate by LDS atomics in an order that
varies between runs, so the bar is FP64 rounding (1e-13 relative), not bit
equality.  A bin count beyond the scratch array's maximum is refused with an
error (GSQG_ERR_ARG, raised as ValueError)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


def _spectrum(pg, N, seed):
    rs = np.random.default_rng(seed)
    g = pg.make_grid(N, TWO_PI)
    return g, pg.fft2(torch.from_numpy(rs.standard_normal((N, N))).cuda(), g)


def _edges(N, n_bins):
    return np.linspace(0.0, N / 2, n_bins + 1)


def test_noalloc_spectral_diagnostics_matches_allocating_entry_point():
    import pygsquig_amd as pg
    from pygsquig_amd.plan import transform_plan

    for N, n_bins, alpha in [(512, 128, 1.0), (64, 16, 0.5)]:  # second call smaller
        print(f"[noalloc N={N} bins={n_bins}]", flush=True)
        g, X = _spectrum(pg, N, seed=N)
        plan = transform_plan(g)
        edges = _edges(N, n_bins)
        want = plan.spectral_diagnostics(X, alpha, edges, 1.0)
        got = plan.spectral_diagnostics(X, alpha, edges, 1.0, noalloc=True)
        for a, b in zip(got, want):
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * np.abs(b).max())
        assert np.all(np.isfinite(got[0])) and got[2][0] > 0


def test_noalloc_refuses_too_many_bins():
    import pygsquig_amd as pg
    from pygsquig_amd.plan import transform_plan

    g, X = _spectrum(pg, 64, seed=1)
    plan = transform_plan(g)
    with pytest.raises(ValueError, match="n_bins must be in"):
        plan.spectral_diagnostics(X, 1.0, _edges(64, 257), 1.0, noalloc=True)
