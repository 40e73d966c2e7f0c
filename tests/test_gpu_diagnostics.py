"""GPU diagnostics (utils/diagnostics.py, gSQGSolver.get_diagnostics) against
the oracle: binned energy spectrum, integer shells, total energy, enstrophy,
palinstrophy, energy flux and the physical-space diagnostics."""

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    return pg


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def random_spectrum(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


@pytest.mark.parametrize("N,alpha", [(64, 1.0), (256, 0.5), (256, 2.0), (128, -1.0)])
def test_spectral_diagnostics_vs_oracle(pg, N, alpha):
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    X = random_spectrum(N, N + int(10 * alpha))
    Xd = torch.from_numpy(X).cuda()
    kb, ek = pg.compute_energy_spectrum(Xd, g, alpha)
    kbo, eko = R.energy_spectrum(X, go, alpha)
    np.testing.assert_array_equal(kb, kbo)
    assert rel(ek, eko) < 1e-12
    assert np.max(np.abs(ek - eko) / np.maximum(np.abs(eko), 1e-300)) < 1e-12  # every bin
    assert abs(pg.compute_total_energy(Xd, g, alpha) / R.total_energy(X, go, alpha) - 1) < 1e-12
    assert abs(pg.compute_enstrophy(Xd, g, alpha) / R.enstrophy(X, go, alpha) - 1) < 1e-12
    assert abs(pg.compute_palinstrophy(Xd, g, alpha) / R.palinstrophy(X, go, alpha) - 1) < 1e-12
    assert rel(pg.compute_shell_spectrum(Xd, g, alpha), R.shell_spectrum(X, go, alpha)) < 1e-12
    F = random_spectrum(N, 7)
    assert abs(pg.compute_energy_flux(Xd, torch.from_numpy(F).cuda(), g)
               - R.energy_flux(X, F, go)) < 1e-12 * abs(R.energy_flux(X, F, go)) + 1e-9
    assert pg.compute_energy_flux(Xd, None, g) == 0.0


def test_spectrum_modes_on_bin_edges(pg):
    """Modes on the axes sit at |k| = k1d[i], near-integer radii that are the
    bin edges (width 2 at N = 256): each lands in the oracle's bin."""
    N = 256
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    X = np.zeros((N, N), complex)
    X[:, 0] = 1.0 + np.arange(N)  # (i, 0) for every kx
    X[0, :] = 1.0 + 0.5 * np.arange(N)  # (0, j) for every ky
    kb, ek = pg.compute_energy_spectrum(torch.from_numpy(X).cuda(), g, 1.0)
    _, eko = R.energy_spectrum(X, go, 1.0)
    np.testing.assert_allclose(ek, eko, rtol=1e-13, atol=0)


@pytest.mark.parametrize("alpha", [1.0, 0.5])
def test_get_diagnostics_vs_oracle(pg, alpha):
    N = 128
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    theta0 = ics.kolmogorov_53(N, seed=4, alpha=alpha)
    s = pg.gSQGSolver(g, alpha, 0.0, 8)
    st = s.initialize(torch.from_numpy(theta0).cuda())
    d = s.get_diagnostics(st)
    so = R.Solver(go, alpha, 0.0, 8)
    do = so.get_diagnostics(so.initialize(theta0))
    for k in ("kinetic_energy", "enstrophy", "max_theta"):
        assert abs(d[k] / do[k] - 1) < 1e-12, (k, d[k], do[k])
    assert d["step"] == 0 and d["time"] == 0.0
