"""Uneven ensembles: every member of a batched plan follows the
single-realization solver exactly (bit for bit for the same call sequence)
and matches the oracle."""

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


@pytest.mark.parametrize("scheme,nu", [("if", 1e-4), ("explicit", 1e-12)])
def test_members_match_single_solver_and_oracle(pg, scheme, nu):
    members = [(32, 0.0), (64, 1.0), (32, 1.5), (64, 0.5), (32, 2.0), (64, -1.0)]
    dt, n = 1e-3, 6
    ens = pg.EnsembleSolver(members, nu_p=nu, p=8, scheme=scheme)
    thetas = [ics.kolmogorov_53(N, seed=20 + m, alpha=a, rms=0.3) for m, (N, a) in enumerate(members)]
    st = ens.initialize([torch.from_numpy(t) for t in thetas])
    st = ens.multistep(st, n, dt)
    st = ens.step(st, dt)
    assert int(st["step"]) == n + 1 and abs(float(st["time"]) - (n + 1) * dt) < 1e-15
    for m, ((N, a), th0) in enumerate(zip(members, thetas)):
        g = pg.make_grid(N, TWO_PI)
        s = pg.gSQGSolver(g, a, nu, 8, scheme=scheme)
        s0 = s.initialize(torch.from_numpy(th0).cuda())
        ss = s.step(s.multistep(s0, n, dt), dt)  # the ensemble's call sequence
        got = st["theta_hat"][m]
        assert torch.equal(got, ss["theta_hat"]), f"member {m} differs from the single solver"
        # A loop of single steps runs each step's first-stage prologue as a
        # kernel of its own instead of fused into the previous epilogue: the
        # same arithmetic in another instantiation, equal to rounding.
        sl = s0
        for _ in range(n + 1):
            sl = s.step(sl, dt)
        rel = float(torch.linalg.norm(got - sl["theta_hat"]) / torch.linalg.norm(sl["theta_hat"]))
        assert rel < 1e-14, (m, rel)
        go = R.make_grid(N, TWO_PI)
        so = R.Solver(go, a, nu, 8, scheme=scheme)
        sto = so.initialize(th0)
        for _ in range(n + 1):
            sto = so.step(sto, dt)
        err = np.linalg.norm(got.cpu().numpy() - sto["theta_hat"]) / np.linalg.norm(sto["theta_hat"])
        assert err < 1e-12, (m, err)


def test_ensemble_validation(pg):
    with pytest.raises(ValueError, match="alpha must be in"):
        pg.EnsembleSolver([(32, 3.0)])
    with pytest.raises(ValueError, match="p must be"):
        pg.EnsembleSolver([(32, 1.0)], p=3)
    ens = pg.EnsembleSolver([(32, 1.0), (32, 0.5)])
    with pytest.raises(ValueError, match="expected 2 initial fields"):
        ens.initialize([np.zeros((32, 32))])
