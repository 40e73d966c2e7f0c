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


ALPHAS = (0.0, 0.5, 1.0, 1.5, 2.0)


def _rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


# explicit RK4 is stable for dt nu |k|^8 < 2.78 at every mode the grids hold,
# |k| up to 724 at N = 1024: nu = 1e-20 (1e-12 would amplify roundoff in the
# empty modes above |k| ~ 70 by ~7 per step)
@pytest.mark.parametrize("scheme,nu", [("if", 1e-4), ("explicit", 1e-20)])
def test_one_launch_batch_matches_streams_and_oracle(pg, scheme, nu):
    """launch="batch": one column and one row launch per RK4 stage for every
    grid size.  Each member matches the per-group-stream run with the same
    (default radix-8) tiles to rounding, and the oracle to 1e-12 (N <= 512;
    N = 1024 against the streams run only).  21 steps: one replayed 16-step
    graph, eager steps and the last step without a prologue.

    The two runs execute the same per-mode arithmetic compiled into different
    kernels, where the compiler may fuse another multiply-add pair into an
    fma; they are compared to 1e-13, not bit for bit."""
    sizes = (64, 128, 256, 512, 1024)
    members = [(N, ALPHAS[(2 * k + r) % 5]) for k, N in enumerate(sizes) for r in range(2)]
    dt, n = 1e-3, 21
    thetas = [ics.kolmogorov_53(N, seed=40 + m, alpha=a, rms=0.3) for m, (N, a) in enumerate(members)]
    outs = {}
    for launch in ("batch", "streams"):
        ens = pg.EnsembleSolver(members, nu_p=nu, p=8, scheme=scheme, launch=launch)
        st = ens.initialize([torch.from_numpy(t) for t in thetas])
        st = ens.multistep(st, n, dt)
        torch.cuda.synchronize()
        assert int(st["step"]) == n
        outs[launch] = st["theta_hat"]
    for m, ((N, a), th0) in enumerate(zip(members, thetas)):
        got = outs["batch"][m]
        assert _rel(got, outs["streams"][m]) < 1e-13, (m, N)
        if N > 512:
            continue
        go = R.make_grid(N, TWO_PI)
        so = R.Solver(go, a, nu, 8, scheme=scheme)
        sto = so.initialize(th0)
        for _ in range(n):
            sto = so.step(sto, dt)
        err = np.linalg.norm(got.cpu().numpy() - sto["theta_hat"]) / np.linalg.norm(sto["theta_hat"])
        assert err < 1e-12, (m, N, err)


def test_one_launch_batch_repeated_calls_and_config_shape(pg):
    """The batch workload's shape (members at N = 128, 256, 512, alpha
    cycling) over two calls with different step counts, the second replaying
    the cached graph: equal to the streams run to rounding, and each call's
    result finite and distinct from the input."""
    members = [(N, ALPHAS[m % 5]) for m, N in enumerate([128, 256, 512] * 3)]
    dt = 1e-3
    thetas = [torch.from_numpy(ics.kolmogorov_53(N, seed=60 + m, alpha=a, rms=0.5))
              for m, (N, a) in enumerate(members)]
    res = {}
    for launch in ("batch", "streams"):
        ens = pg.EnsembleSolver(members, nu_p=1e-4, p=8, scheme="if", launch=launch)
        st = ens.initialize(thetas)
        st = ens.multistep(ens.multistep(st, 18, dt), 35, dt)
        torch.cuda.synchronize()
        res[launch] = st["theta_hat"]
        assert abs(float(st["time"]) - 53 * dt) < 1e-14
    for m in range(len(members)):
        assert _rel(res["batch"][m], res["streams"][m]) < 1e-13, m
        assert bool(torch.isfinite(res["batch"][m]).all())


def test_one_launch_batch_validation(pg):
    with pytest.raises(ValueError, match="launch must be"):
        pg.EnsembleSolver([(64, 1.0)], launch="fused")
    with pytest.raises(ValueError, match="needs grid sizes"):
        pg.EnsembleSolver([(32, 1.0), (64, 1.0)], launch="batch")
    with pytest.raises(ValueError, match="needs grid sizes"):
        pg.EnsembleSolver([(2048, 1.0)], launch="batch")
    from pygsquig_amd import _lib
    from pygsquig_amd.plan import Plan
    import ctypes

    lib = _lib.load()
    g = pg.make_grid(64, TWO_PI)
    forced = Plan(g, 1.0, 1e-4, 8, forcing=pg.RingForcing(kf=8, dk=2, epsilon=0.1))
    arr = (ctypes.c_void_p * 1)(forced.handle.value)
    h = ctypes.c_void_p()
    rc = lib.gsqg_batch_create(arr, 1, ctypes.byref(h))
    assert rc == _lib.GSQG_ERR_UNSUPPORTED and "no forcing" in _lib.last_error()
    plain = Plan(g, 1.0, 1e-4, 8)
    arr2 = (ctypes.c_void_p * 2)(plain.handle.value, plain.handle.value)
    assert lib.gsqg_batch_create(arr2, 2, ctypes.byref(h)) == _lib.GSQG_ERR_ARG
    assert "twice" in _lib.last_error()
