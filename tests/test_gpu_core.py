"""GPU parity of the core path against the NumPy oracle: transforms,
velocity inversion, compute_rhs and short RK4 trajectories.  Tolerances are
relative L2 in FP64."""

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def rel(a, b):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    return pg


@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512, 1024])
def test_fft2_ifft2(pg, N):
    print(f"[fft2/ifft2 N={N}]", flush=True)
    g = pg.make_grid(N, TWO_PI)
    rs = np.random.default_rng(N)
    f = rs.standard_normal((N, N))
    X = pg.fft2(torch.from_numpy(f).cuda(), g)
    assert rel(X, np.fft.fft2(f)) < 1e-14
    Y = rs.standard_normal((N, N)) + 1j * rs.standard_normal((N, N))  # non-Hermitian
    y = pg.ifft2(torch.from_numpy(Y).cuda(), g)
    assert rel(y, np.fft.ifft2(Y).real) < 1e-14


@pytest.mark.parametrize("N", [256, 512])
def test_fft2_ifft2_every_shape(pg, N):
    """The API transforms in every tile shape of N = 256 and 512 against
    numpy's c2c.  The radix-4 shapes there hold one transform per wave
    (N = 256 columns, N = 512 half-length rows), whose last exchange is the
    row-swap transpose instead of LDS."""
    from pygsquig_amd.plan import Plan

    g = pg.make_grid(N, TWO_PI)
    plan = Plan(g, 1.0, 0.0, 8)
    rs = np.random.default_rng(N + 1)
    f = rs.standard_normal((N, N))
    Y = rs.standard_normal((N, N)) + 1j * rs.standard_normal((N, N))
    Xn, yn = np.fft.fft2(f), np.fft.ifft2(Y).real
    nv = plan.shape()[2]
    for c in range(nv):
        for r in range(nv):
            print(f"[fft2/ifft2 N={N} shapes ({c},{r})]", flush=True)
            plan.set_shape(c, r)
            X = torch.empty((N, N), dtype=torch.complex128, device="cuda")
            plan.fft2(torch.from_numpy(f).cuda(), X)
            y = torch.empty((N, N), dtype=torch.float64, device="cuda")
            plan.ifft2(torch.from_numpy(Y).cuda(), y)
            assert rel(X, Xn) < 1e-14, (c, r)
            assert rel(y, yn) < 1e-14, (c, r)


@pytest.mark.parametrize("N,alpha", [(32, 1.0), (64, 0.5), (128, 2.0), (256, 1.0), (64, 0.0)])
def test_velocity(pg, N, alpha):
    print(f"[velocity N={N} alpha={alpha}]", flush=True)
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=3, alpha=alpha))
    u, v = pg.compute_velocity_from_theta(torch.from_numpy(th).cuda(), g, alpha)
    uo, vo = R.compute_velocity_from_theta(th, go, alpha)
    assert rel(u, uo) < 1e-13 and rel(v, vo) < 1e-13


@pytest.mark.parametrize("N,alpha,nu,p", [(32, 1.0, 1e-3, 2), (64, 0.5, 1e-6, 4), (256, 1.0, 1e-16, 8),
                                          (128, 2.0, 0.0, 8)])
def test_rhs(pg, N, alpha, nu, p):
    print(f"[rhs N={N}]", flush=True)
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=4, alpha=alpha))
    s = pg.gSQGSolver(g, alpha, nu, p)
    r = s.compute_rhs(torch.from_numpy(th).cuda())
    ro = R.core_rhs(th, go, alpha, nu, p)
    assert rel(r, ro) < 1e-12


@pytest.mark.parametrize("N,nsteps", [(32, 5), (64, 20), (256, 10)])
def test_step_explicit_short(pg, N, nsteps):
    print(f"[step N={N} n={nsteps}]", flush=True)
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    theta0 = 0.1 * ics.kolmogorov_53(N, seed=5)
    s = pg.gSQGSolver(g, 1.0, 1e-16, 8)
    so = R.Solver(go, 1.0, 1e-16, 8)
    st = s.initialize(torch.from_numpy(theta0).cuda())
    sto = so.initialize(theta0)
    for _ in range(nsteps):
        st = s.step(st, 1e-3)
        sto = so.step(sto, 1e-3)
    assert rel(st["theta_hat"], sto["theta_hat"]) < 1e-13
    assert int(st["step"]) == nsteps
    assert abs(float(st["time"]) - sto["time"]) < 1e-15


def test_multistep_matches_step_loop(pg):
    N = 64
    g = pg.make_grid(N, TWO_PI)
    theta0 = 0.1 * ics.kolmogorov_53(N, seed=6)
    s = pg.gSQGSolver(g, 1.0, 1e-16, 8)
    st = s.initialize(torch.from_numpy(theta0).cuda())
    a = s.multistep(st, 12, 1e-3)
    b = st
    for _ in range(12):
        b = s.step(b, 1e-3)
    assert rel(a["theta_hat"], b["theta_hat"].cpu().numpy()) < 1e-14
