"""Passive scalars with source terms on the GPU, against the oracle.

The per-stage right-hand side comes from gsqg_scalar_rhs (the fused scalar
kernels' advection and diffusion) with the source added on the device; the
oracle restates PassiveScalarEvolver.step (scalars/passive_scalar.py:172-258),
MultiSpeciesEvolver's coupled sources (:378-403) and the sources of
scalars/source_terms.py.  Bar: relative L2 <= 1e-12."""

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R
from oracle import sources as OS

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    return pg


def _rel(a, b):
    a = a.cpu().numpy() if torch.is_tensor(a) else a
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(N=64, seed=3):
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(0.5 * ics.kolmogorov_53(N, seed=seed))
    u, v = R.compute_velocity_from_theta(th, go, 1.0)
    s0 = ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 8) + 0.3 * np.cos(go.x)
    return go, u, v, s0


def _pairs(pg):
    from pygsquig_amd import sources as S

    return {
        "exponential": (S.ExponentialGrowth(-0.5), OS.ExponentialGrowth(-0.5)),
        "localized": (S.LocalizedSource(1.2, 0.3, 5.9, 0.6, time_dependent=lambda t: 1 + t),
                      OS.LocalizedSource(1.2, 0.3, 5.9, 0.6, time_dependent=lambda t: 1 + t)),
        "chemical": (S.ChemicalReaction(0.8, threshold=0.0), OS.ChemicalReaction(0.8, 0.0)),
        "periodic": (S.TimePeriodicSource(0.7, 4.0, 0.2, spatial_pattern=lambda g: torch.sin(g.y)),
                     OS.TimePeriodicSource(0.7, 4.0, 0.2, spatial_pattern=lambda g: np.sin(g.y))),
        "combination": (S.LinearCombination([(1.0, S.ExponentialGrowth(0.2)),
                                             (0.5, S.ChemicalReaction(0.3))]),
                        OS.LinearCombination([(1.0, OS.ExponentialGrowth(0.2)),
                                              (0.5, OS.ChemicalReaction(0.3))])),
    }


def test_scalar_rhs_vs_oracle(pg):
    from pygsquig_amd.plan import Plan

    N, kappa = 64, 0.02
    go, u, v, s0 = _setup(N)
    g = pg.make_grid(N, TWO_PI)
    plan = Plan(g, 1.0, 0.0, 8, kappas=(kappa,))
    sh = R.fft2(s0)
    out = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    plan.scalar_rhs(0, torch.from_numpy(sh).cuda(), torch.from_numpy(u).cuda(),
                    torch.from_numpy(v).cuda(), out)
    assert _rel(out, R.scalar_rhs(sh, u, v, go, kappa)) < 1e-12


@pytest.mark.parametrize("name", ["exponential", "localized", "chemical", "periodic",
                                  "combination"])
def test_evolver_with_source_vs_oracle(pg, name):
    from pygsquig_amd.scalars import PassiveScalarEvolver, PassiveScalarState

    N, kappa, dt, n = 64, 0.01, 2e-3, 3
    go, u, v, s0 = _setup(N)
    g = pg.make_grid(N, TWO_PI)
    src, osrc = _pairs(pg)[name]
    ev = PassiveScalarEvolver(grid=g, kappa=kappa, source_fn=src, name="c")
    st = ev.initialize(scalar0=torch.from_numpy(s0))
    st = PassiveScalarState(scalar_hat=st.scalar_hat, time=0.25)
    uu, vv = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
    sh, t, tsrc = R.fft2(s0), 0.25, 0.0
    for _ in range(n):
        st = ev.step(st, dt, uu, vv)
        srch = R.scalar_source_hat(sh, osrc, go, t)
        tsrc = tsrc + dt * np.mean(R.ifft2(sh) * R.ifft2(srch))
        sh = R.scalar_step(sh, u, v, go, kappa, dt, osrc, t=t)
        t = t + dt
    assert _rel(st.scalar_hat, sh) < 1e-12, name
    assert abs(st.time - t) < 1e-15
    assert abs(st.total_source - tsrc) <= 1e-12 * max(abs(tsrc), 1e-30), (st.total_source, tsrc)


def test_multispecies_with_coupled_sources_vs_oracle(pg):
    from pygsquig_amd import sources as S
    from pygsquig_amd.scalars import MultiSpeciesEvolver

    N, dt, n, k = 64, 2e-3, 3, 0.6
    go, u, v, s0 = _setup(N)
    g = pg.make_grid(N, TWO_PI)
    b0 = 0.2 * np.sin(2 * go.y)

    def conv_t(phys, grid, t):
        return {"A": -k * phys["A"], "B": k * phys["A"]}

    def conv_n(phys, grid, t):
        return {"A": -k * phys["A"], "B": k * phys["A"]}

    ms = MultiSpeciesEvolver(grid=g, species={"A": {"kappa": 0.01,
                                                    "source": S.ExponentialGrowth(-0.3)},
                                              "B": {"kappa": 0.0}},
                             coupled_sources=[conv_t])
    st = ms.initialize({"A": torch.from_numpy(s0), "B": torch.from_numpy(b0)})
    uu, vv = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
    sc = {"A": R.fft2(s0), "B": R.fft2(b0)}
    t = 0.0
    for _ in range(n):
        st = ms.step(st, dt, uu, vv)
        sc = {"A": R.scalar_step(sc["A"], u, v, go, 0.01, dt, OS.ExponentialGrowth(-0.3), t=t),
              "B": R.scalar_step(sc["B"], u, v, go, 0.0, dt)}
        sc = R.apply_coupled_sources(sc, [conv_n], go, t + dt, dt)
        t = t + dt
    for name in ("A", "B"):
        assert _rel(st.scalars[name], sc[name]) < 1e-12, name


@pytest.mark.parametrize("scheme,nu", [("explicit", 1e-12), ("if", 1e-6)])
def test_solver_with_sourced_scalar_vs_oracle(pg, scheme, nu):
    """gSQGSolverWithScalars: theta step, velocity of the new theta_hat, then
    the sourced scalar step (core/solver_with_scalars.py:131-178); the IF
    solver's scalars integrate their diffusion exactly."""
    from pygsquig_amd import sources as S
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt, n, kappa = 64, 1e-3, 3, 0.02
    go = R.make_grid(N, TWO_PI)
    theta0 = 0.5 * ics.kolmogorov_53(N, seed=11)
    s0 = ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 16)
    g = pg.make_grid(N, TWO_PI)
    solver = gSQGSolverWithScalars(
        g, 1.0, nu, 8, scheme=scheme,
        passive_scalars={"c": {"kappa": kappa, "source": S.ChemicalReaction(0.5, 0.0)},
                         "d": {"kappa": 0.0}})
    st = solver.initialize(torch.from_numpy(theta0).cuda(),
                           scalar_init={"c": torch.from_numpy(s0),
                                        "d": torch.from_numpy(np.cos(go.x))})
    st = solver.multistep(solver.step(st, dt), n - 1, dt)
    so = R.Solver(go, 1.0, nu, 8, scheme=scheme)
    sto = so.initialize(theta0)
    sc = {"c": R.fft2(s0), "d": R.fft2(np.cos(go.x))}
    t = 0.0
    for _ in range(n):
        sto = so.step(sto, dt)
        u, v = so.compute_velocity(sto["theta_hat"])
        sc = {"c": R.scalar_step(sc["c"], u, v, go, kappa, dt, OS.ChemicalReaction(0.5, 0.0),
                                 t=t, scheme=scheme),
              "d": R.scalar_step(sc["d"], u, v, go, 0.0, dt, scheme=scheme)}
        t = t + dt
    assert _rel(st["theta_hat"], sto["theta_hat"]) < 1e-12
    for name in ("c", "d"):
        assert _rel(st["scalar_state"].scalars[name], sc[name]) < 1e-12, name
    assert int(st["step"]) == n and abs(float(st["time"]) - n * dt) < 1e-15


def test_add_passive_scalar_with_source(pg):
    from pygsquig_amd import sources as S
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt = 64, 1e-3
    g = pg.make_grid(N, TWO_PI)
    solver = gSQGSolverWithScalars(g, 1.0, 1e-12, 8, passive_scalars={"a": {"kappa": 0.0}})
    solver.add_passive_scalar("h", kappa=0.0, source=S.make_heating_source(np.pi, np.pi, 1.0))
    assert solver.scalar_evolver.has_sources()
    st = solver.initialize(torch.zeros((N, N), dtype=torch.float64).cuda())
    st = solver.step(st, dt)
    # zero theta: no flow, the heated scalar grows by exactly dt * S
    go = R.make_grid(N, TWO_PI)
    want = dt * R.fft2(OS.LocalizedSource(1.0, np.pi, np.pi, 0.5)(None, go, 0.0))
    assert _rel(st["scalar_state"].scalars["h"], want) < 1e-12
    with pytest.raises(Exception, match="SourceTerm instance"):
        solver.add_passive_scalar("bad", source=lambda s, g, t: s)
