"""GPU parity of the forced / damped / passive-scalar / integrating-factor
paths against the oracle: short trajectories computed on the fly and long
ones against the committed golden fixtures (tests/golden/make_golden.py).
Tolerances are relative L2 on theta_hat (1e-12 unless stated)."""

import os

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R
from oracle import rng as orng

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else np.asarray(a)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def half(a):
    a = a.detach().cpu().numpy() if torch.is_tensor(a) else a
    return a[:, : a.shape[0] // 2 + 1]


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    return pg


def test_ring_forcing_call(pg):
    """RingForcing.__call__ returns the reference's rescaled F."""
    N = 64
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=1))
    F = pg.RingForcing(kf=10, dk=2, epsilon=0.1)(torch.from_numpy(th).cuda(), key=99, grid=g)
    Fo = R.RingForcing(kf=10, dk=2, epsilon=0.1)(th, key=99, grid=go)
    assert rel(F, Fo) < 1e-13
    rate = np.mean(R.ifft2(th) * R.ifft2(F.cpu().numpy()))
    assert abs(rate - 0.1) < 1e-13


def test_combined_damping_call(pg):
    N = 32
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=2))
    d = pg.CombinedDamping(mu=0.2, kf=8, nu_p=1e-5, p=4)(torch.from_numpy(th).cuda(), grid=g)
    do = R.CombinedDamping(mu=0.2, kf=8, nu_p=1e-5, p=4)(th, go)
    assert rel(d, do) < 1e-14


@pytest.mark.parametrize("N", [32, 64])
def test_forced_damped_rhs(pg, N):
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    th = R.fft2(ics.kolmogorov_53(N, seed=3))
    s = pg.gSQGSolver(g, 1.0, 1e-8, 8)
    f = pg.RingForcing(kf=N / 8, dk=2, epsilon=0.1)
    d = pg.CombinedDamping(mu=0.1, kf=N / 8, nu_p=1e-9, p=8)
    r = s.compute_rhs(torch.from_numpy(th).cuda(), forcing=f, damping=d, key=5)
    fo = R.RingForcing(kf=N / 8, dk=2, epsilon=0.1)
    do = R.CombinedDamping(mu=0.1, kf=N / 8, nu_p=1e-9, p=8)
    ro = R.Solver(go, 1.0, 1e-8, 8).compute_rhs(th, forcing=lambda y: fo(y, key=5, grid=go),
                                                damping=lambda y: do(y, go))
    assert rel(r, ro) < 1e-12


@pytest.mark.parametrize("scheme", ["explicit", "if"])
def test_forced_scalar_steps_vs_oracle(pg, scheme):
    """Forced + damped theta and two passive scalars (kappa 0 and 1e-3),
    per-step explicit keys through step(), and derived keys through
    multistep(); full spectra including the anti-Hermitian ring part."""
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt, n = 64, 5e-4, 8
    nu_d = 1e-8 * (512 / N) ** 8 if scheme == "if" else 1e-14
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    theta0 = ics.kolmogorov_53(N, seed=7)
    s0 = ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 16)
    s1 = ics.gaussian_blob(N, (np.pi / 2, 3 * np.pi / 2), TWO_PI / 16)
    kap = {"a": 0.0, "b": 1e-3}
    solver = gSQGSolverWithScalars(g, 1.0, 0.0, 8, scheme=scheme,
                                   passive_scalars={k: {"kappa": v} for k, v in kap.items()})
    f = pg.RingForcing(kf=8, dk=2, epsilon=0.1, seed=42)
    d = pg.CombinedDamping(mu=0.1, kf=8, nu_p=nu_d, p=8)
    st = solver.initialize(torch.from_numpy(theta0).cuda(),
                           scalar_init={"a": torch.from_numpy(s0), "b": torch.from_numpy(s1)})
    st_ms = solver.multistep(dict(st), n, dt, forcing=f, damping=d)
    for i in range(n):
        st = solver.step(st, dt, forcing=f, damping=d, key=orng.step_key(42, i))
    so = R.Solver(go, 1.0, 0.0, 8, scheme=scheme)
    fo = R.RingForcing(kf=8, dk=2, epsilon=0.1)
    do = R.CombinedDamping(mu=0.1, kf=8, nu_p=nu_d, p=8)
    sto = so.initialize(theta0)
    sc = {"a": R.fft2(s0), "b": R.fft2(s1)}
    for i in range(n):
        k = orng.step_key(42, i)
        sto = so.step(sto, dt, forcing=lambda y, k=k: fo(y, key=k, grid=go), damping=do)
        sc = so.step_scalars(sc, sto["theta_hat"], dt, kap)
    for got in (st, st_ms):
        assert rel(got["theta_hat"], sto["theta_hat"]) < 1e-12
        for name in kap:
            assert rel(got["scalar_state"].scalars[name], sc[name]) < 1e-12
    assert int(st_ms["step"]) == n and abs(float(st_ms["time"]) - n * dt) < 1e-15


def test_p0_listed_unstable_config_is_nonfinite(pg):
    """Config 1 exactly as listed (nu=1e-4, p=8, dt=1e-3) is outside explicit
    RK4's stability region: both the oracle and the GPU go non-finite."""
    N = 256
    g = pg.make_grid(N, TWO_PI)
    s = pg.gSQGSolver(g, 1.0, 1e-4, 8)
    st = s.initialize(torch.from_numpy(ics.sin4(N)).cuda())
    st = s.multistep(st, 100, 1e-3)
    assert not torch.isfinite(st["theta_hat"]).all()
    assert not s.check_finite()
    go = R.make_grid(N, TWO_PI)
    so = R.Solver(go, 1.0, 1e-4, 8)
    sto = so.initialize(ics.sin4(N))
    with np.errstate(all="ignore"):
        for _ in range(100):
            sto = so.step(sto, 1e-3)
    assert not np.isfinite(sto["theta_hat"]).all()


def test_p1_single_shell_closed_form(pg):
    """Config 1 with the reference's default nu=1e-16: sin(4x)cos(4y) is a
    steady single shell, so theta_hat(n) = R(-dt nu |k|^8)^n theta_hat(0)."""
    N, dt, n, nu = 256, 1e-3, 1000, 1e-16
    g = pg.make_grid(N, TWO_PI)
    s = pg.gSQGSolver(g, 1.0, nu, 8)
    st = s.initialize(torch.from_numpy(ics.sin4(N)).cuda())
    th0 = st["theta_hat"].cpu().numpy()
    st = s.multistep(st, n, dt)
    z = -dt * nu * 32.0**4
    fac = (1 + z + z**2 / 2 + z**3 / 6 + z**4 / 24) ** n
    assert rel(st["theta_hat"], th0 * fac) < 1e-12


def test_p2_golden_1000_steps(pg):
    gold = np.load(os.path.join(GOLD, "p2_n256.npz"))
    N = 256
    g = pg.make_grid(N, TWO_PI)
    s = pg.gSQGSolver(g, 1.0, 1e-16, 8)
    st = s.initialize(torch.from_numpy(0.1 * ics.kolmogorov_53(N, seed=5)).cuda())
    st = s.multistep(st, 1000, 1e-3)
    assert rel(half(st["theta_hat"]), gold["theta_hat"]) < 1e-12


def test_p3_golden_forced_scalars_200_steps(pg):
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    gold = np.load(os.path.join(GOLD, "p3_n256.npz"))
    N = 256
    g = pg.make_grid(N, TWO_PI)
    solver = gSQGSolverWithScalars(g, 1.0, 0.0, 8, passive_scalars={"s0": {"kappa": 0.0},
                                                                    "s1": {"kappa": 0.0}})
    init = {f"s{i}": torch.from_numpy(ics.gaussian_blob(N, c, TWO_PI / 16))
            for i, c in enumerate(ics.SCALAR_CENTRES[:2])}
    st = solver.initialize(torch.from_numpy(ics.kolmogorov_53(N, seed=2)).cuda(), scalar_init=init)
    f = pg.RingForcing(kf=40, dk=2, epsilon=0.1, seed=42)
    d = pg.CombinedDamping(mu=0.1, kf=40, nu_p=1e-16, p=8)
    st = solver.multistep(st, 200, 5e-4, forcing=f, damping=d)
    assert rel(half(st["theta_hat"]), gold["theta_hat"]) < 1e-12
    for k in ("s0", "s1"):
        assert rel(half(st["scalar_state"].scalars[k]), gold[k]) < 1e-12
    mask = torch.abs(torch.sqrt(g.k2) - 40) <= 1.0
    assert rel(st["theta_hat"][mask], gold["ring_full"]) < 1e-12


def test_if_golden_stiff_hyperviscosity(pg):
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    gold = np.load(os.path.join(GOLD, "ifs_n256.npz"))
    N = 256
    g = pg.make_grid(N, TWO_PI)
    solver = gSQGSolverWithScalars(g, 1.0, 0.0, 8, scheme="if",
                                   passive_scalars={"s0": {"kappa": 0.0}})
    init = {"s0": torch.from_numpy(ics.gaussian_blob(N, ics.SCALAR_CENTRES[0], TWO_PI / 16))}
    st = solver.initialize(torch.from_numpy(ics.kolmogorov_53(N, seed=2)).cuda(), scalar_init=init)
    f = pg.RingForcing(kf=40, dk=2, epsilon=0.1, seed=42)
    d = pg.CombinedDamping(mu=0.1, kf=40, nu_p=1e-8 * 2.0**8, p=8)
    st = solver.multistep(st, 100, 5e-4, forcing=f, damping=d)
    assert torch.isfinite(st["theta_hat"]).all()
    assert rel(half(st["theta_hat"]), gold["theta_hat"]) < 1e-12
    assert rel(half(st["scalar_state"].scalars["s0"]), gold["s0"]) < 1e-12


def test_if_linear_mode_decays_exactly(pg):
    """IF: a single shell decays by exactly exp(-dt nu |k|^p) per step, at a
    nu where explicit RK4 is unstable."""
    N, dt, nu = 128, 1e-2, 1e-6
    g = pg.make_grid(N, TWO_PI)
    s = pg.gSQGSolver(g, 1.0, nu, 8, scheme="if")
    st = s.initialize(torch.from_numpy(ics.sin4(N)).cuda())
    th0 = st["theta_hat"].cpu().numpy()
    st = s.multistep(st, 10, dt)
    fac = np.exp(-dt * nu * 32.0**4 * 10)
    assert rel(st["theta_hat"], th0 * fac) < 1e-12


@pytest.mark.parametrize("N,shapes", [(256, (0, 0)), (1024, (3, 3))])
def test_graph_steps_bit_identical_to_eager(pg, N, shapes):
    """Forced, damped theta with two scalars, 37 steps (two 16-step graphs, a
    one-step graph, eager first and last steps): inside the captured graphs
    the scalar step runs as a branch beside theta's stages 2-4; every array
    comes out bit for bit as in the eager loop that runs the two in one
    stream, on two graph runs in a row."""
    from pygsquig_amd import _lib
    from pygsquig_amd.plan import Plan

    g = pg.make_grid(N, TWO_PI)
    kf = N // 16
    f = pg.RingForcing(kf=kf, dk=2, epsilon=0.1, seed=42)
    d = pg.CombinedDamping(mu=0.1, kf=kf, nu_p=1e-8 * (512 / N) ** 8, p=8)
    plan = Plan(g, 1.0, 0.0, 8, scheme=_lib.SCHEME_IF, forcing=f, damping=d,
                kappas=(0.0, 1e-3))
    plan.set_shape(*shapes)
    th0 = torch.from_numpy(np.fft.fft2(ics.kolmogorov_53(N, seed=5))).cuda()
    s0 = [torch.from_numpy(np.fft.fft2(ics.gaussian_blob(N, c, TWO_PI / 16))).cuda()
          for c in ics.SCALAR_CENTRES[:2]]
    runs = []
    for use_graph in (False, True, True):
        print(f"[pipeline N={N} graph={use_graph}]", flush=True)
        plan.load_theta(th0)
        for i, s in enumerate(s0):
            plan.load_scalar(i, s)
        plan.advance(5e-4, 37, _lib.KEY_DERIVED, 42, 0, use_graph=use_graph)
        out = [torch.empty_like(th0) for _ in range(3)]
        plan.store_theta(out[0])
        for i in range(2):
            plan.store_scalar(i, out[1 + i])
        torch.cuda.synchronize()
        runs.append(out)
    for r in runs[1:]:
        for a, b in zip(r, runs[0]):
            assert torch.equal(a, b)
    assert bool(torch.isfinite(runs[0][0]).all())


def test_forced_scalars_1000_steps_vs_oracle(pg):
    """1000 steps at N = 128 of the full path: IF scheme, ring forcing (kf = 16,
    phases from step keys derived from the seed, as multistep derives them),
    CombinedDamping and two passive scalars, one graph-replayed multistep call
    against the oracle stepped one step at a time.  theta_hat and both scalars
    to relative L2 1e-12 after 1000 steps."""
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt, n = 128, 5e-4, 1000
    nu_d = 1e-8 * (512 / N) ** 8
    g = pg.make_grid(N, TWO_PI)
    go = R.make_grid(N, TWO_PI)
    theta0 = ics.kolmogorov_53(N, seed=3)
    s0 = ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 16)
    s1 = ics.gaussian_blob(N, (np.pi / 2, 3 * np.pi / 2), TWO_PI / 16)
    kap = {"a": 0.0, "b": 1e-3}
    solver = gSQGSolverWithScalars(g, 1.0, 0.0, 8, scheme="if",
                                   passive_scalars={k: {"kappa": v} for k, v in kap.items()})
    f = pg.RingForcing(kf=16, dk=2, epsilon=0.1, seed=7)
    d = pg.CombinedDamping(mu=0.1, kf=16, nu_p=nu_d, p=8)
    st = solver.initialize(torch.from_numpy(theta0).cuda(),
                           scalar_init={"a": torch.from_numpy(s0), "b": torch.from_numpy(s1)})
    print("[1000 steps N=128] gpu", flush=True)
    st = solver.multistep(st, n, dt, forcing=f, damping=d)
    so = R.Solver(go, 1.0, 0.0, 8, scheme="if")
    fo = R.RingForcing(kf=16, dk=2, epsilon=0.1)
    do = R.CombinedDamping(mu=0.1, kf=16, nu_p=nu_d, p=8)
    sto = so.initialize(theta0)
    sc = {"a": R.fft2(s0), "b": R.fft2(s1)}
    print("[1000 steps N=128] oracle", flush=True)
    for i in range(n):
        k = orng.step_key(7, i)
        sto = so.step(sto, dt, forcing=lambda y, k=k: fo(y, key=k, grid=go), damping=do)
        sc = so.step_scalars(sc, sto["theta_hat"], dt, kap)
    err = rel(st["theta_hat"], sto["theta_hat"])
    print(f"[1000 steps N=128] theta rel {err:.3e}", flush=True)
    assert err < 1e-12
    for name in kap:
        e = rel(st["scalar_state"].scalars[name], sc[name])
        print(f"[1000 steps N=128] scalar {name} rel {e:.3e}", flush=True)
        assert e < 1e-12, name
    assert int(st["step"]) == n
