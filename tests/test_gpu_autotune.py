"""Shape variants and the autotuner: every (column, row) shape pair advances
a plan to the same state as the default shapes -- bit for bit between
radix-8 tiles without forcing; to rounding where a radix-16 or radix-4
kernel takes part (their transforms round differently) or with ring forcing
(the <theta f> partials are then summed in another order) -- that state
matches the oracle, and gsqg_autotune restores the state it times with."""

import numpy as np
import pytest
import torch

from oracle import ics
from oracle import reference_sqg as R

pytestmark = pytest.mark.gpu
TWO_PI = 2 * np.pi
# shape variants per grid size (gsqg.h: radix-8 tiles, plus radix 16 and radix 4
# from N = 64, and the field-group column tile from N = 64 to 512)
N_VARIANTS = {16: 3, 32: 3, 64: 6, 128: 6, 256: 6, 512: 6, 1024: 4, 2048: 3, 4096: 3, 8192: 3}


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return pg


def _ensemble_plan(pg, N, alphas, nu=1e-4):
    from pygsquig_amd import _lib
    from pygsquig_amd.plan import Plan

    g = pg.make_grid(N, TWO_PI)
    return Plan(g, alphas[0], nu, 8, scheme=_lib.SCHEME_IF, member_alphas=alphas)


def _run(plan, ths, dt, n, key_mode=0, key=0):
    if len(ths) == 1:  # a single state: load_theta keeps the ring side state
        plan.load_theta(ths[0])
    else:
        for m, th in enumerate(ths):
            plan.load_member(m, th)
    plan.advance(dt, n, key_mode, key, 0, use_graph=True)
    outs = [torch.empty_like(th) for th in ths]
    if len(ths) == 1:
        plan.store_theta(outs[0])
    else:
        for m, o in enumerate(outs):
            plan.store_member(m, o)
    return outs


@pytest.mark.parametrize("N", [64, 256, 512, 1024])
def test_shapes_identical_and_match_oracle(pg, N):
    alphas = [0.0, 1.0, 1.5]
    dt, n, nu = 1e-3, 21, 1e-4  # 21 steps: one 16-step graph, a one-step graph, eager ends
    plan = _ensemble_plan(pg, N, alphas, nu)
    th0 = [ics.kolmogorov_53(N, seed=30 + m, alpha=a, rms=0.3) for m, a in enumerate(alphas)]
    ths = [torch.from_numpy(np.fft.fft2(t)).cuda() for t in th0]
    col, row, nv = plan.shape()
    assert (col, row) == (0, 0) and nv == N_VARIANTS[N]
    radix = [plan.shape_radix(k) for k in range(nv)]
    assert radix[0] == 8 and set(radix) == {4, 8, 16}
    base = _run(plan, ths, dt, n)
    for c in range(nv):
        for r in range(nv):
            plan.set_shape(c, r)
            assert plan.shape()[:2] == (c, r)
            got = _run(plan, ths, dt, n)
            for m in range(len(alphas)):
                if radix[c] == radix[r] == 8:
                    assert torch.equal(got[m], base[m]), f"shape ({c},{r}) member {m}"
                else:
                    rel = float(torch.linalg.norm(got[m] - base[m]) / torch.linalg.norm(base[m]))
                    assert rel < 1e-13, (c, r, m, rel)
    go = R.make_grid(N, TWO_PI)
    for m, a in enumerate(alphas):
        so = R.Solver(go, a, nu, 8, scheme="if")
        st = so.initialize(th0[m])
        for _ in range(n):
            st = so.step(st, dt)
        err = np.linalg.norm(base[m].cpu().numpy() - st["theta_hat"]) / np.linalg.norm(
            st["theta_hat"])
        assert err < 1e-12, (m, err)


def test_forced_shapes_match_oracle(pg):
    """Ring forcing and damping at N = 128 (each shape pair sums the
    <theta f> partials over another block count).  The configuration keeps
    the forcing's normalisation well conditioned: every stage's rate above
    the reference's 1e-10 threshold is at most ~80x smaller than the mean
    |theta f| it sums (checked with the oracle), so rounding stays rounding."""
    from oracle import rng as orng
    from pygsquig_amd import _lib
    from pygsquig_amd.plan import Plan

    N, dt, n = 128, 5e-4, 8
    nu_d = 1e-8 * (512 / N) ** 8
    g = pg.make_grid(N, TWO_PI)
    f = pg.RingForcing(kf=16, dk=2, epsilon=0.1)
    d = pg.CombinedDamping(mu=0.1, kf=16, nu_p=nu_d, p=8)
    plan = Plan(g, 1.0, 0.0, 8, scheme=_lib.SCHEME_IF, forcing=f, damping=d)
    theta0 = ics.kolmogorov_53(N, seed=7)
    th = torch.from_numpy(np.fft.fft2(theta0)).cuda()
    base = _run(plan, [th], dt, n, _lib.KEY_DERIVED, 42)[0]
    go = R.make_grid(N, TWO_PI)
    so = R.Solver(go, 1.0, 0.0, 8, scheme="if")
    fo = R.RingForcing(kf=16, dk=2, epsilon=0.1)
    do = R.CombinedDamping(mu=0.1, kf=16, nu_p=nu_d, p=8)
    sto = so.initialize(theta0)
    for i in range(n):
        k = orng.step_key(42, i)
        sto = so.step(sto, dt, forcing=lambda y, k=k: fo(y, key=k, grid=go), damping=do)
    want = sto["theta_hat"]
    nv = plan.shape()[2]
    assert nv == N_VARIANTS[N]
    for c in range(nv):
        for r in range(nv):
            plan.set_shape(c, r)
            got = _run(plan, [th], dt, n, _lib.KEY_DERIVED, 42)[0]
            rel = float(torch.linalg.norm(got - base) / torch.linalg.norm(base))
            err = np.linalg.norm(got.cpu().numpy() - want) / np.linalg.norm(want)
            assert rel < 1e-13 and err < 1e-12, ((c, r), rel, err)


def test_autotune_restores_state_and_picks_fastest(pg):
    N, alphas = 256, [0.5, 2.0]
    plan = _ensemble_plan(pg, N, alphas)
    ths = [torch.from_numpy(np.fft.fft2(ics.kolmogorov_53(N, seed=40 + m, rms=0.3))).cuda()
           for m in range(len(alphas))]
    for m, th in enumerate(ths):
        plan.load_member(m, th)
    before = [torch.empty_like(t) for t in ths]
    for m, o in enumerate(before):
        plan.store_member(m, o)
    best, times = plan.autotune(1e-3, n_steps=8, repeats=2)
    print("autotune N=256:", best, {k: round(v, 4) for k, v in times.items()})
    assert plan.shape()[:2] == best
    assert len(times) == N_VARIANTS[N] ** 2
    assert all(np.isfinite(v) and v > 0 for v in times.values())
    assert times[best] == min(times.values())
    for m, o in enumerate(before):
        now = torch.empty_like(o)
        plan.store_member(m, now)
        assert torch.equal(now, o), f"member {m} changed by the autotuner"


def test_ensemble_autotune_and_validation(pg):
    members = [(64, 0.0), (4096, 1.0), (64, 1.5)]
    ens = pg.EnsembleSolver(members, nu_p=1e-4, p=8, scheme="if")
    th0 = [ics.kolmogorov_53(N, seed=50 + m, alpha=a, rms=0.3) for m, (N, a) in enumerate(members)]
    st = ens.initialize([torch.from_numpy(t) for t in th0])
    ref = ens.multistep(st, 5, 1e-3)
    tuned = ens.autotune(1e-3, n_steps=4, repeats=1)
    assert set(tuned) == {64, 4096}
    assert len(tuned[64][1]) == 36
    assert len(tuned[4096][1]) == 9  # radix 8, 16 and 4 above N = 1024
    got = ens.multistep(st, 5, 1e-3)
    for a, b in zip(got["theta_hat"], ref["theta_hat"]):
        rel = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        assert rel < 1e-13, rel
    with pytest.raises(ValueError, match="shape variant out of range"):
        ens.plans[4096].set_shape(3, 0)


def test_solver_autotune_with_scalars(pg):
    """gSQGSolverWithScalars.autotune tunes the plan its multistep runs; the
    tuned solver's forced steps agree with an untuned solver's."""
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt, n = 128, 5e-4, 6
    g = pg.make_grid(N, TWO_PI)
    f = pg.RingForcing(kf=16, dk=2, epsilon=0.1, seed=42)
    d = pg.CombinedDamping(mu=0.1, kf=16, nu_p=1e-8 * (512 / N) ** 8, p=8)
    sc = {"a": {"kappa": 0.0}, "b": {"kappa": 1e-3}}
    theta0 = torch.from_numpy(ics.kolmogorov_53(N, seed=7)).cuda()
    init = {"a": torch.from_numpy(ics.gaussian_blob(N, (np.pi, np.pi), TWO_PI / 16)),
            "b": torch.from_numpy(ics.gaussian_blob(N, (1.0, 2.0), TWO_PI / 16))}
    out = []
    for tune in (False, True):
        s = gSQGSolverWithScalars(g, 1.0, 0.0, 8, scheme="if", passive_scalars=sc)
        st = s.initialize(theta0, scalar_init=init)
        if tune:
            best, times = s.autotune(dt, forcing=f, damping=d, n_steps=4, repeats=1)
            assert s._plan(f, d).shape()[:2] == best and len(times) == 36
        out.append(s.multistep(st, n, dt, forcing=f, damping=d))
    a, b = out
    rel = float(torch.linalg.norm(a["theta_hat"] - b["theta_hat"]) / torch.linalg.norm(b["theta_hat"]))
    assert rel < 1e-13, rel
    for k in sc:
        x, y = a["scalar_state"].scalars[k], b["scalar_state"].scalars[k]
        assert float(torch.linalg.norm(x - y) / torch.linalg.norm(y)) < 1e-13, k
