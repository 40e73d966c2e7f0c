"""GPU parity of the learned closure against the CPU oracle
(oracle/closure_cnn.py): the network's forward pass, its loss and every
parameter gradient (bf16 activations: agreement to rounding flips, a few
1e-3 relative), the Adam step, the sharp spectral filter and subgrid flux
(float64, 1e-12), and the closed solver's right-hand side and steps."""

import numpy as np
import pytest
import torch

from oracle import closure_cnn as O
from oracle import ics
from oracle import reference_sqg as R

pytestmark = pytest.mark.gpu

FWD_TOL = 5e-3    # relative L2 of tau: bf16 rounding flips from fp32 summation order
GRAD_TOL = 2e-2   # relative L2 of each gradient tensor


@pytest.fixture(scope="module")
def pg():
    import pygsquig_amd as pg

    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    return pg


def _data(H, W, n, seed):
    rng = np.random.default_rng(seed)
    th, u, v = (rng.standard_normal((n, H, W)) for _ in range(3))
    tg = 0.3 * rng.standard_normal((n, H, W))
    return th, u, v, tg


def _params(seed):
    p = O.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p += (0.02 * rng.standard_normal(p.size)).astype(np.float32)  # non-zero biases
    return p


def _cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


@pytest.mark.parametrize("H,W,sel", [(16, 64, [2, 0]), (64, 64, [1]), (8, 128, [0, 1, 2])])
def test_forward_vs_oracle(pg, H, W, sel):
    from pygsquig_amd.closure import ClosureNet

    th, u, v, _ = _data(H, W, 3, seed=H + W)
    params = _params(1)
    net = ClosureNet(H, W, max_batch=len(sel), params=params)
    scale = (0.7, 1.3, 0.9)
    s = torch.tensor(sel, dtype=torch.int32, device="cuda")
    got = net.forward(_cuda(th), _cuda(u), _cuda(v), scale, 2.5, sel=s).cpu().numpy()
    want, _ = O.forward(params, np.stack([th[sel], u[sel], v[sel]], axis=1), scale, 2.5)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"forward H={H} W={W} B={len(sel)}: rel L2 {err:.2e}")
    assert err < FWD_TOL


def test_gradient_and_adam_vs_oracle(pg):
    from pygsquig_amd.closure import ClosureNet

    H, W, sel = 16, 64, [1, 2]
    th, u, v, tg = _data(H, W, 3, seed=5)
    params = _params(2)
    net = ClosureNet(H, W, max_batch=2, params=params)
    scale, tscale = (0.8, 1.2, 1.1), 1.7
    s = torch.tensor(sel, dtype=torch.int32, device="cuda")
    args = (_cuda(th), _cuda(u), _cuda(v), _cuda(tg))
    g, loss = net.gradient(*args, scale, tscale, sel=s)
    g = g.cpu().numpy()
    tau, cache = O.forward(params, np.stack([th[sel], u[sel], v[sel]], axis=1), scale, 1.0)
    ol, dtau = O.loss_grad(tau, (tg[sel] * tscale).astype(np.float32))
    og = O.backward(params, cache, dtau)
    assert abs(loss - ol) <= 1e-4 * ol, (loss, ol)
    gu, ou = O.unflatten(g), O.unflatten(og)
    for k in ou:
        e = np.linalg.norm(gu[k] - ou[k]) / np.linalg.norm(ou[k])
        print(f"grad {k}: rel L2 {e:.2e}")
        assert e < GRAD_TOL, k
    # one Adam step from these gradients
    net.train_step(*args, scale, tscale, sel=s, lr=1e-3)
    newp = net.get_params().cpu().numpy()
    w1, _, _ = O.adam(params, g, np.zeros_like(params), np.zeros_like(params), 1, lr=1e-3)
    assert np.abs(newp - w1).max() <= 1e-6, np.abs(newp - w1).max()


def test_training_reduces_loss(pg):
    from pygsquig_amd.closure import ClosureNet

    H = W = 64
    th, u, v, _ = _data(H, W, 8, seed=9)
    target = th * u  # a local nonlinear map the network can fit
    net = ClosureNet(H, W, max_batch=4, seed=3)
    args = (_cuda(th), _cuda(u), _cuda(v), _cuda(target))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    losses = []
    for s in range(60):
        sel = torch.randint(0, 8, (4,), generator=gen, device="cuda", dtype=torch.int32)
        losses.append(net.train_step(*args, (1, 1, 1), 1.0, sel=sel, lr=2e-3, want_loss=True))
    print("loss", losses[0], "->", losses[-1])
    assert np.isfinite(losses).all()
    assert np.mean(losses[-5:]) < 0.5 * losses[0]


def test_spectral_filter_and_subgrid_flux_vs_oracle(pg):
    from pygsquig_amd.closure import SubgridSampler, spectral_filter

    Nf, Nc, L, alpha = 128, 64, 2 * np.pi, 1.0
    gf, gc = pg.make_grid(Nf, L), pg.make_grid(Nc, L)
    theta = 0.5 * ics.kolmogorov_53(Nf, seed=12)
    X = np.fft.fft2(theta)
    ogf, ogc = R.make_grid(Nf, L), R.make_grid(Nc, L)
    idx = np.r_[0:Nc // 2, Nf - Nc // 2:Nf]

    def filt(Y):
        return ogc.dealias_mask * (Y[np.ix_(idx, idx)] * (Nc / Nf) ** 2)

    got = spectral_filter(_cuda(X, torch.complex128), gc).cpu().numpy()
    assert np.abs(got - filt(X)).max() <= 1e-13 * np.abs(filt(X)).max()
    # subgrid flux: F J_f - J_c(F theta), physical
    samp = SubgridSampler(gf, gc, alpha)
    thb, th, u, v, tau = (t.cpu().numpy() for t in samp.sample(_cuda(X, torch.complex128)))
    Jf = -R.core_rhs(X, ogf, alpha, 0.0, 8)
    Xb = filt(X)
    Jc = -R.core_rhs(Xb, ogc, alpha, 0.0, 8)
    tau_o = R.ifft2(filt(Jf) - ogc.dealias_mask * Jc)
    uo, vo = R.compute_velocity_from_theta(Xb, ogc, alpha)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    print("tau rel", rel(tau, tau_o), "rms tau / rms J_c", np.std(tau_o) / np.std(R.ifft2(Jc)))
    assert rel(thb, Xb) < 1e-13
    assert rel(th, R.ifft2(Xb)) < 1e-13
    assert rel(u, uo) < 1e-12 and rel(v, vo) < 1e-12
    assert rel(tau, tau_o) < 1e-12


def _closed_setup(pg, N=64, scheme="if", every=0):
    from pygsquig_amd.closure import ClosureNet, attach_closure
    from pygsquig_amd.plan import Plan
    from pygsquig_amd import _lib

    L = 2 * np.pi
    g = pg.make_grid(N, L)
    params = O.init_params(9)
    net = ClosureNet(N, N, max_batch=1, params=params)
    plan = Plan(g, 1.0, 1e-12, 8, scheme=_lib.SCHEME_IF if scheme == "if" else _lib.SCHEME_EXPLICIT)
    in_scale, out_scale = (2.0, 1.5, 1.5), 0.8
    attach_closure(plan, net, in_scale, out_scale, every)
    return g, plan, net, params, in_scale, out_scale


def test_closed_rhs_vs_oracle(pg):
    from pygsquig_amd.closure import attach_closure

    N, L = 64, 2 * np.pi
    g, plan, net, params, in_scale, out_scale = _closed_setup(pg, N)
    theta0 = 0.5 * ics.kolmogorov_53(N, seed=4)
    th = _cuda(np.fft.fft2(theta0), torch.complex128)
    out = torch.empty_like(th)
    plan.rhs(th, out)
    closed = out.cpu().numpy()
    attach_closure(plan, None)
    plan.rhs(th, out)
    plain = out.cpu().numpy()
    go = R.make_grid(N, L)
    X = th.cpu().numpy()
    ro = R.Solver(go, 1.0, 1e-12, 8).compute_rhs(X)
    cl = O.closure_rhs(X, go, 1.0, params, in_scale, out_scale)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    print("closure share", np.linalg.norm(cl) / np.linalg.norm(ro),
          "plain rel", rel(plain, ro), "closure term rel", rel(closed - plain, cl))
    assert rel(plain, ro) < 1e-12
    assert rel(closed - plain, cl) < FWD_TOL


@pytest.mark.parametrize("scheme,every,n", [("if", 0, 5), ("explicit", 0, 5), ("if", 1, 5),
                                            ("if", 4, 21)])
def test_closed_steps_vs_oracle(pg, scheme, every, n):
    N, L, dt = 64, 2 * np.pi, 1e-3
    g, plan, net, params, in_scale, out_scale = _closed_setup(pg, N, scheme, every)
    theta0 = 0.5 * ics.kolmogorov_53(N, seed=4)
    th = _cuda(np.fft.fft2(theta0), torch.complex128)
    from pygsquig_amd import _lib

    plan.load_theta(th)
    plan.advance(dt, n, _lib.KEY_EXPLICIT, 0, 0, use_graph=True)
    out = torch.empty_like(th)
    plan.store_theta(out)
    got = out.cpu().numpy()
    go = R.make_grid(N, L)
    so = R.Solver(go, 1.0, 1e-12, 8, scheme=scheme)
    cl = lambda y: O.closure_rhs(y, go, 1.0, params, in_scale, out_scale)  # noqa: E731
    st = so.initialize(theta0)
    plain = so.initialize(theta0)
    held = None
    for s in range(n):
        if every == 0:
            st = so.step(st, dt, forcing=cl)
        else:  # tau of the first stage of steps 0, 1, 1 + M, ..., held in between
            if s == 0 or (s - 1) % every == 0:
                held = cl(st["theta_hat"])
            st = so.step(st, dt, forcing=lambda y, h=held: h)
        plain = so.step(plain, dt)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    e, eff = rel(got, st["theta_hat"]), rel(st["theta_hat"], plain["theta_hat"])
    print(f"{scheme} every={every}: closed steps rel {e:.2e}; closure effect {eff:.2e}")
    assert e < 1e-2 * eff  # the closure's contribution is reproduced to 1%


def test_closed_solver_api(pg):
    from pygsquig_amd.closure import ClosedSolver, ClosureNet

    N = 64
    g = pg.make_grid(N, 2 * np.pi)
    net = ClosureNet(N, N, params=O.init_params(2))
    s = ClosedSolver(g, 1.0, 1e-12, 8, net=net, in_scale=(1, 1, 1), out_scale=0.5)
    st = s.initialize(torch.from_numpy(0.5 * ics.kolmogorov_53(N, seed=1)))
    f = pg.RingForcing(kf=8, dk=2, epsilon=0.1)
    d = pg.CombinedDamping(mu=0.1, kf=8, nu_p=1e-10, p=8)
    st2 = s.multistep(st, 20, 1e-3, forcing=f, damping=d)
    assert torch.isfinite(torch.view_as_real(st2["theta_hat"])).all()
    base = pg.gSQGSolver(g, 1.0, 1e-12, 8, scheme="if").multistep(st, 20, 1e-3, forcing=f,
                                                                   damping=d)
    diff = torch.linalg.norm(st2["theta_hat"] - base["theta_hat"]) / torch.linalg.norm(
        base["theta_hat"])
    assert 1e-8 < float(diff) < 0.5
    with pytest.raises(ValueError):
        ClosedSolver(pg.make_grid(128, 2 * np.pi), 1.0, net=net, in_scale=(1, 1, 1), out_scale=1)
