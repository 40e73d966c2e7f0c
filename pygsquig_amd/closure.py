"""Learned subgrid closure for coarse gSQG runs.

pygSQuiG has no closure; this module adds one.  A coarse run at N_c misses
the part of the Jacobian that the filtered-out scales carry.  For a sharp
spectral filter F onto the coarse grid's dealiased modes, the filtered
dynamics are exactly

    d(F theta)/dt = -F J(theta) + (linear terms of F theta)
                  = -J_c(F theta) - tau,   tau = F J(theta) - J_c(F theta),

with J_c the coarse dealiased Jacobian.  A CNN learns tau from the coarse
fields (theta, u, v); the coarse solver then adds the network's tau to its
physical Jacobian before the dealiased transform in every RK4 stage.

Everything that computes runs on the GPU: the filter
(``gsqg_spectral_filter``), the Jacobians (``gsqg_rhs`` of plans without
forcing, damping or viscosity), the transforms, and the network's forward,
backward and Adam kernels (bf16 MFMA, float32 accumulation; see
include/gsqg_closure.h).
"""

from __future__ import annotations

import ctypes
import math
import weakref
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import Grid, default_device, make_grid
from .plan import Plan, _ptr, _stream
from .solver import gSQGSolver

N_PARAMS = 150081
N_CH = 64


def param_shapes():
    """(name, shape) of every parameter in the flat vector's order."""
    s = [("W0", (9, 3, N_CH)), ("b0", (N_CH,))]
    for l in range(1, 5):
        s += [(f"W{l}", (9, N_CH, N_CH)), (f"b{l}", (N_CH,))]
    s += [("W5", (9, N_CH)), ("b5", (1,))]
    return s


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform He initialisation of the convolutions (bound sqrt(6 / fan_in),
    fan_in = 9 * input channels), zero biases, and an output layer 30x
    smaller so the untrained closure starts near zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, sh in param_shapes():
        if name.startswith("b"):
            parts.append(np.zeros(sh, np.float32).ravel())
            continue
        fan_in = 9 * sh[1]
        lim = math.sqrt(6.0 / fan_in) if name != "W5" else 0.1 * math.sqrt(3.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=sh).astype(np.float32).ravel())
    out = np.concatenate(parts)
    assert out.size == N_PARAMS
    return out


class ClosureNet:
    """The network and its training state on the GPU (``gsqg_cnn_*``).

    Images are (h, w) float64 planes; a batch reads ``theta[sel[b]]``,
    ``u[sel[b]]``, ``v[sel[b]]`` (and ``target[sel[b]]``) from (n, h, w)
    device tensors.
    """

    def __init__(self, h: int, w: Optional[int] = None, max_batch: int = 1,
                 device: Optional[torch.device] = None, params=None, seed: int = 0):
        self.lib = _lib.load()
        self.h, self.w = int(h), int(w if w is not None else h)
        self.max_batch = int(max_batch)
        self.device = device or default_device()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.gsqg_cnn_create(self.h, self.w, self.max_batch,
                                                ctypes.byref(handle)), "gsqg_cnn_create")
        self.handle = handle
        self._finalizer = weakref.finalize(self, ClosureNet._destroy, self.lib, handle)
        assert self.lib.gsqg_cnn_num_params() == N_PARAMS
        self.set_params(init_params(seed) if params is None else params)

    @staticmethod
    def _destroy(lib, handle):
        try:
            lib.gsqg_cnn_destroy(handle)
        except Exception:  # pragma: no cover
            pass

    # ---------------------------------------------------------- parameters
    def set_params(self, params) -> None:
        """Copy a flat float32 parameter vector in (resets the Adam state)."""
        p = torch.as_tensor(params).to(device=self.device, dtype=torch.float32).contiguous()
        if p.numel() != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.numel()}")
        _lib.check(self.lib.gsqg_cnn_set_params(self.handle, _ptr(p), _stream()),
                   "gsqg_cnn_set_params")
        torch.cuda.current_stream(self.device).synchronize()

    def get_params(self) -> torch.Tensor:
        out = torch.empty(N_PARAMS, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.gsqg_cnn_get_params(self.handle, _ptr(out), _stream()),
                   "gsqg_cnn_get_params")
        return out

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        _lib.check(self.lib.gsqg_cnn_device_bytes(self.handle, ctypes.byref(b)),
                   "gsqg_cnn_device_bytes")
        return int(b.value)

    # -------------------------------------------------------------- batches
    def _batch(self, theta, u, v, sel, in_scale, target=None, tgt_scale=1.0):
        planes = []
        for t in (theta, u, v) + ((target,) if target is not None else ()):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise TypeError("planes must be contiguous float64 device tensors")
            if tuple(t.shape[-2:]) != (self.h, self.w):
                raise ValueError(f"planes must be (n, {self.h}, {self.w}), got {tuple(t.shape)}")
            planes.append(t)
        n = 1 if theta.dim() == 2 else theta.shape[0]
        for t in planes[1:]:
            if t.shape != theta.shape:
                raise ValueError("theta, u, v (and target) must have one shape")
        b = _lib.GsqgCnnBatch()
        b.theta, b.u, b.v = theta.data_ptr(), u.data_ptr(), v.data_ptr()
        b.img_stride = self.h * self.w
        if sel is not None:
            if sel.dtype != torch.int32 or not sel.is_cuda:
                raise TypeError("sel must be an int32 device tensor")
            b.sel_dev = sel.data_ptr()
            b.batch = int(sel.numel())
        else:
            b.sel_dev = None
            b.batch = n
        for c in range(3):
            b.in_scale[c] = float(in_scale[c])
        b.target = target.data_ptr() if target is not None else None
        b.tgt_scale = float(tgt_scale)
        return b

    def forward(self, theta, u, v, in_scale, out_scale: float = 1.0, sel=None) -> torch.Tensor:
        """tau = out_scale * net(theta * s0, u * s1, v * s2): (batch, h, w) float32."""
        b = self._batch(theta, u, v, sel, in_scale)
        tau = torch.empty((b.batch, self.h, self.w), dtype=torch.float32, device=self.device)
        _lib.check(self.lib.gsqg_cnn_forward(self.handle, ctypes.byref(b), float(out_scale),
                                             _ptr(tau), _stream()), "gsqg_cnn_forward")
        return tau

    def gradient(self, theta, u, v, target, in_scale, tgt_scale: float = 1.0, sel=None):
        """(gradient of the mean squared error in every parameter, loss)."""
        b = self._batch(theta, u, v, sel, in_scale, target, tgt_scale)
        g = torch.empty(N_PARAMS, dtype=torch.float32, device=self.device)
        loss = ctypes.c_double(0.0)
        _lib.check(self.lib.gsqg_cnn_gradient(self.handle, ctypes.byref(b), _ptr(g),
                                              ctypes.byref(loss), _stream()), "gsqg_cnn_gradient")
        return g, loss.value

    def train_step(self, theta, u, v, target, in_scale, tgt_scale: float = 1.0, sel=None,
                   lr: float = 1e-3, want_loss: bool = False) -> Optional[float]:
        """One Adam step on the batch; returns the batch loss before the step
        when ``want_loss`` (that synchronises)."""
        b = self._batch(theta, u, v, sel, in_scale, target, tgt_scale)
        loss = ctypes.c_double(0.0)
        _lib.check(self.lib.gsqg_cnn_train_step(self.handle, ctypes.byref(b), float(lr),
                                                ctypes.byref(loss) if want_loss else None,
                                                _stream()), "gsqg_cnn_train_step")
        return loss.value if want_loss else None


# ------------------------------------------------------------- filtering
def spectral_filter(X: torch.Tensor, grid_c: Grid, scale: Optional[float] = None,
                    add: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Sharp filter of a full fine spectrum onto the coarse grid's dealiased
    modes (|k| < (2/3) pi N_c / L): the unnormalised transform of the
    filtered field on the coarse grid, plus ``add`` on those modes."""
    lib = _lib.load()
    nf, nc = X.shape[0], grid_c.N
    if scale is None:
        scale = (nc / nf) ** 2
    mask = _mask_u8(grid_c)
    out = torch.empty((nc, nc), dtype=torch.complex128, device=grid_c.device)
    _lib.check(lib.gsqg_spectral_filter(_ptr(X), nf, _ptr(out), nc, float(scale), _ptr(mask),
                                        _ptr(add) if add is not None else None, _stream()),
               "gsqg_spectral_filter")
    return out


_MASKS: dict = {}


def _mask_u8(grid: Grid) -> torch.Tensor:
    key = (grid.N, grid.L, str(grid.device))
    m = _MASKS.get(key)
    if m is None:
        m = grid.dealias_mask.to(torch.uint8).contiguous()
        _MASKS[key] = m
    return m


class SubgridSampler:
    """Filtered fields and the subgrid flux of fine states.

    ``sample(theta_hat_f)`` returns the coarse spectrum F theta and the
    physical float64 (theta, u, v, tau) on the coarse grid, with
    tau = ifft2(F J_f - J_c(F theta)) (both Jacobians dealiased on their own
    grids)."""

    def __init__(self, grid_f: Grid, grid_c: Grid, alpha: float):
        self.grid_f, self.grid_c, self.alpha = grid_f, grid_c, alpha
        self.jac_f = Plan(grid_f, alpha, 0.0, 8)
        self.jac_c = Plan(grid_c, alpha, 0.0, 8)
        self._rf = None

    def sample(self, theta_hat_f: torch.Tensor):
        gc = self.grid_c
        nc, nf = gc.N, self.grid_f.N
        if self._rf is None:
            self._rf = torch.empty((nf, nf), dtype=torch.complex128, device=gc.device)
        thb = spectral_filter(theta_hat_f, gc)
        self.jac_f.rhs(theta_hat_f, self._rf)          # -mask_f J_f
        rc = torch.empty_like(thb)
        self.jac_c.rhs(thb, rc)                        # -mask_c J_c(F theta)
        tau_hat = spectral_filter(self._rf, gc, scale=-(nc / nf) ** 2, add=rc)
        th = torch.empty((nc, nc), dtype=torch.float64, device=gc.device)
        u, v, tau = torch.empty_like(th), torch.empty_like(th), torch.empty_like(th)
        self.jac_c.ifft2(thb, th)
        self.jac_c.velocity(thb, u, v)
        self.jac_c.ifft2(tau_hat, tau)
        return thb, th, u, v, tau


class SnapshotStore:
    """Device-resident training set: (capacity, N_c, N_c) float64 planes."""

    def __init__(self, capacity: int, n: int, device: torch.device):
        self.capacity = capacity
        self.theta = torch.empty((capacity, n, n), dtype=torch.float64, device=device)
        self.u = torch.empty_like(self.theta)
        self.v = torch.empty_like(self.theta)
        self.tau = torch.empty_like(self.theta)
        self.count = 0

    def add(self, th, u, v, tau) -> None:
        if self.count >= self.capacity:
            raise ValueError("snapshot store is full")
        i = self.count
        self.theta[i].copy_(th)
        self.u[i].copy_(u)
        self.v[i].copy_(v)
        self.tau[i].copy_(tau)
        self.count += 1

    def scales(self):
        """(input scales 1/rms of theta, u, v; target scale 1/rms tau)."""
        n = self.count
        rms = [float(torch.sqrt(torch.mean(t[:n] ** 2))) for t in (self.theta, self.u, self.v,
                                                                     self.tau)]
        return tuple(1.0 / max(r, 1e-300) for r in rms[:3]), 1.0 / max(rms[3], 1e-300)


def train(net: ClosureNet, store: SnapshotStore, steps: int, batch: int = 32, lr: float = 1e-3,
          in_scale=None, tgt_scale=None, seed: int = 0, log_every: int = 0, log=print):
    """Adam on random batches of the store; returns [(step, loss)] at the
    logged steps (every ``log_every`` steps, and the last)."""
    if in_scale is None or tgt_scale is None:
        s_in, s_t = store.scales()
        in_scale = s_in if in_scale is None else in_scale
        tgt_scale = s_t if tgt_scale is None else tgt_scale
    gen = torch.Generator(device=net.device)
    gen.manual_seed(seed)
    hist = []
    n = store.count
    th, u, v, tau = (t[:n] for t in (store.theta, store.u, store.v, store.tau))
    for s in range(steps):
        sel = torch.randint(0, n, (batch,), generator=gen, device=net.device, dtype=torch.int32)
        want = (log_every and (s % log_every == 0)) or s == steps - 1
        loss = net.train_step(th, u, v, tau, in_scale, tgt_scale, sel=sel, lr=lr, want_loss=want)
        if want:
            hist.append((s, loss))
            if log is not None and log_every:
                log(f"closure train step {s}: loss {loss:.4e}")
    return hist


class ClosedSolver(gSQGSolver):
    """gSQGSolver whose every plan carries the learned closure:
    tau = out_scale * net(theta * s0, u * s1, v * s2) is added to the
    physical Jacobian before its dealiased transform, fresh in every RK4
    stage (``every=0``) or from the first stage of every ``every``-th step
    (1, 2, 4, 8 or 16) and held in between."""

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8, *,
                 net: ClosureNet, in_scale: Sequence[float], out_scale: float,
                 scheme: str = "if", every: int = 0):
        super().__init__(grid, alpha, nu_p, p, scheme=scheme)
        if net.h != grid.N or net.w != grid.N:
            raise ValueError(f"network is {net.h} x {net.w}, grid is {grid.N} x {grid.N}")
        self.net = net
        self.in_scale = tuple(float(s) for s in in_scale)
        self.out_scale = float(out_scale)
        self.every = int(every)

    def _step_generic(self, state, dt, forcing, damping, kwargs):
        raise NotImplementedError("the closed solver runs the fused kernels only: pass a "
                                  "RingForcing with tau_f = 0 and a CombinedDamping")

    def _plan(self, forcing, damping, combine=None, kappas=None) -> Plan:
        plan = super()._plan(forcing, damping, combine, kappas)
        if getattr(plan, "closure", None) is not self.net:
            attach_closure(plan, self.net, self.in_scale, self.out_scale, self.every)
        return plan


def attach_closure(plan: Plan, net: Optional[ClosureNet], in_scale=(1.0, 1.0, 1.0),
                   out_scale: float = 1.0, every: int = 0) -> None:
    """Attach (or with ``net=None`` detach) the closure to a plan."""
    lib = _lib.load()
    sc = (ctypes.c_double * 3)(*[float(s) for s in in_scale])
    _lib.check(lib.gsqg_plan_attach_closure(plan.handle, net.handle if net is not None else None,
                                            sc, float(out_scale), int(every)),
               "gsqg_plan_attach_closure")
    plan.closure = net  # keeps the network alive while attached
