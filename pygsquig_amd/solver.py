"""gSQG solvers (pygsquig/core/solver.py, solver_optimized.py,
solver_with_scalars.py counterparts).

``step`` keeps the reference's functional contract -- it takes a state dict
and returns a new one with a full (N, N) complex128 ``theta_hat`` -- while the
arithmetic runs in the fused HIP kernels on a device-resident copy of the
state.  When the ``theta_hat`` passed in is the very tensor a previous call
returned (same object, unmodified), the device copy is reused and the
full-to-half conversion is skipped.

Forcing and damping: our ``RingForcing`` and ``CombinedDamping`` objects are
lowered into the kernels.  Any other callable takes the generic path: the
RK4 stages run with full spectra, the core RHS from the fused kernels and the
callable's output added per stage, exactly as ``compute_rhs`` does in the
reference (core/solver.py:150-180).
"""

from __future__ import annotations

import warnings
import weakref
from typing import Any, Callable, Optional

import numpy as np
import torch

from . import _lib
from .forcing import CombinedDamping, RingForcing
from .grid import Grid, fft2, ifft2
from .plan import Plan
from .rng import key_to_u64, step_key

State = dict

_SCHEMES = {"explicit": _lib.SCHEME_EXPLICIT, "if": _lib.SCHEME_IF}
# Anti-Hermitian input energy off the forcing ring that is treated as
# round-off (relative to |theta_hat|^2).
_OFFRING_TOL = 1e-24


def _to_device_c128(a, grid: Grid) -> torch.Tensor:
    t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
    t = t.to(device=grid.device, dtype=torch.complex128)
    if t.shape != (grid.N, grid.N):
        raise ValueError(f"theta_hat must have shape {(grid.N, grid.N)}, got {tuple(t.shape)}")
    return t.contiguous()


def _scalar_tensor(v, dtype):
    if isinstance(v, torch.Tensor):
        return v.detach().to("cpu", dtype)
    return torch.tensor(v, dtype=dtype)


class gSQGSolver:
    """core/solver.py:61-245.

    Extra keyword ``scheme``: "explicit" (the reference's RK4 with the
    hyperviscous term inside every stage) or "if" (integrating-factor RK4,
    exact for the linear terms; not in the reference)."""

    _combine = _lib.COMBINE_STEP
    _hv_dc_zero = True  # core/operators.py:205

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8, *,
                 scheme: str = "explicit"):
        self.grid = grid
        self.alpha = alpha
        self.nu_p = nu_p
        self.p = p
        if not -2 <= alpha <= 2:
            raise ValueError(f"alpha must be in [-2, 2], got {alpha}")
        if p not in [2, 4, 8]:
            raise ValueError(f"p must be 2, 4, or 8, got {p}")
        if scheme not in _SCHEMES:
            raise ValueError(f"scheme must be 'explicit' or 'if', got {scheme!r}")
        self.scheme = scheme
        self._plans: dict = {}

    # ------------------------------------------------------------ plans
    def _kappas(self):
        return ()

    def _plan(self, forcing: Optional[RingForcing], damping: Optional[CombinedDamping],
              combine: Optional[int] = None, kappas=None) -> Plan:
        kap = self._kappas() if kappas is None else tuple(kappas)
        comb = self._combine if combine is None else combine
        key = (forcing.signature() if forcing is not None else None,
               damping.signature() if damping is not None else None, comb, kap)
        plan = self._plans.get(key)
        if plan is None:
            plan = Plan(self.grid, self.alpha, self.nu_p, self.p, scheme=_SCHEMES[self.scheme],
                        combine=comb, hv_dc_zero=self._hv_dc_zero, damping=damping,
                        forcing=forcing,
                        kappas=kap)
            self._plans[key] = plan
        return plan

    def autotune(self, dt: float, forcing=None, damping=None, n_steps: int = 32,
                 repeats: int = 3):
        """Pick the column/row tile shapes of the plan that ``multistep`` with
        these forcing and damping arguments runs (``Plan.autotune``: every
        shape pair timed on ``n_steps`` graphed steps, best of ``repeats``;
        the plan's state is restored).  Returns ((col, row), {(col, row): ms})."""
        return self._plan(forcing, damping).autotune(dt, n_steps, repeats)

    # ------------------------------------------------------------- state
    def initialize(self, theta0=None, seed: Optional[int] = None) -> State:
        """core/solver.py:96-124.  The seed path draws a normal field from
        ``numpy.random.default_rng(seed)`` (the reference uses jax.random) and
        keeps |k| < (N/8) 2 pi / L."""
        g = self.grid
        if theta0 is not None:
            th = fft2(theta0, g)
        elif seed is not None:
            rs = np.random.default_rng(seed)
            th = fft2(rs.standard_normal((g.N, g.N)), g)
            kc = g.N // 8
            th = th * (torch.sqrt(g.k2) < kc * 2 * np.pi / g.L)
        else:
            th = torch.zeros((g.N, g.N), dtype=torch.complex128, device=g.device)
        return {"theta_hat": th, "time": torch.tensor(0.0, dtype=torch.float64),
                "step": torch.tensor(0, dtype=torch.int64)}

    def compute_velocity(self, theta_hat):
        """core/solver.py:126-136."""
        plan = self._plan(None, None)
        th = _to_device_c128(theta_hat, self.grid)
        u = torch.empty((self.grid.N, self.grid.N), dtype=torch.float64, device=self.grid.device)
        v = torch.empty_like(u)
        plan.velocity(th, u, v)
        return u, v

    def compute_hyperviscosity(self, theta_hat):
        """core/solver.py:138-148."""
        from .operators import hyperviscosity

        return hyperviscosity(_to_device_c128(theta_hat, self.grid), self.grid, self.nu_p, self.p)

    def _split_plugins(self, forcing, damping, kwargs):
        fused_f = forcing if isinstance(forcing, RingForcing) and forcing.tau_f == 0 else None
        fused_d = damping if isinstance(damping, CombinedDamping) else None
        generic = (forcing is not None and fused_f is None) or (damping is not None and fused_d is None)
        return fused_f, fused_d, generic

    def compute_rhs(self, theta_hat, forcing: Optional[Callable] = None,
                    damping: Optional[Callable] = None, **kwargs):
        """core/solver.py:150-180."""
        th = _to_device_c128(theta_hat, self.grid)
        fused_f, fused_d, generic = self._split_plugins(forcing, damping, kwargs)
        if generic:
            out = self._core_rhs_full(th)
            if forcing is not None:
                out = out + self._call_plugin(forcing, th, kwargs)
            if damping is not None:
                out = out + self._call_plugin(damping, th, kwargs)
            return out
        key = kwargs.get("key")
        k = key_to_u64(key) if key is not None else (fused_f.seed if fused_f else 0)
        out = torch.empty_like(th)
        self._plan(fused_f, fused_d).rhs(th, out, k)
        return out

    def _call_plugin(self, fn, th, kwargs, dt=None, step=None):
        """One per-stage call of a forcing or damping plug-in.  Our RingForcing
        and CombinedDamping take the arguments the reference's call sites
        bind (scripts/run.py:380-382): the key (``step_key(seed, step)`` when
        none is given, as in the fused path), dt, the grid.  Any other
        callable gets ``fn(theta_hat, **kwargs)`` (core/solver.py:174,178)."""
        if isinstance(fn, RingForcing):
            key = kwargs.get("key")
            if key is None:
                key = step_key(fn.seed, step) if step is not None else fn.seed
            return fn(th, key=key, dt=kwargs.get("dt", dt), grid=self.grid)
        if isinstance(fn, CombinedDamping):
            return fn(th, grid=self.grid)
        return fn(th, **kwargs)

    def _advection_plan(self) -> Plan:
        """Plan of the bare advection term fft2(-J) (no dissipation)."""
        plan = self._plans.get("advection")
        if plan is None:
            plan = Plan(self.grid, self.alpha, 0.0, self.p)
            self._plans["advection"] = plan
        return plan

    def _hv_operator(self) -> torch.Tensor:
        """-nu_p |k|^p over the full plane, 0 at k = 0 (core/operators.py:202-205)."""
        op = self._plans.get("hv_operator")
        if op is None:
            op = -self.nu_p * (self.grid.k2 ** (self.p / 2))
            op[0, 0] = 0.0
            self._plans["hv_operator"] = op
        return op

    def _core_rhs_full(self, th: torch.Tensor) -> torch.Tensor:
        """advection_hat + hyperviscosity(theta_hat) of a full spectrum that
        may carry an anti-Hermitian part (core/solver.py:28-57): the
        advection sees Re ifft2(theta_hat) only, the hyperviscosity acts on
        every mode."""
        out = torch.empty_like(th)
        self._advection_plan().rhs(th, out)
        if self.nu_p != 0:
            out = out + self._hv_operator() * th
        return out

    def _if_factors(self, dt: float, damping: Optional[CombinedDamping]):
        """exp(L dt / 2), exp(L dt) of the integrating-factor scheme's linear
        operator L: the solver's hyperviscosity plus the CombinedDamping's
        terms, 0 at k = 0 except the large-scale drag."""
        key = ("if", float(dt), damping.signature() if damping is not None else None)
        e = self._plans.get(key)
        if e is None:
            g = self.grid
            lin = self._hv_operator().clone()
            if damping is not None:
                ld = -damping.nu_p * (g.k2 ** (damping.p / 2))
                ld[0, 0] = 0.0
                if damping.mu != 0:
                    # the mask as float64: a Python float times a bool tensor
                    # would be promoted to float32
                    drag = (torch.sqrt(g.k2) < damping.kf / 2).to(torch.float64)
                    ld = ld - damping.mu * drag
                lin = lin + ld
            e = (torch.exp(lin * (0.5 * dt)), torch.exp(lin * dt))
            self._plans[key] = e
        return e

    def _step_generic(self, state, dt, forcing, damping, kwargs):
        """One RK4 step with per-stage plug-in calls on full spectra.  The
        explicit scheme is core/solver.py:182-214 with compute_rhs; the IF
        scheme is the integrating-factor (Lawson) RK4 of the fused kernels,
        with the solver's hyperviscosity and a CombinedDamping in its linear
        operator and the forcing called on each stage state."""
        th = _to_device_c128(state["theta_hat"], self.grid)
        step = int(_scalar_tensor(state.get("step", 0), torch.int64))
        if self.scheme == "if":
            if damping is not None and not isinstance(damping, CombinedDamping):
                raise ValueError("the IF scheme needs a CombinedDamping object: a generic "
                                 "damping callable has no linear operator to integrate")
            e1, e2 = self._if_factors(dt, damping)

            def nonlin(y):
                out = torch.empty_like(y)
                self._advection_plan().rhs(y, out)
                if forcing is not None:
                    out = out + self._call_plugin(forcing, y, kwargs, dt, step)
                return out

            k1 = nonlin(th)
            k2 = nonlin(e1 * (th + (0.5 * dt) * k1))
            k3 = nonlin(e1 * th + (0.5 * dt) * k2)
            k4 = nonlin(e2 * th + dt * (e1 * k3))
            acc = e2 * k1
            acc = acc + 2.0 * (e1 * k2)
            acc = acc + 2.0 * (e1 * k3)
            return e2 * th + (dt / 6.0) * (acc + k4)

        def rhs(y):
            out = self._core_rhs_full(y)
            if forcing is not None:
                out = out + self._call_plugin(forcing, y, kwargs, dt, step)
            if damping is not None:
                out = out + self._call_plugin(damping, y, kwargs, dt, step)
            return out

        k1 = rhs(th)
        k2 = rhs(th + 0.5 * dt * k1)
        k3 = rhs(th + 0.5 * dt * k2)
        k4 = rhs(th + dt * k3)
        return th + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6

    def _load(self, plan: Plan, th: torch.Tensor) -> None:
        ref = plan.state_ref() if plan.state_ref is not None else None
        if ref is th and plan.state_version == th._version:
            return
        off = plan.load_theta(th, check=True)
        if off > 0:
            tot = float(torch.sum(torch.abs(th) ** 2))
            if tot > 0 and off > _OFFRING_TOL * tot:
                warnings.warn(
                    f"theta_hat has an anti-Hermitian part off the forcing ring "
                    f"(relative energy {off / tot:.3e}); it is dropped", RuntimeWarning)

    def _remember(self, plan: Plan, out: torch.Tensor) -> None:
        plan.state_ref = weakref.ref(out)
        plan.state_version = out._version

    def _advance_fused(self, state, dt, n_steps, forcing, damping, key, combine=None,
                       use_graph=True) -> torch.Tensor:
        plan = self._plan(forcing, damping, combine)
        th = _to_device_c128(state["theta_hat"], self.grid)
        self._load(plan, th)
        if key is not None:
            mode, k, step0 = _lib.KEY_EXPLICIT, key_to_u64(key), 0
        else:
            step0 = int(_scalar_tensor(state.get("step", 0), torch.int64))
            mode, k = _lib.KEY_DERIVED, (forcing.seed if forcing is not None else 0)
        plan.advance(dt, n_steps, mode, k, step0, use_graph)
        out = torch.empty_like(th)
        plan.store_theta(out)
        self._remember(plan, out)
        return out

    def _combine_for(self, forcing, damping) -> int:
        """The RK4 combine of a step: the solver's own without plug-ins; the
        reference's ``rk4_step`` order with them (core/solver_optimized.py:152-164
        falls back to it)."""
        if forcing is None and damping is None:
            return self._combine
        return _lib.COMBINE_STEP

    def step(self, state: State, dt: float, forcing: Optional[Callable] = None,
             damping: Optional[Callable] = None, **kwargs) -> State:
        """core/solver.py:182-214."""
        fused_f, fused_d, generic = self._split_plugins(forcing, damping, kwargs)
        if generic:
            new = self._step_generic(state, dt, forcing, damping, kwargs)
        else:
            new = self._advance_fused(state, dt, 1, fused_f, fused_d, kwargs.get("key"),
                                      combine=self._combine_for(forcing, damping))
        return {"theta_hat": new, "time": state["time"] + dt, "step": state["step"] + 1}

    def multistep(self, state: State, n_steps: int, dt: float, forcing=None, damping=None,
                  **kwargs) -> State:
        """n_steps device-resident steps (core/solver_optimized.py:168-199
        for the unforced case; forcing/damping objects are also accepted,
        with per-step keys step_key(forcing.seed, step), and every step then
        combines as ``step`` does, so ``multistep(n)`` equals n ``step`` calls)."""
        fused_f, fused_d, generic = self._split_plugins(forcing, damping, kwargs)
        if generic:
            s = state
            for _ in range(n_steps):
                s = self.step(s, dt, forcing, damping, **kwargs)
            return s
        new = self._advance_fused(state, dt, n_steps, fused_f, fused_d, kwargs.get("key"),
                                  combine=self._combine_for(forcing, damping))
        t = _scalar_tensor(state["time"], torch.float64)
        for _ in range(n_steps):
            t = t + dt
        return {"theta_hat": new, "time": t, "step": state["step"] + n_steps}

    def capture_graphs(self, dt: float, forcing=None, damping=None, key=None) -> None:
        """Capture the HIP graphs that ``multistep(state, n, dt, forcing,
        damping[, key=key])`` replays, without running them
        (gsqg_graph_capture): a later call with the same arguments only
        replays.  The reference pays the equivalent cost, jit compilation of
        its lax.scan, on its first multistep call.  Generic callables run no
        graphs; nothing is captured for them."""
        fused_f, fused_d, generic = self._split_plugins(forcing, damping, {})
        if generic:
            return
        plan = self._plan(fused_f, fused_d, self._combine_for(forcing, damping))
        if key is not None:
            mode, k = _lib.KEY_EXPLICIT, key_to_u64(key)
        else:
            mode, k = _lib.KEY_DERIVED, (fused_f.seed if fused_f is not None else 0)
        plan.graph_capture(dt, mode, k)

    def get_diagnostics(self, state: State) -> dict:
        """core/solver.py:216-245: kinetic energy 0.5 <u^2 + v^2>, enstrophy
        0.5 <theta^2> and max |theta| in physical space, reduced on the GPU
        (gsqg_physical_diagnostics)."""
        from .operators import velocity_plan

        th = _to_device_c128(state["theta_hat"], self.grid).contiguous()
        ke, ens, mx = velocity_plan(self.grid, self.alpha).physical_diagnostics(th)
        return {"kinetic_energy": ke, "enstrophy": ens, "max_theta": mx,
                "time": float(state["time"]), "step": int(state["step"])}

    def check_finite(self, forcing=None, damping=None) -> bool:
        """True if no non-finite state was produced since the last check."""
        return not self._plan(forcing, damping).nonfinite()


class gSQGSolverOptimized(gSQGSolver):
    """core/solver_optimized.py:83-226: the unforced step combines with
    ``theta_hat + (dt/6.0)*(...)`` (:62) and hyperviscosity is not zeroed at
    k=0 (:42-45); with forcing or damping the step falls back to the
    ``rk4_step`` combine (:152-164), still without the k=0 zeroing."""

    _combine = _lib.COMBINE_MULTISTEP
    _hv_dc_zero = False  # core/solver_optimized.py:42-43

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8, **kw):
        super().__init__(grid, alpha, nu_p, p, **kw)
        self.params = {"alpha": alpha, "nu_p": nu_p, "p": float(p)}
