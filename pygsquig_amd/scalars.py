"""Passive scalars (pygsquig/scalars/state.py, passive_scalar.py,
core/solver_with_scalars.py counterparts).

Scalar RK4 stages run in the same fused column/row kernels as theta, with the
velocity of the new theta frozen over the four stages
(core/solver_with_scalars.py:163-171).  A scalar with a source term
(scalars/source_terms.py, pygsquig_amd/sources.py) and coupled sources take a
per-stage path instead: the advection-diffusion right-hand side from the
kernels (gsqg_scalar_rhs), the source added on the device in torch, and the
RK4 combination of the reference.
"""

from __future__ import annotations

from collections.abc import Mapping
from dataclasses import dataclass, field
from typing import Any, NamedTuple, Optional

import numpy as np
import torch

from .grid import Grid, fft2, ifft2
from .plan import Plan
from .solver import _to_device_c128, gSQGSolver
from .sources import (ChemicalReaction, ExponentialGrowth, LinearCombination,  # noqa: F401
                      LocalizedSource, PassiveScalarError, SourceTerm, SourceTermError,
                      TimePeriodicSource, make_cooling_source, make_heating_source,
                      make_reaction_source)


class PassiveScalarState(NamedTuple):
    """scalars/state.py:13-56."""

    scalar_hat: torch.Tensor
    time: float
    total_source: float = 0.0
    total_dissipation: float = 0.0

    def to_dict(self) -> dict:
        return {"scalar_hat": self.scalar_hat, "time": self.time,
                "total_source": self.total_source, "total_dissipation": self.total_dissipation}

    @classmethod
    def from_dict(cls, d: dict) -> "PassiveScalarState":
        return cls(scalar_hat=d["scalar_hat"], time=d.get("time", 0.0),
                   total_source=d.get("total_source", 0.0),
                   total_dissipation=d.get("total_dissipation", 0.0))


class MultiScalarState(NamedTuple):
    """scalars/state.py:59-118."""

    scalars: dict
    time: float
    diagnostics: Any = {}

    def get_scalar(self, name: str) -> Optional[torch.Tensor]:
        return self.scalars.get(name)

    def update_scalar(self, name: str, scalar_hat) -> "MultiScalarState":
        s = dict(self.scalars)
        s[name] = scalar_hat
        return MultiScalarState(scalars=s, time=self.time, diagnostics=self.diagnostics)

    def to_dict(self) -> dict:
        return {"scalars": self.scalars, "time": self.time, "diagnostics": self.diagnostics}


def scalar_diagnostics(name: str, scalar_hat, grid: Grid, total_source=0.0,
                       total_dissipation=0.0) -> dict:
    """PassiveScalarEvolver.get_diagnostics (scalars/passive_scalar.py:260-288)."""
    s = ifft2(scalar_hat, grid)
    return {
        f"{name}_mean": float(torch.mean(s)),
        f"{name}_variance": float(torch.var(s, unbiased=False)),
        f"{name}_max": float(torch.max(s)),
        f"{name}_min": float(torch.min(s)),
        f"{name}_total": float(torch.sum(s) * (grid.L / grid.N) ** 2),
        f"{name}_source_integrated": total_source,
        f"{name}_dissipation_integrated": total_dissipation,
    }


class LazyScalarDiagnostics(Mapping):
    """``MultiScalarState.diagnostics`` computed on first access: the
    reference computes them eagerly every step with host syncs
    (scalars/passive_scalar.py:368); here they cost nothing unless read."""

    def __init__(self, scalars: dict, grid: Grid):
        self._scalars = scalars
        self._grid = grid
        self._cache: dict = {}

    def _get(self, name):
        if name not in self._cache:
            self._cache[name] = scalar_diagnostics(name, self._scalars[name], self._grid)
        return self._cache[name]

    def __getitem__(self, name):
        if name not in self._scalars:
            raise KeyError(name)
        return self._get(name)

    def __iter__(self):
        return iter(self._scalars)

    def __len__(self):
        return len(self._scalars)


def _check_source(source):
    """passive_scalar.py:134-136."""
    if source is not None and not isinstance(source, SourceTerm):
        raise PassiveScalarError(
            f"source_fn must be a SourceTerm instance, got {type(source).__name__}")


def _phys(f, g: Grid) -> torch.Tensor:
    return torch.as_tensor(f).to(device=g.device, dtype=torch.float64).contiguous()


def _if_rk4(y, nonlin, lin, dt):
    """Integrating-factor (Lawson) RK4 for dy/dt = lin y + nonlin(y), lin
    diagonal: the scheme of the fused kernels, in the same operation order."""
    e1 = torch.exp(lin * (0.5 * dt))
    e2 = torch.exp(lin * dt)
    k1 = nonlin(y)
    k2 = nonlin(e1 * (y + (0.5 * dt) * k1))
    k3 = nonlin(e1 * y + (0.5 * dt) * k2)
    k4 = nonlin(e2 * y + dt * (e1 * k3))
    acc = e2 * k1
    acc = acc + 2.0 * (e1 * k2)
    acc = acc + 2.0 * (e1 * k3)
    return e2 * y + (dt / 6.0) * (acc + k4)


@dataclass
class PassiveScalarEvolver:
    """scalars/passive_scalar.py:106-288."""

    grid: Grid
    kappa: float = 0.0
    source_fn: Any = None
    name: str = "scalar"

    def __post_init__(self):
        if self.kappa < 0:
            raise ValueError(f"kappa must be non-negative, got {self.kappa}")
        _check_source(self.source_fn)
        self._plan: Optional[Plan] = None
        self._plan_adv: Optional[Plan] = None  # kappa 0: advection alone (IF, sourced)
        self._plan_if: Optional[Plan] = None   # IF scheme, fused (IF, no source)
        # "if": integrating-factor RK4 with the diffusion exact, as the fused
        # kernels run the scalars of an IF solver (set by gSQGSolverWithScalars)
        self.scheme = "explicit"

    def plan(self) -> Plan:
        if self._plan is None:
            self._plan = Plan(self.grid, 1.0, 0.0, 8, kappas=(self.kappa,))
        return self._plan

    def compute_source(self, scalar_hat, t: float):
        """:172-190: fft2(S(ifft2(s_hat), grid, t)); zeros without a source."""
        g = self.grid
        sh = _to_device_c128(scalar_hat, g)
        if self.source_fn is None:
            return torch.zeros_like(sh)
        return fft2(self.source_fn(ifft2(sh, g), g, t), g)

    def compute_rhs(self, scalar_hat, u, v, t: float = 0.0):
        """:192-210: fft2(-J(s; u, v)) - kappa k^2 s_hat + source."""
        g = self.grid
        sh = _to_device_c128(scalar_hat, g)
        uu, vv = _phys(u, g), _phys(v, g)
        out = torch.empty_like(sh)
        self.plan().scalar_rhs(0, sh, uu, vv, out)
        if self.source_fn is not None:
            out = out + self.compute_source(sh, t)
        return out

    def _step_sourced(self, sh, dt, uu, vv, t):
        """:212-232 with the source: RK4 (time_integrator.py:14-36) of
        compute_rhs, the source evaluated at the step's start time t in every
        stage; with scheme "if", Lawson RK4 with the diffusion exact."""
        g = self.grid
        if self.scheme == "if":
            if self._plan_adv is None:
                self._plan_adv = Plan(g, 1.0, 0.0, 8, kappas=(0.0,))
            p0 = self._plan_adv

            def nonlin(y):
                out = torch.empty_like(y)
                p0.scalar_rhs(0, y, uu, vv, out)
                return out + fft2(self.source_fn(ifft2(y, g), g, t), g)

            return _if_rk4(sh, nonlin, -self.kappa * g.k2, dt)

        def rhs(y):
            out = torch.empty_like(y)
            self.plan().scalar_rhs(0, y, uu, vv, out)
            return out + fft2(self.source_fn(ifft2(y, g), g, t), g)

        k1 = rhs(sh)
        k2 = rhs(sh + 0.5 * dt * k1)
        k3 = rhs(sh + 0.5 * dt * k2)
        k4 = rhs(sh + dt * k3)
        return sh + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6

    def initialize(self, scalar0=None, seed: Optional[int] = None) -> PassiveScalarState:
        """:139-170 (the seed path draws from numpy.random.default_rng)."""
        g = self.grid
        if scalar0 is not None:
            sh = fft2(scalar0, g)
        elif seed is not None:
            rs = np.random.default_rng(seed)
            sh = fft2(rs.standard_normal((g.N, g.N)), g)
            kc = g.N // 4
            sh = sh * (torch.sqrt(g.k2) < kc * 2 * np.pi / g.L)
        else:
            raise ValueError("Must provide either scalar0 or seed for initialization")
        return PassiveScalarState(scalar_hat=sh, time=0.0, total_source=0.0,
                                  total_dissipation=0.0)

    def step(self, state: PassiveScalarState, dt: float, u, v) -> PassiveScalarState:
        """:212-258 with frozen physical (u, v)."""
        g = self.grid
        sh = _to_device_c128(state.scalar_hat, g)
        uu, vv = _phys(u, g), _phys(v, g)
        tsrc = state.total_source
        if self.source_fn is not None:
            out = self._step_sourced(sh, dt, uu, vv, state.time)
            # dt * mean(s * ifft2(source)) at the step's start (:234-242)
            src = self.compute_source(sh, state.time)
            tsrc = float(tsrc + dt * torch.mean(ifft2(sh, g) * ifft2(src, g)))
        elif self.scheme == "if":
            out = self._step_if_fused(sh, dt, uu, vv)
        else:
            p = self.plan()
            p.load_scalar(0, sh, check=False)
            p.advance_scalars(uu, vv, dt)
            out = torch.empty_like(sh)
            p.store_scalar(0, out)
        tdis = state.total_dissipation
        if self.kappa > 0:
            # dt * mean(s * ifft2(-kappa k2 s_hat)) (:244-249)
            s = ifft2(sh, g)
            d = ifft2(-self.kappa * g.k2 * sh, g)
            tdis = float(tdis + dt * torch.mean(s * d))
        return PassiveScalarState(scalar_hat=out, time=state.time + dt,
                                  total_source=float(tsrc),
                                  total_dissipation=float(tdis))

    def _step_if_fused(self, sh, dt, uu, vv):
        """IF step of an unsourced scalar: the fused kernels of a plan with the
        integrating-factor scheme."""
        from . import _lib

        if self._plan_if is None:
            self._plan_if = Plan(self.grid, 1.0, 0.0, 8, kappas=(self.kappa,),
                                 scheme=_lib.SCHEME_IF)
        p = self._plan_if
        p.load_scalar(0, sh, check=False)
        p.advance_scalars(uu, vv, dt)
        out = torch.empty_like(sh)
        p.store_scalar(0, out)
        return out

    def get_diagnostics(self, state: PassiveScalarState) -> dict:
        return scalar_diagnostics(self.name, state.scalar_hat, self.grid, state.total_source,
                                  state.total_dissipation)


@dataclass
class MultiSpeciesEvolver:
    """scalars/passive_scalar.py:291-403."""

    grid: Grid
    species: dict = field(default_factory=dict)
    coupled_sources: list = field(default_factory=list)

    def __post_init__(self):
        self.evolvers = {}
        for name, cfg in self.species.items():
            self.evolvers[name] = PassiveScalarEvolver(grid=self.grid, kappa=cfg.get("kappa", 0.0),
                                                       source_fn=cfg.get("source", None), name=name)
        self._plan: Optional[Plan] = None

    def kappas(self):
        return tuple(float(ev.kappa) for ev in self.evolvers.values())

    def initialize(self, initial_fields: dict) -> MultiScalarState:
        """:320-339."""
        out = {}
        for name, ev in self.evolvers.items():
            if name in initial_fields:
                out[name] = ev.initialize(scalar0=initial_fields[name]).scalar_hat
            else:
                out[name] = ev.initialize(scalar0=torch.zeros((self.grid.N, self.grid.N),
                                                              dtype=torch.float64)).scalar_hat
        return MultiScalarState(scalars=out, time=0.0, diagnostics={})

    def has_sources(self) -> bool:
        return bool(self.coupled_sources) or any(ev.source_fn is not None
                                                 for ev in self.evolvers.values())

    def set_scheme(self, scheme: str) -> None:
        for ev in self.evolvers.values():
            ev.scheme = scheme

    def _step_each(self, state: MultiScalarState, dt: float, u, v) -> MultiScalarState:
        """:341-376 species by species (the sourced path), then the coupled
        sources (:378-403): s_hat += dt * fft2(S) at time t + dt, every
        coupled source seeing the fields before any of them is added."""
        g = self.grid
        new = {}
        for name, ev in self.evolvers.items():
            st = PassiveScalarState(scalar_hat=state.scalars[name], time=state.time)
            new[name] = ev.step(st, dt, u, v).scalar_hat
        if self.coupled_sources:
            t1 = state.time + dt
            phys = {name: ifft2(sh, g) for name, sh in new.items()}
            for source_fn in self.coupled_sources:
                for name, src in source_fn(phys, g, t1).items():
                    if name in new:
                        new[name] = new[name] + dt * fft2(src, g)
        return MultiScalarState(scalars=new, time=state.time + dt,
                                diagnostics=LazyScalarDiagnostics(new, g))

    def step(self, state: MultiScalarState, dt: float, u, v) -> MultiScalarState:
        """:341-376: every scalar with the same frozen (u, v)."""
        if self.has_sources():
            return self._step_each(state, dt, u, v)
        g = self.grid
        if self._plan is None or self._plan.kappas != self.kappas():
            self._plan = Plan(g, 1.0, 0.0, 8, kappas=self.kappas())
        p = self._plan
        names = list(self.evolvers)
        for i, n in enumerate(names):
            p.load_scalar(i, _to_device_c128(state.scalars[n], g), check=False)
        uu = torch.as_tensor(u).to(device=g.device, dtype=torch.float64).contiguous()
        vv = torch.as_tensor(v).to(device=g.device, dtype=torch.float64).contiguous()
        p.advance_scalars(uu, vv, dt)
        new = {}
        for i, n in enumerate(names):
            t = torch.empty((g.N, g.N), dtype=torch.complex128, device=g.device)
            p.store_scalar(i, t)
            new[n] = t
        return MultiScalarState(scalars=new, time=state.time + dt,
                                diagnostics=LazyScalarDiagnostics(new, g))


class gSQGSolverWithScalars(gSQGSolver):
    """core/solver_with_scalars.py:57-236."""

    def __init__(self, grid: Grid, alpha: float, nu_p: float = 0.0, p: int = 8,
                 passive_scalars: Optional[dict] = None, **kw):
        super().__init__(grid, alpha, nu_p, p, **kw)
        self.scalar_evolver = (MultiSpeciesEvolver(grid=grid, species=passive_scalars)
                               if passive_scalars is not None else None)

    def _kappas(self):
        return self.scalar_evolver.kappas() if self.scalar_evolver is not None else ()

    def initialize(self, theta0=None, seed: Optional[int] = None,
                   scalar_init: Optional[dict] = None) -> dict:
        """:96-129."""
        base = super().initialize(theta0, seed)
        if self.scalar_evolver is None:
            return base
        st = dict(base)
        st["scalar_state"] = self.scalar_evolver.initialize(scalar_init or {})
        return st

    def _load_scalars(self, plan: Plan, scal: dict, names):
        for i, n in enumerate(names):
            t = _to_device_c128(scal[n], self.grid)
            ref = plan.scalar_refs.get(i)
            if ref is not None and ref[0]() is t and ref[1] == t._version:
                continue
            plan.load_scalar(i, t, check=False)

    def _step_sourced(self, state, dt, forcing, damping, kwargs):
        """:131-178 with sourced scalars: the theta step (fused), the velocity
        of the new theta_hat, then the scalar step on the per-stage path."""
        base = {k: v for k, v in state.items() if k != "scalar_state"}
        new = super().step(base, dt, forcing, damping, **kwargs)
        u, v = self.compute_velocity(new["theta_hat"])
        self.scalar_evolver.set_scheme(self.scheme)
        new["scalar_state"] = self.scalar_evolver.step(state["scalar_state"], dt, u, v)
        return new

    def _advance_with_scalars(self, state, dt, n_steps, forcing, damping, kwargs):
        import weakref

        from . import _lib
        from .rng import key_to_u64

        fused_f, fused_d, generic = self._split_plugins(forcing, damping, kwargs)
        if generic:
            raise NotImplementedError("generic forcing/damping callables with passive scalars: "
                                      "use RingForcing/CombinedDamping")
        plan = self._plan(fused_f, fused_d)
        th = _to_device_c128(state["theta_hat"], self.grid)
        self._load(plan, th)
        sst = state["scalar_state"]
        names = list(self.scalar_evolver.evolvers)
        self._load_scalars(plan, sst.scalars, names)
        key = kwargs.get("key")
        if key is not None:
            mode, k, step0 = _lib.KEY_EXPLICIT, key_to_u64(key), 0
        else:
            mode, k = _lib.KEY_DERIVED, (fused_f.seed if fused_f is not None else 0)
            step0 = int(state.get("step", 0))
        plan.advance(dt, n_steps, mode, k, step0, True)
        out = torch.empty_like(th)
        plan.store_theta(out)
        self._remember(plan, out)
        new = {}
        for i, n in enumerate(names):
            t = torch.empty_like(th)
            plan.store_scalar(i, t)
            plan.scalar_refs[i] = (weakref.ref(t), t._version)
            new[n] = t
        t_s = sst.time
        for _ in range(n_steps):
            t_s = t_s + dt
        return out, MultiScalarState(scalars=new, time=t_s,
                                     diagnostics=LazyScalarDiagnostics(new, self.grid))

    def step(self, state: dict, dt: float, forcing=None, damping=None, **kwargs) -> dict:
        """:131-178."""
        if state.get("scalar_state") is None:
            return super().step(state, dt, forcing, damping, **kwargs)
        if self.scalar_evolver.has_sources():
            return self._step_sourced(state, dt, forcing, damping, kwargs)
        base = {k: v for k, v in state.items() if k != "scalar_state"}
        th, sst = self._advance_with_scalars(state, dt, 1, forcing, damping, kwargs)
        out = {"theta_hat": th, "time": base["time"] + dt, "step": base["step"] + 1}
        out["scalar_state"] = sst
        return out

    def multistep(self, state: dict, n_steps: int, dt: float, forcing=None, damping=None,
                  **kwargs) -> dict:
        if state.get("scalar_state") is None:
            return super().multistep(state, n_steps, dt, forcing, damping, **kwargs)
        if self.scalar_evolver.has_sources():
            for _ in range(n_steps):
                state = self._step_sourced(state, dt, forcing, damping, kwargs)
            return state
        th, sst = self._advance_with_scalars(state, dt, n_steps, forcing, damping, kwargs)
        t = state["time"]
        for _ in range(n_steps):
            t = t + dt
        return {"theta_hat": th, "time": t, "step": state["step"] + n_steps, "scalar_state": sst}

    def get_diagnostics(self, state: dict) -> dict:
        """:180-210."""
        base = {k: v for k, v in state.items() if k != "scalar_state"}
        d = super().get_diagnostics(base)
        sst = state.get("scalar_state")
        if sst is not None and self.scalar_evolver is not None:
            for name, sh in sst.scalars.items():
                d.update(scalar_diagnostics(name, sh, self.grid))
        return d

    def add_passive_scalar(self, name: str, kappa: float = 0.0, source=None) -> None:
        """:212-236."""
        _check_source(source)
        cfg = {"kappa": kappa, "source": source}
        if self.scalar_evolver is None:
            self.scalar_evolver = MultiSpeciesEvolver(grid=self.grid, species={name: cfg})
            return
        if name in self.scalar_evolver.evolvers:
            raise ValueError(f"Scalar '{name}' already exists")
        self.scalar_evolver.species[name] = cfg
        self.scalar_evolver.evolvers[name] = PassiveScalarEvolver(grid=self.grid, kappa=kappa,
                                                                  source_fn=source, name=name)
