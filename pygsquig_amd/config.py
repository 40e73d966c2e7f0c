"""Run configuration: the YAML schema of pygSQuiG's runner (io/config.py).

A YAML file maps onto ``RunConfig`` with the same sections, keys, defaults
and validation messages as the reference (io/config.py:16-441), so its
configuration files load unchanged.  One key is added:
``solver.time_integration.scheme`` ("explicit", the reference's RK4 and the
default, or "if", the integrating-factor RK4 that keeps the listed
hyperviscosities stable at the listed time steps).  ``method`` accepts
"RK4" only in practice: "SSP-RK3" validates, as in the reference, and is
refused when a run is built.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Optional, Union

import numpy as np
import yaml


def _require(ok: bool, msg: str) -> None:
    if not ok:
        raise ValueError(msg)


@dataclass
class GridConfig:
    """io/config.py:16-33."""

    N: int = 256
    L: float = 2 * np.pi

    def __post_init__(self):
        _require(self.N > 0 and self.N % 2 == 0, f"N must be a positive even integer, got {self.N}")
        _require(self.L > 0, f"L must be positive, got {self.L}")


@dataclass
class DissipationConfig:
    """io/config.py:36-57: (-Delta)^(p/2) with coefficient nu_p."""

    type: str = "hyperviscosity"
    nu_p: float = 1.0e-16
    p: int = 8

    def __post_init__(self):
        _require(self.type in ("hyperviscosity", "viscosity"), f"Unknown dissipation type: {self.type}")
        _require(self.nu_p >= 0, f"nu_p must be non-negative, got {self.nu_p}")
        _require(self.p in (2, 4, 8), f"p must be 2, 4, or 8, got {self.p}")


@dataclass
class DampingConfig:
    """io/config.py:60-80: linear drag for |k| < kf * k_cutoff_factor."""

    type: str = "linear_drag"
    mu: float = 0.1
    k_cutoff_factor: float = 0.5

    def __post_init__(self):
        _require(self.type in ("linear_drag", "none"), f"Unknown damping type: {self.type}")
        _require(self.mu >= 0, f"mu must be non-negative, got {self.mu}")
        _require(self.k_cutoff_factor > 0,
                 f"k_cutoff_factor must be positive, got {self.k_cutoff_factor}")


@dataclass
class TimeIntegrationConfig:
    """io/config.py:83-108, plus ``scheme``."""

    method: str = "RK4"
    dt: float = 0.001
    adaptive_cfl: bool = True
    cfl_safety: float = 0.8
    dt_max: Optional[float] = None
    scheme: str = "explicit"

    def __post_init__(self):
        _require(self.method in ("RK4", "SSP-RK3"), f"Unknown time integration method: {self.method}")
        _require(self.dt > 0, f"dt must be positive, got {self.dt}")
        _require(0 < self.cfl_safety < 1, f"cfl_safety must be in (0, 1), got {self.cfl_safety}")
        _require(self.dt_max is None or self.dt_max > 0, f"dt_max must be positive, got {self.dt_max}")
        _require(self.scheme in ("explicit", "if"),
                 f"scheme must be 'explicit' or 'if', got {self.scheme!r}")


@dataclass
class SolverConfig:
    """io/config.py:111-129."""

    alpha: float
    dissipation: DissipationConfig = field(default_factory=DissipationConfig)
    damping: Optional[DampingConfig] = None
    time_integration: TimeIntegrationConfig = field(default_factory=TimeIntegrationConfig)

    def __post_init__(self):
        _require(-2 <= self.alpha <= 2, f"alpha must be in [-2, 2], got {self.alpha}")


@dataclass
class ForcingConfig:
    """io/config.py:132-161."""

    type: str = "ring"
    kf: float = 20.0
    dk: float = 1.0
    epsilon: float = 0.1
    tau_f: float = 0.0
    seed: Optional[int] = None

    def __post_init__(self):
        _require(self.type in ("ring", "none"), f"Unknown forcing type: {self.type}")
        _require(self.kf > 0, f"kf must be positive, got {self.kf}")
        _require(self.dk > 0, f"dk must be positive, got {self.dk}")
        _require(self.epsilon >= 0, f"epsilon must be non-negative, got {self.epsilon}")
        _require(self.tau_f >= 0, f"tau_f must be non-negative, got {self.tau_f}")


_FIELDS = ("theta", "vorticity", "streamfunction", "velocity", "scalars")
_DIAGNOSTICS = ("energy_spectrum", "scalar_flux", "enstrophy", "energy_flux")


@dataclass
class OutputConfig:
    """io/config.py:164-189."""

    fields: list = field(default_factory=lambda: ["theta"])
    diagnostics: list = field(default_factory=lambda: ["energy_spectrum", "scalar_flux"])
    save_every_n_steps: Optional[int] = None
    compress: bool = True

    def __post_init__(self):
        bad = set(self.fields) - set(_FIELDS)
        _require(not bad, f"Unknown fields: {bad}")
        bad = set(self.diagnostics) - set(_DIAGNOSTICS)
        _require(not bad, f"Unknown diagnostics: {bad}")


@dataclass
class SimulationConfig:
    """io/config.py:192-222."""

    t_end: float
    output_interval: float = 1.0
    checkpoint_interval: float = 10.0
    wall_time_limit: Optional[float] = None
    log_interval: float = 0.1

    def __post_init__(self):
        _require(self.t_end > 0, f"t_end must be positive, got {self.t_end}")
        for name in ("output_interval", "checkpoint_interval", "log_interval"):
            v = getattr(self, name)
            _require(v > 0, f"{name} must be positive, got {v}")
        _require(self.wall_time_limit is None or self.wall_time_limit > 0,
                 f"wall_time_limit must be positive, got {self.wall_time_limit}")


@dataclass
class ScalarSourceConfig:
    """io/config.py:225-239."""

    type: str
    parameters: dict = field(default_factory=dict)

    def __post_init__(self):
        _require(self.type in ("exponential", "localized", "chemical", "periodic", "none"),
                 f"Unknown source type: {self.type}")


@dataclass
class PassiveScalarConfig:
    """io/config.py:242-265."""

    name: str
    kappa: float = 0.0
    source: Optional[ScalarSourceConfig] = None
    initial_condition: str = "zero"
    initial_params: dict = field(default_factory=dict)

    def __post_init__(self):
        _require(self.kappa >= 0, f"kappa must be non-negative, got {self.kappa}")
        _require(self.initial_condition in ("zero", "random", "gaussian", "uniform"),
                 f"Unknown initial condition: {self.initial_condition}")


@dataclass
class ScalarsConfig:
    """io/config.py:268-284."""

    enabled: bool = False
    species: list = field(default_factory=list)

    def __post_init__(self):
        names = [s.name for s in self.species]
        _require(len(names) == len(set(names)), "Duplicate scalar names found")


@dataclass
class InitialConditionConfig:
    """io/config.py:287-310."""

    type: str = "random"
    checkpoint_path: Optional[str] = None
    amplitude: float = 1.0
    seed: Optional[int] = None

    def __post_init__(self):
        _require(self.type in ("random", "checkpoint", "function", "zero"),
                 f"Unknown initial condition type: {self.type}")
        _require(not (self.type == "checkpoint" and self.checkpoint_path is None),
                 "checkpoint_path required when type='checkpoint'")
        _require(self.amplitude > 0, f"amplitude must be positive, got {self.amplitude}")


def _species(d: dict) -> PassiveScalarConfig:
    src = d.get("source")
    return PassiveScalarConfig(
        name=d["name"], kappa=d.get("kappa", 0.0),
        source=ScalarSourceConfig(**src) if src is not None else None,
        initial_condition=d.get("initial_condition", "zero"),
        initial_params=d.get("initial_params", {}))


def _prune(x):
    """Drop None values and empty dicts, recursively (io/config.py:409-418)."""
    if isinstance(x, dict):
        return {k: _prune(v) for k, v in x.items()
                if v is not None and not (isinstance(v, dict) and not v)}
    return x


@dataclass
class RunConfig:
    """io/config.py:313-437: every section of one run."""

    grid: GridConfig
    solver: SolverConfig
    forcing: Optional[ForcingConfig] = None
    scalars: Optional[ScalarsConfig] = None
    output: OutputConfig = field(default_factory=OutputConfig)
    simulation: SimulationConfig = field(default_factory=lambda: SimulationConfig(t_end=100.0))
    initial_condition: InitialConditionConfig = field(default_factory=InitialConditionConfig)

    @classmethod
    def from_dict(cls, data: dict) -> "RunConfig":
        s = dict(data.get("solver", {}))
        for key, typ in (("dissipation", DissipationConfig), ("damping", DampingConfig),
                         ("time_integration", TimeIntegrationConfig)):
            if key in s:
                s[key] = typ(**s[key])
        scalars = None
        sd = data.get("scalars")
        if sd is not None and sd.get("enabled", False):
            scalars = ScalarsConfig(enabled=True, species=[_species(x) for x in sd.get("species", [])])
        return cls(grid=GridConfig(**data.get("grid", {})), solver=SolverConfig(**s),
                   forcing=ForcingConfig(**data["forcing"]) if "forcing" in data else None,
                   scalars=scalars, output=OutputConfig(**data.get("output", {})),
                   simulation=SimulationConfig(**data.get("simulation", {"t_end": 100.0})),
                   initial_condition=InitialConditionConfig(**data.get("initial_condition", {})))

    @classmethod
    def from_yaml(cls, path: Union[str, Path]) -> "RunConfig":
        with open(path) as f:
            return cls.from_dict(yaml.safe_load(f) or {})

    def to_dict(self) -> dict[Any, Any]:
        return dict(_prune(dataclasses.asdict(self)))

    def to_yaml(self, path: Union[str, Path]) -> None:
        with open(path, "w") as f:
            yaml.safe_dump(self.to_dict(), f, default_flow_style=False, sort_keys=False)


def load_config(path: Union[str, Path]) -> RunConfig:
    """io/config.py:440-452."""
    return RunConfig.from_yaml(path)
