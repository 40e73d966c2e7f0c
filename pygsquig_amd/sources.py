"""Passive-scalar source terms (pygsquig/scalars/source_terms.py).

A source is a callable ``S(scalar, grid, t)`` that returns a physical (N, N)
field from the physical scalar; the scalar right-hand side adds
``fft2(S(ifft2(s_hat), grid, t))`` in every RK4 stage, with ``t`` the time at
the start of the step (scalars/passive_scalar.py:172-210).  The classes keep
the reference's names, arguments, validation and formulas, on torch tensors
of the grid's device.  Any other callable of that signature works as well.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch

from .grid import Grid


class pygSQuiGError(Exception):
    """Base of the errors below (pygsquig/exceptions.py:9)."""


class PassiveScalarError(pygSQuiGError):
    """pygsquig/exceptions.py:45."""


class SourceTermError(PassiveScalarError):
    """pygsquig/exceptions.py:51."""


class SourceTerm(ABC):
    """scalars/source_terms.py:20-52: ``__call__(scalar, grid, t)`` returns
    the source in physical space."""

    @abstractmethod
    def __call__(self, scalar: torch.Tensor, grid: Grid, t: float) -> torch.Tensor:
        ...

    def validate_input(self, scalar, grid: Grid) -> None:
        shape = tuple(scalar.shape)
        if shape != (grid.N, grid.N):
            raise SourceTermError(
                f"Invalid input to source term: scalar has shape {shape}, "
                f"expected {(grid.N, grid.N)}")


@dataclass
class ExponentialGrowth(SourceTerm):
    """S = rate * scalar (:55-82)."""

    rate: float

    def __post_init__(self):
        if not isinstance(self.rate, (float, int)):
            raise SourceTermError(f"Rate must be numeric, got {type(self.rate).__name__}")

    def __call__(self, scalar, grid, t):
        return self.rate * scalar


@dataclass
class LocalizedSource(SourceTerm):
    """S = A(t) exp(-r^2 / sigma^2) around (x0, y0), with the periodic
    distance taken over shifts of 2 pi as the reference does (:85-160)."""

    amplitude: float
    x0: float
    y0: float
    sigma: float
    time_dependent: Optional[Callable[[float], float]] = None

    def __post_init__(self):
        if self.sigma <= 0:
            raise SourceTermError(f"Sigma must be positive, got {self.sigma}")
        self.sigma2 = self.sigma ** 2

    def __call__(self, scalar, grid, t):
        amp = self.amplitude * self.time_dependent(t) if self.time_dependent is not None \
            else self.amplitude
        two_pi = 2 * np.pi
        x, y = grid.x, grid.y
        dx = torch.minimum(torch.abs(x - self.x0), torch.abs(x - self.x0 + two_pi))
        dx = torch.minimum(dx, torch.abs(x - self.x0 - two_pi))
        dy = torch.minimum(torch.abs(y - self.y0), torch.abs(y - self.y0 + two_pi))
        dy = torch.minimum(dy, torch.abs(y - self.y0 - two_pi))
        r2 = dx ** 2 + dy ** 2
        return amp * torch.exp(-r2 / self.sigma2)


@dataclass
class ChemicalReaction(SourceTerm):
    """S = -rate * scalar^2, zero where scalar <= threshold (:163-197)."""

    rate: float
    threshold: Optional[float] = None

    def __post_init__(self):
        if self.rate < 0:
            raise SourceTermError(f"Reaction rate must be non-negative, got {self.rate}")

    def __call__(self, scalar, grid, t):
        source = -self.rate * scalar ** 2
        if self.threshold is not None:
            source = torch.where(scalar > self.threshold, source, torch.zeros_like(source))
        return source


@dataclass
class TimePeriodicSource(SourceTerm):
    """S = amplitude sin(frequency t + phase) f(x, y), f = 1 by default
    (:200-236)."""

    amplitude: float
    frequency: float
    phase: float = 0.0
    spatial_pattern: Optional[Callable[[Grid], torch.Tensor]] = None

    def __call__(self, scalar, grid, t):
        time_factor = self.amplitude * np.sin(self.frequency * t + self.phase)
        if self.spatial_pattern is not None:
            spatial = torch.as_tensor(self.spatial_pattern(grid), dtype=torch.float64,
                                      device=grid.device)
        else:
            spatial = torch.ones((grid.N, grid.N), dtype=torch.float64, device=grid.device)
        return time_factor * spatial


class LinearCombination(SourceTerm):
    """S = sum of weight_i * S_i (:239-277)."""

    def __init__(self, sources: list):
        self.sources = sources
        for _weight, source in sources:
            if not isinstance(source, SourceTerm):
                raise SourceTermError(f"Expected SourceTerm, got {type(source).__name__}")

    def __call__(self, scalar, grid, t):
        result = torch.zeros_like(scalar)
        for weight, source in self.sources:
            result = result + weight * source(scalar, grid, t)
        return result


def make_heating_source(x0: float, y0: float, radius: float = 0.5,
                        power: float = 1.0) -> LocalizedSource:
    """:281-296: 2 sigma ~ radius."""
    return LocalizedSource(amplitude=power, x0=x0, y0=y0, sigma=radius / 2.0)


def make_cooling_source(decay_time: float) -> ExponentialGrowth:
    """:299-309."""
    return ExponentialGrowth(rate=-1.0 / decay_time)


def make_reaction_source(reaction_rate: float, threshold: float = 0.0) -> ChemicalReaction:
    """:312-323."""
    return ChemicalReaction(rate=reaction_rate, threshold=threshold)


def source_from_config(source_config) -> Optional[SourceTerm]:
    """The source a run configuration's ``scalars.species[*].source`` entry
    names (type and parameters as scripts/run.py:62-91 reads them, with the
    same defaults); None for no entry or type "none"."""
    if source_config is None:
        return None
    kind, prm = source_config.type, dict(source_config.parameters or {})
    if kind == "exponential":
        return ExponentialGrowth(rate=prm.get("rate", 0.0))
    if kind == "localized":
        return LocalizedSource(amplitude=prm.get("amplitude", 1.0), x0=prm.get("x0", 0.0),
                               y0=prm.get("y0", 0.0), sigma=prm.get("sigma", 1.0))
    if kind == "chemical":
        return ChemicalReaction(rate=prm.get("rate", 1.0), threshold=prm.get("threshold"))
    if kind == "periodic":
        return TimePeriodicSource(amplitude=prm.get("amplitude", 1.0),
                                  frequency=prm.get("frequency", 1.0), phase=prm.get("phase", 0.0))
    if kind == "none":
        return None
    raise ValueError(f"Unknown source type: {kind}")
