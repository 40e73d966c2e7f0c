"""NumPy restatement of pygSQuiG's passive-scalar source terms
(pygsquig/scalars/source_terms.py).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__``): the sources the oracle's
scalar steps call, on numpy arrays of the oracle's Grid.  Each class follows
the formula of the reference class it names; tests/test_oracle_sources.py
pins them and the steps that use them.
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np


class ExponentialGrowth:
    """S = rate * scalar (source_terms.py:55-82)."""

    def __init__(self, rate: float):
        self.rate = rate

    def __call__(self, scalar, grid, t):
        return self.rate * scalar


class LocalizedSource:
    """S = A(t) exp(-r^2 / sigma^2), periodic distance over shifts of 2 pi
    (source_terms.py:85-160)."""

    def __init__(self, amplitude, x0, y0, sigma, time_dependent: Optional[Callable] = None):
        self.amplitude, self.x0, self.y0, self.sigma = amplitude, x0, y0, sigma
        self.time_dependent = time_dependent

    def __call__(self, scalar, grid, t):
        amp = self.amplitude * self.time_dependent(t) if self.time_dependent else self.amplitude
        tp = 2 * np.pi
        dx = np.minimum(np.abs(grid.x - self.x0), np.abs(grid.x - self.x0 + tp))
        dx = np.minimum(dx, np.abs(grid.x - self.x0 - tp))
        dy = np.minimum(np.abs(grid.y - self.y0), np.abs(grid.y - self.y0 + tp))
        dy = np.minimum(dy, np.abs(grid.y - self.y0 - tp))
        return amp * np.exp(-(dx ** 2 + dy ** 2) / self.sigma ** 2)


class ChemicalReaction:
    """S = -rate * scalar^2, zero where scalar <= threshold
    (source_terms.py:163-197)."""

    def __init__(self, rate, threshold=None):
        self.rate, self.threshold = rate, threshold

    def __call__(self, scalar, grid, t):
        s = -self.rate * scalar ** 2
        if self.threshold is not None:
            s = np.where(scalar > self.threshold, s, 0.0)
        return s


class TimePeriodicSource:
    """S = amplitude sin(frequency t + phase) f(x, y), f = 1 by default
    (source_terms.py:200-236)."""

    def __init__(self, amplitude, frequency, phase=0.0, spatial_pattern=None):
        self.amplitude, self.frequency, self.phase = amplitude, frequency, phase
        self.spatial_pattern = spatial_pattern

    def __call__(self, scalar, grid, t):
        tf = self.amplitude * np.sin(self.frequency * t + self.phase)
        sp = self.spatial_pattern(grid) if self.spatial_pattern else np.ones((grid.N, grid.N))
        return tf * sp


class LinearCombination:
    """S = sum of weight_i S_i (source_terms.py:239-277)."""

    def __init__(self, sources):
        self.sources = sources

    def __call__(self, scalar, grid, t):
        out = np.zeros_like(scalar)
        for w, src in self.sources:
            out = out + w * src(scalar, grid, t)
        return out
