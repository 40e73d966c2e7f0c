"""Source-term classes and their configuration on the host: the reference's
validation messages (scalars/source_terms.py, scalars/passive_scalar.py:134)
and the configuration factory (scripts/run.py:62-91)."""

import pytest

from pygsquig_amd import sources as S
from pygsquig_amd.config import ScalarSourceConfig


def test_validation_messages():
    with pytest.raises(S.SourceTermError, match="Rate must be numeric, got str"):
        S.ExponentialGrowth("fast")
    with pytest.raises(S.SourceTermError, match="Sigma must be positive, got 0"):
        S.LocalizedSource(1.0, 0.0, 0.0, 0)
    with pytest.raises(S.SourceTermError, match="Reaction rate must be non-negative, got -1"):
        S.ChemicalReaction(-1)
    with pytest.raises(S.SourceTermError, match="Expected SourceTerm, got function"):
        S.LinearCombination([(1.0, lambda s, g, t: s)])
    assert issubclass(S.SourceTermError, S.PassiveScalarError)


def test_factories():
    h = S.make_heating_source(1.0, 2.0, radius=0.8, power=3.0)
    assert (h.amplitude, h.x0, h.y0, h.sigma) == (3.0, 1.0, 2.0, 0.4)
    assert S.make_cooling_source(4.0).rate == -0.25
    r = S.make_reaction_source(0.7)
    assert (r.rate, r.threshold) == (0.7, 0.0)


def test_source_from_config():
    mk = lambda t, **p: S.source_from_config(ScalarSourceConfig(type=t, parameters=p))  # noqa: E731
    assert S.source_from_config(None) is None
    assert mk("none") is None
    assert mk("exponential", rate=-0.1).rate == -0.1
    assert mk("exponential").rate == 0.0
    loc = mk("localized", x0=1.0, sigma=0.3)
    assert (loc.amplitude, loc.x0, loc.y0, loc.sigma) == (1.0, 1.0, 0.0, 0.3)
    ch = mk("chemical", threshold=0.2)
    assert (ch.rate, ch.threshold) == (1.0, 0.2)
    per = mk("periodic", frequency=2.0)
    assert (per.amplitude, per.frequency, per.phase) == (1.0, 2.0, 0.0)
    with pytest.raises(ValueError, match="Unknown source type"):
        ScalarSourceConfig(type="volcanic")


def test_evolver_refuses_non_source_callable():
    from pygsquig_amd.scalars import PassiveScalarEvolver

    with pytest.raises(S.PassiveScalarError, match="must be a SourceTerm instance, got function"):
        PassiveScalarEvolver(grid=None, source_fn=lambda s, g, t: s)
