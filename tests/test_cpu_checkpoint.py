"""Checkpoint round trip (the reference's io/hdf5_io.py field names in an
.npz archive): arrays, scalars and metadata come back bit-identical."""

import numpy as np
import torch

from pygsquig_amd.checkpoint import load_checkpoint, save_checkpoint
from pygsquig_amd.scalars import MultiScalarState


def test_round_trip(tmp_path):
    g = torch.Generator().manual_seed(0)
    th = torch.randn(16, 16, dtype=torch.complex128, generator=g)
    sc = {"dye_a": torch.randn(16, 16, dtype=torch.complex128, generator=g),
          "dye_b": torch.randn(16, 16, dtype=torch.complex128, generator=g)}
    st = {"theta_hat": th, "time": torch.tensor(1.25, dtype=torch.float64),
          "step": torch.tensor(40), "forcing_phase": np.arange(5.0),
          "scalar_state": MultiScalarState(scalars=sc, time=1.25)}
    f = tmp_path / "run" / "ck.npz"
    save_checkpoint(st, f, {"N": 16, "alpha": 1.0})
    r = load_checkpoint(f, device=torch.device("cpu"))
    assert torch.equal(r["theta_hat"], th)
    assert float(r["time"]) == 1.25 and int(r["step"]) == 40
    assert r["config"] == {"N": 16, "alpha": 1.0}
    for k, v in sc.items():
        assert torch.equal(r["scalar_state"].scalars[k], v)
    assert r["scalar_state"].time == 1.25
    assert np.array_equal(r["forcing_phase"].numpy(), np.arange(5.0))
    names = set(np.load(f).files)
    assert {"state/theta_hat_real", "state/theta_hat_imag", "state/time", "state/step"} <= names
