"""Checkpoints of solver states in the reference's field layout.

pygSQuiG writes HDF5 (io/hdf5_io.py:49-178): group ``state`` with datasets
``theta_hat_real`` and ``theta_hat_imag``, attributes ``time`` and ``step``,
and every other array in the state as ``<key>`` (or ``<key>_real`` /
``<key>_imag`` when complex).  h5py is not part of this stack, so the same
names are written as the members of a NumPy ``.npz`` archive
(``state/theta_hat_real``, ``state/time`` ...); passive scalars are stored
as ``state/scalar/<name>_real`` / ``_imag``.  Loading never unpickles.
"""

from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np
import torch

FORMAT = "pygsquig_amd.checkpoint/1"


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def save_checkpoint(state: dict, filename, config: Optional[dict] = None) -> None:
    """Write ``state`` ({"theta_hat", "time", "step", ...}) to ``filename``."""
    arrays = {"metadata/format": np.array(FORMAT),
              "metadata/config": np.array(json.dumps(config or {}))}
    th = _np(state["theta_hat"])
    arrays["state/theta_hat_real"] = np.ascontiguousarray(th.real)
    arrays["state/theta_hat_imag"] = np.ascontiguousarray(th.imag)
    arrays["state/time"] = np.array(float(_np(state["time"])))
    arrays["state/step"] = np.array(int(_np(state["step"])))
    for key, value in state.items():
        if key in ("theta_hat", "time", "step"):
            continue
        if key == "scalar_state":
            for name, s in value.scalars.items():
                a = _np(s)
                arrays[f"state/scalar/{name}_real"] = np.ascontiguousarray(a.real)
                arrays[f"state/scalar/{name}_imag"] = np.ascontiguousarray(a.imag)
            arrays["state/scalar_time"] = np.array(float(_np(value.time)))
            continue
        if isinstance(value, (np.ndarray, torch.Tensor)):
            a = _np(value)
            if np.iscomplexobj(a):
                arrays[f"state/{key}_real"] = np.ascontiguousarray(a.real)
                arrays[f"state/{key}_imag"] = np.ascontiguousarray(a.imag)
            else:
                arrays[f"state/{key}"] = a
    d = os.path.dirname(os.path.abspath(filename))
    os.makedirs(d, exist_ok=True)
    with open(filename, "wb") as f:
        np.savez(f, **arrays)


def load_checkpoint(filename, device: Optional[torch.device] = None) -> dict:
    """Read a checkpoint written by ``save_checkpoint``: the state dict with
    ``theta_hat`` as a complex128 tensor on ``device`` (default: cuda:0 when
    available), ``time`` / ``step`` as 0-d tensors, passive scalars (if any)
    as a ``MultiScalarState`` under ``"scalar_state"`` (ready for
    ``gSQGSolverWithScalars.step``), and the run configuration under
    ``"config"``."""
    from .scalars import MultiScalarState

    if device is None:
        device = torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu")
    with np.load(filename, allow_pickle=False) as z:
        if str(z["metadata/format"]) != FORMAT:
            raise ValueError(f"{filename}: not a {FORMAT} file")
        th = z["state/theta_hat_real"] + 1j * z["state/theta_hat_imag"]
        out = {"theta_hat": torch.from_numpy(th).to(device),
               "time": torch.tensor(float(z["state/time"]), dtype=torch.float64),
               "step": torch.tensor(int(z["state/step"]), dtype=torch.int64),
               "config": json.loads(str(z["metadata/config"]))}
        names = sorted({k[len("state/scalar/"):].rsplit("_", 1)[0] for k in z.files
                        if k.startswith("state/scalar/")})
        if names:
            out["scalar_state"] = MultiScalarState(
                scalars={n: torch.from_numpy(z[f"state/scalar/{n}_real"]
                                             + 1j * z[f"state/scalar/{n}_imag"]).to(device)
                         for n in names},
                time=float(z["state/scalar_time"]))
        for k in z.files:
            if not k.startswith("state/") or k.startswith("state/scalar") or k in (
                    "state/theta_hat_real", "state/theta_hat_imag", "state/time", "state/step"):
                continue
            key = k[len("state/"):]
            if key.endswith("_imag"):
                continue
            if key.endswith("_real") and f"state/{key[:-5]}_imag" in z.files:
                out[key[:-5]] = torch.from_numpy(z[k] + 1j * z[f"state/{key[:-5]}_imag"]).to(device)
            else:
                out[key] = torch.from_numpy(z[k]).to(device)
    return out
