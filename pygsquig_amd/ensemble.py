"""Uneven ensembles: independent realizations of different grid size and
gSQG exponent, advanced together.

The reference has no batched path; a user loops ``gSQGSolver.step`` over
realizations (core/solver.py:182-214).  Here the members of one grid size
share one plan: their fields are the y-slices of every launch (one column
and one row launch per RK4 stage for the whole group, each member with its
own alpha), and the steps replay as a HIP graph.  With ``launch="streams"``
groups of different sizes are issued on separate streams, so their kernels
overlap on the device.  With ``launch="batch"`` every group's column pass
of an RK4 stage is one launch and every group's row pass another
(``gsqg_batch_advance``): blocks find their group from a block table.  Each
member follows exactly the single-realization solver's step either way.
"""

from __future__ import annotations

from typing import Optional, Sequence

import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from .grid import default_device, make_grid
from .plan import Plan

_SCHEMES = {"explicit": _lib.SCHEME_EXPLICIT, "if": _lib.SCHEME_IF}


class EnsembleSolver:
    """Advance members ``(N, alpha)`` with shared ``nu_p``, ``p`` and scheme.

    State: ``{"theta_hat": [full (N_m, N_m) complex128 per member], "time",
    "step"}``; ``step``/``multistep`` return a new state.
    """

    def __init__(self, members: Sequence[tuple], L: float = 2 * np.pi, nu_p: float = 0.0,
                 p: int = 8, scheme: str = "explicit", device: Optional[torch.device] = None,
                 launch: str = "streams"):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        if scheme not in _SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(_SCHEMES)}, got {scheme!r}")
        if launch not in ("streams", "batch"):
            raise ValueError(f"launch must be 'streams' or 'batch', got {launch!r}")
        if p not in (2, 4, 8):
            raise ValueError(f"p must be 2, 4, or 8, got {p}")
        self.members = [(int(N), float(a)) for N, a in members]
        for N, a in self.members:
            if not -2 <= a <= 2:
                raise ValueError(f"alpha must be in [-2, 2], got {a}")
        self.L, self.nu_p, self.p, self.scheme = L, nu_p, p, scheme
        self.device = device or default_device()
        # members grouped by grid size, in first-appearance order
        self.groups = {}
        for i, (N, _) in enumerate(self.members):
            self.groups.setdefault(N, []).append(i)
        self.grids = {N: make_grid(N, L, self.device) for N in self.groups}
        self.plans = {N: Plan(self.grids[N], self.members[idx[0]][1], nu_p, p,
                              scheme=_SCHEMES[scheme],
                              member_alphas=[self.members[i][1] for i in idx])
                      for N, idx in self.groups.items()}
        self.streams = {N: torch.cuda.Stream(self.device) for N in self.groups}
        self.launch = launch
        self.batch = None
        if launch == "batch":
            bad = [N for N in self.groups if N not in _BATCH_SIZES]
            if bad:
                raise ValueError(f"launch='batch' needs grid sizes in {_BATCH_SIZES}, got {bad}")
            lib = _lib.load()
            plans = list(self.plans.values())
            arr = (ctypes.c_void_p * len(plans))(*(pl.handle.value for pl in plans))
            handle = ctypes.c_void_p()
            _lib.check(lib.gsqg_batch_create(arr, len(plans), ctypes.byref(handle)),
                       "gsqg_batch_create")
            self.batch = handle
            # the plans outlive the batch: the finalizer holds them
            self._batch_fin = weakref.finalize(self, _destroy_batch, lib, handle, plans)

    def initialize(self, thetas: Sequence) -> dict:
        """Physical initial fields, one (N_m, N_m) array per member."""
        if len(thetas) != len(self.members):
            raise ValueError(f"expected {len(self.members)} initial fields, got {len(thetas)}")
        out = []
        for (N, _), th in zip(self.members, thetas):
            t = torch.as_tensor(th).to(device=self.device, dtype=torch.float64)
            if t.shape != (N, N):
                raise ValueError(f"initial field of shape {tuple(t.shape)} for N={N}")
            out.append(_fft2(t, self.grids[N]))
        return {"theta_hat": out, "time": torch.tensor(0.0, dtype=torch.float64),
                "step": torch.tensor(0, dtype=torch.int64)}

    def multistep(self, state: dict, n_steps: int, dt: float) -> dict:
        """n_steps RK4 steps of every member (core/solver.py:182-214 per member)."""
        ths = state["theta_hat"]
        if len(ths) != len(self.members):
            raise ValueError("state does not match the ensemble's members")
        cur = torch.cuda.current_stream(self.device)
        if self.batch is not None:
            return self._multistep_batch(state, ths, n_steps, dt, cur)
        # inputs and outputs are allocated on the current stream; each group's
        # stream waits for it before use and the current stream waits for the
        # groups before anything here is released or read
        ins = [t.to(self.device, torch.complex128).contiguous() for t in ths]
        new = [torch.empty((N, N), dtype=torch.complex128, device=self.device)
               for N, _ in self.members]
        for N, idx in self.groups.items():
            s, plan = self.streams[N], self.plans[N]
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                for k, i in enumerate(idx):
                    plan.load_member(k, ins[i])
                plan.advance(dt, n_steps, _lib.KEY_EXPLICIT, 0, 0, use_graph=True)
                for k, i in enumerate(idx):
                    plan.store_member(k, new[i])
        for N in self.groups:
            cur.wait_stream(self.streams[N])
        t = torch.as_tensor(state["time"], dtype=torch.float64)
        for _ in range(n_steps):
            t = t + dt
        return {"theta_hat": new, "time": t, "step": torch.as_tensor(state["step"]) + n_steps}

    def _multistep_batch(self, state, ths, n_steps, dt, cur):
        ins = [t.to(self.device, torch.complex128).contiguous() for t in ths]
        new = [torch.empty((N, N), dtype=torch.complex128, device=self.device)
               for N, _ in self.members]
        with torch.cuda.device(self.device):
            for N, idx in self.groups.items():
                for k, i in enumerate(idx):
                    self.plans[N].load_member(k, ins[i])
            lib = _lib.load()
            _lib.check(lib.gsqg_batch_advance(self.batch, float(dt), int(n_steps), 1,
                                              ctypes.c_void_p(cur.cuda_stream)),
                       "gsqg_batch_advance")
            for N, idx in self.groups.items():
                for k, i in enumerate(idx):
                    self.plans[N].store_member(k, new[i])
        t = torch.as_tensor(state["time"], dtype=torch.float64)
        for _ in range(n_steps):
            t = t + dt
        return {"theta_hat": new, "time": t, "step": torch.as_tensor(state["step"]) + n_steps}

    def step(self, state: dict, dt: float) -> dict:
        return self.multistep(state, 1, dt)

    def autotune(self, dt: float, n_steps: int = 32, repeats: int = 3) -> dict:
        """Pick each group's column/row tile shapes by timing every pair on
        the group's plan (``Plan.autotune``; each group alone, on its own
        stream).  Returns ``{N: ((col, row), {(col, row): ms})}``; results of
        later steps are unchanged bit for bit, only their speed.  The
        one-launch batch runs fixed radix-8 tiles, so with ``launch="batch"``
        there is nothing to tune and the result is empty."""
        out = {}
        if self.batch is not None:
            return out
        cur = torch.cuda.current_stream(self.device)
        for N, plan in self.plans.items():
            s = self.streams[N]
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                out[N] = plan.autotune(dt, n_steps, repeats)
            cur.wait_stream(s)
        return out

    def set_shapes(self, shapes: dict) -> None:
        """Set ``{N: (col, row)}`` tile shapes directly (e.g. a saved tuning)."""
        for N, (c, r) in shapes.items():
            self.plans[N].set_shape(c, r)


def _fft2(t: torch.Tensor, grid):
    from .grid import fft2

    return fft2(t, grid)


# grid sizes the one-launch batch compiles (gsqg_batch_create)
_BATCH_SIZES = (64, 128, 256, 512, 1024)


def _destroy_batch(lib, handle, plans):
    try:
        lib.gsqg_batch_destroy(handle)
    except Exception:  # pragma: no cover
        pass
