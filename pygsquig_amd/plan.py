"""Device plans: one ``gsqg_plan`` per (grid, physics) configuration.

A plan owns the device work space of the fused step (half-packed ky-major
fields, twiddle tables, the per-mode flag table, the ring side state) and one
device-resident solver state.  This module builds the host tables the plan
needs with the reference's operation order and wraps the C ABI.
"""

from __future__ import annotations

import ctypes
import weakref
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import Grid, dealias_cutoff, wavenumbers_1d

FLAG_DEALIAS = 1
FLAG_LSDAMP = 2
FLAG_RING = 4


def half_plane_masks(N: int, L: float, damping_kf: Optional[float], ring: Optional[tuple]):
    """Boolean masks over the half plane (N, N/2+1), [kx index][ky index],
    computed as the reference computes them on the full plane:
    dealias ``sqrt(k2) < (2/3) pi N/L`` (core/grid.py:102-105), large-scale
    damping ``sqrt(k2) < kf/2`` (forcing/damping.py:34), ring
    ``|sqrt(k2) - kf| <= dk/2`` (forcing/ring_forcing.py:74-75)."""
    k = wavenumbers_1d(N, L)
    kx = k[:, None]
    ky = k[: N // 2 + 1].copy()
    ky[N // 2] = k[N // 2]
    ky = ky[None, :]
    k2 = kx**2 + ky**2
    kmag = np.sqrt(k2)
    dealias = kmag < dealias_cutoff(N, L)
    ls = (kmag < damping_kf / 2) if damping_kf is not None else np.zeros_like(dealias)
    if ring is not None:
        kf, dk = ring
        rg = np.abs(kmag - kf) <= dk / 2
    else:
        rg = np.zeros_like(dealias)
    return dealias, ls, rg


def pack_flags(dealias, ls, rg) -> np.ndarray:
    """Pack (N, N/2+1) masks into the plan's ky-major (N/2, N) flag bytes."""
    N = dealias.shape[0]
    full = (dealias.astype(np.uint8) | (ls.astype(np.uint8) << 1) | (rg.astype(np.uint8) << 2))
    out = np.ascontiguousarray(full[:, : N // 2].T)
    out[0, :] |= (full[:, N // 2] << 4).astype(np.uint8)
    return out


def _ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Plan:
    """Thin owner of a ``gsqg_plan*``."""

    def __init__(self, grid: Grid, alpha: float = 1.0, nu_p: float = 0.0, p: int = 8, *,
                 scheme: int = _lib.SCHEME_EXPLICIT, combine: int = _lib.COMBINE_STEP,
                 hv_dc_zero: bool = True, damping=None, forcing=None,
                 kappas: Sequence[float] = (), member_alphas: Optional[Sequence[float]] = None):
        self.lib = _lib.load()
        self.grid = grid
        N, L = grid.N, grid.L
        self.N = N
        self.alpha = alpha
        self.nu_p, self.p = nu_p, p
        self.scheme, self.combine = scheme, combine
        self.damping, self.forcing = damping, forcing
        self.kappas = tuple(float(k) for k in kappas)
        dk_kf = None
        if damping is not None and damping.mu != 0:
            dk_kf = damping.kf
        ring = (forcing.kf, forcing.dk) if forcing is not None else None
        dealias, ls, rg = half_plane_masks(N, L, dk_kf, ring)
        self.flags = pack_flags(dealias, ls, rg)
        self.ring_modes = np.ascontiguousarray(np.argwhere(rg).astype(np.int32)) if ring else \
            np.zeros((0, 2), np.int32)
        self.k1d = np.ascontiguousarray(wavenumbers_1d(N, L))
        self._kappa_arr = (ctypes.c_double * max(1, len(self.kappas)))(*(self.kappas or (0.0,)))
        d = _lib.GsqgPlanDesc()
        d.n = N
        d.L = L
        d.alpha = alpha
        d.nu_p = nu_p
        d.p = p
        d.scheme = scheme
        d.combine = combine
        d.hv_dc_zero = 1 if hv_dc_zero else 0
        d.damping = 1 if damping is not None else 0
        if damping is not None:
            d.damp_mu = damping.mu
            d.damp_nu_p = damping.nu_p
            d.damp_p = damping.p
        d.forcing = 1 if forcing is not None else 0
        if forcing is not None:
            if forcing.tau_f != 0:
                raise NotImplementedError("the fused forcing path supports tau_f = 0 only")
            d.epsilon = forcing.epsilon
            d.amplitude = float(forcing.amplitude)
        d.n_scalars = len(self.kappas)
        d.kappa = self._kappa_arr
        d.k1d = self.k1d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.mode_flags = self.flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.ring_modes = self.ring_modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.n_ring = int(self.ring_modes.shape[0])
        self.member_alphas = tuple(float(a) for a in member_alphas) if member_alphas else (alpha,)
        self._malpha = (ctypes.c_double * len(self.member_alphas))(*self.member_alphas)
        d.n_members = len(self.member_alphas)
        d.member_alpha = self._malpha
        handle = ctypes.c_void_p()
        with torch.cuda.device(grid.device):
            _lib.check(self.lib.gsqg_plan_create(ctypes.byref(d), ctypes.byref(handle)),
                       "gsqg_plan_create")
        self.handle = handle
        self._finalizer = weakref.finalize(self, Plan._destroy, self.lib, handle)
        # identity of the full tensor mirroring the device state (see step())
        self.state_ref = None
        self.state_version = None
        self.scalar_refs = {}

    @staticmethod
    def _destroy(lib, handle):
        try:
            lib.gsqg_plan_destroy(handle)
        except Exception:  # pragma: no cover
            pass

    # --------------------------------------------------------------- state
    def load_theta(self, th: torch.Tensor, check: bool = True) -> float:
        off = ctypes.c_double(0.0)
        _lib.check(self.lib.gsqg_load_theta(self.handle, _ptr(th), ctypes.byref(off) if check else None,
                                            _stream()), "gsqg_load_theta")
        self.state_ref = None
        return off.value

    def store_theta(self, out: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_store_theta(self.handle, _ptr(out), _stream()), "gsqg_store_theta")

    def load_member(self, idx: int, th: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_load_member(self.handle, idx, _ptr(th), _stream()),
                   "gsqg_load_member")
        self.state_ref = None

    def store_member(self, idx: int, out: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_store_member(self.handle, idx, _ptr(out), _stream()),
                   "gsqg_store_member")

    def load_scalar(self, idx: int, s: torch.Tensor, check: bool = True) -> float:
        off = ctypes.c_double(0.0)
        _lib.check(self.lib.gsqg_load_scalar(self.handle, idx, _ptr(s),
                                             ctypes.byref(off) if check else None, _stream()),
                   "gsqg_load_scalar")
        self.scalar_refs.pop(idx, None)
        return off.value

    def store_scalar(self, idx: int, out: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_store_scalar(self.handle, idx, _ptr(out), _stream()),
                   "gsqg_store_scalar")

    def advance(self, dt: float, n_steps: int, key_mode: int, key: int, step0: int = 0,
                use_graph: bool = True) -> None:
        _lib.check(self.lib.gsqg_advance(self.handle, float(dt), int(n_steps), int(key_mode),
                                         ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF),
                                         int(step0), 1 if use_graph else 0, _stream()),
                   "gsqg_advance")

    def graph_capture(self, dt: float, key_mode: int, key: int) -> None:
        """Capture the replay graphs ``advance`` uses for (dt, key_mode, key)
        without running them (gsqg_graph_capture); the state is unchanged."""
        with torch.cuda.device(self.grid.device):
            _lib.check(self.lib.gsqg_graph_capture(self.handle, float(dt), int(key_mode),
                                                   ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF)),
                       "gsqg_graph_capture")

    def advance_scalars(self, u: torch.Tensor, v: torch.Tensor, dt: float) -> None:
        _lib.check(self.lib.gsqg_advance_scalars(self.handle, _ptr(u), _ptr(v), float(dt),
                                                 _stream()), "gsqg_advance_scalars")

    # ----------------------------------------------------------- operators
    def scalar_rhs(self, idx: int, sh: torch.Tensor, u: torch.Tensor, v: torch.Tensor,
                   out: torch.Tensor) -> None:
        """Explicit right-hand side of scalar idx (no source) into out; the
        scalar's device state is overwritten."""
        _lib.check(self.lib.gsqg_scalar_rhs(self.handle, int(idx), _ptr(sh), _ptr(u), _ptr(v),
                                            _ptr(out), _stream()), "gsqg_scalar_rhs")

    def rhs(self, th: torch.Tensor, out: torch.Tensor, key: int = 0) -> None:
        _lib.check(self.lib.gsqg_rhs(self.handle, _ptr(th), _ptr(out),
                                     ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF), _stream()),
                   "gsqg_rhs")

    def velocity(self, th: torch.Tensor, u: torch.Tensor, v: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_velocity(self.handle, _ptr(th), _ptr(u), _ptr(v), _stream()),
                   "gsqg_velocity")

    def fft2(self, f: torch.Tensor, X: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_fft2(self.handle, _ptr(f), _ptr(X), _stream()), "gsqg_fft2")

    def ifft2(self, X: torch.Tensor, f: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_ifft2(self.handle, _ptr(X), _ptr(f), _stream()), "gsqg_ifft2")

    def ring_count(self) -> int:
        n = ctypes.c_int(0)
        _lib.check(self.lib.gsqg_plan_ring_count(self.handle, ctypes.byref(n)),
                   "gsqg_plan_ring_count")
        return int(n.value)

    def ring_forcing_ou(self, th: torch.Tensor, key: int, decay: float, vfac: float,
                        state: torch.Tensor, first: bool, out: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_ring_forcing_ou(self.handle, _ptr(th),
                                                 ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF),
                                                 float(decay), float(vfac), _ptr(state),
                                                 1 if first else 0, _ptr(out), _stream()),
                   "gsqg_ring_forcing_ou")

    def ring_forcing(self, th: torch.Tensor, key: int, out: torch.Tensor) -> None:
        _lib.check(self.lib.gsqg_ring_forcing(self.handle, _ptr(th),
                                              ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF),
                                              _ptr(out), _stream()), "gsqg_ring_forcing")

    def spectral_diagnostics(self, X: torch.Tensor, alpha: float, edges: np.ndarray, dk: float,
                             noalloc: bool = False):
        """(bin_sums, shell_sums, sums) of a full spectrum (gsqg_spectral_diagnostics;
        with ``noalloc`` gsqg_spectral_diagnostics_noalloc, which allocates no
        device memory and takes at most 256 bins)."""
        edges = np.ascontiguousarray(edges, dtype=np.float64)
        nb = edges.shape[0] - 1
        bins = np.zeros(nb)
        shells = np.zeros(self.grid.N // 2 + 1)
        sums = np.zeros(3)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        name = "gsqg_spectral_diagnostics_noalloc" if noalloc else "gsqg_spectral_diagnostics"
        _lib.check(getattr(self.lib, name)(self.handle, _ptr(X), float(alpha), dp(edges), nb,
                                           float(dk), dp(bins), dp(shells), dp(sums), _stream()),
                   name)
        return bins, shells, sums

    def physical_diagnostics(self, X: torch.Tensor):
        """(kinetic_energy, enstrophy, max_theta) of a full spectrum."""
        out = (ctypes.c_double * 3)()
        _lib.check(self.lib.gsqg_physical_diagnostics(self.handle, _ptr(X), out, _stream()),
                   "gsqg_physical_diagnostics")
        return out[0], out[1], out[2]

    KERNEL_CLASSES = ("theta_column", "theta_row", "scalar_column", "scalar_row", "other")

    def cfl_extrema(self, X: torch.Tensor):
        """(max |u|, max |X|) of a full spectrum (gsqg_cfl_extrema); synchronises."""
        out = (ctypes.c_double * 2)()
        _lib.check(self.lib.gsqg_cfl_extrema(self.handle, _ptr(X), out, _stream()),
                   "gsqg_cfl_extrema")
        return out[0], out[1]

    def profile_advance(self, dt: float, n_steps: int, key_mode: int, key: int, step0: int = 0):
        """Advance with a HIP event pair around every kernel; returns
        {class: (total_ms, launches)}."""
        ms = (ctypes.c_double * 5)()
        cnt = (ctypes.c_int * 5)()
        _lib.check(self.lib.gsqg_profile_advance(self.handle, float(dt), int(n_steps), int(key_mode),
                                                 ctypes.c_uint64(key & 0xFFFFFFFFFFFFFFFF),
                                                 int(step0), ms, cnt, _stream()),
                   "gsqg_profile_advance")
        return {c: (ms[i], cnt[i]) for i, c in enumerate(self.KERNEL_CLASSES)}

    # ------------------------------------------------------------ autotuner
    def shape(self):
        """(column shape, row shape, number of shape variants) of this plan."""
        c, r, n = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(self.lib.gsqg_plan_shape(self.handle, ctypes.byref(c), ctypes.byref(r),
                                            ctypes.byref(n)), "gsqg_plan_shape")
        return c.value, r.value, n.value

    def shape_radix(self, shape: int) -> int:
        """Elements per lane (the FFT radix: 4, 8 or 16) of shape index
        ``shape``; the row kernel's where the column tile differs (at N = 8192
        shape 2 pairs a radix-4 row tile with the radix-8 column tile)."""
        r = ctypes.c_int(0)
        _lib.check(self.lib.gsqg_plan_shape_radix(self.handle, int(shape), ctypes.byref(r)),
                   "gsqg_plan_shape_radix")
        return r.value

    def set_shape(self, col: int, row: int) -> None:
        _lib.check(self.lib.gsqg_plan_set_shape(self.handle, int(col), int(row)),
                   "gsqg_plan_set_shape")

    def autotune(self, dt: float, n_steps: int = 32, repeats: int = 3):
        """Time every (column, row) tile-shape pair on ``n_steps`` graphed
        steps (gsqg_autotune), keep the fastest; the state is restored.
        Returns ((col, row), {(col, row): ms})."""
        nv = self.shape()[2]
        best = (ctypes.c_int * 2)()
        table = (ctypes.c_double * (nv * nv))()
        _lib.check(self.lib.gsqg_autotune(self.handle, float(dt), int(n_steps), int(repeats), best,
                                          table, _stream()), "gsqg_autotune")
        times = {(c, r): table[c * nv + r] for c in range(nv) for r in range(nv)}
        return (best[0], best[1]), times

    def nonfinite(self, clear: bool = True) -> bool:
        flag = ctypes.c_int(0)
        _lib.check(self.lib.gsqg_nonfinite(self.handle, ctypes.byref(flag), 1 if clear else 0,
                                           _stream()), "gsqg_nonfinite")
        return bool(flag.value)

    def forcing_stats(self):
        r = (ctypes.c_double * 4)()
        s = (ctypes.c_double * 4)()
        _lib.check(self.lib.gsqg_forcing_stats(self.handle, r, s, _stream()), "gsqg_forcing_stats")
        return list(r), list(s)

    def device_bytes(self) -> int:
        b = ctypes.c_size_t(0)
        _lib.check(self.lib.gsqg_plan_device_bytes(self.handle, ctypes.byref(b)), "device_bytes")
        return b.value


_TRANSFORM_PLANS: dict = {}
_PLAN_CACHE_VEL: dict = {}


_FORCING_PLANS: dict = {}


def forcing_plan(grid: Grid, forcing) -> Plan:
    """Plan carrying a RingForcing's ring list (for RingForcing.__call__).
    The plan holds the ring and the rescale parameters only; an OU forcing's
    state stays with the forcing object, so one plan serves both kinds."""
    from .forcing import RingForcing

    key = (grid.N, grid.L, str(grid.device), forcing.signature())
    p = _FORCING_PLANS.get(key)
    if p is None:
        ring = RingForcing(forcing.kf, forcing.dk, forcing.epsilon, amplitude=forcing.amplitude)
        p = Plan(grid, 1.0, 0.0, 8, forcing=ring)
        _FORCING_PLANS[key] = p
    return p


def transform_plan_n(N: int, device: torch.device) -> Plan:
    """Shared plan for fft2/ifft2 of (N, N) fields on ``device``: a pure
    transform depends on N only, so one plan serves every grid of that size."""
    from .grid import make_grid

    dev = torch.device(device)
    if dev.index is None:
        dev = torch.device(dev.type, torch.cuda.current_device())
    key = (int(N), str(dev))
    p = _TRANSFORM_PLANS.get(key)
    if p is None:
        p = Plan(make_grid(int(N), 2 * np.pi, dev), 1.0, 0.0, 8)
        _TRANSFORM_PLANS[key] = p
    return p


def transform_plan(grid: Grid) -> Plan:
    """Shared plan for fft2/ifft2 on ``grid``'s size and device."""
    return transform_plan_n(grid.N, grid.device)
