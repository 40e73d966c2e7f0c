"""ctypes binding of libgsqg.so (the C ABI declared in include/gsqg.h).

The library is built in-tree by ``pygsquig_amd.build.build()``.  There is no
fallback: if the library is missing or cannot be loaded, every entry point
raises ``GsqgLibraryError``.
"""

from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsqg.so")

GSQG_OK = 0
GSQG_ERR_ARG = 1
GSQG_ERR_HIP = 2
GSQG_ERR_NONFINITE = 3
GSQG_ERR_UNSUPPORTED = 4

SCHEME_EXPLICIT = 0
SCHEME_IF = 1
COMBINE_STEP = 0
COMBINE_MULTISTEP = 1
KEY_EXPLICIT = 0
KEY_DERIVED = 1


class GsqgLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or a call failed."""


class GsqgPlanDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("L", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("nu_p", ctypes.c_double),
        ("p", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("combine", ctypes.c_int),
        ("hv_dc_zero", ctypes.c_int),
        ("damping", ctypes.c_int),
        ("damp_mu", ctypes.c_double),
        ("damp_nu_p", ctypes.c_double),
        ("damp_p", ctypes.c_int),
        ("forcing", ctypes.c_int),
        ("epsilon", ctypes.c_double),
        ("amplitude", ctypes.c_double),
        ("n_scalars", ctypes.c_int),
        ("kappa", ctypes.POINTER(ctypes.c_double)),
        ("k1d", ctypes.POINTER(ctypes.c_double)),
        ("mode_flags", ctypes.POINTER(ctypes.c_uint8)),
        ("ring_modes", ctypes.POINTER(ctypes.c_int32)),
        ("n_ring", ctypes.c_int),
        ("n_members", ctypes.c_int),
        ("member_alpha", ctypes.POINTER(ctypes.c_double)),
    ]


class GsqgCnnBatch(ctypes.Structure):
    _fields_ = [
        ("theta", ctypes.c_void_p),
        ("u", ctypes.c_void_p),
        ("v", ctypes.c_void_p),
        ("img_stride", ctypes.c_longlong),
        ("sel_dev", ctypes.c_void_p),
        ("batch", ctypes.c_int),
        ("in_scale", ctypes.c_double * 3),
        ("target", ctypes.c_void_p),
        ("tgt_scale", ctypes.c_double),
    ]


# name -> (restype, argtypes)
_VP = ctypes.c_void_p
_SIGS = {
    "gsqg_version": (ctypes.c_int, []),
    "gsqg_last_error": (ctypes.c_char_p, []),
    "gsqg_plan_create": (ctypes.c_int, [ctypes.POINTER(GsqgPlanDesc), ctypes.POINTER(_VP)]),
    "gsqg_plan_destroy": (ctypes.c_int, [_VP]),
    "gsqg_plan_device_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    "gsqg_load_theta": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_store_theta": (ctypes.c_int, [_VP, _VP, _VP]),
    "gsqg_load_member": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    "gsqg_store_member": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    "gsqg_load_scalar": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_store_scalar": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    "gsqg_advance": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, _VP]),
    "gsqg_graph_capture": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_uint64]),
    "gsqg_advance_scalars": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_double, _VP]),
    "gsqg_rhs": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint64, _VP]),
    "gsqg_velocity": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "gsqg_fft2": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "gsqg_ifft2": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "gsqg_ring_forcing": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64, _VP, _VP]),
    "gsqg_plan_ring_count": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int)]),
    "gsqg_ring_forcing_ou": (ctypes.c_int, [_VP, _VP, ctypes.c_uint64, ctypes.c_double,
                                            ctypes.c_double, _VP, ctypes.c_int, _VP, _VP]),
    "gsqg_profile_advance": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_uint64, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_int), _VP]),
    "gsqg_autotune": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                     _VP]),
    "gsqg_plan_set_shape": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "gsqg_scalar_rhs": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP, _VP, _VP]),
    "gsqg_batch_create": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.POINTER(_VP)]),
    "gsqg_batch_destroy": (ctypes.c_int, [_VP]),
    "gsqg_batch_advance": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_int, ctypes.c_int, _VP]),
    "gsqg_plan_shape_radix": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gsqg_plan_shape": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "gsqg_nonfinite": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _VP]),
    "gsqg_forcing_stats": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_spectral_diagnostics": (ctypes.c_int, [_VP, _VP, ctypes.c_double,
                                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                                 ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_spectral_diagnostics_noalloc": (ctypes.c_int, [_VP, _VP, ctypes.c_double,
                                                         ctypes.POINTER(ctypes.c_double),
                                                         ctypes.c_int, ctypes.c_double,
                                                         ctypes.POINTER(ctypes.c_double),
                                                         ctypes.POINTER(ctypes.c_double),
                                                         ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_physical_diagnostics": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_cfl_extrema": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_double), _VP]),
    # include/gsqg_closure.h
    "gsqg_cnn_num_params": (ctypes.c_int, []),
    "gsqg_cnn_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(_VP)]),
    "gsqg_cnn_destroy": (ctypes.c_int, [_VP]),
    "gsqg_cnn_device_bytes": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    "gsqg_cnn_set_params": (ctypes.c_int, [_VP, _VP, _VP]),
    "gsqg_cnn_get_params": (ctypes.c_int, [_VP, _VP, _VP]),
    "gsqg_cnn_forward": (ctypes.c_int, [_VP, ctypes.POINTER(GsqgCnnBatch), ctypes.c_double, _VP,
                                        _VP]),
    "gsqg_cnn_gradient": (ctypes.c_int, [_VP, ctypes.POINTER(GsqgCnnBatch), _VP,
                                         ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_cnn_train_step": (ctypes.c_int, [_VP, ctypes.POINTER(GsqgCnnBatch), ctypes.c_double,
                                           ctypes.POINTER(ctypes.c_double), _VP]),
    "gsqg_plan_attach_closure": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_double),
                                                ctypes.c_double, ctypes.c_int]),
    "gsqg_spectral_filter": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.c_int, ctypes.c_double,
                                            _VP, _VP, _VP]),
}

# Every symbol include/gsqg.h and include/gsqg_closure.h declare (checked by
# the CPU test suite).
EXPORTED = tuple(_SIGS)

_lock = threading.Lock()
_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library handle."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise GsqgLibraryError(
                f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m pygsquig_amd.build`")
        try:
            lib = ctypes.CDLL(p)
        except OSError as exc:  # pragma: no cover - depends on the host
            raise GsqgLibraryError(f"cannot load {p}: {exc}") from exc
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def last_error() -> str:
    msg = load().gsqg_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    if rc == GSQG_OK:
        return
    msg = last_error()
    if rc in (GSQG_ERR_ARG,):
        raise ValueError(f"{what}: {msg}")
    if rc == GSQG_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise GsqgLibraryError(f"{what} failed (status {rc}): {msg}")
