"""ctypes binding of ``libswm_hip.so`` (the C ABI of include/swm.h).

There is no fallback: if the HIP library is missing or no GPU is visible, every call that
needs it raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswm_hip.so")

# every symbol include/swm.h declares
EXPORTS = ["swm_version", "swm_device_count", "swm_create", "swm_destroy", "swm_last_error",
           "swm_set_static", "swm_set_static_linear", "swm_upload_prog", "swm_download_prog", "swm_step", "swm_stages",
           "swm_download_rk", "swm_reduce", "swm_time_kernel", "swm_ens_create", "swm_ens_destroy",
           "swm_ens_last_error", "swm_ens_tune", "swm_ens_step", "swm_ens_set_persistent",
           "swm_ens_persistent_steps", "swm_ens_phase_ms", "swm_adj_forward", "swm_adj_pullback",
           "swm_adj_time", "swm_reduce_noalloc", "swm_ghost", "swm_dissipation", "swm_tracer",
           "swm_adj_track_Fx", "swm_adj_download_Fx", "swm_set_static_device", "swm_download_static", "swm_time_steps"]

TOPOGRAPHY = {"flat": 0, "ridge": 1, "ridges": 2, "seamount": 3}
WIND = {"channel": 0, "shear": 1, "double_gyre": 2, "constant": 3, "none": 4}


class SwmSetup(ctypes.Structure):
    """``swm_setup``: the scalars the device setup of the static fields reads."""
    _fields_ = [("topography", ctypes.c_int32), ("wind_x", ctypes.c_int32), ("wind_y", ctypes.c_int32),
                ("pad0", ctypes.c_int32)] + [(n, ctypes.c_double) for n in (
                    "Lx", "Ly", "Delta", "H", "topo_height", "topo_width", "Fx0", "Fy0", "rho", "f0", "beta",
                    "eta_refh", "eta_refw", "A", "phik", "wk", "R", "phi")]

_lib = None


class SwmError(RuntimeError):
    """A non-zero status of the C ABI (message from swm_last_error)."""


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library {LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    pi64 = ctypes.POINTER(i64)
    L.swm_version.argtypes = [ctypes.POINTER(i32)]
    L.swm_device_count.argtypes = [ctypes.POINTER(i32)]
    L.swm_create.argtypes = [vp, ctypes.POINTER(vp)]
    L.swm_destroy.argtypes = [vp]
    L.swm_last_error.argtypes = [vp]
    L.swm_last_error.restype = ctypes.c_char_p
    L.swm_set_static.argtypes = [vp] * 7
    L.swm_set_static_linear.argtypes = [vp] * 3
    L.swm_upload_prog.argtypes = [vp] * 5
    L.swm_download_prog.argtypes = [vp] * 5
    L.swm_step.argtypes = [vp, i64, i64, i64, pi64, pi64]
    L.swm_stages.argtypes = [vp, i32, i64, i64]
    L.swm_download_rk.argtypes = [vp] * 7
    L.swm_ghost.argtypes = [vp]
    L.swm_dissipation.argtypes = [vp]
    L.swm_tracer.argtypes = [vp, i64]
    L.swm_reduce.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.swm_reduce_noalloc.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.swm_time_kernel.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_double)]
    L.swm_ens_create.argtypes = [i32, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    L.swm_ens_destroy.argtypes = [vp]
    L.swm_ens_last_error.argtypes = [vp]
    L.swm_ens_last_error.restype = ctypes.c_char_p
    L.swm_ens_tune.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.swm_ens_step.argtypes = [vp, i64, i64, pi64, pi64, pi64]
    L.swm_ens_set_persistent.argtypes = [vp, i32]
    L.swm_ens_phase_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.swm_ens_persistent_steps.argtypes = [vp, pi64, ctypes.POINTER(ctypes.c_char_p)]
    L.swm_adj_forward.argtypes = [vp, i64, i64, i64, i32]
    L.swm_adj_pullback.argtypes = [vp] * 9 + [pi64, ctypes.POINTER(ctypes.c_double)]
    L.swm_adj_time.argtypes = [vp, i32, i32, ctypes.POINTER(ctypes.c_double)]
    L.swm_adj_track_Fx.argtypes = [vp, i32]
    L.swm_adj_download_Fx.argtypes = [vp, vp]
    L.swm_set_static_device.argtypes = [vp, ctypes.POINTER(SwmSetup)]
    L.swm_download_static.argtypes = [vp] * 7
    L.swm_time_steps.argtypes = [vp, i64, i64, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
    for name in EXPORTS:
        if name not in ("swm_last_error", "swm_ens_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = ""):
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = load().swm_last_error(ctx)
            msg = m.decode() if m else ""
        raise SwmError(f"{what} failed with status {rc}: {msg}")


def ptr(a):
    if a is None:
        return None
    assert a.flags.f_contiguous, "host arrays must be in Julia (Fortran) layout"
    return ctypes.c_void_p(a.ctypes.data)
