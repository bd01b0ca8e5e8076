"""ctypes binding of libxrtk.so (the C ABI declared in include/xrtk.h).

The library is loaded from this package directory.  A missing library is an
ImportError: there is no Python or CPU stand-in for the device engine.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxrtk.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libxrtk.so not found at {LIB_PATH}: build it with `make -C xrt-rs_amd` "
        "(or __graft_entry__.build())"
    )

lib = C.CDLL(LIB_PATH)

# ---- status codes (xrtk.h) -------------------------------------------------
OK, ENODEV, EINVAL, ENOMEM, ELAUNCH, EAGAIN, ETIMEDOUT, ESHUTDOWN = range(8)
STATUS_NAMES = {
    OK: "XRTK_OK", ENODEV: "XRTK_ENODEV", EINVAL: "XRTK_EINVAL", ENOMEM: "XRTK_ENOMEM",
    ELAUNCH: "XRTK_ELAUNCH", EAGAIN: "XRTK_EAGAIN", ETIMEDOUT: "XRTK_ETIMEDOUT",
    ESHUTDOWN: "XRTK_ESHUTDOWN",
}
MODE_F32R, MODE_F64 = 0, 1
RAYSRC_AUTO, RAYSRC_SMEM = 0, 1
CHBAR = 1973.2697177417986  # src/core/consts.rs:74


class XrtkError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: {STATUS_NAMES.get(code, code)}")
        self.code = code


def check(code: int, what: str) -> None:
    if code != OK:
        raise XrtkError(code, what)


# ---- structs ----------------------------------------------------------------
class XrtkGpuRay(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "x", "y", "z", "nx", "ny", "nz", "nl", "energy", "es_re", "es_im", "ep_re", "ep_im")]


class XrtkGpuPixel(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "z", "_pad")]


class XrtkRaysF64(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in (
        "x", "y", "z", "nx", "ny", "nz", "nl", "es_re", "es_im", "ep_re", "ep_im", "energy")]


class XrtkWaveParams(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("dx", C.c_double), ("dz", C.c_double),
                ("nx", C.c_uint32), ("nz", C.c_uint32)]


class XrtkOpts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "mode", "pixels_per_lane", "waves_per_group", "tile_rays", "split", "ray_src",
        "persistent", "engine")]


class XrtkJob(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("energy", C.c_double), ("screen", XrtkWaveParams),
                ("out4", C.POINTER(C.c_double))]


class XrtkPlan(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "mode", "pixels_per_lane", "waves_per_group", "tile_rays", "split", "ray_src",
        "persistent", "n_items", "n_fp64", "backend", "inst_flags", "n_mfma")]


class XrtkFresnelScreen(C.Structure):
    _fields_ = [("x_min", C.c_double), ("x_max", C.c_double), ("z_min", C.c_double),
                ("z_max", C.c_double), ("nx", C.c_uint32), ("nz", C.c_uint32)]


class XrtkUndulator(C.Structure):
    _fields_ = [("gamma", C.c_double), ("gamma2", C.c_double), ("beam_current", C.c_double),
                ("kx", C.c_double), ("ky", C.c_double), ("period_mm", C.c_double),
                ("phase_deg", C.c_double), ("n_periods", C.c_uint32), ("n_steps", C.c_uint32)]


class XrtkServerCfg(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int32),
                ("quantum_ms", C.c_double), ("mode", C.c_int32), ("policy", C.c_int32),
                ("opts", C.POINTER(XrtkOpts))]


class XrtkServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("rejected", C.c_uint64), ("launches", C.c_uint64),
                ("max_concurrent", C.c_uint64), ("pairs_done", C.c_double),
                ("busy_ms", C.c_double), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t)]


# numpy views with the exact C layouts
GPU_RAY_DTYPE = np.dtype([(n, np.float32) for n in (
    "x", "y", "z", "nx", "ny", "nz", "nl", "energy", "es_re", "es_im", "ep_re", "ep_im")])
GPU_PIXEL_DTYPE = np.dtype([(n, np.float32) for n in ("x", "y", "z", "_pad")])
assert GPU_RAY_DTYPE.itemsize == C.sizeof(XrtkGpuRay) == 48
assert GPU_PIXEL_DTYPE.itemsize == C.sizeof(XrtkGpuPixel) == 16
GPU_OBS_DTYPE = np.dtype([(n, np.float32) for n in ("energy", "theta", "psi", "_pad")])
assert GPU_OBS_DTYPE.itemsize == 16 and C.sizeof(XrtkUndulator) == 64

# ---- prototypes --------------------------------------------------------------
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_i = C.c_int
_sz = C.c_size_t

PROTOTYPES = {
    "xrtk_opts_default": (None, [C.POINTER(XrtkOpts)]),
    "xrtk_ctx_shared": (_p, []),
    "xrtk_ctx_create": (_i, [_i, C.POINTER(_p)]),
    "xrtk_ctx_destroy": (None, [_p]),
    "xrtk_ctx_device_name": (_i, [_p, C.c_char_p, _sz]),
    "xrtk_kirchhoff_gpu_f32": (_i, [_p, _p, _sz, _p, _sz, _dp, C.POINTER(XrtkOpts)]),
    "xrtk_kirchhoff_gpu_f32_noalloc": (_i, [_p, _p, _sz, _p, _sz, _dp, C.POINTER(XrtkOpts)]),
    "xrtk_kirchhoff_auto_f32": (_i, [_p, _sz, _p, _sz, _dp]),
    "xrtk_diffraction_f64": (_i, [_p, C.POINTER(XrtkRaysF64), _sz, _dp, _dp, _dp, _sz, _dp,
                                  C.POINTER(XrtkOpts)]),
    "xrtk_diffraction_auto_f64": (_i, [C.POINTER(XrtkRaysF64), _sz, _dp, _dp, _dp, _sz, _dp]),
    "xrtk_debug_fail_launches": (None, [C.c_uint32]),
    "xrtk_diffraction_grid_f64": (_i, [_p, C.POINTER(XrtkRaysF64), _sz,
                                       C.POINTER(XrtkWaveParams), _dp, C.POINTER(XrtkOpts)]),
    "xrtk_prepare_wave": (_i, [C.POINTER(XrtkWaveParams), _dp, _dp, _dp]),
    "xrtk_beam_to_rays": (_i, [_dp] * 11 + [_sz, _dp, C.POINTER(_sz), _sz, C.POINTER(_dp)]),
    "xrtk_rayset_upload": (_i, [_p, C.POINTER(XrtkRaysF64), _dp, _sz, C.POINTER(C.c_uint64)]),
    "xrtk_rayset_from_beam": (_i, [_p] + [_dp] * 12 + [_sz, _dp, C.POINTER(_sz), _sz,
                                                   C.POINTER(C.c_uint64)]),
    "xrtk_rayset_release": (_i, [_p, C.c_uint64]),
    "xrtk_rayset_bytes": (_i, [_p, C.c_uint64, C.POINTER(_sz)]),
    "xrtk_solve_jobs": (_i, [_p, C.POINTER(XrtkJob), _sz, C.POINTER(XrtkOpts)]),
    "xrtk_last_timing": (_i, [_p, _dp, _dp]),
    "xrtk_last_plan": (_i, [_p, C.POINTER(XrtkPlan)]),
    "xrtk_auto_last_backend": (_i, []),
    "xrtk_autotune": (_i, [_p, _sz, _sz, _i, C.POINTER(XrtkOpts), _dp]),
    "xrtk_autotune_jobs": (_i, [_p, C.POINTER(XrtkJob), _sz, _i, C.POINTER(XrtkOpts), _dp]),
    "xrtk_undulator_i_map": (_i, [_p, C.POINTER(XrtkUndulator), _dp, _dp, _dp, _sz, _dp, _dp,
                                  _dp]),
    "xrtk_undulator_gpu_f32": (_i, [_p, _p, _sz, C.c_double, C.c_double, C.c_double, _sz,
                                    C.c_double, C.c_double, C.c_uint32, C.c_double, _dp]),
    "xrtk_fresnel_propagate": (_i, [_p, _dp, _dp, _dp, _dp, _sz, C.c_double, C.c_double,
                                    C.POINTER(XrtkFresnelScreen), _dp, _dp, _dp,
                                    C.POINTER(XrtkOpts)]),
    "xrtk_server_create": (_i, [_p, C.POINTER(XrtkServerCfg), C.POINTER(_p)]),
    "xrtk_server_submit": (_i, [_p, C.POINTER(XrtkJob), C.POINTER(C.c_uint64)]),
    "xrtk_server_poll": (_i, [_p, C.c_uint64]),
    "xrtk_server_wait": (_i, [_p, C.c_uint64, C.c_double, _dp]),
    "xrtk_server_stats": (_i, [_p, C.POINTER(XrtkServerStats)]),
    "xrtk_server_destroy": (None, [_p]),
    "xrtk_now": (C.c_double, []),
    "xrtk_host_alloc": (_i, [_sz, C.POINTER(_p)]),
    "xrtk_host_free": (None, [_p]),
}

for _name, (_res, _args) in PROTOTYPES.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def dptr(a: np.ndarray):
    """double* of a C-contiguous float64 array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def opts(mode: int = MODE_F32R, **kw) -> XrtkOpts:
    o = XrtkOpts()
    lib.xrtk_opts_default(C.byref(o))
    o.mode = mode
    for k, v in kw.items():
        setattr(o, k, int(v))
    return o
