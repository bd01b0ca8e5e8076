"""ctypes binding of the oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check or time the device engine.  Nothing in the engine imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    lib.oracle_diffraction_f64.argtypes = [C.POINTER(dp), dp, C.c_size_t, dp, dp, dp,
                                           C.c_size_t, C.c_size_t, C.c_int, dp]
    lib.oracle_diffraction_f64.restype = None
    lib.oracle_diffraction_ld.argtypes = lib.oracle_diffraction_f64.argtypes
    lib.oracle_diffraction_ld.restype = None
    lib.oracle_fallback_f32.argtypes = [fp, C.c_size_t, fp, C.c_size_t, C.c_int, dp]
    lib.oracle_fallback_f32.restype = None
    lib.oracle_reduced_f32.argtypes = [fp, C.c_size_t, fp, C.c_size_t, dp]
    lib.oracle_reduced_f32.restype = None
    lib.oracle_fresnel_grid.argtypes = [C.c_double, C.c_double, C.c_size_t, dp]
    lib.oracle_fresnel_grid.restype = None
    lib.oracle_fresnel_f64.argtypes = [dp, dp, dp, dp, C.c_size_t, C.c_double, C.c_double, dp,
                                       C.c_size_t, dp, C.c_size_t, C.c_size_t, C.c_int, dp]
    lib.oracle_fresnel_f64.restype = None
    lib.oracle_prepare_wave.argtypes = [dp, C.c_double, C.c_double, C.c_size_t, C.c_size_t,
                                        dp, dp, dp]
    lib.oracle_prepare_wave.restype = None
    lib.oracle_max_threads.restype = C.c_int
    return lib


lib = _load()
FIELDS = ("x", "y", "z", "nx", "ny", "nz", "nl", "es_re", "es_im", "ep_re", "ep_im", "energy")


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def diffraction_f64(rays: dict, px, py, pz, path_up=None, stride=1, threads=None):
    """diffraction_integral (diffraction.rs:56-117) on the CPU; returns
    (n_out, 2) complex128 for pixels 0, stride, 2*stride, ..."""
    cols = [np.ascontiguousarray(np.asarray(rays[f], dtype=np.float64)) for f in FIELDS]
    n = cols[0].shape[0]
    px, py, pz = (np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (px, py, pz))
    npx = px.shape[0]
    nout = (npx + stride - 1) // stride
    out = np.zeros((nout, 4))
    arr = (C.POINTER(C.c_double) * 12)(*[_d(c) for c in cols])
    pu = (np.ascontiguousarray(np.asarray(path_up, dtype=np.float64))
          if path_up is not None else None)
    threads = threads or os.cpu_count() or 1
    lib.oracle_diffraction_f64(arr, _d(pu) if pu is not None else None, n, _d(px), _d(py),
                               _d(pz), npx, stride, int(threads), _d(out))
    return out.view(np.complex128).reshape(-1, 2)


def diffraction_ld(rays: dict, px, py, pz, path_up=None, stride=1, threads=None):
    """The same sum with path and phase in x87 extended precision (checker
    for phases ~1e12 rad); returns (n_out, 2) complex128."""
    cols = [np.ascontiguousarray(np.asarray(rays[f], dtype=np.float64)) for f in FIELDS]
    n = cols[0].shape[0]
    px, py, pz = (np.ascontiguousarray(np.asarray(a, dtype=np.float64)) for a in (px, py, pz))
    npx = px.shape[0]
    nout = (npx + stride - 1) // stride
    out = np.zeros((nout, 4))
    arr = (C.POINTER(C.c_double) * 12)(*[_d(c) for c in cols])
    pu = (np.ascontiguousarray(np.asarray(path_up, dtype=np.float64))
          if path_up is not None else None)
    threads = threads or os.cpu_count() or 1
    lib.oracle_diffraction_ld(arr, _d(pu) if pu is not None else None, n, _d(px), _d(py),
                              _d(pz), npx, stride, int(threads), _d(out))
    return out.view(np.complex128).reshape(-1, 2)


def fallback_f32(rays32: np.ndarray, pix32: np.ndarray, threads=None):
    """fallback.rs:29-71 on GpuRay/GpuPixel structured arrays."""
    r = np.ascontiguousarray(rays32)
    p = np.ascontiguousarray(pix32)
    out = np.zeros((p.shape[0], 4))
    threads = threads or os.cpu_count() or 1
    lib.oracle_fallback_f32(r.ctypes.data_as(C.POINTER(C.c_float)), r.shape[0],
                            p.ctypes.data_as(C.POINTER(C.c_float)), p.shape[0], int(threads),
                            _d(out))
    return out.view(np.complex128).reshape(-1, 2)


def reduced_f32(rays32: np.ndarray, pix32: np.ndarray):
    """The reference's f32 phase-reduced GPU method (kirchhoff.rs + wgsl)."""
    r = np.ascontiguousarray(rays32)
    p = np.ascontiguousarray(pix32)
    out = np.zeros((p.shape[0], 4))
    lib.oracle_reduced_f32(r.ctypes.data_as(C.POINTER(C.c_float)), r.shape[0],
                           p.ctypes.data_as(C.POINTER(C.c_float)), p.shape[0], _d(out))
    return out.view(np.complex128).reshape(-1, 2)


def prepare_wave(center, dx, dz, nx, nz):
    c = np.ascontiguousarray(np.asarray(center, dtype=np.float64))
    n = nx * nz
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    lib.oracle_prepare_wave(_d(c), float(dx), float(dz), nx, nz, _d(x), _d(y), _d(z))
    return x, y, z


def fresnel_grid(lo, hi, n):
    v = np.zeros(int(n))
    lib.oracle_fresnel_grid(float(lo), float(hi), int(n), _d(v))
    return v


def fresnel_f64(ray_x, ray_z, ray_es, energy, dist, x_min, x_max, nx, z_min, z_max, nz,
                stride=1, threads=None):
    """fresnel_propagate (src/waves/fresnel.rs:53-106): (n_out,) complex128 for
    pixels 0, stride, 2 stride, ... of the iz * nx + ix order, plus the axes."""
    rx = np.ascontiguousarray(ray_x, dtype=np.float64)
    rz = np.ascontiguousarray(ray_z, dtype=np.float64)
    es = np.ascontiguousarray(ray_es, dtype=np.complex128)
    er, ei = np.ascontiguousarray(es.real), np.ascontiguousarray(es.imag)
    xp, zp = fresnel_grid(x_min, x_max, nx), fresnel_grid(z_min, z_max, nz)
    n_out = (nx * nz + stride - 1) // stride
    out = np.zeros(2 * n_out)
    lib.oracle_fresnel_f64(_d(rx), _d(rz), _d(er), _d(ei), rx.shape[0], float(energy),
                           float(dist), _d(xp), int(nx), _d(zp), int(nz), int(stride),
                           int(threads or lib.oracle_max_threads()), _d(out))
    return out.view(np.complex128), xp, zp
