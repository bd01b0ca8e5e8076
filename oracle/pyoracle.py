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
    lib.oracle_diffraction_f64_sincos.argtypes = lib.oracle_diffraction_f64.argtypes
    lib.oracle_diffraction_f64_sincos.restype = None
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
    lib.oracle_undulator_gamma.argtypes = [C.c_double]
    lib.oracle_undulator_gamma.restype = C.c_double
    lib.oracle_undulator_e1.argtypes = [C.c_double] * 4
    lib.oracle_undulator_e1.restype = C.c_double
    lib.oracle_undulator_i_map.argtypes = [C.c_double] * 6 + [C.c_size_t, C.c_double, C.c_int,
                                                               dp, dp, dp, C.c_size_t, C.c_int,
                                                               dp, dp, dp]
    lib.oracle_undulator_i_map.restype = None
    lib.oracle_undulator_wgsl_f32.argtypes = [C.c_double] * 5 + [C.c_size_t, C.c_double, C.c_int,
                                                                  fp, C.c_size_t, dp]
    lib.oracle_undulator_wgsl_f32.restype = None
    lib.oracle_prepare_wave.argtypes = [dp, C.c_double, C.c_double, C.c_size_t, C.c_size_t,
                                        dp, dp, dp]
    lib.oracle_prepare_wave.restype = None
    lib.oracle_max_threads.restype = C.c_int
    return lib


lib = _load()
FIELDS = ("x", "y", "z", "nx", "ny", "nz", "nl", "es_re", "es_im", "ep_re", "ep_im", "energy")


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def diffraction_f64(rays: dict, px, py, pz, path_up=None, stride=1, threads=None,
                    trig="sin_cos"):
    """diffraction_integral (diffraction.rs:56-117) on the CPU; returns
    (n_out, 2) complex128 for pixels 0, stride, 2*stride, ...  trig
    "sin_cos" calls sin() and cos() per pair, "sincos" one sincos() (the
    same bits)."""
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
    fn = {"sin_cos": lib.oracle_diffraction_f64, "sincos": lib.oracle_diffraction_f64_sincos}[trig]
    fn(arr, _d(pu) if pu is not None else None, n, _d(px), _d(py), _d(pz), npx, stride,
       int(threads), _d(out))
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


def undulator_gamma(electron_energy_gev):
    return lib.oracle_undulator_gamma(float(electron_energy_gev))


def undulator_e1(gamma, kx, ky, period_mm):
    return lib.oracle_undulator_e1(float(gamma), float(kx), float(ky), float(period_mm))


def undulator_i_map(und: dict, energies, thetas, psis, n_steps=64, threads=None):
    """build_i_map (src/sources/undulator.rs:120-244).  und: gamma, gamma2,
    beam_current, kx, ky, period_mm, n_periods, phase_deg.  Returns
    (intensity, amp_s, amp_p) as float64 / complex128 arrays."""
    e = np.ascontiguousarray(energies, dtype=np.float64)
    t = np.ascontiguousarray(thetas, dtype=np.float64)
    p = np.ascontiguousarray(psis, dtype=np.float64)
    n = e.shape[0]
    inten, a_s, a_p = np.zeros(n), np.zeros(2 * n), np.zeros(2 * n)
    lib.oracle_undulator_i_map(und["gamma"], und["gamma2"], und["beam_current"], und["kx"],
                               und["ky"], und["period_mm"], int(und["n_periods"]),
                               und.get("phase_deg", 0.0), int(n_steps), _d(e), _d(t), _d(p), n,
                               int(threads or lib.oracle_max_threads()), _d(inten), _d(a_s),
                               _d(a_p))
    return inten, a_s.view(np.complex128), a_p.view(np.complex128)


def undulator_wgsl_f32(und: dict, energies, thetas, psis, n_steps=64):
    """The reference GPU path (undulator.wgsl) in float32: (intensity,
    amp_s, amp_p)."""
    e = np.asarray(energies, dtype=np.float64)
    obs = np.zeros((e.shape[0], 4), dtype=np.float32)
    obs[:, 0], obs[:, 1], obs[:, 2] = e, thetas, psis
    out = np.zeros(5 * e.shape[0])
    lib.oracle_undulator_wgsl_f32(und["gamma"], und["beam_current"], und["kx"], und["ky"],
                                  und["period_mm"], int(und["n_periods"]),
                                  und.get("phase_deg", 0.0), int(n_steps),
                                  obs.ctypes.data_as(C.POINTER(C.c_float)), e.shape[0], _d(out))
    out = out.reshape(-1, 5)
    return out[:, 4].copy(), out[:, 0] + 1j * out[:, 1], out[:, 2] + 1j * out[:, 3]


def undulator(electron_energy_gev, beam_current, kx, ky, period_mm, n_periods, phase_deg=0.0):
    """The parameters of Undulator::new (src/sources/undulator.rs:64-92) as
    the map functions take them."""
    g = undulator_gamma(electron_energy_gev)
    return {"gamma": g, "gamma2": g * g, "beam_current": float(beam_current), "kx": float(kx),
            "ky": float(ky), "period_mm": float(period_mm), "n_periods": int(n_periods),
            "phase_deg": float(phase_deg)}
