"""xrtk — MI355X (gfx950) Kirchhoff diffraction engine, Python host interface.

Mirrors the Kirchhoff path of physwkim/xrt-rs:

  GpuRay / GpuPixel ............. src/gpu/kirchhoff.rs:19-45   (numpy dtypes)
  GpuContext.new() / .shared() .. src/gpu/context.rs:20-31
  kirchhoff_gpu(ctx, rays, px) .. src/gpu/kirchhoff.rs:77-83
  kirchhoff_auto(rays, px) ...... src/gpu/fallback.rs:19-26
  diffraction_integral(rays, px)  src/waves/diffraction.rs:56-64
  diffraction_integral_rs(...) .. crates/xrt-python/src/lib.rs:321-394
  WaveParams / prepare_wave ..... src/waves/prepare_wave.rs:11-68
  beam_to_diffraction_rays ...... src/waves/prepare_wave.rs:78-111

Everything computes on the device through libxrtk's C ABI.  The only CPU
arithmetic is the auto entries' branch when no gfx950 device exists, the
reference's own behaviour: kirchhoff_auto's f64 loop (fallback.rs:24-25) and,
for diffraction_integral / diffraction_integral_rs in their default f64 mode,
the f64 Kahan loop those functions always run (diffraction.rs:56-117).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import (CHBAR, EAGAIN, GPU_OBS_DTYPE, GPU_PIXEL_DTYPE, GPU_RAY_DTYPE, MODE_F32R,
                   MODE_F64, OK, XrtkError, check, dptr, lib)

__all__ = [
    "CHBAR", "GPU_RAY_DTYPE", "GPU_PIXEL_DTYPE", "GpuContext", "kirchhoff_gpu",
    "kirchhoff_auto", "diffraction_integral", "diffraction_integral_rs", "WaveParams",
    "prepare_wave", "beam_to_diffraction_rays", "RaySet", "Job", "solve_jobs", "autotune_jobs",
    "pinned_result",
    "Server", "FresnelScreen", "FresnelResult", "fresnel_propagate", "Undulator",
    "UndulatorGpuResult", "undulator_gpu", "GPU_OBS_DTYPE",
    "XrtkError", "MODE_F32R", "MODE_F64", "RAY_FIELDS",
]

RAY_FIELDS = ("x", "y", "z", "nx", "ny", "nz", "nl", "es_re", "es_im", "ep_re", "ep_im",
              "energy")

_MODES = {"f32r": MODE_F32R, "f64": MODE_F64, MODE_F32R: MODE_F32R, MODE_F64: MODE_F64}


def _mode(m) -> int:
    try:
        return _MODES[m]
    except KeyError:
        raise ValueError(f"unknown mode {m!r}: use 'f32r' or 'f64'") from None


def gpu_rays(n: int) -> np.ndarray:
    """Zeroed GpuRay array (48-byte records)."""
    return np.zeros(n, dtype=GPU_RAY_DTYPE)


def gpu_pixels(x, y, z) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    out = np.zeros(x.shape[0], dtype=GPU_PIXEL_DTYPE)
    out["x"], out["y"], out["z"] = x, np.asarray(y, np.float32), np.asarray(z, np.float32)
    return out


class GpuContext:
    """A device context (GpuContext, src/gpu/context.rs:6-10)."""

    def __init__(self, handle, owned: bool):
        self._h = handle
        self._owned = owned

    @classmethod
    def new(cls, device: int = 0) -> Optional["GpuContext"]:
        h = C.c_void_p()
        rc = lib.xrtk_ctx_create(device, C.byref(h))
        if rc != OK:
            return None
        return cls(h, True)

    @classmethod
    def shared(cls) -> Optional["GpuContext"]:
        h = lib.xrtk_ctx_shared()
        if not h:
            return None
        return cls(C.c_void_p(h), False)

    @property
    def handle(self):
        return self._h

    def adapter_name(self) -> str:
        buf = C.create_string_buffer(256)
        check(lib.xrtk_ctx_device_name(self._h, buf, 256), "xrtk_ctx_device_name")
        return buf.value.decode()

    def last_timing(self):
        """(pair-kernel ms, whole device pipeline ms) of the last solve."""
        a, b = C.c_double(), C.c_double()
        check(lib.xrtk_last_timing(self._h, C.byref(a), C.byref(b)), "xrtk_last_timing")
        return a.value, b.value

    def last_plan(self) -> dict:
        p = _lib.XrtkPlan()
        check(lib.xrtk_last_plan(self._h, C.byref(p)), "xrtk_last_plan")
        return {f: getattr(p, f) for f, _ in p._fields_}

    def autotune(self, n_rays: int, n_pixels: int, mode="f32r"):
        o = _lib.XrtkOpts()
        ms = C.c_double()
        check(lib.xrtk_autotune(self._h, n_rays, n_pixels, _mode(mode), C.byref(o), C.byref(ms)),
              "xrtk_autotune")
        return {f: getattr(o, f) for f, _ in o._fields_}, ms.value

    def close(self):
        if self._owned and self._h:
            lib.xrtk_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_rays32(rays) -> np.ndarray:
    r = np.ascontiguousarray(rays)
    if r.dtype != GPU_RAY_DTYPE:
        raise TypeError("rays must be a GPU_RAY_DTYPE array (GpuRay layout)")
    return r


def _as_px32(pixels) -> np.ndarray:
    p = np.ascontiguousarray(pixels)
    if p.dtype != GPU_PIXEL_DTYPE:
        raise TypeError("pixels must be a GPU_PIXEL_DTYPE array (GpuPixel layout)")
    return p


def _to_complex(out4: np.ndarray) -> np.ndarray:
    """np x 4 f64 -> np x 2 complex128 [(Es, Ep)]."""
    return out4.view(np.complex128).reshape(-1, 2)


def kirchhoff_gpu(ctx: GpuContext, rays, pixels, mode="f32r", **tune) -> np.ndarray:
    """kirchhoff_gpu (src/gpu/kirchhoff.rs:77): (Es, Ep) per pixel as an
    (n_pixels, 2) complex128 array.  Raises XrtkError where the reference
    panics."""
    r, p = _as_rays32(rays), _as_px32(pixels)
    out = np.zeros((p.shape[0], 4), dtype=np.float64)
    o = _lib.opts(_mode(mode), **tune)
    check(lib.xrtk_kirchhoff_gpu_f32(ctx.handle, r.ctypes.data, r.shape[0], p.ctypes.data,
                                     p.shape[0], dptr(out), C.byref(o)),
          "xrtk_kirchhoff_gpu_f32")
    return _to_complex(out)


def kirchhoff_auto(rays, pixels) -> np.ndarray:
    """kirchhoff_auto (src/gpu/fallback.rs:19): the device when present, else
    the reference's f64 CPU loop (fallback.rs:29-71)."""
    r, p = _as_rays32(rays), _as_px32(pixels)
    out = np.zeros((p.shape[0], 4), dtype=np.float64)
    check(lib.xrtk_kirchhoff_auto_f32(r.ctypes.data, r.shape[0], p.ctypes.data, p.shape[0],
                                      dptr(out)), "xrtk_kirchhoff_auto_f32")
    return _to_complex(out)


def auto_last_backend() -> int:
    """1 if the last kirchhoff_auto ran on the device, 0 on the CPU fallback."""
    return lib.xrtk_auto_last_backend()


class _RayCols:
    """Keeps contiguous f64 columns alive and exposes an XrtkRaysF64."""

    def __init__(self, cols: dict, n: int):
        self.arrays = {}
        for f in RAY_FIELDS:
            a = np.ascontiguousarray(np.asarray(cols[f], dtype=np.float64).reshape(-1))
            if a.shape[0] != n:
                raise ValueError(f"array length mismatch: {f} has {a.shape[0]} elements, "
                                 f"expected {n}")
            self.arrays[f] = a
        self.s = _lib.XrtkRaysF64(*[dptr(self.arrays[f]) for f in RAY_FIELDS])


def _ray_count(rays: dict) -> int:
    return int(np.asarray(rays["x"]).reshape(-1).shape[0])


def diffraction_integral(rays: dict, pixels, ctx: Optional[GpuContext] = None,
                         mode="f64", **tune) -> np.ndarray:
    """diffraction_integral (src/waves/diffraction.rs:56): rays as a dict of
    the DiffractionRay columns (RAY_FIELDS), pixels as (px, py, pz) arrays or a
    WaveParams.  Returns (n_pixels, 2) complex128 [(Es, Ep)].

    mode "f64" (the default, the reference's arithmetic) runs the plain FP64
    device kernel; without a device (and no ctx given) it runs the
    reference's f64 Kahan loop on the CPU, like the reference itself.
    mode "f32r" runs the FP32 phase-reduced kernel (complex rel-L2 ~1e-6 to
    ~1e-5 against FP64) and needs a device."""
    m = _mode(mode)
    if ctx is None:
        ctx = GpuContext.shared()
        if ctx is None and m == MODE_F64 and not tune:
            return _diffraction_auto_f64(rays, pixels)
    if ctx is None:
        raise XrtkError(_lib.ENODEV, "diffraction_integral")
    n = _ray_count(rays)
    rc = _RayCols(rays, n)
    o = _lib.opts(m, **tune)
    if isinstance(pixels, WaveParams):
        w = pixels.to_c()
        out = np.zeros((pixels.nx * pixels.nz, 4), dtype=np.float64)
        check(lib.xrtk_diffraction_grid_f64(ctx.handle, C.byref(rc.s), n, C.byref(w), dptr(out),
                                            C.byref(o)), "xrtk_diffraction_grid_f64")
        return _to_complex(out)
    px, py, pz = (np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
                  for a in pixels)
    if not (px.shape == py.shape == pz.shape):
        raise ValueError("pixel coordinate arrays differ in length")
    out = np.zeros((px.shape[0], 4), dtype=np.float64)
    check(lib.xrtk_diffraction_f64(ctx.handle, C.byref(rc.s), n, dptr(px), dptr(py), dptr(pz),
                                   px.shape[0], dptr(out), C.byref(o)), "xrtk_diffraction_f64")
    return _to_complex(out)


def _diffraction_auto_f64(rays: dict, pixels) -> np.ndarray:
    """xrtk_diffraction_auto_f64: the FP64 device kernel on the shared
    context, or the reference's f64 Kahan loop on the CPU without a device."""
    n = _ray_count(rays)
    rc = _RayCols(rays, n)
    if isinstance(pixels, WaveParams):
        pixels = prepare_wave(pixels)
    px, py, pz = (np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
                  for a in pixels)
    if not (px.shape == py.shape == pz.shape):
        raise ValueError("pixel coordinate arrays differ in length")
    out = np.zeros((px.shape[0], 4), dtype=np.float64)
    check(lib.xrtk_diffraction_auto_f64(C.byref(rc.s), n, dptr(px), dptr(py), dptr(pz),
                                        px.shape[0], dptr(out)), "xrtk_diffraction_auto_f64")
    return _to_complex(out)


def _check_equal_lengths(pairs):
    """check_equal_lengths, crates/xrt-python/src/lib.rs:33-47."""
    if len(pairs) < 2:
        return
    ref_name, ref_len = pairs[0]
    for name, ln in pairs[1:]:
        if ln != ref_len:
            raise ValueError(f"array length mismatch: {name} has {ln} elements but "
                             f"{ref_name} has {ref_len}")


def diffraction_integral_rs(ray_x, ray_y, ray_z, ray_nx, ray_ny, ray_nz, ray_nl, ray_es_re,
                            ray_es_im, ray_ep_re, ray_ep_im, ray_energy, pix_x, pix_y, pix_z,
                            mode="f64"):
    """diffraction_integral_rs (crates/xrt-python/src/lib.rs:321-394): fifteen
    float sequences in, (es_re, es_im, ep_re, ep_im) lists out, ValueError on
    a length mismatch.  The default f64 mode gives the reference's result
    (plain FP64 device kernel, or the reference's f64 Kahan loop on the CPU
    without a device); mode="f32r" opts into the FP32 phase-reduced kernel.
    The library call runs without holding any Python lock (ctypes releases
    the GIL)."""
    names = ["ray_x", "ray_y", "ray_z", "ray_nx", "ray_ny", "ray_nz", "ray_nl", "ray_es_re",
             "ray_es_im", "ray_ep_re", "ray_ep_im", "ray_energy"]
    cols = [ray_x, ray_y, ray_z, ray_nx, ray_ny, ray_nz, ray_nl, ray_es_re, ray_es_im,
            ray_ep_re, ray_ep_im, ray_energy]
    arrs = [np.asarray(c, dtype=np.float64).reshape(-1) for c in cols]
    _check_equal_lengths([(nm, a.shape[0]) for nm, a in zip(names, arrs)])
    pxs = [np.asarray(c, dtype=np.float64).reshape(-1) for c in (pix_x, pix_y, pix_z)]
    _check_equal_lengths([(nm, a.shape[0]) for nm, a in zip(("pix_x", "pix_y", "pix_z"), pxs)])
    rays = dict(zip(RAY_FIELDS, [arrs[0], arrs[1], arrs[2], arrs[3], arrs[4], arrs[5], arrs[6],
                                 arrs[7], arrs[8], arrs[9], arrs[10], arrs[11]]))
    if _mode(mode) == MODE_F64:
        res = _diffraction_auto_f64(rays, tuple(pxs))
    else:
        res = diffraction_integral(rays, tuple(pxs), mode=mode)
    es, ep = res[:, 0], res[:, 1]
    return (es.real.tolist(), es.imag.tolist(), ep.real.tolist(), ep.imag.tolist())


@dataclass
class WaveParams:
    """WaveParams (src/waves/prepare_wave.rs:11-35)."""
    center: Sequence[float]
    dx: float
    dz: float
    nx: int
    nz: int

    def to_c(self) -> _lib.XrtkWaveParams:
        w = _lib.XrtkWaveParams()
        for i in range(3):
            w.center[i] = float(self.center[i])
        w.dx, w.dz, w.nx, w.nz = float(self.dx), float(self.dz), int(self.nx), int(self.nz)
        return w


def prepare_wave(params: WaveParams):
    """prepare_wave (src/waves/prepare_wave.rs:40-68): (x, y, z) arrays of the
    nx*nz pixels, x-major with z fastest."""
    n = params.nx * params.nz
    x, y, z = (np.zeros(n) for _ in range(3))
    w = params.to_c()
    check(lib.xrtk_prepare_wave(C.byref(w), dptr(x), dptr(y), dptr(z)), "xrtk_prepare_wave")
    return x, y, z


def beam_to_diffraction_rays(beam: dict, normals, good) -> dict:
    """beam_to_diffraction_rays (src/waves/prepare_wave.rs:78-111).  beam holds
    x, y, z, a, b, c, e and optionally es/ep (complex); normals is (n, 3)."""
    cols = {k: np.ascontiguousarray(np.asarray(beam[k], dtype=np.float64)) for k in
            ("x", "y", "z", "a", "b", "c", "e")}
    n_beam = cols["x"].shape[0]
    normals = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(n_beam, 3))
    good = np.ascontiguousarray(np.asarray(good, dtype=np.uint64).reshape(-1))
    amps = None
    if beam.get("es") is not None and beam.get("ep") is not None:
        es = np.asarray(beam["es"], dtype=np.complex128)
        ep = np.asarray(beam["ep"], dtype=np.complex128)
        amps = [np.ascontiguousarray(a) for a in (es.real, es.imag, ep.real, ep.imag)]
    ng = good.shape[0]
    out = [np.zeros(ng) for _ in range(12)]
    outp = (C.POINTER(C.c_double) * 12)(*[dptr(a) for a in out])
    nul = C.POINTER(C.c_double)()
    a4 = [dptr(a) for a in amps] if amps else [nul] * 4
    check(lib.xrtk_beam_to_rays(dptr(cols["x"]), dptr(cols["y"]), dptr(cols["z"]),
                                dptr(cols["a"]), dptr(cols["b"]), dptr(cols["c"]),
                                dptr(cols["e"]), *a4, n_beam, dptr(normals),
                                good.ctypes.data_as(C.POINTER(C.c_size_t)), ng, outp),
          "xrtk_beam_to_rays")
    return dict(zip(RAY_FIELDS, out))


class RaySet:
    """A ray set resident in HBM (xrtk_rayset_upload)."""

    def __init__(self, ctx: GpuContext, rays: dict, path_up=None):
        self.ctx = ctx
        n = _ray_count(rays)
        rc = _RayCols(rays, n)
        pu = None
        if path_up is not None:
            pu = np.ascontiguousarray(np.asarray(path_up, dtype=np.float64).reshape(-1))
            if pu.shape[0] != n:
                raise ValueError("path_up length mismatch")
        h = C.c_uint64()
        check(lib.xrtk_rayset_upload(ctx.handle, C.byref(rc.s),
                                     dptr(pu) if pu is not None else C.POINTER(C.c_double)(),
                                     n, C.byref(h)), "xrtk_rayset_upload")
        self.handle = h.value
        self.n = n

    @classmethod
    def from_beam(cls, ctx: GpuContext, beam: dict, normals, good, path=None) -> "RaySet":
        """beam_to_diffraction_rays (src/waves/prepare_wave.rs:78-111) on the
        device into a resident ray set: beam holds x, y, z, a, b, c, e and
        optionally es / ep (complex); normals is (n, 3); path (optional) the
        beam's optical path per beam ray, folded into the phase."""
        cols = {k: np.ascontiguousarray(np.asarray(beam[k], dtype=np.float64)) for k in
                ("x", "y", "z", "a", "b", "c", "e")}
        n_beam = cols["x"].shape[0]
        normals = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(n_beam, 3))
        good = np.ascontiguousarray(np.asarray(good, dtype=np.uint64).reshape(-1))
        nul = C.POINTER(C.c_double)()
        amps = [nul] * 4
        keep = []
        if beam.get("es") is not None and beam.get("ep") is not None:
            es = np.asarray(beam["es"], dtype=np.complex128)
            ep = np.asarray(beam["ep"], dtype=np.complex128)
            keep = [np.ascontiguousarray(a) for a in (es.real, es.imag, ep.real, ep.imag)]
            amps = [dptr(a) for a in keep]
        pth = None
        if path is not None:
            pth = np.ascontiguousarray(np.asarray(path, dtype=np.float64).reshape(-1))
            if pth.shape[0] != n_beam:
                raise ValueError("path length mismatch")
        h = C.c_uint64()
        check(lib.xrtk_rayset_from_beam(ctx.handle, *[dptr(cols[k]) for k in
                                                     ("x", "y", "z", "a", "b", "c", "e")],
                                        *amps, dptr(pth) if pth is not None else nul, n_beam,
                                        dptr(normals), good.ctypes.data_as(C.POINTER(C.c_size_t)),
                                        good.shape[0], C.byref(h)), "xrtk_rayset_from_beam")
        self = cls.__new__(cls)
        self.ctx = ctx
        self.handle = h.value
        self.n = int(good.shape[0])
        return self

    def nbytes(self) -> int:
        b = C.c_size_t()
        check(lib.xrtk_rayset_bytes(self.ctx.handle, self.handle, C.byref(b)), "xrtk_rayset_bytes")
        return b.value

    def release(self):
        if self.handle:
            check(lib.xrtk_rayset_release(self.ctx.handle, self.handle), "xrtk_rayset_release")
            self.handle = 0


@dataclass
class Job:
    rays: RaySet
    energy: float
    screen: WaveParams

    def to_c(self, out4: np.ndarray) -> _lib.XrtkJob:
        j = _lib.XrtkJob()
        j.rays = self.rays.handle
        j.energy = float(self.energy)
        j.screen = self.screen.to_c()
        j.out4 = dptr(out4)
        return j


def pinned_result(n_pixels: int) -> np.ndarray:
    """An (n_pixels, 2) complex128 result buffer [(Es, Ep)] in page-locked
    host memory (xrtk_host_alloc), freed when the array is collected: the
    out= buffer of repeated solves."""
    import weakref
    nbytes = max(1, int(n_pixels)) * 32
    p = C.c_void_p()
    check(lib.xrtk_host_alloc(nbytes, C.byref(p)), "xrtk_host_alloc")
    buf = (C.c_char * nbytes).from_address(p.value)
    arr = np.frombuffer(buf, dtype=np.complex128, count=2 * int(n_pixels)).reshape(-1, 2)
    weakref.finalize(buf, lib.xrtk_host_free, p.value)
    return arr


def _out4(a: np.ndarray, n_pixels: int) -> np.ndarray:
    """The np x 4 f64 view of a caller's (n_pixels, 2) complex128 buffer."""
    if (not isinstance(a, np.ndarray) or a.dtype != np.complex128 or a.shape != (n_pixels, 2)
            or not a.flags.c_contiguous or not a.flags.writeable):
        raise ValueError(f"out: need a writable C-contiguous ({n_pixels}, 2) complex128 array")
    return a.view(np.float64).reshape(-1, 4)


def solve_jobs(ctx: GpuContext, jobs: Sequence[Job], mode="f32r", out=None, **tune):
    """All jobs in one persistent launch (hipGraph-replayed while the batch
    repeats).  Returns one (n_pixels, 2) complex128 array per job; with out=
    (one such array per job, e.g. from pinned_result) the results are written
    there and those arrays are returned."""
    if out is None:
        outs = [np.zeros((j.screen.nx * j.screen.nz, 4)) for j in jobs]
    else:
        if len(out) != len(jobs):
            raise ValueError("out: one buffer per job")
        outs = [_out4(a, j.screen.nx * j.screen.nz) for a, j in zip(out, jobs)]
    arr = (_lib.XrtkJob * len(jobs))(*[j.to_c(o) for j, o in zip(jobs, outs)])
    o = _lib.opts(_mode(mode), **tune)
    check(lib.xrtk_solve_jobs(ctx.handle, arr, len(jobs), C.byref(o)), "xrtk_solve_jobs")
    return list(out) if out is not None else [_to_complex(x) for x in outs]


def autotune_jobs(ctx: GpuContext, jobs: Sequence[Job], mode="f32r"):
    """Times the launch shapes on this batch (each instance cut to 65536
    rays) and fixes the fastest as the context's batched-solve configuration.
    Returns (chosen options, device ms of the winner on the cut batch)."""
    outs = [np.zeros((j.screen.nx * j.screen.nz, 4)) for j in jobs]
    arr = (_lib.XrtkJob * len(jobs))(*[j.to_c(o) for j, o in zip(jobs, outs)])
    o = _lib.XrtkOpts()
    ms = C.c_double()
    check(lib.xrtk_autotune_jobs(ctx.handle, arr, len(jobs), _mode(mode), C.byref(o),
                                 C.byref(ms)), "xrtk_autotune_jobs")
    return {f: getattr(o, f) for f, _ in o._fields_}, ms.value


class Server:
    """Continuous-batching solver (xrtk_server_*)."""

    def __init__(self, ctx: GpuContext, hbm_budget_bytes: int, max_active: int = 64,
                 quantum_ms: float = 10.0, mode="f32r", policy="shortest", **shape):
        """shape: launch-shape options (pixels_per_lane, waves_per_group,
        persistent, ...) instead of the context's batch configuration."""
        pol = {"shortest": 0, "arrival": 1}[policy]
        self._shape = _lib.opts(_mode(mode), **shape) if shape else None
        cfg = _lib.XrtkServerCfg(int(hbm_budget_bytes), int(max_active), float(quantum_ms),
                                 _mode(mode), pol,
                                 C.pointer(self._shape) if self._shape is not None else None)
        h = C.c_void_p()
        check(lib.xrtk_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "xrtk_server_create")
        self._h = h
        self._keep = {}

    def submit(self, job: Job) -> int:
        out = np.zeros((job.screen.nx * job.screen.nz, 4))
        cj = job.to_c(out)
        t = C.c_uint64()
        check(lib.xrtk_server_submit(self._h, C.byref(cj), C.byref(t)), "xrtk_server_submit")
        self._keep[t.value] = out
        return t.value

    def poll(self, ticket: int) -> bool:
        rc = lib.xrtk_server_poll(self._h, ticket)
        if rc == EAGAIN:
            return False
        check(rc, "xrtk_server_poll")
        return True

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """Blocks; returns ((n, 2) complex128 result, [submit, admit, first
        launch, done] timestamps)."""
        ts = np.zeros(4)
        check(lib.xrtk_server_wait(self._h, ticket, float(timeout_ms), dptr(ts)),
              "xrtk_server_wait")
        out = self._keep.pop(ticket)
        return _to_complex(out), ts

    def stats(self) -> dict:
        s = _lib.XrtkServerStats()
        check(lib.xrtk_server_stats(self._h, C.byref(s)), "xrtk_server_stats")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if self._h:
            lib.xrtk_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class FresnelScreen:
    """FresnelScreen (src/waves/fresnel.rs:13-27)."""
    x_min: float
    x_max: float
    nx: int
    z_min: float
    z_max: float
    nz: int

    def to_c(self) -> _lib.XrtkFresnelScreen:
        return _lib.XrtkFresnelScreen(float(self.x_min), float(self.x_max), float(self.z_min),
                                      float(self.z_max), int(self.nx), int(self.nz))


@dataclass
class FresnelResult:
    """FresnelResult (src/waves/fresnel.rs:29-48): amplitude[iz * nx + ix]."""
    amplitude: np.ndarray
    x_pixels: np.ndarray
    z_pixels: np.ndarray
    nx: int
    nz: int

    def intensity(self) -> np.ndarray:
        return np.abs(self.amplitude) ** 2


def fresnel_propagate(ray_x, ray_z, ray_es, energy: float, dist: float, screen: FresnelScreen,
                      ctx: Optional[GpuContext] = None, mode="f32r",
                      engine: int = 0) -> FresnelResult:
    """fresnel_propagate (src/waves/fresnel.rs:53-106) on the device: source
    rays at (x, z) on the plane z = 0 with complex amplitudes ray_es,
    propagated dist mm at energy eV onto the screen grid.  mode "f32r":
    FP32 factors from FP64 phases, FP32 sums per ray slice, FP64 merge; the
    products on FP16 hi + lo parts (engine 0) or on FP32 operands
    (engine 1).  mode "f64": FP64 throughout."""
    rx = np.ascontiguousarray(np.asarray(ray_x, dtype=np.float64).reshape(-1))
    rz = np.ascontiguousarray(np.asarray(ray_z, dtype=np.float64).reshape(-1))
    es = np.asarray(ray_es, dtype=np.complex128).reshape(-1)
    if not (rx.shape == rz.shape == es.shape):
        raise ValueError("ray_x, ray_z and ray_es differ in length")
    if int(screen.nx) < 1 or int(screen.nz) < 1:
        raise ValueError("screen needs nx >= 1 and nz >= 1")
    ctx = ctx or GpuContext.shared()
    if ctx is None:
        raise XrtkError(_lib.ENODEV, "fresnel_propagate")
    er, ei = np.ascontiguousarray(es.real), np.ascontiguousarray(es.imag)
    out = np.zeros(2 * int(screen.nx) * int(screen.nz))
    xp, zp = np.zeros(int(screen.nx)), np.zeros(int(screen.nz))
    sc = screen.to_c()
    o = _lib.opts(_mode(mode), engine=engine)
    check(lib.xrtk_fresnel_propagate(ctx.handle, dptr(rx), dptr(rz), dptr(er), dptr(ei),
                                     rx.shape[0], float(energy), float(dist), C.byref(sc),
                                     dptr(out), dptr(xp), dptr(zp), C.byref(o)),
          "xrtk_fresnel_propagate")
    return FresnelResult(out.view(np.complex128), xp, zp, int(screen.nx), int(screen.nz))


M0C2 = 0.510998928  # src/core/consts.rs:33 [MeV]


@dataclass
class Undulator:
    """The parameters of Undulator::new (src/sources/undulator.rs:64-92) that
    the intensity map reads; gamma from SynchrotronParams::new
    (bending_magnet.rs:38-53)."""
    electron_energy_gev: float
    beam_current: float
    kx: float
    ky: float
    period_mm: float
    n_periods: int
    phase_deg: float = 0.0

    @property
    def gamma(self) -> float:
        return self.electron_energy_gev * 1e3 / M0C2

    def k_squared(self) -> float:
        return self.kx * self.kx + self.ky * self.ky

    def to_c(self, n_steps: int = 64) -> _lib.XrtkUndulator:
        g = self.gamma
        return _lib.XrtkUndulator(g, g * g, float(self.beam_current), float(self.kx),
                                  float(self.ky), float(self.period_mm), float(self.phase_deg),
                                  int(self.n_periods), int(n_steps))

    def build_i_map(self, energies, thetas, psis, ctx: Optional[GpuContext] = None,
                    n_steps: int = 64):
        """build_i_map (src/sources/undulator.rs:120-244) on the device, FP64:
        (intensity, amp_s, amp_p) as float64 / complex128 arrays."""
        e = np.ascontiguousarray(np.asarray(energies, dtype=np.float64).reshape(-1))
        t = np.ascontiguousarray(np.asarray(thetas, dtype=np.float64).reshape(-1))
        p = np.ascontiguousarray(np.asarray(psis, dtype=np.float64).reshape(-1))
        if not (e.shape == t.shape == p.shape):
            raise ValueError("energies, thetas and psis differ in length")
        ctx = ctx or GpuContext.shared()
        if ctx is None:
            raise XrtkError(_lib.ENODEV, "build_i_map")
        n = e.shape[0]
        inten, a_s, a_p = np.zeros(n), np.zeros(2 * n), np.zeros(2 * n)
        u = self.to_c(n_steps)
        check(lib.xrtk_undulator_i_map(ctx.handle, C.byref(u), dptr(e), dptr(t), dptr(p), n,
                                       dptr(inten), dptr(a_s), dptr(a_p)), "xrtk_undulator_i_map")
        return inten, a_s.view(np.complex128), a_p.view(np.complex128)


@dataclass
class UndulatorGpuResult:
    """UndulatorGpuResult (src/gpu/undulator.rs:57-64)."""
    amp_s: np.ndarray
    amp_p: np.ndarray
    intensity: np.ndarray


def undulator_gpu(ctx: GpuContext, obs_points, kx, ky, period_mm, n_periods, gamma,
                  beam_current, n_integration_steps=64, phase_deg=0.0) -> UndulatorGpuResult:
    """undulator_gpu (src/gpu/undulator.rs:66-91): obs_points as a
    GPU_OBS_DTYPE array (GpuObsPoint layout)."""
    o = np.ascontiguousarray(obs_points)
    if o.dtype != GPU_OBS_DTYPE:
        raise TypeError("obs_points must be a GPU_OBS_DTYPE array (GpuObsPoint layout)")
    n = o.shape[0]
    out = np.zeros(5 * n)
    check(lib.xrtk_undulator_gpu_f32(ctx.handle, o.ctypes.data, n, float(kx), float(ky),
                                     float(period_mm), int(n_periods), float(gamma),
                                     float(beam_current), int(n_integration_steps),
                                     float(phase_deg), dptr(out)), "xrtk_undulator_gpu_f32")
    r = out.reshape(-1, 5)
    return UndulatorGpuResult(r[:, 0] + 1j * r[:, 1], r[:, 2] + 1j * r[:, 3], r[:, 4].copy())


def now() -> float:
    return lib.xrtk_now()
