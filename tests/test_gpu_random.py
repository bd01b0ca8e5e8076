"""Seeded random geometries: every launch shape of the FP32-reduced kernel
against the long-double oracle on a pixel sample, and the plain FP64 kernel
against the FP64 oracle (the same formula).  (The FP64 kernel is not the
reference for the reduced one here: at tens of metres its own rounding of
k*path, ~1e12 rad, reaches 1e-4.)

Energies 5-30 keV, screen distances 0.1-30 m, apertures 0.02-0.5 mm,
screens 1-150 pixels a side (square or not) spanning about one Fresnel
zone per few pixels, pixels per lane 2, 4, 8 and 16: the planner's choices
(tile shape, padding, routing, the phase-reduction variants and the slot
line) vary from case to case, and each must stay within the method's bar."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu

REL_F32R = 1e-4
N_CASES = 30


def case(k):
    rng = np.random.default_rng(1000 + k)
    energy = float(rng.uniform(5000.0, 30000.0))
    dist = float(np.exp(rng.uniform(np.log(100.0), np.log(30000.0))))
    half = float(np.exp(rng.uniform(np.log(0.02), np.log(0.5))))
    nx, nz = (int(v) for v in rng.integers(1, 151, size=2))
    if rng.uniform() < 0.5:
        nz = nx
    lam = 2 * np.pi / W.k_of(energy)
    leff = 1.0 / (1.0 / dist + 1.0 / W.SRC_DIST)
    # a screen of a few Fresnel zones of the aperture, as seen from it
    span = float(rng.uniform(0.5, 4.0)) * lam * leff / (2 * half) * max(nx, nz) / 8
    rays = W.point_source_aperture(2048, half=half, energy=energy, seed=50 + k)
    g = xrtk.WaveParams((float(rng.uniform(-0.01, 0.01)), dist, float(rng.uniform(-0.01, 0.01))),
                        span, span * float(rng.uniform(0.5, 2.0)), nx, nz)
    return rays, g, dict(E=energy, L=dist, half=half, nx=nx, nz=nz)


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_geometry(ctx, k):
    rays, g, info = case(k)
    print(f"case {k}: {info}", flush=True)
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    px, py, pz = xrtk.prepare_wave(g)
    stride = max(1, px.shape[0] // 97)
    assert W.rel_l2(f64[::stride], O.diffraction_f64(rays, px, py, pz, stride=stride)) < 1e-9
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e64 = W.rel_l2(f64[::stride], ld)
    print(f"  FP64 kernel vs long double {e64:.2e}", flush=True)
    for P in (2, 4, 8, 16):
        got = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=P)
        plan = ctx.last_plan()
        err = W.rel_l2(got[::stride], ld)
        print(f"  P={P}: flags {plan['inst_flags']:#x} fp64-routed {plan['n_fp64']} "
              f"rel-L2 vs long double {err:.2e}, vs FP64 {W.rel_l2(got, f64):.2e}", flush=True)
        assert err < REL_F32R, (P, info)


@pytest.mark.parametrize("k", range(0, N_CASES, 3))
def test_random_geometry_gpu_types(ctx, k):
    """The reference's own call shape on the same cases: kirchhoff_gpu on
    float32 GpuRay / GpuPixel records (a pixel list, off the slot line by
    float32 rounding), against the long-double oracle evaluated at the
    float32 values."""
    rays, g, info = case(k)
    px, py, pz = xrtk.prepare_wave(g)
    r32 = xrtk.gpu_rays(rays["x"].shape[0])
    for f in xrtk.RAY_FIELDS:
        r32[f] = rays[f]
    p32 = xrtk.gpu_pixels(px, py, pz)
    rays32 = {f: r32[f].astype(np.float64) for f in xrtk.RAY_FIELDS}
    stride = max(1, px.shape[0] // 97)
    ld = O.diffraction_ld(rays32, p32["x"].astype(np.float64), p32["y"].astype(np.float64),
                          p32["z"].astype(np.float64), stride=stride)
    print(f"case {k}: {info}", flush=True)
    # the default plan, and the shapes with the slot line (with per-slot
    # offsets where the list allows it)
    for shape in ({}, {"pixels_per_lane": 8}, {"pixels_per_lane": 16}):
        got = xrtk.kirchhoff_gpu(ctx, r32, p32, **shape)
        plan = ctx.last_plan()
        err = W.rel_l2(got[::stride], ld)
        print(f"  P={plan['pixels_per_lane']} flags {plan['inst_flags']:#x} "
              f"fp64-routed {plan['n_fp64']}: rel-L2 vs long double {err:.2e}", flush=True)
        assert err < REL_F32R, (shape, info)
