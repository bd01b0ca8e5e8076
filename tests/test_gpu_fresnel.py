"""Fresnel propagation on the device (src/waves/fresnel.rs:53-106): the FP64
kernel and both FP32-sum kernels (products on FP16 hi + lo parts, the
default, and on FP32 operands, engine 1) against the oracle, the
reference's unit tests, closed forms, screen axes, ray chunking and the
amplitude scale."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu

REL_F32 = 1e-5   # FP32 factors and products, FP64 phases and accumulation
REL_F64 = 1e-11  # FP64 factors and products
# One term's error, FP32 operands: the factors' FP32 rounding (~2^-23); FP16
# parts: x - hi - lo <= 2^-22 |x| per factor and the dropped lo.lo, up to
# ~3 x 2^-22 = 7e-7 per product.
TERM_F32 = {"f32r": 1e-6, "f32ops": 2e-7}
F32_MODES = ("f32r", "f32ops")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def k_of(e):
    return e / W.CHBAR * 1e7


def run_both(ctx, rx, rz, es, e, d, scr):
    """f64, f32r (FP16 parts) and f32ops (FP32 operands, engine 1)."""
    out = {m: xrtk.fresnel_propagate(rx, rz, es, e, d, scr, ctx=ctx, mode=m)
           for m in ("f32r", "f64")}
    out["f32ops"] = xrtk.fresnel_propagate(rx, rz, es, e, d, scr, ctx=ctx, mode="f32r",
                                           engine=1)
    return out


def test_reference_unit_tests(ctx):
    """fresnel_single_point_source and fresnel_two_slit (fresnel.rs:130-165)."""
    r = xrtk.fresnel_propagate([0.0], [0.0], [1 + 0j], 10000.0, 1000.0,
                               xrtk.FresnelScreen(-0.1, 0.1, 21, -0.1, 0.1, 21), ctx=ctx)
    assert r.amplitude.shape == (21 * 21,)
    # |U| = 1 at every pixel: centre >= corner up to two terms' rounding
    assert abs(r.amplitude[10 * 21 + 10]) >= abs(r.amplitude[0]) * (1 - 2 * TERM_F32["f32r"])
    r = xrtk.fresnel_propagate([-0.01, 0.01], [0.0, 0.0], [1 + 0j, 1 + 0j], 10000.0, 1000.0,
                               xrtk.FresnelScreen(-0.1, 0.1, 101, 0.0, 0.0, 1), ctx=ctx)
    inten = r.intensity()
    assert inten[50] > 0.0 and inten[60] > 0.0


def test_closed_forms(ctx):
    e, d, a = 10000.0, 1000.0, 0.01
    scr = xrtk.FresnelScreen(-0.1, 0.1, 101, 0.0, 0.0, 1)
    for m, r in run_both(ctx, [-a, a], [0.0, 0.0], [1 + 0j, 1 + 0j], e, d, scr).items():
        exp = 2 + 2 * np.cos(2 * k_of(e) * a * r.x_pixels / d)
        tol = 1e-9 if m == "f64" else 2e-6
        assert np.max(np.abs(r.intensity() - exp)) < tol, m
    x0, z0, es = 0.013, -0.007, 0.6 - 0.8j
    scr = xrtk.FresnelScreen(-0.05, 0.05, 9, -0.04, 0.06, 7)
    for m, r in run_both(ctx, [x0], [z0], [es], 12000.0, 750.0, scr).items():
        X, Z = np.meshgrid(r.x_pixels, r.z_pixels)
        exp = es * np.exp(1j * (k_of(12000.0) / (2 * 750.0)) * ((X - x0) ** 2 + (Z - z0) ** 2))
        tol = 1e-12 if m == "f64" else TERM_F32[m]
        assert np.max(np.abs(r.amplitude - exp.ravel())) < tol, m


def test_axes_match_reference_formula(ctx):
    scr = xrtk.FresnelScreen(-0.1, 0.3, 5, 0.25, 9.0, 1)
    r = xrtk.fresnel_propagate([0.0], [0.0], [1 + 0j], 1e4, 1e3, scr, ctx=ctx)
    assert np.array_equal(r.x_pixels, O.fresnel_grid(-0.1, 0.3, 5))
    assert np.array_equal(r.z_pixels, O.fresnel_grid(0.25, 9.0, 1))


def test_zero_rays(ctx):
    r = xrtk.fresnel_propagate([], [], [], 1e4, 1e3, xrtk.FresnelScreen(-1, 1, 3, -1, 1, 2),
                               ctx=ctx)
    assert r.amplitude.shape == (6,) and not np.any(r.amplitude)


def _aperture(n, half, e, seed):
    rays = W.point_source_aperture(n, half, e, seed)
    return rays["x"], rays["z"], rays["es_re"] + 1j * rays["es_im"]


@pytest.mark.parametrize("shape", [(64, 48), (1, 37), (29, 1)])
def test_vs_oracle_small(ctx, shape):
    nx, nz = shape
    rx, rz, es = _aperture(4096, 0.1, 10000.0, 5)
    scr = xrtk.FresnelScreen(-0.06, 0.05, nx, -0.04, 0.07, nz)
    ref, _, _ = O.fresnel_f64(rx, rz, es, 10000.0, 1000.0, scr.x_min, scr.x_max, nx,
                              scr.z_min, scr.z_max, nz)
    got = run_both(ctx, rx, rz, es, 10000.0, 1000.0, scr)
    assert rel_l2(got["f64"].amplitude, ref) < REL_F64
    for m in F32_MODES:
        assert rel_l2(got[m].amplitude, ref) < REL_F32, m


def test_config0_aperture_to_256sq(ctx):
    """65,536 point-source rays on +-0.1 mm at 10 keV, 1 m to a 256^2 screen of
    +-0.05 mm; oracle on a strided pixel sample."""
    rx, rz, es = _aperture(65536, 0.1, 10000.0, 1)
    scr = xrtk.FresnelScreen(-0.05, 0.05, 256, -0.05, 0.05, 256)
    stride = 97
    ref, _, _ = O.fresnel_f64(rx, rz, es, 10000.0, 1000.0, -0.05, 0.05, 256, -0.05, 0.05, 256,
                              stride=stride)
    got = run_both(ctx, rx, rz, es, 10000.0, 1000.0, scr)
    e64 = rel_l2(got["f64"].amplitude[::stride], ref)
    assert e64 < REL_F64
    for m in F32_MODES:
        e32 = rel_l2(got[m].amplitude[::stride], ref)
        print(f"fresnel 65536 x 256^2: f64 {e64:.2e}, {m} {e32:.2e}, "
              f"vs f64 over all pixels {rel_l2(got[m].amplitude, got['f64'].amplitude):.2e}")
        assert e32 < REL_F32, m


def test_ray_chunks(ctx):
    """A wide, short screen has few tiles, so the rays split into many slices
    whose partial sums are merged in FP64."""
    rx, rz, es = _aperture(20000, 0.05, 8000.0, 11)
    scr = xrtk.FresnelScreen(-0.2, 0.2, 8192, -0.01, 0.01, 3)
    stride = 211
    ref, _, _ = O.fresnel_f64(rx, rz, es, 8000.0, 500.0, -0.2, 0.2, 8192, -0.01, 0.01, 3,
                              stride=stride)
    got = run_both(ctx, rx, rz, es, 8000.0, 500.0, scr)
    assert rel_l2(got["f64"].amplitude[::stride], ref) < REL_F64
    for m in F32_MODES:
        assert rel_l2(got[m].amplitude[::stride], ref) < REL_F32, m


def test_ragged_rays_and_tiles(ctx):
    """A ray count that is not a whole number of LDS blocks (zero-padded) on
    a screen one to three pixels past whole 128- and 64-pixel tiles."""
    rx, rz, es = _aperture(4099, 0.08, 12000.0, 17)
    scr = xrtk.FresnelScreen(-0.05, 0.05, 130, -0.05, 0.04, 131)
    ref, _, _ = O.fresnel_f64(rx, rz, es, 12000.0, 800.0, scr.x_min, scr.x_max, 130,
                              scr.z_min, scr.z_max, 131)
    got = run_both(ctx, rx, rz, es, 12000.0, 800.0, scr)
    assert rel_l2(got["f64"].amplitude, ref) < REL_F64
    for m in F32_MODES:
        assert rel_l2(got[m].amplitude, ref) < REL_F32, m


def test_slices_set_by_the_fp32_cap(ctx):
    """A screen of 4096^2 pixels fills the device with tiles alone, so only
    the cap on an FP32 sum splits the rays: 40,000 rays take five slices of
    8,000 with FP16 parts (cap 8,192) and three of 13,344 with FP32 operands
    (cap 16,384), the last one partly zero-padded."""
    rx, rz, es = _aperture(40000, 0.1, 10000.0, 19)
    scr = xrtk.FresnelScreen(-0.1, 0.1, 4096, -0.1, 0.1, 4096)
    stride = 40009
    ref, _, _ = O.fresnel_f64(rx, rz, es, 10000.0, 1000.0, -0.1, 0.1, 4096, -0.1, 0.1, 4096,
                              stride=stride)
    for engine in (0, 1):
        got = xrtk.fresnel_propagate(rx, rz, es, 10000.0, 1000.0, scr, ctx=ctx, engine=engine)
        e32 = rel_l2(got.amplitude[::stride], ref)
        print(f"fresnel 40000 x 4096^2, engine {engine}: {e32:.2e}")
        assert e32 < REL_F32, engine


def test_amplitude_scale_is_exact(ctx):
    """The FP16-part kernel scales the amplitudes by a power of two and back:
    amplitudes times 2^-70 or 2^50 give the same field times that power,
    bit for bit."""
    rx, rz, es = _aperture(3000, 0.05, 10000.0, 23)
    scr = xrtk.FresnelScreen(-0.03, 0.03, 70, -0.03, 0.03, 50)
    base = xrtk.fresnel_propagate(rx, rz, es, 10000.0, 1000.0, scr, ctx=ctx).amplitude
    for p in (-70, 50):
        got = xrtk.fresnel_propagate(rx, rz, es * 2.0 ** p, 10000.0, 1000.0, scr,
                                     ctx=ctx).amplitude
        assert np.array_equal(got, base * 2.0 ** p), p
