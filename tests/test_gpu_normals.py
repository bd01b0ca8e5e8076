"""Per-ray normals and bundles that are not thin along the line of sight, on
the device.

Rays taken off a curved optic carry a normal each (src/oes/reflect.rs:53-56,
src/waves/diffraction.rs:22-25,86).  The FP32-reduced kernel then evaluates
the obliquity n.d per pair from the ray's own normal (the planner leaves
UNIFORM_N off, so neither the separable amplitude nor the slot line runs).
These tests run that form on curved-optic ray sets against the long-double
oracle, force it onto uniform-normal inputs (where it must agree with the
specialised forms to FP32 rounding), and check that a grazing-mirror
footprint, spread 100 mm along the beam, is routed to the FP64 kernel and
matches the oracle there."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu

VAR_CHECK, VAR_TILE, VAR_UNIFORM_N, VAR_NOFRACT, VAR_LIN2 = 1, 2, 8, 16, 32
VAR_PAX, VAR_SEPAMP, VAR_LINE = 64, 128, 256
NO_UNIFORM_N = 0x8000
REL_F32R = 1e-4
STRESS_BAR = 3e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


def solve(ctx, monkeypatch, rays, g, force=0, **kw):
    if force:
        monkeypatch.setenv("XRTK_FORCE_FLAGS", hex(force))
    else:
        monkeypatch.delenv("XRTK_FORCE_FLAGS", raising=False)
    out = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", **kw)
    monkeypatch.delenv("XRTK_FORCE_FLAGS", raising=False)
    return out, ctx.last_plan()["inst_flags"]


# force flags of k_plan: 0x100 (with TILE clear) the centroid reference, 0x800
# no paraxial expansion (rsq/rcp per pair), 0x400 no zero-axis dot product
FORCED = {
    "tile_pax": NO_UNIFORM_N,
    "tile_exact": NO_UNIFORM_N | 0x800,
    "tile_pax_lin3": NO_UNIFORM_N | 0x400,
    "centroid": NO_UNIFORM_N | 0x100,
}


@pytest.mark.parametrize("variant", sorted(FORCED))
@pytest.mark.parametrize("case", ["point", "gaussian"])
def test_forced_per_ray_normal_matches_uniform_forms(ctx, monkeypatch, case, variant):
    """Uniform-normal inputs through the per-ray-normal kernels (UNIFORM_N
    forbidden): the same field as the specialised kernels (separable
    amplitude, slot line) to FP32 rounding, and as the FP64 kernel."""
    if case == "point":
        rays, g = W.point_source_aperture(65536, seed=1), grid(W.screen(1000.0, 0.02, 128))
    else:
        rays, g = W.gaussian_source(65536, seed=2), grid(W.screen(1000.0, 0.02, 96))
    if variant == "centroid":
        # the centroid reference needs a small whole phase: a 20 um aperture
        rays = W.point_source_aperture(16384, half=0.01, seed=1)
    uni, fu = solve(ctx, monkeypatch, rays, g)
    per, fp = solve(ctx, monkeypatch, rays, g, FORCED[variant])
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    print(f"{case}/{variant}: flags {fu:#x} -> {fp:#x}, per-ray vs specialised "
          f"{rel_l2(per, uni):.2e}, vs f64 {rel_l2(per, f64):.2e}")
    assert fu & VAR_UNIFORM_N and not fp & VAR_UNIFORM_N
    assert not fp & (VAR_SEPAMP | VAR_LINE)
    if variant == "centroid":
        assert not fp & VAR_TILE
    else:
        assert fp & VAR_TILE
        assert bool(fp & VAR_PAX) == (variant != "tile_exact")
        assert bool(fp & VAR_LIN2) == (variant != "tile_pax_lin3")
    assert rel_l2(per, uni) < 5e-6
    assert rel_l2(per, f64) < 1e-5


@pytest.mark.parametrize("variant", ["default", "tile_exact", "centroid"])
def test_curved_optic_point_source_shape(ctx, monkeypatch, variant):
    """65,536 rays on a toroidal cap (Rx = 2 m, Rz = 5 m over +-0.1 mm, each
    normal tilted at random by up to 0.2 rad) x 256^2 pixels at 1000 mm,
    10 keV: FP32-reduced with per-ray normals against the long-double oracle
    on a pixel sample and the FP64 kernel on every pixel; the FP64 kernel
    against the FP64 oracle."""
    rays = W.curved_aperture(65536, half=0.1, tilt=0.2, seed=11)
    g = grid(W.screen(1000.0, 0.02, 256))
    force = {"default": 0, "tile_exact": 0x800, "centroid": 0x100}[variant]
    if variant == "centroid":
        rays = W.curved_aperture(16384, half=0.01, tilt=0.2, seed=11)
    f32, flags = solve(ctx, monkeypatch, rays, g, force)
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    px, py, pz = xrtk.prepare_wave(g)
    stride = 1031
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e32, e64 = rel_l2(f32[::stride], ld), rel_l2(f64[::stride], ld)
    full = rel_l2(f32, f64)
    print(f"curved {variant}: flags {flags:#x}, f32r vs ld {e32:.2e}, f64 vs ld {e64:.2e}, "
          f"f32r vs f64 {full:.2e}")
    assert not flags & VAR_UNIFORM_N and not flags & VAR_LIN2  # a sagged, tilted cap
    assert not flags & (VAR_SEPAMP | VAR_LINE)
    assert bool(flags & VAR_TILE) == (variant != "centroid")
    assert e32 < STRESS_BAR and full < REL_F32R
    assert rel_l2(f64[::stride], O.diffraction_f64(rays, px, py, pz, stride=stride)) < 1e-10


def test_curved_optic_gaussian_shape(ctx, monkeypatch):
    """1,048,576 Gaussian-source rays moved onto a toroidal cap with per-ray
    normals (tilted by up to 0.05 rad) x 512^2 pixels at 1000 mm: FP32-reduced
    vs FP64 on every pixel, both vs the long-double oracle on a sample."""
    rays = W.gaussian_source(1048576, seed=2)
    c = W.curved_aperture(1048576, half=0.2, tilt=0.05, seed=12)
    x, z = rays["x"], rays["z"]
    y = x * x / (2 * 2000.0) + z * z / (2 * 5000.0)
    rays = dict(rays, y=y, nx=c["nx"], ny=c["ny"], nz=c["nz"])
    d = np.sqrt(x * x + (y + W.SRC_DIST) ** 2 + z * z)
    rays["nl"] = (rays["nx"] * x + rays["ny"] * (y + W.SRC_DIST) + rays["nz"] * z) / d
    g = grid(W.screen(1000.0, 0.02, 512))
    f32, flags = solve(ctx, monkeypatch, rays, g)
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    full = rel_l2(f32, f64)
    px, py, pz = xrtk.prepare_wave(g)
    stride = 4099
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e32, e64 = rel_l2(f32[::stride], ld), rel_l2(f64[::stride], ld)
    print(f"curved gaussian 1M x 512^2: flags {flags:#x}, f32r vs f64 {full:.2e}, "
          f"f32r vs ld {e32:.2e}, f64 vs ld {e64:.2e}, intensity err "
          f"{W.intensity_error(f32, f64)}")
    assert not flags & VAR_UNIFORM_N and flags & VAR_TILE
    assert full < REL_F32R and e32 < STRESS_BAR


@pytest.mark.parametrize("bend", [None, 50000.0])
def test_thick_bundle_routes_to_fp64(ctx, bend):
    """A grazing mirror's footprint (3 mrad, 100 mm long, flat or bent to
    50 m so the normals differ per ray) is not thin along the line of sight:
    the reduced phase would exceed FP32's range, so the solve runs on the
    plain FP64 kernel (n_fp64 = 1), which matches the oracle there.  The
    screen sits on the reflected beam 1000 mm downstream."""
    rays = W.grazing_mirror(32768, length=100.0, width=0.5, theta=3e-3, bend_radius=bend, seed=13)
    zc = 1000.0 * np.tan(6e-3)
    g = xrtk.WaveParams((0.0, 1000.0, zc), 0.05, 0.05, 64, 64)
    got = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    plan = ctx.last_plan()
    assert plan["n_fp64"] == 1 and plan["mode"] == xrtk.MODE_F64, plan
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    assert np.array_equal(got, f64)
    px, py, pz = xrtk.prepare_wave(g)
    stride = 67
    ref = O.diffraction_f64(rays, px, py, pz, stride=stride)
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    print(f"grazing bend={bend}: f64 vs oracle {rel_l2(got[::stride], ref):.2e}, "
          f"vs ld {rel_l2(got[::stride], ld):.2e}")
    assert rel_l2(got[::stride], ref) < 1e-10
    assert rel_l2(got[::stride], ld) < 1e-5
    # the batched and served paths take the same route
    rs = xrtk.RaySet(ctx, rays)
    try:
        (b,) = xrtk.solve_jobs(ctx, [xrtk.Job(rs, 10000.0, g)])
        assert ctx.last_plan()["n_fp64"] == 1
        assert np.array_equal(b, got)
    finally:
        rs.release()
