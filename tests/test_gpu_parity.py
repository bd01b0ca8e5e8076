"""Device parity: the FP64 pair kernel against the oracle, the FP32
phase-reduced kernel against both, the reference's GPU tests restated, edge
cases and the workload shapes (gfx950 only)."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W
from helpers import f32_exact, fixture_case, mp_cases, ray, rays_of, to_gpu_pixels, to_gpu_rays

pytestmark = pytest.mark.gpu

REL_F32R = 1e-4   # complex rel-L2 bar of the FP32-reduced kernel vs FP64
REL_F64 = 1e-11   # FP64 kernel vs the oracle (same formula, same rounding but sin/cos ulps)
# A handful of rays on a few scattered pixels (the reference's own GPU tests
# use 1e-3 there): the in-tile residual phase reaches ~50 rev and nothing
# averages the FP32 rounding out.
REL_TINY = 5e-5
STRESS_BAR = 3e-5  # FP32-reduced vs the long-double oracle in the precision sweep


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


# ---- golden vector ------------------------------------------------------------
def test_f64_kernel_fixture(ctx):
    """golden_diffraction_numerical bar (1e-10 abs) on the FP64 device kernel."""
    rays, pix, exp = fixture_case()
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f64")
    assert np.max(np.abs(got - exp)) < 1e-10
    assert xrtk.GpuContext.last_plan(ctx)["backend"] == 1


def test_f32r_kernel_fixture(ctx):
    """golden_gpu_long_distance_vs_fixture bar (rel < 1e-3 per pixel); the
    reduced kernel with FP64 centroid/base phase is far inside it."""
    rays, pix, exp = fixture_case()
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f32r")
    rel = np.abs(got - exp).max(axis=1) / np.abs(exp).max(axis=1)
    assert np.all(rel < 1e-3)
    assert rel_l2(got, exp) < REL_TINY


@pytest.mark.parametrize("case", mp_cases(), ids=lambda c: c[0])
def test_f64_kernel_vs_oracle_mp_cases(ctx, case):
    name, rays, pix, exp = case
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f64")
    ref = O.diffraction_f64(rays, *pix)
    assert rel_l2(got, ref) < REL_F64


# The reduced phase k (path - |p - c|) reaches ~1e3 revolutions in the 2 um
# near-field case (offsets comparable to the distance), where FP32 keeps
# ~1e-4 rad per pair; elsewhere it stays below ~100 revolutions.
F32R_MP_TOL = {"micro_1kev_2um": 3e-4}


@pytest.mark.parametrize("case", mp_cases(), ids=lambda c: c[0])
def test_f32r_kernel_vs_mpmath(ctx, case):
    name, rays, pix, exp = case
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f32r")
    assert rel_l2(got, exp) < F32R_MP_TOL.get(name, REL_TINY)


# ---- the reference's GPU tests (tests/golden_gpu_kirchhoff.rs) restated -------------
def test_long_distance_vs_cpu(ctx):
    """golden_gpu_long_distance_vs_cpu (:103-145): f32 inputs, rel < 1e-3."""
    rays, pix, _ = fixture_case()
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(pix)
    got = xrtk.kirchhoff_gpu(ctx, r32, p32)
    cpu = O.fallback_f32(r32, p32)
    assert np.all(np.abs(got[:, 0]) > 1e-10)
    assert np.all(np.abs(got[:, 0] - cpu[:, 0]) / np.abs(cpu[:, 0]) < 1e-3)
    assert rel_l2(got, cpu) < REL_TINY


def test_short_distance_vs_cpu(ctx):
    """golden_gpu_short_distance_vs_cpu (:184-247): 1 keV, 0.1 mm, rel < 5e-2."""
    rays = rays_of([ray(energy=1000.0, es=1 + 0j, ep=0.5 + 0.1j),
                    ray(x=0.01, energy=1000.0, es=0.8 - 0.1j, ep=0.6 + 0j)])
    pix = ([0.0, 0.005], [0.0, 0.0], [0.1, 0.1])
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(pix)
    got = xrtk.kirchhoff_gpu(ctx, r32, p32)
    cpu = O.fallback_f32(r32, p32)
    assert np.all(np.abs(got[:, 0] - cpu[:, 0]) / np.abs(cpu[:, 0]) < 5e-2)
    assert rel_l2(got, cpu) < 1e-4


def test_symmetry(ctx):
    """golden_gpu_symmetry (:249-311): equal |Es| within 1e-4 relative."""
    rays = rays_of([ray(es=1 + 0j, ep=1 + 0j)])
    got = xrtk.kirchhoff_gpu(ctx, to_gpu_rays(rays),
                             to_gpu_pixels(([-0.01, 0.01], [0.0, 0.0], [1000.0, 1000.0])))
    a, b = abs(got[0, 0]), abs(got[1, 0])
    assert a > 1e-10 and b > 1e-10
    assert abs(a - b) < 1e-4 * a


def test_auto_dispatch_uses_device(ctx):
    """golden_gpu_auto_dispatch (:313-327) — and that the device ran it."""
    rays, pix, _ = fixture_case()
    got = xrtk.kirchhoff_auto(to_gpu_rays(rays), to_gpu_pixels(pix))
    assert xrtk.auto_last_backend() == 1
    assert np.all(np.isfinite(got)) and np.all(np.abs(got[:, 0]) > 1e-10)


def test_extreme_distance(ctx):
    """golden_gpu_extreme_distance (:329-391): 10 m, rel < 0.05."""
    rays = rays_of([ray(es=1 + 0j, ep=0.5 + 0j), ray(x=0.01, es=0.8 + 0.1j, ep=0.6 - 0.1j)])
    pix = ([0.0], [0.0], [10000.0])
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(pix)
    got = xrtk.kirchhoff_gpu(ctx, r32, p32)
    cpu = O.fallback_f32(r32, p32)
    assert abs(got[0, 0]) > 1e-10
    assert abs(got[0, 0] - cpu[0, 0]) / abs(cpu[0, 0]) < 0.05
    assert rel_l2(got, cpu) < REL_TINY


# ---- edge cases ------------------------------------------------------------------------
def test_zero_rays_gives_zeros(ctx):
    rays = rays_of([ray()])
    empty = {k: v[:0] for k, v in rays.items()}
    for mode in ("f32r", "f64"):
        got = xrtk.diffraction_integral(empty, ([0.0, 1.0], [0.0, 0.0], [1.0, 1.0]), ctx=ctx,
                                        mode=mode)
        assert got.shape == (2, 2) and np.all(got == 0)
    got = xrtk.kirchhoff_gpu(ctx, xrtk.gpu_rays(0), to_gpu_pixels(([0.0], [0.0], [1.0])))
    assert np.all(got == 0)


def test_zero_pixels(ctx):
    rays, _, _ = fixture_case()
    got = xrtk.diffraction_integral(rays, ([], [], []), ctx=ctx)
    assert got.shape == (0, 2)


def test_ray_at_pixel_skipped(ctx):
    """ray_at_pixel_location_skipped: the near-singular path skips path == 0."""
    rays = rays_of([ray(x=1.0, y=2.0, z=3.0, es=1 + 0j, ep=1 + 0j),
                    ray(x=1.5, y=2.0, z=3.0, es=0.5 + 0j, ep=0j, energy=10000.0)])
    pix = ([1.0, 1.2], [2.0, 2.0], [3.0, 3.0])
    ref = O.diffraction_f64(rays, *pix)
    for mode in ("f64", "f32r"):
        got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode=mode)
        assert np.all(np.isfinite(got))
        assert rel_l2(got, ref) < REL_F64
        assert ctx.last_plan()["n_fp64"] == 1  # a ray on a pixel: the FP64 kernel runs
    one = rays_of([ray(x=1.0, y=2.0, z=3.0, es=1 + 0j, ep=1 + 0j)])
    for mode in ("f64", "f32r"):
        got = xrtk.diffraction_integral(one, ([1.0], [2.0], [3.0]), ctx=ctx, mode=mode)
        assert np.all(got == 0)


def test_polychromatic_groups(ctx):
    """Two energies run as two monochromatic instances (each with its own
    base phase); the reference GPU path uses one mean energy here."""
    rng = np.random.default_rng(3)
    lst = [ray(x=rng.uniform(-.1, .1), z=rng.uniform(-.1, .1), nx=0, ny=1, nz=0,
               energy=8000.0 if i % 2 else 12000.0, es=complex(rng.normal(), rng.normal()),
               ep=complex(rng.normal(), rng.normal())) for i in range(600)]
    rays = rays_of(lst)
    pix = (np.linspace(-.02, .02, 37), np.full(37, 1000.0), np.linspace(.01, -.01, 37))
    ref = O.diffraction_f64(rays, *pix)
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f32r")
    assert rel_l2(got, ref) < REL_F32R
    got64 = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f64")
    assert rel_l2(got64, ref) < REL_F64


def test_broadband_routes_to_fp64(ctx):
    rng = np.random.default_rng(4)
    lst = [ray(x=rng.uniform(-.05, .05), y=rng.uniform(-.05, .05), energy=5000.0 + 10.0 * i)
           for i in range(100)]
    rays = rays_of(lst)
    pix = ([0.0, 0.001], [0.0, 0.0], [100.0, 100.0])
    got = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f32r")
    assert ctx.last_plan()["mode"] == xrtk.MODE_F64
    assert rel_l2(got, O.diffraction_f64(rays, *pix)) < REL_F64


# ---- launch-shape invariance (every instantiated kernel) ------------------------------------
SHAPES_F32 = [(P, W, 1, per) for P in (2, 4, 8, 16) for W in (1, 4) for per in (0, 1)]
SHAPES_F64 = [(P, W, 1, per) for P in (1, 2) for W in (1, 4) for per in (0, 1)]


def test_f32r_result_independent_of_launch_shape(ctx):
    """Waves per group and persistence only move work between waves: for one
    pixels-per-lane P the results are bit-identical.  P sets the pixel tile
    (64 P pixels) and with it the tile reference point, so results across P
    agree to FP32 rounding."""
    rays = W.point_source_aperture(20000, seed=9)
    scr = W.screen(1000.0, 0.02, 40)
    by_p = {}
    for P, Wv, s, per in SHAPES_F32:
        got = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f32r", pixels_per_lane=P,
                                        waves_per_group=Wv, ray_src=s, persistent=per,
                                        tile_rays=1024, split=3)
        if P not in by_p:
            by_p[P] = got
        else:
            assert np.array_equal(got, by_p[P]), (P, Wv, s, per)
    ps = sorted(by_p)
    for P in ps[1:]:
        assert rel_l2(by_p[P], by_p[ps[0]]) < 1e-5


@pytest.mark.parametrize("nx,nz", [(300, 70), (33, 100), (1, 257), (257, 1)])
def test_grid_shapes_match_pixel_list(ctx, nx, nz):
    """Non-square and one-pixel-wide prepare_wave grids (edge tiles padded
    along one or both axes; 32 x 32 or 16 x 64 tiles at P = 16, chosen per
    grid): the grid generated on the device gives the FP64 kernel the same
    bits as the same points passed as a pixel list, the oracle agrees on a
    sample, and the FP32-reduced solve stays within its bar."""
    rays = W.point_source_aperture(4096, seed=23)
    g = xrtk.WaveParams((0.001, 1000.0, -0.002), 0.02, 0.015, nx, nz)
    px, py, pz = xrtk.prepare_wave(g)
    f64_grid = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    f64_list = xrtk.diffraction_integral(rays, (px, py, pz), ctx=ctx, mode="f64")
    assert np.array_equal(f64_grid, f64_list)
    stride = 37
    ref = O.diffraction_f64(rays, px, py, pz, stride=stride)
    assert rel_l2(f64_grid[::stride], ref) < REL_F64
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    assert ctx.last_plan()["n_fp64"] == 0
    assert rel_l2(f32, f64_grid) < 1e-5


def test_f64_result_independent_of_launch_shape(ctx):
    rays = W.point_source_aperture(3000, seed=10)
    scr = W.screen(1000.0, 0.02, 24)
    base = None
    for P, Wv, s, per in SHAPES_F64:
        got = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f64", pixels_per_lane=P,
                                        waves_per_group=Wv, persistent=per, split=2)
        if base is None:
            base = got
        else:
            assert np.array_equal(got, base), (P, Wv, s, per)


def test_split_moves_only_chunk_rounding(ctx):
    """The ray split sets where FP32 chunk sums end (then merged in FP64)."""
    rays = W.point_source_aperture(30000, seed=11)
    scr = W.screen(1000.0, 0.02, 32)
    a = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f32r", split=1)
    b = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f32r", split=7)
    assert rel_l2(b, a) < 1e-5
    c = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f64", split=1)
    d = xrtk.diffraction_integral(rays, grid(scr), ctx=ctx, mode="f64", split=7)
    assert rel_l2(d, c) < 1e-13


# ---- workload shapes -------------------------------------------------------------------------------
def test_point_source_full_size(ctx):
    """65,536 rays x 256^2 pixels at 1000 mm, 10 keV: FP32-reduced vs the FP64
    kernel over every pixel, and the FP64 kernel vs the oracle on a strided
    pixel subsample."""
    rays = W.point_source_aperture(65536, seed=1)
    scr = W.screen(1000.0, 0.02, 256)
    g = grid(scr)
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    err = rel_l2(f32, f64)
    print(f"point source 65536 x 256^2: rel-L2(f32r, f64) = {err:.3e}, "
          f"intensity err (max, p99) = {W.intensity_error(f32, f64)}")
    assert err < REL_F32R
    px, py, pz = xrtk.prepare_wave(g)
    stride = 1031
    ref = O.diffraction_f64(rays, px, py, pz, stride=stride)
    assert rel_l2(f64[::stride], ref) < REL_F64


def test_gaussian_source_full_size(ctx):
    """1,048,576 rays x 512^2 at 1000 mm: FP32-reduced vs FP64 on every pixel."""
    rays = W.gaussian_source(1048576, seed=2)
    g = grid(W.screen(1000.0, 0.02, 512))
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    err = rel_l2(f32, f64)
    print(f"gaussian source 1M x 512^2: rel-L2 = {err:.3e}, "
          f"intensity err = {W.intensity_error(f32, f64)}")
    assert err < REL_F32R
    px, py, pz = xrtk.prepare_wave(g)
    stride = 4099
    ref = O.diffraction_f64(rays, px, py, pz, stride=stride)
    assert rel_l2(f64[::stride], ref) < 1e-9


@pytest.mark.parametrize("energy", W.STRESS_ENERGIES)
@pytest.mark.parametrize("distance", W.STRESS_DISTANCES)
def test_stress_cells(ctx, energy, distance):
    """262,144 rays x 256^2 for every (E, L) cell.  The FP32-reduced kernel is
    checked against the extended-precision oracle on a pixel subsample (the
    plain FP64 sum itself rounds each phase k*path, up to 4.6e12 rad, by
    ~phase * 2^-53); FP32-reduced vs the FP64 kernel over every pixel is
    reported and held to the FP64 kernel's own error scale."""
    rays, scr, info = W.stress_cell(energy, distance)
    g = grid(scr)
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    plan = ctx.last_plan()
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    px, py, pz = xrtk.prepare_wave(g)
    stride = 1031
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e32 = rel_l2(f32[::stride], ld)
    e64 = rel_l2(f64[::stride], ld)
    full = rel_l2(f32, f64)
    phase = W.k_of(energy) * (distance + 1.0)
    print(f"E={energy:.0f} L={distance:.0f} a={info['aperture_mm']:.3f} "
          f"NF={info['fresnel_number']:.1f} P={plan['pixels_per_lane']} "
          f"fp64-routed={plan['n_fp64']}: "
          f"f32r-vs-ld {e32:.2e}  f64-vs-ld {e64:.2e}  f32r-vs-f64 {full:.2e}")
    # the untuned plan takes smaller tiles where larger ones would route the
    # instance to the FP64 kernel: every cell runs reduced
    assert plan["n_fp64"] == 0
    assert e32 < REL_F32R
    # regression bar (the L = 100 mm cells, where ray-to-pixel directions fan
    # out by ~1e-3 rad, are the worst)
    assert e32 < STRESS_BAR
    assert full < max(REL_F32R, 2 * e64 + 8 * phase * 2.0 ** -53)


@pytest.mark.parametrize("variant", ["centroid", "tile"])
def test_phase_reference_variants(ctx, monkeypatch, variant):
    """Both reduced-phase forms (centroid reference; per-tile reference with
    the tile phase folded into the amplitudes) against the FP64 kernel."""
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x100" if variant == "centroid" else "0x102")
    rays = W.point_source_aperture(65536, seed=1)
    g = grid(W.screen(1000.0, 0.02, 128))
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    monkeypatch.delenv("XRTK_FORCE_FLAGS")
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    err = rel_l2(f32, f64)
    print(f"{variant}: rel-L2 = {err:.3e}")
    assert err < REL_F32R


def test_zero_axis_variant_is_exact(ctx, monkeypatch):
    """A flat aperture has o = r - c == 0 along y; the variant that drops that
    term of o.m gives the same bits as the full dot product."""
    rays = W.point_source_aperture(20000, seed=3)
    g = grid(W.screen(1000.0, 0.02, 64))
    lin2 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x400")
    full = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    assert np.array_equal(lin2, full)


VAR_TILE, VAR_NOFRACT, VAR_LIN2, VAR_PAX, VAR_SEPAMP, VAR_LINE = 2, 16, 32, 64, 128, 256
VAR_LINEDEV, VAR_EPZERO, VAR_UNIFORM_N = 1024, 2048, 8
VALU = 1  # XRTK_ENGINE_VALU: the VALU kernels only


@pytest.mark.parametrize("form", ["line", "linedev", "per_slot", "per_ray_normal", "exact"])
def test_ep_zero_variant_is_exact(ctx, monkeypatch, form):
    """Rays with Ep = 0 everywhere: the kernels that skip the Ep sums give the
    same bits as the general sums (EPZERO forbidden with 0x10000), in each
    form that has them: the slot line, the line with float32 offsets, the
    per-slot separable form, the per-ray-normal form; the form without the
    paraxial expansion keeps the general sums (the same bits trivially)."""
    rays = W.gaussian_source(65536, seed=2)
    g = grid(W.screen(1000.0, 0.02, 96))
    extra = {"line": 0, "linedev": 0, "per_slot": 0x2000, "per_ray_normal": 0x8000,
             "exact": 0x800}[form]

    def run(force):
        if force:
            monkeypatch.setenv("XRTK_FORCE_FLAGS", hex(force))
        else:
            monkeypatch.delenv("XRTK_FORCE_FLAGS", raising=False)
        if form == "linedev":
            px = xrtk.prepare_wave(g)
            out = xrtk.kirchhoff_gpu(ctx, to_gpu_rays(rays), to_gpu_pixels(px), pixels_per_lane=8)
        else:
            out = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=8)
        monkeypatch.delenv("XRTK_FORCE_FLAGS", raising=False)
        return out, ctx.last_plan()["inst_flags"]

    fast, f1 = run(extra)
    full, f2 = run(extra | 0x10000)
    print(f"{form}: {f1:#x} vs {f2:#x}")
    assert f1 & VAR_EPZERO and not f2 & VAR_EPZERO
    want = {"line": VAR_LINE, "linedev": VAR_LINEDEV, "per_slot": VAR_SEPAMP,
            "per_ray_normal": VAR_PAX, "exact": VAR_TILE}[form]
    assert f1 & want
    if form == "per_ray_normal":
        assert not f1 & VAR_UNIFORM_N
    if form in ("per_slot", "exact"):
        assert not f1 & VAR_LINE
    assert np.array_equal(fast, full)
    assert np.all(fast[:, 1] == 0)


@pytest.mark.parametrize("distance", [100.0, 1000.0, 30000.0])
def test_paraxial_expansion_matches_exact_form(ctx, monkeypatch, distance):
    """The tile variant's paraxial 1/path and 1/(path + R) (one FMA each)
    against rsq/rcp of the exact path; both against the long-double oracle."""
    rays, scr, _ = W.stress_cell(20000.0, distance, n_rays=32768, side=96)
    g = grid(scr)
    pax = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    flags = ctx.last_plan()["inst_flags"]
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x800")
    exact = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    assert not ctx.last_plan()["inst_flags"] & VAR_PAX
    monkeypatch.delenv("XRTK_FORCE_FLAGS")
    px, py, pz = xrtk.prepare_wave(g)
    stride = 37
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e_pax, e_ex = rel_l2(pax[::stride], ld), rel_l2(exact[::stride], ld)
    print(f"L={distance}: flags {flags:#x}, paraxial {e_pax:.2e}, exact form {e_ex:.2e}, "
          f"between {rel_l2(pax, exact):.2e}")
    assert flags & VAR_TILE and flags & VAR_PAX and flags & VAR_LIN2
    assert e_pax < STRESS_BAR and e_ex < STRESS_BAR
    assert e_pax < 1.5 * e_ex + 1e-6


def test_gaussian_case_variant(ctx):
    """The headline case runs the tile variant with every specialisation."""
    rays = W.gaussian_source(65536, seed=2)
    xrtk.diffraction_integral(rays, grid(W.screen(1000.0, 0.02, 128)), ctx=ctx, mode="f32r")
    f = ctx.last_plan()["inst_flags"]
    assert f & VAR_TILE and f & VAR_NOFRACT and f & VAR_LIN2 and f & VAR_PAX, hex(f)
    assert f & VAR_SEPAMP, hex(f)


@pytest.mark.parametrize("case", ["gaussian", "L100", "L1000", "L10000", "L30000"])
def test_separable_amplitude_matches_exact_form(ctx, monkeypatch, case):
    """The amplitude as a ray factor times a slot factor (no per-pair
    amplitude) against the per-pair form, both against the long-double
    oracle.  The planner allows it when its bound on the relative amplitude
    error is <= 1e-6: not at L = 100 mm, where the rays' paths to one pixel
    differ by ~1e-6 of their length."""
    if case == "gaussian":
        rays, g = W.gaussian_source(65536, seed=2), grid(W.screen(1000.0, 0.02, 96))
    else:
        rays, scr, _ = W.stress_cell(20000.0, float(case[1:]), n_rays=32768, side=96)
        g = grid(scr)
    # both on the VALU kernel, whose amplitude forms these are (grids within
    # the block bound would otherwise take the matrix cores)
    sep = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", engine=VALU)
    flags = ctx.last_plan()["inst_flags"]
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x1000")
    per_pair = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", engine=VALU)
    assert not ctx.last_plan()["inst_flags"] & VAR_SEPAMP
    monkeypatch.delenv("XRTK_FORCE_FLAGS")
    between = rel_l2(sep, per_pair)
    print(f"{case}: flags {flags:#x}, separable vs per-pair {between:.2e}")
    if case == "L100":
        assert not flags & VAR_SEPAMP and between == 0.0
        return
    assert flags & VAR_SEPAMP and flags & VAR_PAX
    px, py, pz = xrtk.prepare_wave(g)
    stride = 37
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e_sep, e_pp = rel_l2(sep[::stride], ld), rel_l2(per_pair[::stride], ld)
    print(f"   vs ld: separable {e_sep:.2e}, per-pair {e_pp:.2e}")
    assert between < 2e-6
    assert e_sep < STRESS_BAR and e_sep < e_pp + 2e-6


def _line_case(case):
    if case == "gaussian":
        return W.gaussian_source(65536, seed=2), grid(W.screen(1000.0, 0.02, 96))
    rays, scr, _ = W.stress_cell(20000.0, float(case[1:]), n_rays=32768, side=96)
    return rays, grid(scr)


@pytest.mark.parametrize("P", [4, 8, 16])
@pytest.mark.parametrize("case", ["gaussian", "L1000", "L10000", "L30000"])
def test_slot_line_rotation_matches_per_slot_phase(ctx, monkeypatch, case, P):
    """A lane's slots on a line of the grid: the phase at slot 0 and the step
    to the next slot, then one rotation and the three-term recurrence along
    the lane, against every slot's own phase (both with the separable
    amplitude) and the long-double oracle.  At L = 1 m and P = 16 (15 steps
    of 2 pixels, 0.47 um at 20 keV) the bound on the phase's departure from
    the line exceeds 1e-7 rev and the line form is refused: the same bits
    as the per-slot form."""
    rays, g = _line_case(case)
    line = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=P)
    flags = ctx.last_plan()["inst_flags"]
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x2000")
    per_slot = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=P)
    f2 = ctx.last_plan()["inst_flags"]
    monkeypatch.delenv("XRTK_FORCE_FLAGS")
    assert not f2 & VAR_LINE and f2 & VAR_SEPAMP, hex(f2)
    if case == "L1000" and P == 16:
        assert not flags & VAR_LINE, hex(flags)
        assert np.array_equal(line, per_slot)
        return
    assert flags & VAR_LINE and flags & VAR_SEPAMP, hex(flags)
    px, py, pz = xrtk.prepare_wave(g)
    stride = 37
    ld = O.diffraction_ld(rays, px, py, pz, stride=stride)
    e_line, e_ps = rel_l2(line[::stride], ld), rel_l2(per_slot[::stride], ld)
    between = rel_l2(line, per_slot)
    print(f"{case} P={P}: flags {flags:#x}, line vs per-slot {between:.2e}, "
          f"vs ld: line {e_line:.2e}, per-slot {e_ps:.2e}")
    assert between < 3e-6
    assert e_line < STRESS_BAR and e_line < e_ps + 3e-6


@pytest.mark.parametrize("P", [4, 8, 16])
def test_slot_line_with_float32_offsets(ctx, monkeypatch, P):
    """kirchhoff_gpu's float32 pixels: a grid in prepare_wave order, each
    point off the lane's line (the secant to its last slot) by float32
    rounding (~2e-9 mm, ~1e-5 rev of phase at the aperture's edge).  The
    line form takes each slot's offset into the phase to first order (at
    P = 16 in two passes of 8 slots, each anchored at its first slot);
    against the per-slot form (the offsets exact) and the FP64 kernel on the
    same float32 inputs."""
    rays = W.gaussian_source(65536, seed=2)
    g = grid(W.screen(1000.0, 0.02, 256))
    px, py, pz = xrtk.prepare_wave(g)
    r32 = xrtk.gpu_rays(rays["x"].shape[0])
    for f in xrtk.RAY_FIELDS:
        r32[f] = rays[f]
    p32 = xrtk.gpu_pixels(px, py, pz)
    line = xrtk.kirchhoff_gpu(ctx, r32, p32, pixels_per_lane=P)
    flags = ctx.last_plan()["inst_flags"]
    monkeypatch.setenv("XRTK_FORCE_FLAGS", "0x4000")
    per_slot = xrtk.kirchhoff_gpu(ctx, r32, p32, pixels_per_lane=P)
    f2 = ctx.last_plan()["inst_flags"]
    monkeypatch.delenv("XRTK_FORCE_FLAGS")
    ref = xrtk.kirchhoff_gpu(ctx, r32, p32, mode="f64")
    between, e_line, e_ps = rel_l2(line, per_slot), rel_l2(line, ref), rel_l2(per_slot, ref)
    print(f"P={P}: flags {flags:#x}; line vs per-slot {between:.2e}; vs f64: line {e_line:.2e}, "
          f"per-slot {e_ps:.2e}")
    assert flags & VAR_LINE and flags & VAR_LINEDEV, hex(flags)
    assert not f2 & VAR_LINE, hex(f2)
    assert between < 3e-6 and e_line < e_ps + 3e-6


def test_slot_line_needs_a_line(ctx):
    """Pixel lists: a lane holds P consecutive list entries.  In
    prepare_wave order those are neighbours along a row of the grid (a
    line, allowed) unless a lane straddles the end of a row: rows of 98
    entries at P = 4 (refused), and a shuffled list has no line at all
    (refused).  At L = 100 mm the separable amplitude, and with it the line,
    is refused; at P = 2 there is no line variant."""
    rays = W.gaussian_source(65536, seed=2)
    for side, half, line in ((256, 0.005, True), (256, 0.02, True), (96, 0.02, True),
                             (98, 0.02, False)):
        g = grid(W.screen(1000.0, half, side))
        px, py, pz = xrtk.prepare_wave(g)
        got = xrtk.diffraction_integral(rays, (px, py, pz), ctx=ctx, mode="f32r", pixels_per_lane=4)
        assert bool(ctx.last_plan()["inst_flags"] & VAR_LINE) == line, (side, half)
        ref = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
        assert rel_l2(got, ref) < 1e-5
    perm = np.random.default_rng(5).permutation(px.shape[0])
    shuffled = xrtk.diffraction_integral(rays, (px[perm], py[perm], pz[perm]), ctx=ctx, mode="f32r",
                                         pixels_per_lane=4)
    assert not ctx.last_plan()["inst_flags"] & VAR_LINE
    assert rel_l2(shuffled, ref[perm]) < 1e-5
    rays, g = _line_case("L100")
    xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=8)
    assert not ctx.last_plan()["inst_flags"] & (VAR_LINE | VAR_SEPAMP)
    rays, g = _line_case("gaussian")
    xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r", pixels_per_lane=2)
    f = ctx.last_plan()["inst_flags"]
    assert f & VAR_SEPAMP and not f & VAR_LINE


def _transform(rays, pix, M):
    """Applies the orthogonal matrix M to positions, normals and pixels."""
    out = dict(rays)
    for a, b, c in (("x", "y", "z"), ("nx", "ny", "nz")):
        v = M @ np.stack([rays[a], rays[b], rays[c]])
        out[a], out[b], out[c] = v[0].copy(), v[1].copy(), v[2].copy()
    p = M @ np.stack(pix)
    return out, (p[0].copy(), p[1].copy(), p[2].copy())


@pytest.mark.parametrize("frame", ["yxz", "xzy", "rotated"])
def test_geometry_frames(ctx, frame):
    """The same aperture and screen in another coordinate frame: with the zero
    axis of o along x or z (axis order of the records), and rotated so that
    no axis of o is zero.  The integrand depends only on distances and
    n.d, so every frame gives the field of the original one."""
    rays = W.point_source_aperture(16384, seed=9)
    pix = xrtk.prepare_wave(grid(W.screen(1000.0, 0.02, 48)))
    if frame == "yxz":
        M = np.array([[0.0, 1, 0], [1, 0, 0], [0, 0, 1]])
    elif frame == "xzy":
        M = np.array([[1.0, 0, 0], [0, 0, 1], [0, 1, 0]])
    else:
        a, b = np.radians(30.0), np.radians(20.0)
        Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
        Rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
        M = Rx @ Rz
    r2, p2 = _transform(rays, pix, M)
    base = xrtk.diffraction_integral(rays, pix, ctx=ctx, mode="f32r")
    got = xrtk.diffraction_integral(r2, p2, ctx=ctx, mode="f32r")
    ref = xrtk.diffraction_integral(r2, p2, ctx=ctx, mode="f64")
    e_ref, e_base = rel_l2(got, ref), rel_l2(got, base)
    print(f"{frame}: vs f64 {e_ref:.2e}, vs original frame {e_base:.2e}")
    assert e_ref < REL_F32R
    if frame != "rotated":
        # a permutation of axes keeps the FP32 records up to roundings of the
        # FP64 |q|, whose terms are summed in another order
        assert e_base < 1e-7
    else:
        assert e_base < REL_F32R


def test_polarisation_forms(ctx):
    """Rays with Es only, Ep only, both (random amplitudes and phases) and
    neither: the FP32 kernel carries each ray as a magnitude, a phase offset
    and a polarisation vector."""
    n = 8192
    rng = np.random.default_rng(5)
    rays = W.point_source_aperture(n, seed=5)
    es = rng.uniform(0.2, 1.0, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    ep = rng.uniform(0.2, 1.0, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    q = n // 4
    ep[:q] = 0.0
    es[q:2 * q] = 0.0
    es[2 * q:2 * q + 64] = 0.0
    ep[2 * q:2 * q + 64] = 0.0
    rays.update(es_re=es.real.copy(), es_im=es.imag.copy(), ep_re=ep.real.copy(),
                ep_im=ep.imag.copy())
    g = grid(W.screen(1000.0, 0.02, 64))
    f32 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    f64 = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f64")
    for c in range(2):
        assert rel_l2(f32[:, c], f64[:, c]) < REL_F32R
    # exchanging Es and Ep exchanges the two fields
    sw = dict(rays, es_re=rays["ep_re"], es_im=rays["ep_im"], ep_re=rays["es_re"],
              ep_im=rays["es_im"])
    f32s = xrtk.diffraction_integral(sw, g, ctx=ctx, mode="f32r")
    assert rel_l2(f32s[:, ::-1], f64) < REL_F32R
    # dark rays alone give exact zeros
    dark = {k: v[2 * q:2 * q + 64] for k, v in rays.items()}
    assert not np.any(xrtk.diffraction_integral(dark, g, ctx=ctx, mode="f32r"))


# ---- resident ray sets, upstream path and batches ----------------------------------------------
def test_path_up_phase(ctx):
    rng = np.random.default_rng(12)
    rays = W.point_source_aperture(4000, seed=12)
    pu = 10000.0 + rng.uniform(0.0, 0.01, 4000)
    for k in ("es_re", "es_im"):
        rays[k] = rays[k].copy()
    rays["es_re"][:] = 1.0
    rays["es_im"][:] = 0.0
    scr = W.screen(1000.0, 0.02, 16)
    px, py, pz = xrtk.prepare_wave(grid(scr))
    ref = O.diffraction_f64(rays, px, py, pz, path_up=pu)
    rs = xrtk.RaySet(ctx, rays, path_up=pu)
    try:
        for mode, tol in (("f32r", REL_F32R), ("f64", 1e-9)):
            (got,) = xrtk.solve_jobs(ctx, [xrtk.Job(rs, 10000.0, grid(scr))], mode=mode)
            assert rel_l2(got, ref) < tol, mode
    finally:
        rs.release()


def test_batch_equals_standalone_bitwise(ctx):
    specs = W.batch_instances(n=6, seed=21, max_rays=60000, max_side=96)
    sets = [xrtk.RaySet(ctx, W.instance_rays(s)) for s in specs]
    try:
        jobs = [xrtk.Job(rs, s["energy"], grid(W.instance_screen(s))) for rs, s in zip(sets, specs)]
        mixed = xrtk.solve_jobs(ctx, jobs)
        for j, m in zip(jobs, mixed):
            (alone,) = xrtk.solve_jobs(ctx, [j])
            assert np.array_equal(alone, m)
        # graph replays of the same batch, with eager work in between
        for _ in range(3):
            again = xrtk.solve_jobs(ctx, jobs)
            xrtk.diffraction_integral(W.point_source_aperture(500, seed=1),
                                      grid(W.screen(1000.0, 0.02, 8)), ctx=ctx, mode="f32r")
            for a, m in zip(again, mixed):
                assert np.array_equal(a, m)
        # and each job against the FP64 kernel
        for j, m, s in zip(jobs, mixed, specs):
            ref = xrtk.diffraction_integral(W.instance_rays(s), j.screen, ctx=ctx, mode="f64")
            assert rel_l2(m, ref) < REL_F32R
    finally:
        for rs in sets:
            rs.release()


def test_batch_config3_full_size(ctx):
    """The 64 uneven instances at their stated sizes (16k-2M rays, 64^2-1024^2
    screens, 6-25 keV, 200-20,000 mm) in one persistent launch: every
    instance against the long-double oracle on a strided pixel sample (the
    plain FP64 sum itself rounds phases of up to ~1e12 rad here), and its
    first instances bit for bit against their standalone solves."""
    specs = W.batch_instances()
    assert len(specs) == 64
    sets = [xrtk.RaySet(ctx, W.instance_rays(s)) for s in specs]
    try:
        jobs = [xrtk.Job(rs, s["energy"], grid(W.instance_screen(s))) for rs, s in zip(sets, specs)]
        outs = xrtk.solve_jobs(ctx, jobs)
        plan = ctx.last_plan()
        assert plan["n_fp64"] == 0, plan
        # most instances take the matrix cores; the near-field ones (block
        # remainder over its bound, no separable amplitude) the VALU kernel
        assert 0 < plan["n_mfma"] < len(specs), plan
        errs = []
        for i, s in enumerate(specs):
            px, py, pz = xrtk.prepare_wave(jobs[i].screen)
            stride = max(1, px.shape[0] // 12)
            ld = O.diffraction_ld(W.instance_rays(s), px, py, pz, stride=stride)
            errs.append(rel_l2(outs[i][::stride], ld))
        worst = int(np.argmax(errs))
        print(f"config 3: {len(specs)} instances ({plan['n_mfma']} on the matrix cores), max "
              f"rel-L2 vs long double {max(errs):.2e} (instance {worst}: {specs[worst]}), "
              f"median {np.median(errs):.2e}")
        assert max(errs) < STRESS_BAR
        for j in range(3):
            (alone,) = xrtk.solve_jobs(ctx, [jobs[j]])
            assert np.array_equal(alone, outs[j])
    finally:
        for rs in sets:
            rs.release()
