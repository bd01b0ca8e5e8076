"""kirchhoff_gpu (src/gpu/kirchhoff.rs:77) on float32 pixel lists that lie on
a grid, as prepare_wave's screen does once rounded to float32
(src/waves/prepare_wave.rs:40-68): the list is laid out as that grid and,
where the planner allows it, runs on the matrix cores with each row's and
column's float32 offset from the grid applied.  Parity with the long-double
oracle evaluated at the float32 values and with the list path
(XRTK_ENGINE_VALU), and the lists that must keep the list path (gfx950
only)."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W
from helpers import to_gpu_rays

pytestmark = pytest.mark.gpu

VAR_MFMA = 4096     # XRTK_VAR_MFMA
VALU = 1            # XRTK_ENGINE_VALU
# The block kernel against the long-double oracle on a pixel sample (its
# error sources as on grid screens: FP32 factor phases, FP16 hi + lo parts,
# FP32 block sums, the block model's remainder <= 1e-6 rev, and here the
# first-order turn by the offsets' difference, held below 1e-7 rev).
REL_BLOCK = 1e-5
REL_KERNELS = 1e-5  # block kernel vs the VALU list path, every pixel
REL_F32R = 1e-4     # the list path's bar (test_gpu_random.py)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def grid_of(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


def ld_at(r32, p32, n=97):
    """(stride, long-double oracle at the float32 rays and pixels on every
    stride-th pixel)."""
    rays = {f: r32[f].astype(np.float64) for f in xrtk.RAY_FIELDS}
    stride = max(1, p32.shape[0] // n)
    ld = O.diffraction_ld(rays, p32["x"].astype(np.float64), p32["y"].astype(np.float64),
                          p32["z"].astype(np.float64), stride=stride)
    return stride, ld


def jittered(g, amp, seed):
    """prepare_wave's grid with each row's x and each column's z moved by up
    to `amp` of a pixel (a grid of uneven spacing), rounded to float32."""
    px, py, pz = xrtk.prepare_wave(g)
    rng = np.random.default_rng(seed)
    jx = rng.uniform(-amp, amp, g.nx) * (2.0 * g.dx / (g.nx - 1))
    jz = rng.uniform(-amp, amp, g.nz) * (2.0 * g.dz / (g.nz - 1))
    px = px + np.repeat(jx, g.nz)
    pz = pz + np.tile(jz, g.nx)
    return xrtk.gpu_pixels(px, py, pz)


def run(ctx, r32, p32, **kw):
    got = xrtk.kirchhoff_gpu(ctx, r32, p32, **kw)
    return got, ctx.last_plan()


@pytest.mark.parametrize("case", ["point_64k_256", "gaussian_1M_512"])
def test_prepare_wave_list_on_matrix_cores(ctx, case):
    """The point-source (65,536 x 256^2) and Gaussian-source (1,048,576 x
    512^2) cases through kirchhoff_gpu on float32 records: on the matrix
    cores, within REL_BLOCK of the long-double oracle and REL_KERNELS of the
    list path on every pixel."""
    if case == "point_64k_256":
        rays, scr = W.point_source_aperture(65536, seed=1), W.screen(1000.0, 0.02, 256)
    else:
        rays, scr = W.gaussian_source(1048576, seed=2), W.screen(1000.0, 0.02, 512)
    r32 = to_gpu_rays(rays)
    p32 = xrtk.gpu_pixels(*xrtk.prepare_wave(grid_of(scr)))
    got, plan = run(ctx, r32, p32)
    assert plan["n_mfma"] == 1 and plan["inst_flags"] & VAR_MFMA, plan
    ref, plan_v = run(ctx, r32, p32, engine=VALU)
    assert plan_v["n_mfma"] == 0
    stride, ld = ld_at(r32, p32)
    e_ld, e_v = rel_l2(got[::stride], ld), rel_l2(got, ref)
    print(f"{case}: vs long double {e_ld:.2e}, vs list path {e_v:.2e}")
    assert e_ld < REL_BLOCK
    assert e_v < REL_KERNELS


@pytest.mark.parametrize("center,amp", [((0.0, 0.0), 5e-4), ((0.5, -0.5), 3e-4)])
def test_uneven_grid_offsets(ctx, center, amp):
    """Rows and columns up to 5e-4 of a pixel off an even grid (and a screen
    0.5 mm off axis, where float32 rounding adds up to 2e-4 of a pixel): at
    ~0.13 rev of phase per pixel the offsets are worth up to 6e-5 rev, a
    field error of 4e-4 if they were dropped; with them applied the block
    kernel stays within REL_BLOCK of the oracle."""
    rays = W.point_source_aperture(65536, seed=3)
    g = xrtk.WaveParams((center[0], 1000.0, center[1]), 0.02, 0.02, 256, 256)
    p32 = jittered(g, amp, seed=4)
    r32 = to_gpu_rays(rays)
    got, plan = run(ctx, r32, p32)
    assert plan["n_mfma"] == 1, plan
    ref, _ = run(ctx, r32, p32, engine=VALU)
    stride, ld = ld_at(r32, p32)
    e_ld, e_v = rel_l2(got[::stride], ld), rel_l2(got, ref)
    print(f"center {center}: vs long double {e_ld:.2e}, vs list path {e_v:.2e}")
    assert e_ld < REL_BLOCK
    assert e_v < REL_KERNELS


@pytest.mark.parametrize("sx,sz", [(-1.0, 1.0), (1.0, -1.0), (-1.0, -1.0)])
def test_mirrored_grids(ctx, sx, sz):
    """prepare_wave screens with negative half-widths (x or z descending
    along the list): still grids, on the matrix cores, against the oracle."""
    rays = W.point_source_aperture(32768, seed=11)
    r32 = to_gpu_rays(rays)
    g = xrtk.WaveParams((0.003, 1000.0, -0.002), 0.02 * sx, 0.015 * sz, 200, 136)
    p32 = xrtk.gpu_pixels(*xrtk.prepare_wave(g))
    got, plan = run(ctx, r32, p32)
    assert plan["n_mfma"] == 1, plan
    stride, ld = ld_at(r32, p32)
    e = rel_l2(got[::stride], ld)
    print(f"mirrored ({sx}, {sz}): vs long double {e:.2e}")
    assert e < REL_BLOCK


def test_ep_pass_with_offsets(ctx):
    """Rays with Ep != 0 on an uneven grid: both passes of the offsets
    kernel, against the oracle."""
    rays = W.point_source_aperture(32768, seed=5)
    rng = np.random.default_rng(6)
    rays["ep_re"] = 0.3 * rays["es_re"] - 0.2 * rays["es_im"]
    rays["ep_im"] = rng.uniform(-0.1, 0.1, rays["x"].shape[0])
    g = xrtk.WaveParams((0.0, 1000.0, 0.0), 0.01, 0.015, 128, 192)
    p32 = jittered(g, 3e-4, seed=7)
    r32 = to_gpu_rays(rays)
    got, plan = run(ctx, r32, p32)
    assert plan["n_mfma"] == 1 and not plan["inst_flags"] & 2048, plan  # not EPZERO
    stride, ld = ld_at(r32, p32)
    e = rel_l2(got[::stride], ld)
    print(f"Ep != 0: vs long double {e:.2e}")
    assert e < REL_BLOCK
    assert np.abs(got[:, 1]).max() > 0


def test_lists_off_the_grid_keep_the_list_path(ctx):
    """Lists that are not a prepare_wave grid run as lists, within the list
    path's bar: one point moved, the transposed order, rows further than
    1/100 of a pixel off an even grid, and a single column."""
    rays = W.point_source_aperture(8192, seed=8)
    r32 = to_gpu_rays(rays)
    g = xrtk.WaveParams((0.0, 1000.0, 0.0), 0.64, 0.64, 64, 64)
    px, py, pz = xrtk.prepare_wave(g)
    moved = pz.copy()
    moved[1000] += 1e-3
    tx = px.reshape(64, 64).T.ravel()
    tz = pz.reshape(64, 64).T.ravel()
    cases = {
        "moved": xrtk.gpu_pixels(px, py, moved),
        "transposed": xrtk.gpu_pixels(tx, py, tz),
        "uneven": jittered(g, 0.05, seed=9),
        "column": xrtk.gpu_pixels(*xrtk.prepare_wave(
            xrtk.WaveParams((0.0, 1000.0, 0.0), 0.0, 0.64, 1, 64))),
    }
    for name, p32 in cases.items():
        got, plan = run(ctx, r32, p32)
        stride, ld = ld_at(r32, p32)
        e = rel_l2(got[::stride], ld)
        print(f"{name}: n_mfma {plan['n_mfma']}, vs long double {e:.2e}")
        assert plan["n_mfma"] == 0, name
        assert e < REL_F32R, name


def test_grid_list_equals_itself_through_kirchhoff_auto(ctx):
    """kirchhoff_auto (src/gpu/fallback.rs:19-26) takes the same device
    route: the same bits as kirchhoff_gpu."""
    rays = W.point_source_aperture(16384, seed=10)
    r32 = to_gpu_rays(rays)
    p32 = xrtk.gpu_pixels(*xrtk.prepare_wave(
        xrtk.WaveParams((0.0, 1000.0, 0.0), 0.0075, 0.0075, 96, 96)))
    got, plan = run(ctx, r32, p32)
    assert plan["n_mfma"] == 1
    auto = xrtk.kirchhoff_auto(r32, p32)
    assert xrtk.auto_last_backend() == 1
    assert np.array_equal(auto, got)
