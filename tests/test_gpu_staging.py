"""kirchhoff_gpu (src/gpu/kirchhoff.rs:77) on host buffers large enough to go
through the page-locked staging block in 8 MiB chunks, each chunk's DMA queued
as it lands: ray and pixel lists whose last chunk is ragged, rays sent before
the host has planned, rays of two energies (regrouped by energy after that
first upload and sent again), against the long-double oracle on pixel
samples (gfx950 only)."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W
from helpers import to_gpu_rays

pytestmark = pytest.mark.gpu

VAR_MFMA = 4096  # XRTK_VAR_MFMA
REL_F32R = 1e-4  # the FP32-reduced bar, every route


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def ld_at(r32, p32, n=61):
    rays = {f: r32[f].astype(np.float64) for f in xrtk.RAY_FIELDS}
    stride = max(1, p32.shape[0] // n)
    ld = O.diffraction_ld(rays, p32["x"].astype(np.float64), p32["y"].astype(np.float64),
                          p32["z"].astype(np.float64), stride=stride)
    return stride, ld


def test_ragged_chunks_rays_and_pixels(ctx):
    """200,003 float32 rays (9.6 MB: one 8 MiB chunk and a ragged one) and an
    800 x 800 prepare_wave list (10.2 MB, likewise), twice in a row (the
    second call reuses the block)."""
    rays = W.point_source_aperture(200003, seed=31)
    r32 = to_gpu_rays(rays)
    p32 = xrtk.gpu_pixels(*xrtk.prepare_wave(
        xrtk.WaveParams((0.0, 1000.0, 0.0), 0.02, 0.02, 800, 800)))
    assert r32.nbytes > (8 << 20) and r32.nbytes % (8 << 20)
    assert p32.nbytes > (8 << 20) and p32.nbytes % (8 << 20)
    first = xrtk.kirchhoff_gpu(ctx, r32, p32)
    plan = ctx.last_plan()
    again = xrtk.kirchhoff_gpu(ctx, r32, p32)
    assert np.array_equal(first, again)
    # the variant flags come back with the result
    assert plan["n_mfma"] == 1 and plan["inst_flags"] & VAR_MFMA, plan
    stride, ld = ld_at(r32, p32)
    e = rel_l2(first[::stride], ld)
    print(f"ragged chunks: vs long double {e:.2e}")
    assert e < REL_F32R


def test_two_energies_sent_again_regrouped(ctx):
    """Rays of two energies, interleaved: the first upload (in the given
    order) is replaced by the energy-grouped one; each group is its own
    instance against the oracle's per-ray energies."""
    rays = W.point_source_aperture(180001, seed=32)
    rays["energy"] = np.where(np.arange(rays["x"].shape[0]) % 3 == 0, 12000.0, 10000.0)
    r32 = to_gpu_rays(rays)
    p32 = xrtk.gpu_pixels(*xrtk.prepare_wave(
        xrtk.WaveParams((0.0, 1000.0, 0.0), 0.02, 0.02, 96, 96)))
    got = xrtk.kirchhoff_gpu(ctx, r32, p32)
    stride, ld = ld_at(r32, p32)
    e = rel_l2(got[::stride], ld)
    print(f"two energies: vs long double {e:.2e}")
    assert e < REL_F32R
    # the same rays sorted by energy beforehand give the same field
    order = np.argsort(r32["energy"], kind="stable")
    got2 = xrtk.kirchhoff_gpu(ctx, r32[order], p32)
    assert rel_l2(got2, got) < 1e-6
