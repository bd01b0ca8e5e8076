"""The allocation-free one-shot entry point (xrtk_kirchhoff_gpu_f32_noalloc)
through the C ABI: the same bits as xrtk_kirchhoff_gpu_f32 on the same
inputs, on two calls in a row (the second smaller, over what the first left
in the shared scratch block), and a clear refusal of a solve that needs more
than the block."""
import ctypes as C

import numpy as np
import pytest

import xrtk
from xrtk import _lib
from xrtk import workloads as W
from helpers import to_gpu_pixels, to_gpu_rays

pytestmark = pytest.mark.gpu


def _call(fn, ctx, r32, p32):
    out = np.zeros((p32.shape[0], 4))
    o = _lib.opts(xrtk.MODE_F32R)
    rc = fn(ctx.handle, r32.ctypes.data, r32.shape[0], p32.ctypes.data, p32.shape[0],
            _lib.dptr(out), C.byref(o))
    return rc, out


def _inputs(n_rays, side, seed):
    rays = W.point_source_aperture(n_rays, seed=seed)
    g = xrtk.WaveParams((0.0, 1000.0, 0.0), 0.02, 0.02, side, side)
    return to_gpu_rays(rays), to_gpu_pixels(xrtk.prepare_wave(g))


def test_noalloc_equals_oneshot_on_two_calls(ctx):
    lib = _lib.lib
    # 65,536 rays x 256^2 pixels (the point-source workload), then a smaller
    # 4,096 x 48^2 solve over what the first left in the block
    for n_rays, side, seed in ((65536, 256, 1), (4096, 48, 2)):
        r32, p32 = _inputs(n_rays, side, seed)
        rc, got = _call(lib.xrtk_kirchhoff_gpu_f32_noalloc, ctx, r32, p32)
        assert rc == _lib.OK, (n_rays, side, rc)
        rc, ref = _call(lib.xrtk_kirchhoff_gpu_f32, ctx, r32, p32)
        assert rc == _lib.OK
        assert np.array_equal(got, ref), (n_rays, side)
        assert np.all(np.isfinite(got)) and np.abs(got).max() > 0


def test_noalloc_refuses_a_solve_beyond_the_block(ctx):
    lib = _lib.lib
    r32, p32 = _inputs(1 << 20, 64, 3)  # 1 M rays: ~240 MiB of buffers
    rc, _ = _call(lib.xrtk_kirchhoff_gpu_f32_noalloc, ctx, r32, p32)
    assert rc == _lib.ENOMEM
    # the block and the context stay usable
    r32, p32 = _inputs(2048, 16, 4)
    rc, got = _call(lib.xrtk_kirchhoff_gpu_f32_noalloc, ctx, r32, p32)
    assert rc == _lib.OK
    rc, ref = _call(lib.xrtk_kirchhoff_gpu_f32, ctx, r32, p32)
    assert rc == _lib.OK and np.array_equal(got, ref)
    assert lib.xrtk_kirchhoff_gpu_f32_noalloc(None, None, 0, None, 0, None, None) == _lib.ENODEV
