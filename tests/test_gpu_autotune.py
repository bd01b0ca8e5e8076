"""Autotuners: the per-shape table used by one-shot solves and the batch
configuration used by xrtk_solve_jobs and the server.  Each test tunes its
own context, so the shared context keeps the default shapes."""
import numpy as np
import pytest

import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu

REL_F32R = 1e-4


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


@pytest.fixture
def fresh_ctx():
    c = xrtk.GpuContext.new()
    assert c is not None
    yield c
    c.close()


def test_autotune_bucket_used_by_oneshot(fresh_ctx):
    c = fresh_ctx
    opts, ms = c.autotune(200000, 65536, mode="f32r")
    print(f"bucket 2^17 rays x 2^16 pixels: {opts}, {ms:.2f} ms")
    assert opts["pixels_per_lane"] in (2, 4, 8, 16) and opts["waves_per_group"] in (1, 4)
    assert ms > 0
    rays = W.point_source_aperture(200000, seed=5)
    wave = grid(W.screen(1000.0, 0.02, 256))
    out = xrtk.diffraction_integral(rays, wave, ctx=c, mode="f32r")
    plan = c.last_plan()
    assert plan["pixels_per_lane"] == opts["pixels_per_lane"]
    assert plan["waves_per_group"] == opts["waves_per_group"]
    assert plan["persistent"] == opts["persistent"]
    ref = xrtk.diffraction_integral(rays, wave, ctx=c, mode="f64")
    assert W.rel_l2(out, ref) < REL_F32R


def test_autotune_jobs_fixes_batch_config(fresh_ctx):
    c = fresh_ctx
    specs = W.batch_instances(n=5, seed=41, max_rays=50000, max_side=128)
    sets = [xrtk.RaySet(c, W.instance_rays(s)) for s in specs]
    srv = None
    try:
        jobs = [xrtk.Job(rs, s["energy"], grid(W.instance_screen(s)))
                for rs, s in zip(sets, specs)]
        opts, ms = xrtk.autotune_jobs(c, jobs)
        print(f"batch of {len(jobs)}: {opts}, {ms:.2f} ms")
        assert opts["pixels_per_lane"] in (2, 4, 8, 16) and opts["waves_per_group"] in (1, 4)
        assert ms > 0
        mixed = xrtk.solve_jobs(c, jobs)
        plan = c.last_plan()
        assert plan["pixels_per_lane"] == opts["pixels_per_lane"]
        assert plan["waves_per_group"] == opts["waves_per_group"]
        for j, m, s in zip(jobs, mixed, specs):
            (alone,) = xrtk.solve_jobs(c, [j])
            assert np.array_equal(alone, m)
            ref = xrtk.diffraction_integral(W.instance_rays(s), j.screen, ctx=c, mode="f64")
            assert W.rel_l2(m, ref) < REL_F32R
        # a server created after the tune runs the same configuration
        srv = xrtk.Server(c, hbm_budget_bytes=8 << 30)
        got, _ = srv.wait(srv.submit(jobs[0]), timeout_ms=60000)
        assert np.array_equal(got, mixed[0])
    finally:
        if srv is not None:
            srv.close()
        for rs in sets:
            rs.release()
