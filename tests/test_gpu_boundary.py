"""The drop-in boundary on the device: the reference callers' own bars through
the default entry points, the auto entry points against the oracle, and the
error paths and environment switches of the library."""
import json
import os

import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import _lib
from xrtk import workloads as W
from helpers import fixture_case, mp_cases, to_gpu_pixels, to_gpu_rays

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


def test_diffraction_integral_rs_reference_test(ctx):
    """validation/test_diffraction.py:36-59 with the device present: the
    default (f64) mode meets the fixture at 1e-10 absolute per component, on
    the device (the FP64 kernel)."""
    rays, pix, exp = fixture_case()
    args = [rays[k] for k in xrtk.RAY_FIELDS] + list(pix)
    es_re, es_im, ep_re, ep_im = xrtk.diffraction_integral_rs(*args)
    assert xrtk.auto_last_backend() == 1
    assert ctx.last_plan()["mode"] == xrtk.MODE_F64
    assert np.max(np.abs(np.array(es_re) - exp[:, 0].real)) < 1e-10
    assert np.max(np.abs(np.array(es_im) - exp[:, 0].imag)) < 1e-10
    assert np.max(np.abs(np.array(ep_re) - exp[:, 1].real)) < 1e-10
    assert np.max(np.abs(np.array(ep_im) - exp[:, 1].imag)) < 1e-10
    # diffraction_integral's default is the same FP64 kernel
    d = xrtk.diffraction_integral(rays, pix)
    assert np.array_equal(d[:, 0], np.array(es_re) + 1j * np.array(es_im))
    # the FP32-reduced form is opt-in
    f = xrtk.diffraction_integral(rays, pix, mode="f32r")
    assert ctx.last_plan()["mode"] == xrtk.MODE_F32R
    assert rel_l2(f, exp) < 5e-5


@pytest.mark.parametrize("case", mp_cases(), ids=lambda c: c[0])
def test_diffraction_integral_default_vs_oracle(ctx, case):
    name, rays, pix, exp = case
    got = xrtk.diffraction_integral(rays, pix)
    p = ctx.last_plan()
    assert p["backend"] == 1 and p["mode"] == xrtk.MODE_F64
    assert rel_l2(got, O.diffraction_f64(rays, *pix)) < 1e-11


def test_kirchhoff_auto_device_vs_oracle(ctx):
    """kirchhoff_auto (src/gpu/fallback.rs:19-26) with a device: the device
    ran it, and its result equals the reference's CPU branch
    (kirchhoff_cpu_from_gpu_types, fallback.rs:29-71, restated by the oracle
    on the same float32 inputs) to FP32-reduced accuracy, on the fixture
    and on a 4096-ray point-source aperture over 32^2 pixels."""
    rays, pix, _ = fixture_case()
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(pix)
    got = xrtk.kirchhoff_auto(r32, p32)
    assert xrtk.auto_last_backend() == 1
    ref = O.fallback_f32(r32, p32)
    assert rel_l2(got, ref) < 5e-5
    big = W.point_source_aperture(4096, seed=9)
    g = grid(W.screen(1000.0, 0.02, 32))
    r32, p32 = to_gpu_rays(big), to_gpu_pixels(xrtk.prepare_wave(g))
    got = xrtk.kirchhoff_auto(r32, p32)
    assert xrtk.auto_last_backend() == 1
    ref = O.fallback_f32(r32, p32)
    e = rel_l2(got, ref)
    print(f"kirchhoff_auto vs fallback oracle: {e:.2e}")
    assert e < 1e-5


# ---- error paths -------------------------------------------------------------------------
def test_injected_launch_failure_oneshot(ctx):
    """A failed pair-kernel launch surfaces as XRTK_ELAUNCH (the reference
    panics on a failed map, kirchhoff.rs:268-271) and the context keeps
    working: the next call gives the normal result."""
    rays = W.point_source_aperture(2048, seed=3)
    g = grid(W.screen(1000.0, 0.02, 16))
    good = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    _lib.lib.xrtk_debug_fail_launches(1)
    try:
        with pytest.raises(xrtk.XrtkError) as ei:
            xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    finally:
        _lib.lib.xrtk_debug_fail_launches(0)
    assert ei.value.code == _lib.ELAUNCH
    again = xrtk.diffraction_integral(rays, g, ctx=ctx, mode="f32r")
    assert np.array_equal(again, good)
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(xrtk.prepare_wave(g))
    _lib.lib.xrtk_debug_fail_launches(1)
    try:
        with pytest.raises(xrtk.XrtkError) as ei:
            xrtk.kirchhoff_gpu(ctx, r32, p32)
    finally:
        _lib.lib.xrtk_debug_fail_launches(0)
    assert ei.value.code == _lib.ELAUNCH
    assert np.all(np.isfinite(xrtk.kirchhoff_gpu(ctx, r32, p32)))


def test_injected_launch_failure_batched_and_served(ctx):
    """The batched solve returns XRTK_ELAUNCH on a failed first launch and
    succeeds afterwards; a served request whose launch fails completes with
    XRTK_ELAUNCH (its partial sums are not merged), and the server goes on
    serving the next request correctly."""
    specs = W.batch_instances(n=3, seed=51, max_rays=40000, max_side=64)
    sets = [xrtk.RaySet(ctx, W.instance_rays(s)) for s in specs]
    try:
        jobs = [xrtk.Job(rs, s["energy"], grid(W.instance_screen(s)))
                for rs, s in zip(sets, specs)]
        _lib.lib.xrtk_debug_fail_launches(1)
        try:
            with pytest.raises(xrtk.XrtkError) as ei:
                xrtk.solve_jobs(ctx, jobs)
        finally:
            _lib.lib.xrtk_debug_fail_launches(0)
        assert ei.value.code == _lib.ELAUNCH
        ref = xrtk.solve_jobs(ctx, jobs)
        srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, quantum_ms=100.0)
        try:
            _lib.lib.xrtk_debug_fail_launches(1)
            t = srv.submit(jobs[0])
            with pytest.raises(xrtk.XrtkError) as ei:
                srv.wait(t, timeout_ms=60000)
            assert ei.value.code == _lib.ELAUNCH
            _lib.lib.xrtk_debug_fail_launches(0)
            got, _ = srv.wait(srv.submit(jobs[1]), timeout_ms=60000)
            assert np.array_equal(got, ref[1])
        finally:
            _lib.lib.xrtk_debug_fail_launches(0)
            srv.close()
    finally:
        for rs in sets:
            rs.release()


def test_release_right_after_submit_is_refused(ctx):
    """A request holds its ray set from the moment submit returns: an
    immediate release is XRTK_EAGAIN, never a free under the request."""
    rays = W.point_source_aperture(30000, seed=4)
    rs = xrtk.RaySet(ctx, rays)
    g = grid(W.screen(1000.0, 0.02, 64))
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30)
    try:
        for _ in range(5):
            t = srv.submit(xrtk.Job(rs, 10000.0, g))
            with pytest.raises(xrtk.XrtkError) as ei:
                _lib.check(_lib.lib.xrtk_rayset_release(ctx.handle, rs.handle), "release")
            assert ei.value.code == _lib.EAGAIN
            srv.wait(t, timeout_ms=60000)
    finally:
        srv.close()
    rs.release()


def test_oversized_served_screen_rejected(ctx):
    """A screen of more than 0x7fff0000 pixels (32-bit slot indices) is
    XRTK_EINVAL at submit."""
    import ctypes as C
    rs = xrtk.RaySet(ctx, W.point_source_aperture(64, seed=4))
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30)
    try:
        # through the C ABI: the request is refused before its output is used
        out = np.zeros(4)
        job = xrtk.Job(rs, 10000.0, xrtk.WaveParams((0.0, 1000.0, 0.0), 0.02, 0.02,
                                                    70000, 70000)).to_c(out)
        t = C.c_uint64()
        assert _lib.lib.xrtk_server_submit(srv._h, C.byref(job), C.byref(t)) == _lib.EINVAL
    finally:
        srv.close()
        rs.release()


# ---- environment switches ----------------------------------------------------------------
def test_tune_cache_round_trip(tmp_path, monkeypatch):
    """XRTK_TUNE_CACHE: a context's autotune writes its table to the file; a
    context created later reads it and its one-shot solves of that bucket
    take the tuned shape."""
    path = tmp_path / "tune.json"
    monkeypatch.setenv("XRTK_TUNE_CACHE", str(path))
    a = xrtk.GpuContext.new()
    try:
        opts, _ = a.autotune(100000, 16384, mode="f32r")
    finally:
        a.close()
    entries = json.loads(path.read_text())
    assert any(e["log2_rays"] == 16 and e["log2_pixels"] == 14 and e["P"] == opts["pixels_per_lane"]
               for e in entries)
    # a hand-written entry for another bucket is read back too
    entries.append({"mode": 0, "log2_rays": 12, "log2_pixels": 12, "P": 4, "W": 4, "ray_src": 1,
                    "persistent": 0, "split": 0, "tile": 8192})
    path.write_text("[\n" + ",\n".join("  " + json.dumps(e) for e in entries) + "\n]\n")
    b = xrtk.GpuContext.new()
    try:
        rays = W.point_source_aperture(100000, seed=6)
        xrtk.diffraction_integral(rays, grid(W.screen(1000.0, 0.02, 128)), ctx=b, mode="f32r")
        p = b.last_plan()
        assert (p["pixels_per_lane"], p["waves_per_group"], p["persistent"]) == \
            (opts["pixels_per_lane"], opts["waves_per_group"], opts["persistent"])
        small = W.point_source_aperture(4096, seed=6)
        xrtk.diffraction_integral(small, grid(W.screen(1000.0, 0.02, 64)), ctx=b, mode="f32r")
        p = b.last_plan()
        assert (p["pixels_per_lane"], p["waves_per_group"], p["persistent"]) == (4, 4, 0)
    finally:
        b.close()
    monkeypatch.delenv("XRTK_TUNE_CACHE")
    c = xrtk.GpuContext.new()  # without the switch: the default shapes
    try:
        xrtk.diffraction_integral(small, grid(W.screen(1000.0, 0.02, 64)), ctx=c, mode="f32r")
        assert c.last_plan()["persistent"] == 1
    finally:
        c.close()


def test_no_graph_switch_gives_the_same_bits(monkeypatch):
    """XRTK_NO_GRAPH=1: batched solves replay eagerly instead of from a
    hipGraph, with the same results."""
    c = xrtk.GpuContext.new()
    specs = W.batch_instances(n=4, seed=61, max_rays=30000, max_side=96)
    sets = [xrtk.RaySet(c, W.instance_rays(s)) for s in specs]
    try:
        jobs = [xrtk.Job(rs, s["energy"], grid(W.instance_screen(s)))
                for rs, s in zip(sets, specs)]
        graph = [xrtk.solve_jobs(c, jobs) for _ in range(3)][-1]
        monkeypatch.setenv("XRTK_NO_GRAPH", "1")
        jobs2 = [xrtk.Job(rs, s["energy"] + 1.0, grid(W.instance_screen(s)))
                 for rs, s in zip(sets, specs)]
        for _ in range(3):
            eager = xrtk.solve_jobs(c, jobs2)
        monkeypatch.delenv("XRTK_NO_GRAPH")
        ref = [xrtk.solve_jobs(c, jobs2) for _ in range(3)][-1]
        for a, b in zip(eager, ref):
            assert np.array_equal(a, b)
        for a, b in zip(graph, xrtk.solve_jobs(c, jobs)):
            assert np.array_equal(a, b)
    finally:
        for rs in sets:
            rs.release()
        c.close()


def test_graph_replays_time_their_pair_kernels(ctx):
    """Replays of a captured batch record the pair-kernel events inside the
    graph: every replay reports a pair time of its own, above zero and
    within the replay's whole device time (and the same bits each time)."""
    rays = W.point_source_aperture(65536, seed=62)
    rs = xrtk.RaySet(ctx, rays)
    try:
        job = xrtk.Job(rs, 10000.0, grid(W.screen(1000.0, 0.02, 256)))
        first = xrtk.solve_jobs(ctx, [job])[0]
        for _ in range(4):
            again = xrtk.solve_jobs(ctx, [job])[0]
            pair_ms, solve_ms = ctx.last_timing()
            assert 0.0 < pair_ms <= solve_ms + 1e-3, (pair_ms, solve_ms)
            assert np.array_equal(again, first)
    finally:
        rs.release()


# ---- beam_to_diffraction_rays on the device ---------------------------------------------
def _beam(n, seed, amps=True):
    """A Beam's SoA columns after a curved optic: positions on a cap,
    directions fanning out from a point 10 m upstream, energies at 10 keV,
    per-ray normals, complex amplitudes (or none)."""
    rng = np.random.default_rng(seed)
    x, z = rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n)
    y = (x * x + z * z) / 4000.0
    d = np.sqrt(x * x + (y + 1e4) ** 2 + z * z)
    beam = {"x": x, "y": y, "z": z, "a": x / d, "b": (y + 1e4) / d, "c": z / d,
            "e": np.full(n, 10000.0)}
    if amps:
        beam["es"] = np.exp(1j * rng.uniform(0, 6.28, n)) * rng.uniform(0.5, 1.0, n)
        beam["ep"] = 0.3 * np.exp(1j * rng.uniform(0, 6.28, n))
    t = rng.uniform(-0.05, 0.05, (n, 2))
    nrm = np.stack([t[:, 0], np.ones(n), t[:, 1]], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return beam, nrm, d - 1e4


@pytest.mark.parametrize("amps,with_path", [(True, True), (False, False), (True, False)])
def test_rayset_from_beam_equals_host_gather(ctx, amps, with_path):
    """The device gather (xrtk_rayset_from_beam) writes the records the host
    gather (beam_to_diffraction_rays, prepare_wave.rs:78-111) and upload
    write: the same bits through the FP64 and the FP32-reduced solves, with
    the good indices in any order; with a path column its phase is folded in
    as the upload's path_up would, matching the oracle."""
    n = 20000
    beam, nrm, path = _beam(n, 7, amps)
    good = np.random.default_rng(8).permutation(n)[: n - 1234]
    pth = path if with_path else None
    dev = xrtk.RaySet.from_beam(ctx, beam, nrm, good, path=pth)
    rays = xrtk.beam_to_diffraction_rays(beam, nrm, good)
    host = xrtk.RaySet(ctx, rays, path_up=path[good] if with_path else None)
    g = xrtk.WaveParams((0.0, 1000.0, 0.0), 0.02, 0.02, 48, 48)
    try:
        assert dev.n == host.n == good.shape[0]
        for mode in ("f64", "f32r"):
            (a,) = xrtk.solve_jobs(ctx, [xrtk.Job(dev, 10000.0, g)], mode=mode)
            (b,) = xrtk.solve_jobs(ctx, [xrtk.Job(host, 10000.0, g)], mode=mode)
            assert np.array_equal(a, b), mode
        px, py, pz = xrtk.prepare_wave(g)
        ref = O.diffraction_f64(rays, px, py, pz, path_up=path[good] if with_path else None,
                                stride=37)
        (a,) = xrtk.solve_jobs(ctx, [xrtk.Job(dev, 10000.0, g)], mode="f64")
        assert rel_l2(a[::37], ref) < 1e-9
    finally:
        dev.release()
        host.release()


def test_rayset_from_beam_rejects_bad_index(ctx):
    beam, nrm, _ = _beam(100, 9)
    with pytest.raises(xrtk.XrtkError) as ei:
        xrtk.RaySet.from_beam(ctx, beam, nrm, [0, 5, 100])
    assert ei.value.code == _lib.EINVAL
