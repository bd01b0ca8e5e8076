"""The block pair kernel on the matrix cores (XRTK_VAR_MFMA): parity with the
long-double oracle and with the FP32 VALU pair kernel on grid screens, block
edges (screens that are not multiples of the 64-pixel block), the
power-of-two amplitude scale, the planner's refusals, the upstream path
phase, energy groups of the one-shot grid entry and graph replays
(gfx950 only)."""
import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu

VAR_MFMA = 4096     # XRTK_VAR_MFMA
VALU = 1            # XRTK_ENGINE_VALU
# The block kernel against the long-double oracle on a pixel sample.  Its
# error sources: the FP32 factor phases (~5e-7 rev at most), the FP16 hi+lo
# operands (~2^-21 per product), FP32 block sums and the block model's
# remainder (<= 1e-6 rev by the planner's bound, i.e. <= 6.3e-6 relative).
REL_BLOCK = 1e-5
# The two FP32 kernels against each other (each ~2-5e-6 from the oracle).
REL_KERNELS = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


def solve(ctx, rays, wave, energy=10000.0, path_up=None, **kw):
    """One batched solve of a resident copy of `rays`: (result, plan)."""
    rs = xrtk.RaySet(ctx, rays, path_up=path_up)
    try:
        (o,) = xrtk.solve_jobs(ctx, [xrtk.Job(rs, energy, wave)], **kw)
        plan = ctx.last_plan()
    finally:
        rs.release()
    return o, plan


def ld_sample(rays, wave, n=96, path_up=None):
    """(stride, long-double oracle on every stride-th pixel)."""
    px, py, pz = xrtk.prepare_wave(wave)
    stride = max(1, px.shape[0] // n)
    return stride, O.diffraction_ld(rays, px, py, pz, path_up=path_up, stride=stride)


@pytest.mark.parametrize("case", ["point_64k_256", "gaussian_1M_512"])
def test_workload_shapes_on_matrix_cores(ctx, case):
    """The point-source (65,536 x 256^2) and Gaussian-source (1,048,576 x
    512^2) solves at 1000 mm, 10 keV run on the matrix cores; they agree with
    the VALU kernel over every pixel and with the long-double oracle on a
    sample."""
    if case == "point_64k_256":
        rays, wave = W.point_source_aperture(65536, seed=1), grid(W.screen(1000.0, 0.02, 256))
    else:
        rays, wave = W.gaussian_source(1048576, seed=2), grid(W.screen(1000.0, 0.02, 512))
    om, pm = solve(ctx, rays, wave)
    ov, pv = solve(ctx, rays, wave, engine=VALU)
    assert pm["inst_flags"] & VAR_MFMA and pm["n_mfma"] == 1, pm
    assert not pv["inst_flags"] & VAR_MFMA and pv["n_mfma"] == 0, pv
    stride, ld = ld_sample(rays, wave)
    e_ld, e_v = rel_l2(om[::stride], ld), rel_l2(om, ov)
    print(f"{case}: block kernel vs long double {e_ld:.2e}, vs VALU kernel {e_v:.2e}")
    assert e_ld < REL_BLOCK
    assert e_v < REL_KERNELS


@pytest.mark.parametrize("nx,nz", [(100, 37), (1, 200), (200, 1), (65, 65), (63, 130),
                                   (129, 64)])
def test_block_edges(ctx, nx, nz):
    """Screens that are not multiples of the 64 x 64 block, and one-pixel-wide
    screens: the blocks past the edge compute but write nothing; every real
    pixel agrees with the VALU kernel and the oracle."""
    rays = W.point_source_aperture(65536, seed=5)
    pitch = 0.04 / 511  # the Gaussian-source screen's pixel pitch [mm]
    wave = xrtk.WaveParams((0.0, 1000.0, 0.0), 0.5 * pitch * max(nx - 1, 1),
                           0.5 * pitch * max(nz - 1, 1), nx, nz)
    om, pm = solve(ctx, rays, wave)
    ov, _ = solve(ctx, rays, wave, engine=VALU)
    assert pm["n_mfma"] == 1, pm
    stride, ld = ld_sample(rays, wave, n=160)
    assert np.all(np.isfinite(om))
    assert rel_l2(om[::stride], ld) < REL_BLOCK
    assert rel_l2(om, ov) < REL_KERNELS


def test_amplitude_scale_is_exact(ctx):
    """Amplitudes scaled by 2^-80 or 2^60 give the same result scaled by the
    same power of two, bit for bit: the kernel's FP16 factors are normalised
    by the largest amplitude's power of two, which cancels exactly."""
    rays = W.point_source_aperture(65536, seed=6)
    wave = grid(W.screen(1000.0, 0.02, 128))
    base, pb = solve(ctx, rays, wave)
    assert pb["n_mfma"] == 1
    for e in (-80, 60):
        f = 2.0 ** e
        scaled = dict(rays, es_re=rays["es_re"] * f, es_im=rays["es_im"] * f)
        o, p = solve(ctx, scaled, wave)
        assert p["n_mfma"] == 1
        assert np.array_equal(o, base * f), e


def test_uneven_amplitudes(ctx):
    """Ray amplitudes spread log-uniformly over six decades, some rays at
    zero: small amplitudes keep their FP16 factors in range relative to the
    largest one, and the result matches the oracle."""
    rays = W.point_source_aperture(65536, seed=7)
    rng = np.random.default_rng(3)
    a = 10.0 ** rng.uniform(-6.0, 0.0, rays["x"].shape[0])
    a[::97] = 0.0
    rays = dict(rays, es_re=rays["es_re"] * a, es_im=rays["es_im"] * a)
    wave = grid(W.screen(1000.0, 0.02, 192))
    o, p = solve(ctx, rays, wave)
    assert p["n_mfma"] == 1
    stride, ld = ld_sample(rays, wave)
    assert rel_l2(o[::stride], ld) < REL_BLOCK


def test_planner_refusals(ctx):
    """Inputs the block kernel does not take run on the VALU kernels, with
    the same bits as a solve that forbids the matrix cores: a near-field cell
    (100 mm, 30 keV: the block remainder exceeds its bound) and rays with
    per-ray normals (curved optic: no separable amplitude)."""
    near_rays, near_scr, _ = W.stress_cell(30000.0, 100.0, n_rays=65536)
    curved = W.curved_aperture(65536, tilt=0.2, seed=2)
    cases = [("near field", near_rays, grid(near_scr), 30000.0),
             ("per-ray normals", curved, grid(W.screen(1000.0, 0.02, 128)), 10000.0)]
    for what, rays, wave, e in cases:
        o, p = solve(ctx, rays, wave, energy=e)
        ov, pv = solve(ctx, rays, wave, energy=e, engine=VALU)
        assert p["n_mfma"] == 0 and not p["inst_flags"] & VAR_MFMA, (what, p)
        assert np.array_equal(o, ov), what


@pytest.mark.parametrize("case", ["10m_20keV", "20m_25keV", "config3_instance7"])
def test_steep_phase_steps(ctx, case):
    """Grids whose phase moves 0.8-1.3 rev per pixel (more than half a
    revolution: the kernel takes the steps mod 1) while the block remainder
    stays within its bound: on the matrix cores, against the long-double
    oracle and the VALU kernel."""
    if case == "config3_instance7":
        spec = W.batch_instances()[7]
        rays, scr, e = W.instance_rays(spec), W.instance_screen(spec), spec["energy"]
    else:
        dist, e, half = {"10m_20keV": (10000.0, 20000.0, 1.0),
                         "20m_25keV": (20000.0, 25000.0, 1.5)}[case]
        rays = W.point_source_aperture(65536, energy=e, seed=21)
        scr = W.screen(dist, half, 256)
    wave = grid(scr)
    o, p = solve(ctx, rays, wave, energy=e)
    ov, _ = solve(ctx, rays, wave, energy=e, engine=VALU)
    assert p["n_mfma"] == 1, (case, p)
    stride, ld = ld_sample(rays, wave)
    e_ld, e_v = rel_l2(o[::stride], ld), rel_l2(o, ov)
    print(f"{case}: block kernel vs long double {e_ld:.2e}, vs VALU {e_v:.2e}")
    assert e_ld < REL_BLOCK
    assert e_v < 2 * REL_KERNELS


def test_ep_pass(ctx):
    """Rays with Ep != 0 take both passes of the block kernel: the Es pass
    gives the same bits as for the same rays with Ep = 0, and Ep (here
    0.5 e^{i pi/5} Es on half the rays and an independent field on the rest)
    matches the oracle and the VALU kernel."""
    rays = W.gaussian_source(262144, seed=14)
    wave = grid(W.screen(1000.0, 0.02, 256))
    es = rays["es_re"] + 1j * rays["es_im"]
    rng = np.random.default_rng(14)
    ep = np.where(np.arange(es.shape[0]) % 2 == 0, 0.5 * np.exp(1j * np.pi / 5) * es,
                  0.3 * np.exp(2j * np.pi * rng.uniform(size=es.shape[0])))
    rays_ep = dict(rays, ep_re=ep.real, ep_im=ep.imag)
    o0, p0 = solve(ctx, rays, wave)
    o, p = solve(ctx, rays_ep, wave)
    ov, _ = solve(ctx, rays_ep, wave, engine=VALU)
    assert p0["n_mfma"] == 1 and p["n_mfma"] == 1, p
    assert np.array_equal(o[:, 0], o0[:, 0]) or rel_l2(o[:, 0], o0[:, 0]) < 1e-6
    stride, ld = ld_sample(rays_ep, wave)
    e_ld = rel_l2(o[::stride], ld)
    print(f"Ep != 0: block kernel vs long double {e_ld:.2e}, vs VALU {rel_l2(o, ov):.2e}")
    assert e_ld < REL_BLOCK
    assert rel_l2(o, ov) < REL_KERNELS


def test_upstream_path_phase(ctx):
    """Rays carrying an upstream optical path (beam.path): its phase joins
    each ray's offset before the block kernel's factors are formed."""
    rays = W.point_source_aperture(65536, seed=9)
    rng = np.random.default_rng(9)
    path_up = 1e4 + rng.uniform(0.0, 1e-3, rays["x"].shape[0])
    wave = grid(W.screen(1000.0, 0.02, 160))
    o, p = solve(ctx, rays, wave, path_up=path_up)
    ov, _ = solve(ctx, rays, wave, path_up=path_up, engine=VALU)
    assert p["n_mfma"] == 1
    stride, ld = ld_sample(rays, wave, path_up=path_up)
    assert rel_l2(o[::stride], ld) < REL_BLOCK
    assert rel_l2(o, ov) < REL_KERNELS


def test_one_shot_grid_energy_groups(ctx):
    """The one-shot grid entry (diffraction_integral on a WaveParams) with
    rays at two energies: one instance per energy group, both on the matrix
    cores, summed into one result."""
    r1 = W.point_source_aperture(32768, energy=9000.0, seed=10)
    r2 = W.point_source_aperture(32768, energy=11000.0, seed=11)
    rays = {k: np.concatenate([r1[k], r2[k]]) for k in r1}
    wave = grid(W.screen(1000.0, 0.02, 128))
    o = xrtk.diffraction_integral(rays, wave, ctx=ctx, mode="f32r")
    p = ctx.last_plan()
    ov = xrtk.diffraction_integral(rays, wave, ctx=ctx, mode="f32r", engine=VALU)
    assert p["n_mfma"] == 2, p
    stride, ld = ld_sample(rays, wave)
    assert rel_l2(o[::stride], ld) < REL_BLOCK
    assert rel_l2(o, ov) < REL_KERNELS


def test_graph_replays_identical(ctx, monkeypatch):
    """Repeated batched solves (the first eager, then graph replays) and
    eager replays without a graph give the same bits."""
    rays = W.gaussian_source(262144, seed=12)
    rs = xrtk.RaySet(ctx, rays)
    try:
        jobs = [xrtk.Job(rs, 10000.0, grid(W.screen(1000.0, 0.02, 256))),
                xrtk.Job(rs, 12000.0, grid(W.screen(800.0, 0.02, 300)))]
        first = xrtk.solve_jobs(ctx, jobs)
        assert ctx.last_plan()["n_mfma"] == 2
        for _ in range(3):
            again = xrtk.solve_jobs(ctx, jobs)
            for a, b in zip(first, again):
                assert np.array_equal(a, b)
        monkeypatch.setenv("XRTK_NO_GRAPH", "1")
        jobs2 = [xrtk.Job(rs, 10000.0, grid(W.screen(1000.0, 0.02, 256)))]
        e1 = xrtk.solve_jobs(ctx, jobs2)
        e2 = xrtk.solve_jobs(ctx, jobs2)
        assert np.array_equal(e1[0], e2[0])
        assert np.array_equal(e1[0], first[0])
    finally:
        rs.release()


def test_pinned_result_buffers(ctx):
    """solve_jobs(out=...) writes into the caller's page-locked buffers and
    returns them; a buffer of the wrong shape is refused."""
    rays = W.point_source_aperture(16384, seed=13)
    rs = xrtk.RaySet(ctx, rays)
    try:
        wave = grid(W.screen(1000.0, 0.02, 96))
        job = xrtk.Job(rs, 10000.0, wave)
        (ref,) = xrtk.solve_jobs(ctx, [job])
        buf = xrtk.pinned_result(96 * 96)
        (got,) = xrtk.solve_jobs(ctx, [job], out=[buf])
        assert got is buf and np.array_equal(buf, ref)
        with pytest.raises(ValueError):
            xrtk.solve_jobs(ctx, [job], out=[xrtk.pinned_result(95 * 96)])
    finally:
        rs.release()
