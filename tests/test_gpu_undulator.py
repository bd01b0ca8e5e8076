"""Undulator intensity map on the device (build_i_map, undulator_gpu)
against the oracle and the reference's own tests."""
import numpy as np
import pytest

import pyoracle as O
import xrtk

pytestmark = pytest.mark.gpu

FINE_STR = 1.0 / 137.03599976
SIE0 = 1.602176565e-19


def und(kx=0.0, ky=2.0, phase=0.0):
    return xrtk.Undulator(3.0, 0.3, kx, ky, 30.0, 50, phase)


def o_und(u):
    return O.undulator(u.electron_energy_gev, u.beam_current, u.kx, u.ky, u.period_mm,
                       u.n_periods, u.phase_deg)


def e1_of(u):
    return O.undulator_e1(u.gamma, u.kx, u.ky, u.period_mm)


def points(u, n, seed):
    rng = np.random.default_rng(seed)
    e1, g = e1_of(u), u.gamma
    e = e1 * rng.uniform(0.5, 9.5, n)
    e[: n // 8] = e1 * np.round(rng.uniform(0.5, 9.5, n // 8))  # harmonics exactly
    th = rng.uniform(-1.5, 1.5, n) / g
    ps = rng.uniform(-1.5, 1.5, n) / g
    th[: n // 16] = 0.0
    ps[: n // 16] = 0.0
    return e, th, ps


@pytest.mark.parametrize("kx,ky,phase", [(0.0, 2.0, 0.0), (1.0, 1.0, 90.0), (0.7, 1.9, 30.0)])
def test_map_vs_oracle(ctx, kx, ky, phase):
    u = und(kx, ky, phase)
    e, th, ps = points(u, 4096, 3)
    di, ds, dp = u.build_i_map(e, th, ps, ctx=ctx)
    oi, os_, op = O.undulator_i_map(o_und(u), e, th, ps)
    scale = oi.max()
    assert np.all(np.abs(di - oi) <= 1e-11 * scale + 1e-9 * oi)
    amp = np.sqrt(scale)
    assert np.all(np.abs(ds - os_) <= 1e-11 * amp + 1e-9 * np.abs(os_))
    assert np.all(np.abs(dp - op) <= 1e-11 * amp + 1e-9 * np.abs(op))


@pytest.mark.parametrize("n_steps", [16, 128])
def test_step_count(ctx, n_steps):
    u = und(0.4, 1.7, 45.0)
    e, th, ps = points(u, 512, 7)
    di, _, _ = u.build_i_map(e, th, ps, ctx=ctx, n_steps=n_steps)
    oi, _, _ = O.undulator_i_map(o_und(u), e, th, ps, n_steps=n_steps)
    assert np.all(np.abs(di - oi) <= 1e-11 * oi.max() + 1e-9 * oi)


def test_undulator_gpu_axes(ctx):
    """tests/undulator_axes.rs:12-42 through the undulator_gpu entry point."""
    obs = np.zeros(1, dtype=xrtk.GPU_OBS_DTYPE)
    obs["energy"] = 1000.0
    planar = xrtk.undulator_gpu(ctx, obs, 0.0, 2.0, 30.0, 50, 5870.85, 0.3, 64, 0.0)
    s, p = abs(planar.amp_s[0]), abs(planar.amp_p[0])
    assert s > 0 and p < 1e-6 * s
    rot = xrtk.undulator_gpu(ctx, obs, 2.0, 0.0, 30.0, 50, 5870.85, 0.3, 64, 0.0)
    s, p = abs(rot.amp_s[0]), abs(rot.amp_p[0])
    assert p > 0 and s < 1e-6 * p


def test_undulator_gpu_matches_cpu_at_odd_harmonics(ctx):
    """tests/undulator_gpu_parity.rs:12-64 (bar 1e-4 for the f32 shader);
    the FP64 device map meets it with many digits to spare."""
    u = und()
    e1 = e1_of(u)
    ratios = np.array([0.95, 1.0, 3.0, 5.0, 7.0, 9.0])
    e = ratios * e1
    z = np.zeros(6)
    cpu, cs, _ = O.undulator_i_map(o_und(u), e, z, z)
    obs = np.zeros(6, dtype=xrtk.GPU_OBS_DTYPE)
    obs["energy"] = e.astype(np.float32)
    g = xrtk.undulator_gpu(ctx, obs, u.kx, u.ky, u.period_mm, u.n_periods, u.gamma,
                           u.beam_current, 64, 0.0)
    # the f32 observation points carry the energies rounded to f32
    cpu32, cs32, _ = O.undulator_i_map(o_und(u), obs["energy"].astype(np.float64), z, z)
    assert np.all(np.abs(g.intensity - cpu) / cpu < 1e-4)
    assert np.all(np.abs(np.abs(g.amp_s) - np.abs(cs)) / np.abs(cs) < 1e-4)
    assert np.all(np.abs(g.intensity - cpu32) / cpu32 < 1e-9)


def test_spectrum_and_red_shift(ctx):
    """the_spectrum_peaks... and the_line_red_shifts... (undulator.rs
    :635-706) on the device map."""
    u = und()
    e1, n, g = e1_of(u), 50.0, u.gamma
    r = 0.9 + 0.2 * np.arange(201) / 200.0
    z = np.zeros(201)
    scan = u.build_i_map(r * e1, z, z, ctx=ctx)[0]
    peak = u.build_i_map([e1], [0.0], [0.0], ctx=ctx)[0][0]
    assert np.all(scan <= peak * (1 + 1e-12))
    i = u.build_i_map(e1 * np.array([1 - 1 / n, 1 + 1 / n, 3, 5, 2, 4]), np.zeros(6),
                      np.zeros(6), ctx=ctx)[0]
    assert np.all(i[:2] < 1e-6 * peak)
    assert np.all(i[2:4] > 1e-3 * peak)
    assert np.all(i[4:] < 1e-6 * peak)
    for f in (0.5, 1.0):
        r = 0.6 + 0.5 * np.arange(401) / 400.0
        i = u.build_i_map(r * e1, np.full(401, f / g), np.zeros(401), ctx=ctx)[0]
        exp = 1.0 / (1.0 + f * f / (1.0 + 0.5 * u.k_squared()))
        assert abs(r[np.argmax(i)] - exp) < 0.5 / n


def test_helical_closed_form(ctx):
    e1 = e1_of(und(1.0, 1.0, 0.0))
    ims = []
    for ph in (90.0, -90.0):
        u = und(1.0, 1.0, ph)
        _, s, p = u.build_i_map([e1], [0.0], [0.0], ctx=ctx)
        ims.append((s[0] * np.conj(p[0])).imag)
    assert ims[0] < 0 < ims[1]
    u = und(1.0, 1.0, 90.0)
    amp2flux = FINE_STR * 0.3 / SIE0 * 50
    exp = 0.25 * 50 * 50 * u.gamma ** 2 * amp2flux / e1
    assert abs(abs(ims[0]) - exp) < 1e-6 * exp


def test_argument_errors(ctx):
    u = und()
    with pytest.raises(ValueError):
        u.build_i_map([1.0, 2.0], [0.0], [0.0], ctx=ctx)
    with pytest.raises(xrtk.XrtkError):
        u.build_i_map([1.0], [0.0], [0.0], ctx=ctx, n_steps=0)
    i, s, p = u.build_i_map([], [], [], ctx=ctx)
    assert i.shape == (0,)
