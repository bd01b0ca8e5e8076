"""The oracle pinned against the reference's golden vectors, mpmath brute
force, closed forms and invariants (CPU only)."""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O
from helpers import GOLDEN, f32_exact, fixture_case, mp_cases, ray, rays_of, to_gpu_pixels, \
    to_gpu_rays

CHBAR = 1973.2697177417986


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# ---- reference golden vector -------------------------------------------------
def test_fixture_numerical():
    """golden_diffraction_numerical (tests/golden_diffraction.rs:119-189):
    1e-10 absolute per component."""
    rays, pix, exp = fixture_case()
    got = O.diffraction_f64(rays, *pix)
    assert got.shape == exp.shape
    assert np.max(np.abs(got.real - exp.real)) < 1e-10
    assert np.max(np.abs(got.imag - exp.imag)) < 1e-10


def test_sincos_form_gives_the_same_bits():
    """The sincos() form of the loop (timed beside sin()/cos() for the CPU
    baseline) gives the same bits on the fixture and the mpmath cases."""
    cases = [fixture_case()[:2]] + [(r, p) for _, r, p, _ in mp_cases()]
    for rays, pix in cases:
        a = O.diffraction_f64(rays, *pix)
        b = O.diffraction_f64(rays, *pix, trig="sincos")
        assert np.array_equal(a, b)


def test_fixture_finite_nonzero():
    """golden_diffraction_finite_results (golden_diffraction.rs:18-77)."""
    rays, pix, _ = fixture_case()
    got = O.diffraction_f64(rays, *pix)
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got[:, 0]) > 1e-30)


def test_chbar_matches_reference_constant():
    """CHBAR (src/core/consts.rs:74), pinned at 1e-8 by golden_constants.rs:64."""
    with open(os.path.join(GOLDEN, "constants.json")) as f:
        assert abs(json.load(f)["CHBAR"] - CHBAR) < 1e-8


# ---- mpmath brute force ------------------------------------------------------
def fp64_bound(rays, pix):
    """Error bound of any FP64 evaluation of the sum: each term's phase k*path
    is rounded by a few ulp (k, path, product), i.e. <= 4 * phase * 2^-53
    radians, amplified by the cancellation ratio sum|t| / |sum t|."""
    px, py, pz = (np.asarray(a)[:, None] for a in pix)
    dx, dy, dz = px - rays["x"], py - rays["y"], pz - rays["z"]
    path = np.sqrt(dx * dx + dy * dy + dz * dz)
    k = rays["energy"] / CHBAR * 1e7
    ns = (rays["nx"] * dx + rays["ny"] * dy + rays["nz"] * dz) / path
    amp = np.abs(k / (4 * math.pi) * (rays["nl"] + ns) / path)
    mag = np.hypot(rays["es_re"], rays["es_im"]) * amp + np.hypot(rays["ep_re"], rays["ep_im"]) * amp
    return mag, 4.0 * float(np.max(k * path)) * 2.0 ** -53


@pytest.mark.parametrize("case", mp_cases(), ids=lambda c: c[0])
def test_oracle_vs_mpmath(case):
    """The oracle against 50-digit sums, within the FP64 bound of the case
    (about 1e-12 relative where the phases are ~1e4 rad)."""
    name, rays, pix, exp = case
    got = O.diffraction_f64(rays, *pix)
    mag, eps = fp64_bound(rays, pix)
    err = np.abs(got - exp).sum(axis=1)
    bound = (eps + 1e-14) * mag.sum(axis=1)
    assert np.all(err <= bound), (err / bound).max()
    if name.startswith("micro"):
        assert rel_l2(got, exp) < 1e-11


def _numpy_formula(rays, pix, sign=1.0, obliq=True, four=4.0, with_i=True):
    px, py, pz = (np.asarray(a)[:, None] for a in pix)
    dx, dy, dz = px - rays["x"], py - rays["y"], pz - rays["z"]
    path = np.sqrt(dx * dx + dy * dy + dz * dz)
    ns = (rays["nx"] * dx + rays["ny"] * dy + rays["nz"] * dz) / path
    k = rays["energy"] / CHBAR * 1e7
    amp = k / (four * math.pi) * (rays["nl"] + (ns if obliq else 0.0)) / path
    u = (1j if with_i else 1.0) * amp * np.exp(sign * 1j * k * path)
    es = ((rays["es_re"] + 1j * rays["es_im"]) * u).sum(axis=1)
    ep = ((rays["ep_re"] + 1j * rays["ep_im"]) * u).sum(axis=1)
    return np.stack([es, ep], axis=1)


@pytest.mark.parametrize("mutation", ["phase_sign", "no_obliquity_ns", "two_pi", "no_i"])
def test_mp_vectors_detect_planted_mistakes(mutation):
    """The mpmath vectors pin the formula: a restatement with one plausible
    mistake planted misses them by far more than the oracle's tolerance."""
    name, rays, pix, exp = mp_cases()[0]
    assert name.startswith("micro")
    kw = {"phase_sign": dict(sign=-1.0), "no_obliquity_ns": dict(obliq=False),
          "two_pi": dict(four=2.0), "no_i": dict(with_i=False)}[mutation]
    assert rel_l2(_numpy_formula(rays, pix), exp) < 1e-10  # the unmutated form agrees
    assert rel_l2(_numpy_formula(rays, pix, **kw), exp) > 1e-2


@pytest.mark.parametrize("name", ["curved_10kev_1m", "curved_8kev_50mm"])
def test_mp_vectors_pin_per_ray_normals(name):
    """The curved-optic vectors carry a normal per ray (diffraction.rs:22-25,
    86): the oracle meets them, and a restatement that gives every ray the
    first ray's normal (or the flat aperture's (0, 1, 0)) misses them."""
    (case,) = [c for c in mp_cases() if c[0] == name]
    _, rays, pix, exp = case
    assert np.ptp(rays["nx"]) > 0.1 and np.ptp(rays["nz"]) > 0.1
    # (FP64 rounding of phases ~5e10 rad leaves ~2.5e-6 at 1 m; within the
    # bound test_oracle_vs_mpmath holds it to)
    assert rel_l2(O.diffraction_f64(rays, *pix), exp) < 1e-5
    for n in ((rays["nx"][0], rays["ny"][0], rays["nz"][0]), (0.0, 1.0, 0.0)):
        uni = dict(rays, nx=np.full_like(rays["nx"], n[0]), ny=np.full_like(rays["ny"], n[1]),
                   nz=np.full_like(rays["nz"], n[2]))
        assert rel_l2(_numpy_formula(uni, pix), exp) > 1e-3


def test_single_ray_closed_form():
    """One ray, one pixel: U = i k/(4 pi) (nl + n.d/r) e^{ikr} / r (diffraction.rs:91-98),
    evaluated in 50-digit mpmath."""
    mp.mp.dps = 50
    r = ray(x=0.01, y=0.0, z=0.02, nx=0.0, ny=1.0, nz=0.0, nl=0.9, es=0.7 - 0.2j, ep=0.3 + 0.1j,
            energy=2000.0)
    p = (0.03, 0.4, -0.01)
    d = [mp.mpf(p[0]) - mp.mpf(r["x"]), mp.mpf(p[1]) - mp.mpf(r["y"]), mp.mpf(p[2]) - mp.mpf(r["z"])]
    path = mp.sqrt(sum(v * v for v in d))
    ns = (mp.mpf(r["nx"]) * d[0] + mp.mpf(r["ny"]) * d[1] + mp.mpf(r["nz"]) * d[2]) / path
    k = mp.mpf(2000.0 / CHBAR * 1e7)  # the reference's f64 k (diffraction.rs:89)
    u = mp.mpc(0, 1) * k / (4 * mp.pi) * (mp.mpf(r["nl"]) + ns) * mp.expj(k * path) / path
    es = complex(mp.mpc(r["es_re"], r["es_im"]) * u)
    ep = complex(mp.mpc(r["ep_re"], r["ep_im"]) * u)
    got = O.diffraction_f64(rays_of([r]), [p[0]], [p[1]], [p[2]])
    tol = 4.0 * float(k * path) * 2.0 ** -53 + 1e-14  # FP64 phase rounding
    assert abs(got[0, 0] - es) / abs(es) < tol
    assert abs(got[0, 1] - ep) / abs(ep) < tol


# ---- invariants and the reference's unit tests --------------------------------
def test_symmetry():
    """golden_diffraction_symmetry (golden_diffraction.rs:79-117)."""
    rays = rays_of([ray(es=1 + 0j, ep=1 + 0j)])
    got = O.diffraction_f64(rays, [-0.01, 0.01], [0.0, 0.0], [1000.0, 1000.0])
    assert abs(abs(got[0, 0]) - abs(got[1, 0])) < 1e-10


def test_ray_at_pixel_skipped():
    """ray_at_pixel_location_skipped (diffraction.rs:196-225): exactly zero."""
    rays = rays_of([ray(x=1.0, y=2.0, z=3.0, es=1 + 0j, ep=1 + 0j)])
    got = O.diffraction_f64(rays, [1.0], [2.0], [3.0])
    assert got[0, 0] == 0 and got[0, 1] == 0


def test_single_ray_single_pixel_ep_zero():
    """single_ray_single_pixel (diffraction.rs:123-149)."""
    rays = rays_of([ray(es=1 + 0j, ep=0j)])
    got = O.diffraction_f64(rays, [0.0], [1000.0], [0.0])
    assert abs(got[0, 0]) > 0
    assert abs(got[0, 1]) < 1e-30


def test_linearity_in_amplitudes():
    rng = np.random.default_rng(5)
    n = 40
    base = rays_of([ray(x=rng.uniform(-.1, .1), z=0.0, y=rng.uniform(-.1, .1)) for _ in range(n)])
    a, b = 0.3 - 0.7j, -1.1 + 0.2j
    es1 = rng.normal(size=n) + 1j * rng.normal(size=n)
    es2 = rng.normal(size=n) + 1j * rng.normal(size=n)
    pix = ([0.0, 0.01], [0.0, -0.02], [500.0, 500.0])

    def run(es):
        r = dict(base)
        r["es_re"], r["es_im"] = es.real.copy(), es.imag.copy()
        return O.diffraction_f64(r, *pix)[:, 0]

    lhs = run(a * es1 + b * es2)
    rhs = a * run(es1) + b * run(es2)
    assert np.max(np.abs(lhs - rhs)) / np.max(np.abs(rhs)) < 1e-12


def test_joint_translation_invariance():
    rng = np.random.default_rng(6)
    n = 50
    rays = rays_of([ray(x=rng.uniform(-.05, .05), y=rng.uniform(-.05, .05), energy=1000.0,
                        es=complex(rng.normal(), rng.normal())) for _ in range(n)])
    pix = (np.array([0.0, 0.01]), np.array([0.005, 0.0]), np.array([1.0, 1.0]))
    shift = (0.25, -0.5, 2.0)
    moved = dict(rays)
    moved["x"] = rays["x"] + shift[0]
    moved["y"] = rays["y"] + shift[1]
    moved["z"] = rays["z"] + shift[2]
    a = O.diffraction_f64(rays, *pix)
    b = O.diffraction_f64(moved, pix[0] + shift[0], pix[1] + shift[1], pix[2] + shift[2])
    assert rel_l2(b, a) < 1e-9


def test_parallel_pixels_center_brighter():
    """parallel_pixels (diffraction.rs:151-194)."""
    rays = rays_of([ray(x=(i - 50.0) * 0.01) for i in range(100)])
    px = [(i - 25.0) * 0.02 for i in range(50)]
    got = O.diffraction_f64(rays, px, [1000.0] * 50, [0.0] * 50)
    assert abs(got[25, 0]) ** 2 > abs(got[0, 0]) ** 2


def test_kahan_stability_and_large_count():
    """kahan_summation_stability and large_ray_count_diffraction
    (diffraction.rs:227-285, 307-351)."""
    n = 10000
    rays = rays_of([ray(x=(i - n / 2) * 0.001, es=complex(1.0 / n, 0)) for i in range(n)])
    big = O.diffraction_f64(rays, [0.0], [0.0], [100.0])
    small_rays = rays_of([ray(x=(i - 50.0) * 0.001, es=complex(1.0 / 100, 0)) for i in range(100)])
    small = O.diffraction_f64(small_rays, [0.0], [0.0], [100.0])
    assert np.isfinite(big[0, 0].real) and abs(big[0, 0]) > 0
    assert 0.01 < abs(big[0, 0]) / abs(small[0, 0]) < 100.0
    n = 50000
    rays = rays_of([ray(x=(i - n / 2) * 0.0001, es=complex(1.0 / n, 0)) for i in range(n)])
    got = O.diffraction_f64(rays, [0.0, 0.1], [0.0, 0.0], [100.0, 100.0])
    assert abs(got[0, 0]) > abs(got[1, 0]) * 0.5


def test_polychromatic_finite():
    """polychromatic_diffraction (diffraction.rs:353-397)."""
    rays = rays_of([ray(energy=8000.0), ray(x=0.01, energy=12000.0)])
    got = O.diffraction_f64(rays, [0.0], [0.0], [100.0])
    assert np.isfinite(got[0, 0].real) and abs(got[0, 0]) > 0


def test_stride_subsample_matches_full():
    rays, pix, _ = fixture_case()
    px = np.linspace(-0.05, 0.05, 9)
    full = O.diffraction_f64(rays, px, np.zeros(9), np.full(9, 1000.0))
    sub = O.diffraction_f64(rays, px, np.zeros(9), np.full(9, 1000.0), stride=4)
    assert np.array_equal(sub, full[::4])


# ---- fallback and reduced-method restatements -----------------------------------
def test_fallback_matches_diffraction_on_f32_inputs():
    """fallback.rs:29-71 is the same sum without Kahan; on f32-exact inputs it
    agrees with diffraction_integral to rounding."""
    rays, pix, _ = fixture_case()
    r32 = to_gpu_rays(rays)
    p32 = to_gpu_pixels(pix)
    fb = O.fallback_f32(r32, p32)
    ref = O.diffraction_f64(f32_exact(rays), *(np.asarray(a, np.float32).astype(np.float64)
                                              for a in pix))
    assert rel_l2(fb, ref) < 1e-12


def test_reference_gpu_method_meets_its_tolerance():
    """The reference GPU path (kirchhoff.rs + kirchhoff.wgsl, f32) against the
    fixture, at the bar of golden_gpu_long_distance_vs_fixture
    (golden_gpu_kirchhoff.rs:147-182): rel < 1e-3 per pixel."""
    rays, pix, exp = fixture_case()
    got = O.reduced_f32(to_gpu_rays(rays), to_gpu_pixels(pix))
    rel = np.abs(got[:, 0] - exp[:, 0]) / np.abs(exp[:, 0])
    assert np.all(rel < 1e-3)


# ---- prepare_wave ---------------------------------------------------------------
def test_prepare_wave_oracle_closed_form():
    """prepare_wave_pixel_count / _center / _grid_extent (prepare_wave.rs:118-143)."""
    x, y, z = O.prepare_wave([0.0, 1000.0, 0.0], 1.0, 1.0, 10, 10)
    assert x.shape[0] == 100
    x, y, z = O.prepare_wave([5.0, 1000.0, 3.0], 1.0, 1.0, 1, 1)
    assert abs(x[0] - 5.0) < 1e-10 and abs(y[0] - 1000.0) < 1e-10 and abs(z[0] - 3.0) < 1e-10
    x, y, z = O.prepare_wave([0.0, 1000.0, 0.0], 5.0, 3.0, 11, 7)
    assert abs(x[0] + 5.0) < 1e-10
    assert abs(x[(11 - 1) * 7] - 5.0) < 1e-10
    assert abs(z[6] - 3.0) < 1e-10 and abs(z[0] + 3.0) < 1e-10


# ---- extended-precision checker -------------------------------------------------
@pytest.mark.parametrize("case", mp_cases(), ids=lambda c: c[0])
def test_ld_oracle_vs_mpmath(case):
    """The long-double restatement against the 50-digit sums: its phase
    rounding is 2^-11 of the FP64 one, so it meets the FP64 bound with room."""
    name, rays, pix, exp = case
    got = O.diffraction_ld(rays, *pix)
    mag, eps = fp64_bound(rays, pix)
    err = np.abs(got - exp).sum(axis=1)
    assert np.all(err <= (eps / 64 + 1e-14) * mag.sum(axis=1))


def test_ld_oracle_beats_fp64_at_large_phase():
    """At 30 keV and 30 m (phase ~4.6e12 rad) the FP64 oracle's error is
    ~1e-4 per term; the long-double checker agrees with mpmath to ~1e-7."""
    mp.mp.dps = 50
    rng = np.random.default_rng(8)
    n = 12
    rays = rays_of([ray(x=rng.uniform(-.1, .1), y=0.0, z=rng.uniform(-.1, .1), nx=0.0, ny=1.0,
                        nz=0.0, energy=30000.0) for _ in range(n)])
    pix = ([0.01, -0.02], [30000.0, 30000.0], [0.0, 0.005])
    exp = []
    for p in range(2):
        es = mp.mpc(0)
        for j in range(n):
            d = [mp.mpf(pix[0][p]) - mp.mpf(rays["x"][j]), mp.mpf(pix[1][p]) - mp.mpf(rays["y"][j]),
                 mp.mpf(pix[2][p]) - mp.mpf(rays["z"][j])]
            path = mp.sqrt(sum(v * v for v in d))
            ns = d[1] / path
            k = mp.mpf(30000.0 / CHBAR * 1e7)  # the reference's f64 k
            es += mp.mpc(0, 1) * k / (4 * mp.pi) * (1 + ns) * mp.expj(k * path) / path
        exp.append(complex(es))
    exp = np.array(exp)
    ld = O.diffraction_ld(rays, *pix)[:, 0]
    f64 = O.diffraction_f64(rays, *pix)[:, 0]
    e_ld = np.abs(ld - exp) / np.abs(exp)
    e_64 = np.abs(f64 - exp) / np.abs(exp)
    assert np.all(e_ld < 1e-6)
    assert np.max(e_64) > 10 * np.max(e_ld)
