"""The Fresnel oracle (src/waves/fresnel.rs:53-106) pinned by the
reference's own tests, closed forms, mpmath brute force and planted mistakes
(CPU only)."""
import math

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O

CHBAR = 1973.2697177417986


def k_of(e):
    return e / CHBAR * 1e7


def test_reference_single_point_source():
    """fresnel_single_point_source (fresnel.rs:130-149)."""
    u, _, _ = O.fresnel_f64([0.0], [0.0], [1 + 0j], 10000.0, 1000.0, -0.1, 0.1, 21, -0.1, 0.1, 21)
    assert u.shape == (21 * 21,)
    assert abs(u[10 * 21 + 10]) >= abs(u[0])


def test_reference_two_slit():
    """fresnel_two_slit (fresnel.rs:151-165)."""
    u, _, _ = O.fresnel_f64([-0.01, 0.01], [0.0, 0.0], [1 + 0j, 1 + 0j], 10000.0, 1000.0,
                            -0.1, 0.1, 101, 0.0, 0.0, 1)
    inten = np.abs(u) ** 2
    assert inten[50] > 0.0 and inten[60] > 0.0


def test_point_source_closed_form():
    """One ray at (x0, z0): U = Es exp(i k r^2 / (2 d)) exactly."""
    x0, z0, es = 0.013, -0.007, 0.6 - 0.8j
    e, d = 12000.0, 750.0
    u, xp, zp = O.fresnel_f64([x0], [z0], [es], e, d, -0.05, 0.05, 9, -0.04, 0.06, 7)
    X, Z = np.meshgrid(xp, zp)  # rows iz, columns ix: index iz * nx + ix
    exp = es * np.exp(1j * (k_of(e) / (2 * d)) * ((X - x0) ** 2 + (Z - z0) ** 2))
    assert np.max(np.abs(u - exp.ravel())) < 1e-12


def test_two_slit_closed_form():
    """Slits at +-a on the x axis: |U|^2 = 2 + 2 cos(2 k a x / d)."""
    a, e, d = 0.01, 10000.0, 1000.0
    u, xp, _ = O.fresnel_f64([-a, a], [0.0, 0.0], [1 + 0j, 1 + 0j], e, d, -0.1, 0.1, 101,
                             0.0, 0.0, 1)
    exp = 2 + 2 * np.cos(2 * k_of(e) * a * xp / d)
    assert np.max(np.abs(np.abs(u) ** 2 - exp)) < 1e-9


def test_grid_formula():
    v = O.fresnel_grid(-0.1, 0.3, 5)
    assert np.array_equal(v, -0.1 + (0.3 - -0.1) * np.arange(5) / 4.0)
    assert np.array_equal(O.fresnel_grid(0.25, 9.0, 1), [0.25])  # max(n - 1, 1)


def _mp_fresnel(rx, rz, es, e, d, xp, zp):
    mp.mp.dps = 40
    k = mp.mpf(e) / mp.mpf(CHBAR) * 10 ** 7
    pre = k / (2 * mp.mpf(d))
    out = []
    for z in zp:
        for x in xp:
            s = mp.mpc(0)
            for j in range(len(rx)):
                r2 = (mp.mpf(x) - mp.mpf(rx[j])) ** 2 + (mp.mpf(z) - mp.mpf(rz[j])) ** 2
                s += mp.mpc(es[j].real, es[j].imag) * mp.expj(pre * r2)
            out.append(complex(s))
    return np.array(out)


def _case():
    rng = np.random.default_rng(21)
    n = 24
    rx = rng.uniform(-0.05, 0.05, n)
    rz = rng.uniform(-0.05, 0.05, n)
    es = rng.normal(size=n) + 1j * rng.normal(size=n)
    return rx, rz, es, 9000.0, 800.0


def test_vs_mpmath():
    rx, rz, es, e, d = _case()
    u, xp, zp = O.fresnel_f64(rx, rz, es, e, d, -0.08, 0.08, 6, -0.06, 0.07, 5)
    ref = _mp_fresnel(rx, rz, es, e, d, xp, zp)
    # phases reach k r^2 / (2 d) ~ 1e2 rad: FP64 rounds each by ~1e2 * 2^-53
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.parametrize("mutation", ["phase_sign", "k_over_d", "drop_es_phase"])
def test_mpmath_vector_detects_planted_mistakes(mutation):
    rx, rz, es, e, d = _case()
    _, xp, zp = O.fresnel_f64(rx, rz, es, e, d, -0.08, 0.08, 6, -0.06, 0.07, 5)
    ref = _mp_fresnel(rx, rz, es, e, d, xp, zp)
    X, Z = np.meshgrid(xp, zp)
    r2 = (X.ravel()[:, None] - rx[None, :]) ** 2 + (Z.ravel()[:, None] - rz[None, :]) ** 2
    pre = k_of(e) / (2 * d)
    good = (es[None, :] * np.exp(1j * pre * r2)).sum(1)
    assert np.linalg.norm(good - ref) / np.linalg.norm(ref) < 1e-10
    if mutation == "phase_sign":
        bad = (es[None, :] * np.exp(-1j * pre * r2)).sum(1)
    elif mutation == "k_over_d":
        bad = (es[None, :] * np.exp(1j * 2 * pre * r2)).sum(1)
    else:
        bad = (np.abs(es)[None, :] * np.exp(1j * pre * r2)).sum(1)
    assert np.linalg.norm(bad - ref) / np.linalg.norm(ref) > 1e-2


def test_stride_and_threads():
    rx, rz, es, e, d = _case()
    full, _, _ = O.fresnel_f64(rx, rz, es, e, d, -0.08, 0.08, 33, -0.06, 0.07, 17, threads=1)
    sub, _, _ = O.fresnel_f64(rx, rz, es, e, d, -0.08, 0.08, 33, -0.06, 0.07, 17, stride=7,
                              threads=4)
    assert np.array_equal(full[::7], sub)
