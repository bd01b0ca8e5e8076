"""The undulator oracle (src/sources/undulator.rs:120-244, build_i_map)
pinned by the reference's own tests restated, the closed form of a helical
undulator's fundamental and a 40-digit evaluation of one point (CPU only)."""
import math

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O

FINE_STR = 1.0 / 137.03599976
SIE0 = 1.602176565e-19


def planar(kx=0.0, ky=2.0, phase=0.0):
    return O.undulator(3.0, 0.3, kx, ky, 30.0, 50, phase)


def e1_of(u):
    return O.undulator_e1(u["gamma"], u["kx"], u["ky"], u["period_mm"])


def on_axis(u, e):
    return O.undulator_i_map(u, [e], [0.0], [0.0])


def test_gamma_and_fundamental():
    """undulator_fundamental_energy (undulator.rs:469-486); gamma as
    SynchrotronParams::new (bending_magnet.rs:40)."""
    assert O.undulator_gamma(3.0) == 3000.0 / 0.510998928
    u = O.undulator(6.0, 0.2, 0.0, 1.0, 20.0, 100)
    e1 = e1_of(u)
    assert 1000.0 < e1 < 100000.0
    # E1 = 2 gamma^2 h c / (lambda_u (1 + K^2/2)) / e
    exp = 2 * u["gamma"] ** 2 * 6.626069573e-34 * 299792458.0 / (0.02 * 1.5) / SIE0
    assert abs(e1 - exp) < 1e-12 * exp


def test_nonzero_at_fundamental():
    """undulator_build_i_map_nonzero / undulator_on_axis_nonzero (:488-514)."""
    u = planar()
    i, _, _ = on_axis(u, e1_of(u))
    assert i[0] > 0.0 and np.isfinite(i[0])


def test_planar_polarisation():
    """a_planar_undulator_radiates_sigma_polarized_on_axis (:555-578)."""
    u = O.undulator(3.0, 0.3, 0.0, 2.0, 30.0, 50)
    e1 = e1_of(u)
    _, s, p = on_axis(u, e1)
    assert abs(s[0]) > 0 and abs(p[0]) < 1e-6 * abs(s[0])
    r = O.undulator(3.0, 0.3, 2.0, 0.0, 30.0, 50)
    _, s, p = on_axis(r, e1)
    assert abs(p[0]) > 0 and abs(s[0]) < 1e-6 * abs(p[0])


def test_helical_handedness_and_closed_form():
    """a_helical_undulator_keeps_its_sense_of_rotation (:580-633): the sign
    of Im(Es conj(Ep)) and its magnitude K^2 N^2 gamma^2 / 4 x amp2flux / E1."""
    h = lambda ph: O.undulator(3.0, 0.3, 1.0, 1.0, 30.0, 50, ph)
    e1 = e1_of(h(0.0))
    _, s, p = on_axis(h(90.0), e1)
    right = s[0] * np.conj(p[0])
    _, s, p = on_axis(h(-90.0), e1)
    left = s[0] * np.conj(p[0])
    assert right.imag < 0 < left.imag
    assert abs(right.imag + left.imag) < 1e-9 * abs(right.imag)
    u = h(90.0)
    amp2flux = FINE_STR * 0.3 / SIE0 * 50
    exp = 0.25 * 1.0 * 50 * 50 * u["gamma2"] * amp2flux / e1
    assert abs(abs(right.imag) - exp) < 1e-6 * exp


def test_spectrum_peaks_at_odd_harmonics():
    """the_spectrum_peaks_at_the_fundamental_and_its_odd_harmonics (:635-679)."""
    u = planar()
    e1 = e1_of(u)
    n = 50.0
    peak = on_axis(u, e1)[0][0]
    scan = O.undulator_i_map(u, e1 * (0.9 + 0.2 * np.arange(201) / 200.0), np.zeros(201),
                             np.zeros(201))[0]
    assert np.all(scan <= peak)
    for r in (1 - 1 / n, 1 + 1 / n):
        assert on_axis(u, r * e1)[0][0] < 1e-6 * peak
    for m in (3.0, 5.0):
        assert on_axis(u, m * e1)[0][0] > 1e-3 * peak
    for m in (2.0, 4.0):
        assert on_axis(u, m * e1)[0][0] < 1e-6 * peak


def test_line_red_shifts_off_axis():
    """the_line_red_shifts_off_axis_by_the_undulator_equation (:681-706)."""
    u = planar()
    e1, g = e1_of(u), u["gamma"]
    for f in (0.5, 1.0):
        theta = f / g
        exp = 1.0 / (1.0 + f * f / (1.0 + 0.5 * 4.0))
        r = 0.6 + 0.5 * np.arange(401) / 400.0
        i = O.undulator_i_map(u, r * e1, np.full(401, theta), np.zeros(401))[0]
        assert abs(r[np.argmax(i)] - exp) < 0.5 / 50


def test_reference_gpu_method_meets_its_parity_bar():
    """the_shader_reproduces_the_cpu_spectrum_at_every_odd_harmonic
    (tests/undulator_gpu_parity.rs:12-64) on the two restatements: the
    float32 shader within 1e-4 of build_i_map at a sidelobe and the odd
    harmonics."""
    u = O.undulator(3.0, 0.3, 0.0, 2.0, 30.0, 50)
    e1 = e1_of(u)
    e = e1 * np.array([0.95, 1.0, 3.0, 5.0, 7.0, 9.0])
    z = np.zeros(e.shape[0])
    cpu, cs, _ = O.undulator_i_map(u, e, z, z)
    gpu, gs, _ = O.undulator_wgsl_f32(u, e, z, z)
    assert np.all(np.abs(gpu - cpu) / cpu < 1e-4)
    assert np.all(np.abs(np.abs(gs) - np.abs(cs)) / np.abs(cs) < 1e-4)


def _mp_point(u, e, theta, psi, n_steps=64):
    mp.mp.dps = 40
    f = mp.mpf
    pi2 = 2 * mp.pi
    gamma, gamma2 = f(u["gamma"]), f(u["gamma2"])
    kx, ky, lu = f(u["kx"]), f(u["ky"]), f(u["period_mm"]) * f("1e-3")
    ph = f(u["phase_deg"]) * mp.pi / 180
    q = (1 + f("0.5") * (kx * kx + ky * ky)) / (2 * gamma2)
    wig_amp = lu / (8 * pi2 * gamma2)
    sic = f("2.99792458e10") * f("1e-2")
    wc = f(e) * f("1519267514747457.9") / (sic * (1 - q))
    th, ps = f(theta), f(psi)
    off = th * th + ps * ps
    dirz = mp.sqrt(1 - off)
    slip = off / (1 + dirz) + q * dirz
    ax = az = mp.mpc(0)
    dt = f(1) / n_steps
    for j in range(n_steps):
        t = (j + f("0.5")) * dt
        phi = pi2 * t
        bx = ky / gamma * mp.sin(phi)
        bz = -kx / gamma * mp.sin(phi + ph)
        xe = -(ky * lu / (pi2 * gamma)) * mp.cos(phi)
        ze = (kx * lu / (pi2 * gamma)) * mp.cos(phi + ph)
        wig = wig_amp * (ky * ky * mp.sin(2 * phi) + kx * kx * mp.sin(2 * phi + 2 * ph))
        tp = wc * (lu * t * slip - dirz * wig - xe * th - ze * ps)
        ax += (bx - th) * mp.expj(tp) * dt
        az += (bz - ps) * mp.expj(tp) * dt
    # exact geometric sum of the N periods
    n = u["n_periods"]
    dphi = wc * lu * slip
    res = sum(mp.expj(k * dphi) for k in range(n))
    amp2flux = f(1) / f("137.03599976") * f(u["beam_current"]) / f("1.602176565e-19") * n
    sf = mp.sqrt(amp2flux / f(e))
    return complex(ax * res * gamma2 * sf), complex(az * res * gamma2 * sf)


@pytest.mark.parametrize("ratio,theta,psi", [(1.0, 0.0, 0.0), (3.0, 0.0, 0.0),
                                             (0.8, 1e-4, -5e-5)])
def test_vs_mpmath(ratio, theta, psi):
    """One point against a 40-digit evaluation with the N-period sum done
    term by term (no sinc folding)."""
    u = O.undulator(3.0, 0.3, 0.4, 1.7, 30.0, 50, 30.0)
    e = ratio * e1_of(u)
    _, s, p = O.undulator_i_map(u, [e], [theta], [psi])
    es, ep = _mp_point(u, e, theta, psi)
    ref = math.hypot(abs(es), abs(ep))
    assert abs(s[0] - es) < 1e-9 * ref and abs(p[0] - ep) < 1e-9 * ref
