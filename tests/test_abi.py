"""The C ABI library: loads, exports every declared symbol, layouts, host-side
entry points and the no-device behaviour of the auto entry point (CPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as O
import xrtk
from xrtk import _lib
from helpers import fixture_case, to_gpu_pixels, to_gpu_rays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "xrtk.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xrtk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(_lib.lib, s), s
        assert s in _lib.PROTOTYPES, f"{s} has no ctypes prototype"


def test_layouts_match_reference_structs():
    """GpuRay 48 B, GpuPixel 16 B (src/gpu/kirchhoff.rs:19-45)."""
    assert C.sizeof(_lib.XrtkGpuRay) == 48
    assert C.sizeof(_lib.XrtkGpuPixel) == 16
    assert [f for f, _ in _lib.XrtkGpuRay._fields_] == [
        "x", "y", "z", "nx", "ny", "nz", "nl", "energy", "es_re", "es_im", "ep_re", "ep_im"]


def test_chbar_constant():
    assert _lib.CHBAR == 1973.2697177417986


def test_prepare_wave_matches_oracle_bit_exact():
    w = xrtk.WaveParams((0.1, 1000.0, -0.3), 0.02, 0.035, 17, 9)
    x, y, z = xrtk.prepare_wave(w)
    ox, oy, oz = O.prepare_wave(w.center, w.dx, w.dz, w.nx, w.nz)
    assert np.array_equal(x, ox) and np.array_equal(y, oy) and np.array_equal(z, oz)
    x, y, z = xrtk.prepare_wave(xrtk.WaveParams((5.0, 1000.0, 3.0), 1.0, 1.0, 1, 1))
    assert (x[0], y[0], z[0]) == (5.0, 1000.0, 3.0)


def test_beam_to_diffraction_rays():
    """beam_to_diffraction_rays_basic (prepare_wave.rs:145-166) and the
    nl = n . (a, b, c) rule."""
    n = 3
    beam = {"x": [1.0, 4.0, 7.0], "y": [2.0, 5.0, 8.0], "z": [3.0, 6.0, 9.0],
            "a": [0.0, 0.6, 0.0], "b": [1.0, 0.8, 0.0], "c": [0.0, 0.0, 1.0],
            "e": [10000.0, 9000.0, 8000.0]}
    normals = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    rays = xrtk.beam_to_diffraction_rays(beam, normals, [0])
    assert rays["x"].shape == (1,)
    assert abs(rays["x"][0] - 1.0) < 1e-10 and abs(rays["energy"][0] - 10000.0) < 1e-10
    assert rays["es_re"][0] == 1.0 and rays["es_im"][0] == 0.0 and rays["ep_re"][0] == 0.0
    rays = xrtk.beam_to_diffraction_rays(beam, normals, [1, 2])
    assert rays["nl"][0] == 0.8 and rays["nl"][1] == 1.0
    beam["es"] = np.array([0.5 + 0.5j] * n)
    beam["ep"] = np.array([0.1 - 0.2j] * n)
    rays = xrtk.beam_to_diffraction_rays(beam, normals, [2])
    assert rays["es_im"][0] == 0.5 and rays["ep_im"][0] == -0.2


def test_diffraction_integral_rs_value_error():
    """check_equal_lengths semantics (crates/xrt-python/src/lib.rs:33-47)."""
    good = [0.0, 1.0]
    with pytest.raises(ValueError, match="array length mismatch"):
        xrtk.diffraction_integral_rs(good, good, [0.0], good, good, good, good, good, good, good,
                                     good, good, [0.0], [0.0], [1.0])
    with pytest.raises(ValueError, match="array length mismatch"):
        xrtk.diffraction_integral_rs(good, good, good, good, good, good, good, good, good, good,
                                     good, good, [0.0], [0.0, 1.0], [1.0])


def test_null_and_length_errors():
    lib = _lib.lib
    out = np.zeros(4)
    assert lib.xrtk_kirchhoff_gpu_f32(None, None, 0, None, 1, _lib.dptr(out), None) == _lib.ENODEV
    assert lib.xrtk_kirchhoff_auto_f32(None, 3, None, 0, None) == _lib.EINVAL
    assert lib.xrtk_prepare_wave(None, None, None, None) == _lib.EINVAL


@pytest.mark.skipif(xrtk.GpuContext.shared() is not None, reason="a device is present")
def test_auto_falls_back_to_cpu_without_device():
    """kirchhoff_auto with no device runs fallback.rs:29-71 (the reference's
    CPU branch); equal to the oracle's restatement of that loop."""
    rays, pix, exp = fixture_case()
    r32, p32 = to_gpu_rays(rays), to_gpu_pixels(pix)
    got = xrtk.kirchhoff_auto(r32, p32)
    assert xrtk.auto_last_backend() == 0
    ref = O.fallback_f32(r32, p32)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-13
    rel = np.abs(got[:, 0] - exp[:, 0]) / np.abs(exp[:, 0])
    assert np.all(rel < 1e-3)  # golden_gpu_auto_dispatch / long_distance bar


@pytest.mark.skipif(xrtk.GpuContext.shared() is not None, reason="a device is present")
def test_require_gpu_turns_fallback_into_enodev(monkeypatch):
    monkeypatch.setenv("XRTK_REQUIRE_GPU", "1")
    rays, pix, _ = fixture_case()
    with pytest.raises(xrtk.XrtkError) as ei:
        xrtk.kirchhoff_auto(to_gpu_rays(rays), to_gpu_pixels(pix))
    assert ei.value.code == _lib.ENODEV


def _run_no_device(code: str, **env):
    """Runs `code` in a fresh interpreter with the device hidden
    (XRTK_FORCE_NO_DEVICE=1: the shared context reports no device), so the
    auto entry points take their CPU branch on any machine."""
    import subprocess
    import sys
    e = dict(os.environ, XRTK_FORCE_NO_DEVICE="1", **env)
    e["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "xrt-rs_amd"),
                                       os.path.join(ROOT, "oracle"),
                                       os.path.join(ROOT, "tests")])
    r = subprocess.run([sys.executable, "-c", code], env=e, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_diffraction_integral_rs_reference_semantics_without_device():
    """The Python caller's own test (validation/test_diffraction.py:36-59):
    the fixture to 1e-10 absolute per component.  Without a device the f64
    entry runs the reference's f64 Kahan loop on the CPU (what
    diffraction_integral_rs always runs, crates/xrt-python/src/lib.rs:386);
    it equals the oracle's restatement of that loop."""
    out = _run_no_device("""
import numpy as np, xrtk, pyoracle as O
from helpers import fixture_case
assert xrtk.GpuContext.shared() is None
rays, pix, exp = fixture_case()
args = [rays[k] for k in xrtk.RAY_FIELDS] + list(pix)
es_re, es_im, ep_re, ep_im = xrtk.diffraction_integral_rs(*args)
assert xrtk.auto_last_backend() == 0
got = np.array(es_re) + 1j * np.array(es_im)
assert np.max(np.abs(got.real - exp[:, 0].real)) < 1e-10
assert np.max(np.abs(got.imag - exp[:, 0].imag)) < 1e-10
gp = np.array(ep_re) + 1j * np.array(ep_im)
assert np.max(np.abs(gp - exp[:, 1])) < 1e-10
ref = O.diffraction_f64(rays, *pix)
both = np.stack([got, gp], axis=1)
assert np.max(np.abs(both - ref) / np.abs(ref)) < 1e-14
# the same through diffraction_integral's default mode
d = xrtk.diffraction_integral(rays, pix)
assert np.array_equal(d, both)
print("ok")
""")
    assert out.strip() == "ok"


def test_require_gpu_turns_f64_fallback_into_enodev():
    out = _run_no_device("""
import xrtk
from helpers import fixture_case
rays, pix, _ = fixture_case()
try:
    xrtk.diffraction_integral_rs(*([rays[k] for k in xrtk.RAY_FIELDS] + list(pix)))
except xrtk.XrtkError as e:
    print(e.code)
""", XRTK_REQUIRE_GPU="1")
    assert out.strip() == str(_lib.ENODEV)


def test_f32r_mode_needs_a_device():
    out = _run_no_device("""
import xrtk
from helpers import fixture_case
rays, pix, _ = fixture_case()
try:
    xrtk.diffraction_integral(rays, pix, mode="f32r")
except xrtk.XrtkError as e:
    print(e.code)
""")
    assert out.strip() == str(_lib.ENODEV)


def test_device_entry_points_report_enodev_without_context():
    lib = _lib.lib
    h = C.c_uint64()
    assert lib.xrtk_rayset_upload(None, None, None, 0, C.byref(h)) == _lib.ENODEV
    assert lib.xrtk_solve_jobs(None, None, 0, None) == _lib.ENODEV
    srv = C.c_void_p()
    assert lib.xrtk_server_create(None, None, C.byref(srv)) in (_lib.ENODEV, _lib.EINVAL)
    sc = _lib.XrtkFresnelScreen(-0.1, 0.1, -0.1, 0.1, 4, 4)
    out = np.zeros(32)
    assert lib.xrtk_fresnel_propagate(None, None, None, None, None, 0, 1e4, 1e3, C.byref(sc),
                                      _lib.dptr(out), None, None, None) == _lib.ENODEV


def test_fresnel_argument_errors():
    """ValueError on mismatched ray columns and an empty screen axis, before
    any device is needed (the reference's max(n - 1, 1) underflows at n = 0)."""
    import xrtk
    scr = xrtk.FresnelScreen(-0.1, 0.1, 4, -0.1, 0.1, 4)
    with pytest.raises(ValueError):
        xrtk.fresnel_propagate([0.0, 1.0], [0.0], [1 + 0j, 1 + 0j], 1e4, 1e3, scr)
    with pytest.raises(ValueError):
        xrtk.fresnel_propagate([0.0], [0.0], [1 + 0j], 1e4, 1e3,
                               xrtk.FresnelScreen(-0.1, 0.1, 0, -0.1, 0.1, 4))


def test_workload_generators_deterministic():
    from xrtk import workloads as W
    a = W.point_source_aperture(1000, seed=7)
    b = W.point_source_aperture(1000, seed=7)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert np.all(np.abs(a["x"]) <= 0.1) and np.all(a["y"] == 0)
    assert np.allclose(np.abs(a["es_re"] + 1j * a["es_im"]), 1.0)
    g = W.gaussian_source(20000, seed=3)
    assert g["x"].shape == (20000,) and np.all(np.abs(g["z"]) <= 0.2)
    assert 0.08 < np.std(g["x"]) < 0.12 and 0.04 < np.std(g["z"]) < 0.06
    inst = W.batch_instances()
    assert len(inst) == 64
    assert all(16384 <= s["n_rays"] <= 2097152 and 64 <= s["side"] <= 1024 for s in inst)
    c = W.curved_aperture(5000, tilt=0.01, seed=3)
    nn = np.sqrt(c["nx"] ** 2 + c["ny"] ** 2 + c["nz"] ** 2)
    assert np.allclose(nn, 1.0) and np.ptp(c["nx"]) > 0.01 and np.ptp(c["y"]) > 0
    assert np.all(np.abs(c["nl"]) <= 1.0) and np.all(c["nl"] > 0.99)
    m = W.grazing_mirror(5000, length=100.0, bend_radius=50000.0, seed=3)
    assert np.ptp(m["y"]) > 99.0 and np.ptp(m["nz"]) > 0 and np.all(m["nl"] > 0)
    u = W.SplitMix64(1).uniform(100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
