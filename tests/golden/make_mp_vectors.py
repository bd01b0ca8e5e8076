"""Writes mp_small.json: seeded small Kirchhoff inputs and their integral
evaluated with mpmath at 50 significant digits (no rounding of any kind in
the sum), the brute-force check of the oracle.

The formula is the one of physwkim/xrt-rs src/waves/diffraction.rs:75-98:
  U = i k/(4 pi) (nl + n.(p - r)/|p - r|) exp(i k |p - r|) / |p - r|,
  Es += es U, Ep += ep U,  k = E / CHBAR * 1e7 [mm^-1].
Inputs are float64 values; k is the float64 value E / CHBAR * 1e7 evaluated
in IEEE double in that order (the reference's k, diffraction.rs:89), then
taken as exact.

Run:  python tests/golden/make_mp_vectors.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
CHBAR = 1973.2697177417986


def integral(rays, pix):
    out = []
    for p in range(len(pix["x"])):
        es = mp.mpc(0)
        ep = mp.mpc(0)
        for j in range(len(rays["x"])):
            dx = mp.mpf(pix["x"][p]) - mp.mpf(rays["x"][j])
            dy = mp.mpf(pix["y"][p]) - mp.mpf(rays["y"][j])
            dz = mp.mpf(pix["z"][p]) - mp.mpf(rays["z"][j])
            path = mp.sqrt(dx * dx + dy * dy + dz * dz)
            if path < mp.mpf("1e-30"):
                continue
            ns = (mp.mpf(rays["nx"][j]) * dx + mp.mpf(rays["ny"][j]) * dy
                  + mp.mpf(rays["nz"][j]) * dz) / path
            k = mp.mpf(float(rays["energy"][j]) / CHBAR * 1e7)
            amp = k / (4 * mp.pi) * (mp.mpf(rays["nl"][j]) + ns) / path
            u = mp.mpc(0, 1) * amp * mp.expj(k * path)
            es += mp.mpc(rays["es_re"][j], rays["es_im"][j]) * u
            ep += mp.mpc(rays["ep_re"][j], rays["ep_im"][j]) * u
        out.append([float(es.real), float(es.imag), float(ep.real), float(ep.imag)])
    return out


def case(name, seed, nr, npx, energy, dist, half, phalf, tilt=False):
    rng = np.random.default_rng(seed)
    rays = {
        "x": rng.uniform(-half, half, nr), "y": np.zeros(nr) if not tilt else rng.uniform(-half, half, nr) * 0.1,
        "z": rng.uniform(-half, half, nr),
        "nx": np.zeros(nr), "ny": np.ones(nr), "nz": np.zeros(nr),
        "nl": rng.uniform(0.8, 1.0, nr),
        "es_re": rng.uniform(0.5, 1.0, nr), "es_im": rng.uniform(-0.1, 0.1, nr),
        "ep_re": rng.uniform(0.5, 1.0, nr), "ep_im": rng.uniform(-0.1, 0.1, nr),
        "energy": np.full(nr, energy),
    }
    pix = {"x": rng.uniform(-phalf, phalf, npx), "y": np.full(npx, dist),
           "z": rng.uniform(-phalf, phalf, npx)}
    rays = {k: [float(v) for v in a] for k, a in rays.items()}
    pix = {k: [float(v) for v in a] for k, a in pix.items()}
    return {"id": name, "rays": rays, "pixels": pix, "expected": integral(rays, pix)}


def curved_case(name, seed, nr, npx, energy, dist, half, phalf, radius, tilt):
    """Rays on a toroidal cap y = x^2/(2 Rx) + z^2/(2 Rz), each with its own
    normal: the cap's normal at its point tilted at random by up to `tilt`
    rad about x and z (rays off a curved optic carry per-ray normals,
    src/oes/reflect.rs:53-56), nl = n . (beam direction from a point source
    10 m upstream) as beam_to_diffraction_rays forms it
    (src/waves/prepare_wave.rs:86-87)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-half, half, nr)
    z = rng.uniform(-half, half, nr)
    y = x * x / (2 * radius[0]) + z * z / (2 * radius[1])
    tx = -x / radius[0] + rng.uniform(-tilt, tilt, nr)
    tz = -z / radius[1] + rng.uniform(-tilt, tilt, nr)
    m = np.sqrt(tx * tx + 1.0 + tz * tz)
    nx, ny, nz = tx / m, 1.0 / m, tz / m
    d = np.sqrt(x * x + (y + 10000.0) ** 2 + z * z)
    rays = {
        "x": x, "y": y, "z": z, "nx": nx, "ny": ny, "nz": nz,
        "nl": (nx * x + ny * (y + 10000.0) + nz * z) / d,
        "es_re": rng.uniform(0.5, 1.0, nr), "es_im": rng.uniform(-0.5, 0.5, nr),
        "ep_re": rng.uniform(-0.3, 0.3, nr), "ep_im": rng.uniform(-0.3, 0.3, nr),
        "energy": np.full(nr, energy),
    }
    pix = {"x": rng.uniform(-phalf, phalf, npx), "y": np.full(npx, dist),
           "z": rng.uniform(-phalf, phalf, npx)}
    rays = {k: [float(v) for v in a] for k, a in rays.items()}
    pix = {k: [float(v) for v in a] for k, a in pix.items()}
    return {"id": name, "rays": rays, "pixels": pix, "expected": integral(rays, pix)}


if __name__ == "__main__":
    cases = [
        case("micro_1kev_2um", 10, 12, 6, 1000.0, 0.002, 0.001, 0.001),
        case("near_1kev_0p5mm", 11, 16, 8, 1000.0, 0.5, 0.05, 0.05),
        case("mid_5kev_20mm", 12, 24, 8, 5000.0, 20.0, 0.05, 0.02, tilt=True),
        case("far_10kev_1m", 13, 32, 8, 10000.0, 1000.0, 0.1, 0.02),
        curved_case("curved_10kev_1m", 14, 32, 8, 10000.0, 1000.0, 0.05, 0.02,
                    (50.0, 200.0), 0.3),
        curved_case("curved_8kev_50mm", 15, 24, 8, 8000.0, 50.0, 0.05, 0.01,
                    (20.0, 80.0), 0.2),
    ]
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "mp_small.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_mp_vectors.py", "mp_dps": 50, "cases": cases},
                  f, indent=1)
