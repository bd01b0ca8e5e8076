"""Seeded synthetic inputs of the Kirchhoff workloads.

Geometry follows xrt's convention: the beam travels along +y, the aperture is
the y = 0 plane with normal (0, 1, 0), screens are prepare_wave grids normal
to y.  Random numbers: SplitMix64 -> 53-bit uniform doubles, Box-Muller for
normals, so a (seed, draw) pair gives the same value everywhere.

  point_source_aperture  flat aperture lit by a point source 10 m upstream,
                         Es = exp(i k d), Ep = 0, nl = cos(theta)
                         (65,536 rays, +-0.1 mm, 256^2 screen +-20 um at
                         1000 mm, 10 keV is the 4.29e9-pair case)
  gaussian_source        Gaussian source 10x5 um, divergence 10x5 urad,
                         propagated 10 m, clipped to +-0.2 mm
                         (1,048,576 rays x 512^2 at 1000 mm, 10 keV)
  stress_cell            262,144 rays x 256^2, aperture scaled to ~32
                         Fresnel zones for (E, L)
  batch_instances        64 uneven instances (16k-2M rays, 64^2-1024^2,
                         6-25 keV, 200-20,000 mm)
  curved_aperture        rays on a toroidal cap, each with its own surface
                         normal (the input a curved optic hands on)
  grazing_mirror         the footprint of a grazing-incidence mirror, spread
                         along the beam (a bundle that is not thin along the
                         line of sight)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

CHBAR = 1973.2697177417986
SRC_DIST = 10000.0  # mm, source upstream of the aperture
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def k_of(energy_ev: float) -> float:
    """k [mm^-1] = E / CHBAR * 1e7 (src/waves/diffraction.rs:89)."""
    return energy_ev / CHBAR * 1e7


class SplitMix64:
    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1)."""
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        t = 2.0 * np.pi * u2
        return np.concatenate([r * np.cos(t), r * np.sin(t)])[:n]


@dataclass
class Screen:
    center: tuple
    dx: float
    dz: float
    nx: int
    nz: int


def _rays_from_positions(x, z, sx, sz, energy):
    n = x.shape[0]
    dxs, dzs = x - sx, z - sz
    d = np.sqrt(dxs * dxs + SRC_DIST * SRC_DIST + dzs * dzs)
    k = k_of(energy)
    ph = k * d
    zeros = np.zeros(n)
    return {
        "x": x, "y": zeros.copy(), "z": z,
        "nx": zeros.copy(), "ny": np.ones(n), "nz": zeros.copy(),
        "nl": SRC_DIST / d,
        "es_re": np.cos(ph), "es_im": np.sin(ph),
        "ep_re": zeros.copy(), "ep_im": zeros.copy(),
        "energy": np.full(n, float(energy)),
    }


def point_source_aperture(n_rays=65536, half=0.1, energy=10000.0, seed=1):
    """Rays uniform on |x|, |z| <= half on y = 0, lit by a point source at
    (0, -10 m, 0)."""
    rng = SplitMix64(seed)
    x = (rng.uniform(n_rays) * 2.0 - 1.0) * half
    z = (rng.uniform(n_rays) * 2.0 - 1.0) * half
    return _rays_from_positions(x, z, 0.0, 0.0, energy)


def gaussian_source(n_rays=1048576, half=0.2, energy=10000.0, sigma=(0.010, 0.005),
                    div=(10e-6, 5e-6), seed=2):
    """Gaussian source, propagated 10 m to the aperture plane, rays outside
    +-half rejected until n_rays are accepted."""
    rng = SplitMix64(seed)
    xs, zs, xa, za = [], [], [], []
    got = 0
    while got < n_rays:
        m = int((n_rays - got) * 1.1) + 1024
        sx = rng.normal(m) * sigma[0]
        sz = rng.normal(m) * sigma[1]
        ax = rng.normal(m) * div[0]
        az = rng.normal(m) * div[1]
        x = sx + SRC_DIST * ax
        z = sz + SRC_DIST * az
        keep = (np.abs(x) <= half) & (np.abs(z) <= half)
        xs.append(sx[keep]); zs.append(sz[keep]); xa.append(x[keep]); za.append(z[keep])
        got += int(keep.sum())
    sx = np.concatenate(xs)[:n_rays]
    sz = np.concatenate(zs)[:n_rays]
    x = np.concatenate(xa)[:n_rays]
    z = np.concatenate(za)[:n_rays]
    return _rays_from_positions(x, z, sx, sz, energy)


def _point_source_rays(x, y, z, n, dirn, energy):
    """Rays at (x, y, z) with normals n (3, N) lit by a point source 10 m
    upstream: Es = exp(i k d), Ep = 0, nl = n . dirn (dirn (3, N) the beam
    direction the obliquity takes, prepare_wave.rs:86-87)."""
    nr = x.shape[0]
    d = np.sqrt(x * x + (y + SRC_DIST) ** 2 + z * z)
    ph = k_of(energy) * d
    zeros = np.zeros(nr)
    return {
        "x": x, "y": y, "z": z, "nx": n[0], "ny": n[1], "nz": n[2],
        "nl": n[0] * dirn[0] + n[1] * dirn[1] + n[2] * dirn[2],
        "es_re": np.cos(ph), "es_im": np.sin(ph),
        "ep_re": zeros.copy(), "ep_im": zeros.copy(),
        "energy": np.full(nr, float(energy)),
    }


def curved_aperture(n_rays=65536, half=0.1, energy=10000.0, radius=(2000.0, 5000.0), tilt=0.0,
                    seed=1):
    """Rays uniform on |x|, |z| <= half of the toroidal cap
    y = x^2 / (2 Rx) + z^2 / (2 Rz), each carrying the surface normal at its
    point, (-x/Rx, 1, -z/Rz) normalised (per-ray normals, as rays taken off a
    curved optic carry, src/oes/reflect.rs:53-56), tilted further at random
    by up to `tilt` rad about x and z (so that the obliquity term n.d/|d|
    differs from ray to ray by ~tilt^2 / 2).  Lit by a point source 10 m
    upstream."""
    rng = SplitMix64(seed)
    x = (rng.uniform(n_rays) * 2.0 - 1.0) * half
    z = (rng.uniform(n_rays) * 2.0 - 1.0) * half
    y = x * x / (2.0 * radius[0]) + z * z / (2.0 * radius[1])
    tx = -x / radius[0] + (rng.uniform(n_rays) * 2.0 - 1.0) * tilt
    tz = -z / radius[1] + (rng.uniform(n_rays) * 2.0 - 1.0) * tilt
    m = np.sqrt(tx * tx + 1.0 + tz * tz)
    n = (tx / m, 1.0 / m, tz / m)
    d = np.sqrt(x * x + (y + SRC_DIST) ** 2 + z * z)
    return _point_source_rays(x, y, z, n, (x / d, (y + SRC_DIST) / d, z / d), energy)


def grazing_mirror(n_rays=65536, length=100.0, width=0.5, theta=3e-3, energy=10000.0,
                   bend_radius=None, seed=1):
    """The footprint of a mirror at grazing angle theta (rad) to the beam,
    which runs along +y: rays at (x, s cos theta, s sin theta), s uniform on
    +-length/2 and x on +-width/2, so the bundle spans `length` mm along the
    line of sight.  Normals (0, -sin theta, cos theta), or, with
    `bend_radius` (mm), those of the mirror bent to that radius along s
    (per ray).  nl = n . (0, cos 2 theta, sin 2 theta), the reflected beam
    direction.  Lit by a point source 10 m upstream.  The reflected beam
    reaches y = L at z = L tan(2 theta)."""
    rng = SplitMix64(seed)
    s = (rng.uniform(n_rays) * 2.0 - 1.0) * (0.5 * length)
    x = (rng.uniform(n_rays) * 2.0 - 1.0) * (0.5 * width)
    c, sn = math.cos(theta), math.sin(theta)
    y, z = s * c, s * sn
    a = s / bend_radius if bend_radius else np.zeros(n_rays)  # local slope of the bend
    n = (np.zeros(n_rays), -np.sin(theta + a), np.cos(theta + a))
    refl = (0.0, math.cos(2 * theta), math.sin(2 * theta))
    return _point_source_rays(x, y, z, n, refl, energy)


def screen(distance=1000.0, half=0.02, side=256) -> Screen:
    return Screen((0.0, float(distance), 0.0), float(half), float(half), int(side), int(side))


def stress_cell(energy, distance, n_rays=262144, side=256, seed=3):
    """One (E, L) cell: aperture a = clamp(sqrt(32 lambda L_eff), 0.05, 0.5) mm
    (~32 Fresnel zones ~ 256 px / 8), screen half-width a (L + 10 m) / 10 m."""
    lam = 2.0 * math.pi / k_of(energy)
    l_eff = 1.0 / (1.0 / distance + 1.0 / SRC_DIST)
    a = min(max(math.sqrt(32.0 * lam * l_eff), 0.05), 0.5)
    nf = a * a / (lam * l_eff)
    rays = point_source_aperture(n_rays, a, energy, seed)
    scr = screen(distance, a * (distance + SRC_DIST) / SRC_DIST, side)
    return rays, scr, {"aperture_mm": a, "fresnel_number": nf}


STRESS_ENERGIES = (5000.0, 10000.0, 20000.0, 30000.0)
STRESS_DISTANCES = (100.0, 1000.0, 10000.0, 30000.0)


def batch_instances(n=64, seed=4, min_rays=16384, max_rays=2097152, min_side=64,
                    max_side=1024):
    """Uneven instances: ray count and screen side log-uniform, E ~ U(6, 25)
    keV, L log-uniform on [200, 20000] mm; each built like the point-source
    aperture (+-0.1 mm) with its own seed; screen half-width
    0.2 * 0.1 mm * (L + 10 m) / 10 m (20 um at 1 m)."""
    rng = SplitMix64(seed)
    u = rng.uniform(4 * n).reshape(n, 4)
    out = []
    for i in range(n):
        nr = int(round(math.exp(math.log(min_rays) + u[i, 0] * math.log(max_rays / min_rays))))
        side = int(round(math.exp(math.log(min_side) + u[i, 1] * math.log(max_side / min_side))))
        e = 6000.0 + u[i, 2] * 19000.0
        dist = math.exp(math.log(200.0) + u[i, 3] * math.log(20000.0 / 200.0))
        out.append({"n_rays": nr, "side": side, "energy": e, "distance": dist,
                    "seed": 1000 + i,
                    "half": 0.2 * 0.1 * (dist + SRC_DIST) / SRC_DIST})
    return out


def instance_rays(spec: dict):
    return point_source_aperture(spec["n_rays"], 0.1, spec["energy"], spec["seed"])


def instance_screen(spec: dict) -> Screen:
    return screen(spec["distance"], spec["half"], spec["side"])


def rel_l2(a: np.ndarray, ref: np.ndarray) -> float:
    """Complex relative L2 over [Es; Ep]."""
    num = np.linalg.norm((a - ref).reshape(-1))
    den = np.linalg.norm(ref.reshape(-1))
    return float(num / den) if den > 0 else float(num)


def intensity_error(a: np.ndarray, ref: np.ndarray):
    """(max, p99) of |I - I0| / max I0 over pixels with I0 >= 1e-3 max I0,
    I = |Es|^2 + |Ep|^2."""
    i0 = (np.abs(ref) ** 2).sum(axis=1)
    i1 = (np.abs(a) ** 2).sum(axis=1)
    m = i0.max() if i0.size else 0.0
    if m <= 0:
        return 0.0, 0.0
    sel = i0 >= 1e-3 * m
    e = np.abs(i1 - i0)[sel] / m
    return float(e.max()), float(np.percentile(e, 99))
