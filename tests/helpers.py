"""Shared input builders of the tests."""
import json
import os

import numpy as np

import xrtk

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = xrtk.RAY_FIELDS


def fixture_case():
    """small_deterministic of diffraction_ref.json: (rays dict, (px, py, pz),
    expected (3, 2) complex)."""
    with open(os.path.join(GOLDEN, "diffraction_ref.json")) as f:
        tc = json.load(f)["test_cases"][0]
    rays = {k: np.asarray(tc["rays"][k], dtype=np.float64) for k in FIELDS}
    pix = tuple(np.asarray(tc["pixels"][k], dtype=np.float64) for k in ("x", "y", "z"))
    exp = np.array([[complex(e["es_re"], e["es_im"]), complex(e["ep_re"], e["ep_im"])]
                    for e in tc["expected"]])
    return rays, pix, exp


def mp_cases():
    with open(os.path.join(GOLDEN, "mp_small.json")) as f:
        d = json.load(f)
    out = []
    for c in d["cases"]:
        rays = {k: np.asarray(c["rays"][k], dtype=np.float64) for k in FIELDS}
        pix = tuple(np.asarray(c["pixels"][k], dtype=np.float64) for k in ("x", "y", "z"))
        exp = np.array([[complex(e[0], e[1]), complex(e[2], e[3])] for e in c["expected"]])
        out.append((c["id"], rays, pix, exp))
    return out


def to_gpu_rays(rays: dict) -> np.ndarray:
    n = rays["x"].shape[0]
    r = xrtk.gpu_rays(n)
    for f in FIELDS:
        r[f] = rays[f]
    return r


def to_gpu_pixels(pix) -> np.ndarray:
    return xrtk.gpu_pixels(*pix)


def ray(x=0.0, y=0.0, z=0.0, nx=0.0, ny=0.0, nz=1.0, nl=1.0, es=1 + 0j, ep=0j, energy=10000.0):
    return dict(x=x, y=y, z=z, nx=nx, ny=ny, nz=nz, nl=nl, es_re=es.real, es_im=es.imag,
                ep_re=ep.real, ep_im=ep.imag, energy=energy)


def rays_of(lst) -> dict:
    return {f: np.array([r[f] for r in lst], dtype=np.float64) for f in FIELDS}


def f32_exact(rays: dict) -> dict:
    """Rays rounded to float32 (the GpuRay inputs), as float64."""
    return {k: v.astype(np.float32).astype(np.float64) for k, v in rays.items()}
