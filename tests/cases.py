"""Seeded synthetic cases shared by the tests (geometry + holograms + mask).

Geometry follows the conventions the reference assumes: pixel size
pix = NXMAX*lambda/(NA*N) so that rayon = round(NXMAX*n0/NA) is the Ewald
radius, carrier at 0.6 of Nyquist per axis in the reference's quadrant
(CIRCLE_CX = N/2 - round(0.3N), CIRCLE_CY = N/2 + round(0.3N)), V = 4*NXMAX.
"""
from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

from mtd_transmission_amd import synth
from oracle import ref

LAM, N0, NA = 6.33e-7, 1.333, 1.2


@dataclass
class Case:
    name: str
    g: "ref.Geom"            # oracle geometry
    acq: "synth.Acquisition"
    params: dict

    def lib_geom(self):
        import mtd_transmission_amd as m
        p = self.params
        gl = m.geom_derive(p["n"], p["nxmax"], p["v"], p["cx"], p["cy"], NA, N0, LAM,
                           p["pix"], 1.0, 1.0, 1.0, p.get("coef", 1.0), p.get("born", 1),
                           p.get("c_aber", 1), p.get("deroul", 0))
        return gl


SPECS = {
    # name: N, NXMAX, V, A, n_ellipsoids, noise, dtype, scan
    # tiny: C=16 is too small for a meaningful degree-4 fit (16 samples for
    # 15 unknowns); it exercises the FFT, scatter and volume stages only.
    "tiny": (64, 8, 32, 6, 2, "gauss", np.float32, "circle"),
    # nomask: no Image_mask.pgm (all pixels background, every 5th sampled)
    "nomask": (128, 16, 64, 10, 2, "gauss", np.float32, "circle"),
    "small": (128, 16, 64, 16, 3, "gauss", np.float32, "circle"),
    "u16": (128, 16, 64, 12, 3, "poisson", np.uint16, "fermat"),
    "cfg1": (512, 64, 256, 100, 3, "gauss", np.float32, "circle"),
}


@functools.lru_cache(maxsize=None)
def make_case(name, born=1, c_aber=1, seed=7, deroul=0):
    n, nx, v, A, n_ell, noise, dtype, scan = SPECS[name]
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    g = ref.derive_geom(n, nx, v, cx, cy, NA, N0, LAM, pix, 1, 1, 1, born=born,
                        c_aber=c_aber, deroul=deroul)
    fov = n * pix
    # small fields of view get a proportionally smaller phantom so the mask
    # keeps enough background samples for the degree-4 fit
    shrink = 1.0 if n >= 512 else 0.45
    ph = synth.scale_phantom(synth.make_phantom(seed, n_ell), shrink * fov / 33.8e-6)
    if scan == "circle":
        kx, ky = synth.scan_circle(A, nx, 0.9)
    else:
        kx, ky = synth.scan_fermat(A, nx, 0.95)
    acq = synth.simulate(n, nx, cx, cy, g.pix_holo, g.lam, g.n0, kx, ky, ph,
                         seed=seed + 1, noise=noise, dtype=dtype)
    if name in ("tiny", "nomask"):
        acq.mask = None  # the absent-mask path (Correction_aberration.cpp:48-52)
    params = dict(n=n, nxmax=nx, v=v, cx=cx, cy=cy, pix=pix, born=born, c_aber=c_aber,
                  deroul=deroul)
    return Case(name, g, acq, params)


def rel_l2(a, b):
    a = np.asarray(a).astype(np.complex128).ravel()
    b = np.asarray(b).astype(np.complex128).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
