"""Seeded synthetic acquisitions: ellipsoid phantoms, illumination scans and
off-axis holograms, plus the aberration mask ``Image_mask.pgm``.

This is input generation for tests and benchmarks (the reference ships no
sample data). Physics: each hologram records |R + U|^2 where R is the tilted
reference wave that puts the +1 order at (CIRCLE_CX, CIRCLE_CY) of the centred
spectrum, and U is the illumination plane wave times the projection-model
transmission of the phantom along the illumination direction, low-passed by
the objective pupil (|f| < NXMAX bins). Illumination frequencies are integer
bins, so every specular lands on a pixel of the crop.

Units: metres for lengths; pixel (x, y) of an N x N hologram sits at
((x - N/2) * pix, (y - N/2) * pix); the phantom is centred on the field of
view at z = 0.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Ellipsoid:
    centre: np.ndarray      # (3,) metres
    axes: np.ndarray        # (3,) semi-axes, metres
    rot: np.ndarray         # (3,3) rotation, columns = ellipsoid axes
    dn: float               # index contrast n - n0

    @property
    def quad(self):
        """M such that (p-c)^T M (p-c) <= 1 inside."""
        return self.rot @ np.diag(1.0 / self.axes ** 2) @ self.rot.T


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def make_phantom(seed, n_ell, n0=1.333, n_range=(1.36, 1.42),
                 axes_um=(2.0, 6.0), lateral_um=6.0, axial_um=3.0):
    """A seeded phantom of ``n_ell`` ellipsoids (index in ``n_range``,
    semi-axes uniform in ``axes_um``), centres within +-lateral/axial."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_ell):
        c = np.array([rng.uniform(-lateral_um, lateral_um),
                      rng.uniform(-lateral_um, lateral_um),
                      rng.uniform(-axial_um, axial_um)]) * 1e-6
        ax = rng.uniform(axes_um[0], axes_um[1], size=3) * 1e-6
        out.append(Ellipsoid(c, ax, random_rotation(rng),
                             float(rng.uniform(*n_range) - n0)))
    return out


def scale_phantom(phantom, factor):
    """Shrink or grow every ellipsoid (centres and axes) by ``factor``."""
    return [Ellipsoid(e.centre * factor, e.axes * factor, e.rot, e.dn)
            for e in phantom]


def scan_circle(A, nxmax, frac=0.9):
    """Speculars equispaced on a circle of radius frac*NXMAX (integer bins)."""
    t = 2 * np.pi * np.arange(A) / A
    kx = np.rint(frac * nxmax * np.cos(t)).astype(np.int64)
    ky = np.rint(frac * nxmax * np.sin(t)).astype(np.int64)
    return kx, ky


def scan_fermat(A, nxmax, frac=0.95):
    """Fermat spiral r_j = frac*NXMAX*sqrt(j/A), theta_j = j*pi*(3-sqrt 5)
    (the pattern of Simul_holo/src/OTF.cpp:97-114), rounded to bins."""
    j = np.arange(A)
    r = frac * nxmax * np.sqrt(j / A)
    th = j * np.pi * (3 - math.sqrt(5))
    kx = np.rint(r * np.cos(th)).astype(np.int64)
    ky = np.rint(r * np.sin(th)).astype(np.int64)
    return kx, ky


def _directions(kx, ky, n, pix, lam, n0):
    sx = kx * lam / (n * pix * n0)
    sy = ky * lam / (n * pix * n0)
    sz = np.sqrt(np.maximum(1 - sx * sx - sy * sy, 0.0))
    return np.stack([sx, sy, sz], axis=-1)


def _chords(X, Y, d, phantom):
    """Per-ellipsoid chord lengths of the rays (X, Y, 0) + t d."""
    P = np.stack([X, Y, np.zeros_like(X)], axis=-1)
    out = []
    for e in phantom:
        M = e.quad
        q = P - e.centre
        a = d @ M @ d
        b = (q @ M) @ d
        c = np.einsum("...i,ij,...j->...", q, M, q) - 1.0
        disc = b * b - a * c
        out.append(2.0 * np.sqrt(np.maximum(disc, 0.0)) / a)
    return out


def projected_phase(n, pix, lam, n0, kx, ky, phantom, coords=None):
    """k_vac * sum_k dn_k * chord_k on the hologram grid (or ``coords``)."""
    if coords is None:
        y, x = np.mgrid[0:n, 0:n]
        X, Y = (x - n / 2) * pix, (y - n / 2) * pix
    else:
        X, Y = coords
    d = _directions(np.float64(kx), np.float64(ky), n, pix, lam, n0)
    k = 2 * np.pi / lam
    phi = np.zeros(X.shape)
    for e, L in zip(phantom, _chords(X, Y, d, phantom)):
        phi += k * e.dn * L
    return phi


def pupil_lowpass(u, nxmax):
    """Keep spatial frequencies with |f| < NXMAX bins (the objective pupil)."""
    n = u.shape[-1]
    f = np.fft.fftfreq(n) * n
    p = (f[None, :] ** 2 + f[:, None] ** 2) < nxmax * nxmax
    return np.fft.ifft2(np.fft.fft2(u) * p)


def reference_wave(n, circle_cx, circle_cy, kx, ky, amp):
    """Reference wave with the carrier that puts the +1 order (R*U) of the
    specular (kx, ky) at (circle_cx + kx, circle_cy + ky) of the centred
    spectrum, phased so the specular value of the crop is real positive."""
    fcx, fcy = n // 2 - circle_cx, n // 2 - circle_cy
    psi = math.pi * ((kx - fcx) + (ky - fcy))
    y, x = np.mgrid[0:n, 0:n]
    return amp * np.exp(1j * (2 * np.pi * (fcx * x + fcy * y) / n + psi))


def hologram(n, nxmax, circle_cx, circle_cy, pix, lam, n0, kx, ky, phantom,
             ref_amp=1.0, lowpass=True):
    """Noise-free intensity |R + U|^2 for one illumination."""
    y, x = np.mgrid[0:n, 0:n]
    phi = projected_phase(n, pix, lam, n0, kx, ky, phantom)
    u = np.exp(1j * (2 * np.pi * (kx * x + ky * y) / n + phi))
    if lowpass:
        u = pupil_lowpass(u, nxmax)
    r = reference_wave(n, circle_cx, circle_cy, kx, ky, ref_amp)
    return np.abs(r + u) ** 2


def support_mask(n, nxmax, pix, lam, n0, kxs, kys, phantom, dilate=2,
                 border=2):
    """``Image_mask.pgm``: 0 where any illumination's projected support of the
    phantom falls (field-frame pixel (x, y) <-> hologram pixel
    (x*N/C, y*N/C)), dilated by ``dilate`` pixels, and on a ``border``-pixel
    frame (where the hologram window tapers the field); 255 elsewhere."""
    C = 2 * nxmax
    y, x = np.mgrid[0:C, 0:C]
    X = (x * n / C - n / 2) * pix
    Y = (y * n / C - n / 2) * pix
    inside = np.zeros((C, C), dtype=bool)
    for kx, ky in zip(kxs, kys):
        d = _directions(np.float64(kx), np.float64(ky), n, pix, lam, n0)
        for L in _chords(X, Y, d, phantom):
            inside |= L > 0
    grown = inside.copy()
    for _ in range(dilate):
        g = grown.copy()
        g[1:, :] |= grown[:-1, :]
        g[:-1, :] |= grown[1:, :]
        g[:, 1:] |= grown[:, :-1]
        g[:, :-1] |= grown[:, 1:]
        grown = g
    if border:
        grown[:border, :] = grown[-border:, :] = True
        grown[:, :border] = grown[:, -border:] = True
    return np.where(grown, 0, 255).astype(np.uint8)


@dataclass
class Acquisition:
    holos: np.ndarray       # (A, N, N) float32 or uint16
    mask: np.ndarray        # (C, C) uint8
    kx: np.ndarray
    ky: np.ndarray
    phantom: list


def simulate(n, nxmax, circle_cx, circle_cy, pix, lam, n0, kx, ky, phantom,
             seed=0, noise="gauss", sigma=8.0, peak=4095.0, photons=20000.0,
             ref_amp=1.0, dtype=np.float32, lowpass=True, blank=False):
    """Seeded hologram stack.

    noise="gauss": 12-bit camera, intensity scaled so the brightest fringe is
    ~0.8*peak counts, Gaussian read noise ``sigma``, rounded, clipped to
    [0, peak]. noise="poisson": shot noise with ``photons`` at the brightest
    fringe, clipped to uint16. noise=None: scaled, no noise. ``blank``
    simulates an empty field (no phantom)."""
    rng = np.random.default_rng(seed)
    A = len(kx)
    out = np.empty((A, n, n), dtype=dtype)
    full = (ref_amp + 1.0) ** 2
    ph = [] if blank else phantom
    for a in range(A):
        I = hologram(n, nxmax, circle_cx, circle_cy, pix, lam, n0,
                     int(kx[a]), int(ky[a]), ph, ref_amp, lowpass)
        if noise == "poisson":
            v = rng.poisson(I * (photons / full))
            out[a] = np.clip(v, 0, 65535).astype(dtype)
        elif noise == "gauss":
            v = I * (0.8 * peak / full) + rng.normal(0.0, sigma, size=I.shape)
            out[a] = np.clip(np.rint(v), 0, peak).astype(dtype)
        else:
            out[a] = (I * (0.8 * peak / full)).astype(dtype)
    mask = support_mask(n, nxmax, pix, lam, n0, kx, ky, phantom)
    return Acquisition(out, mask, np.asarray(kx), np.asarray(ky), phantom)
