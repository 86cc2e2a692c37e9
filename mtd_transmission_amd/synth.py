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

Phase sign: the object wave carries exp(-i phi) with phi = k0 * sum dn * L
the projected delay. That is the reference's convention: its simulator's
Born field of a thin positive-contrast object is -i kv * integral(dn dz)
(calcHolo's factor -i kv / sdz, Simul_holo/src/fonctions.cpp:115-122,136,
the exact inverse of retroPropag_Born), so -i phi, and its Born formation
U_B = |A| cos(phi_c) - 1 + i |A| sin(phi_c) then needs phi_c = -phi.
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
    u = np.exp(1j * (2 * np.pi * (kx * x + ky * y) / n - phi))
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


# --------------------------------------------------------------------------
# GPU (torch) variant of the same generator, for full-size benchmark inputs
# --------------------------------------------------------------------------
def _torch_chords(P, d, phantom, torch):
    """Chord lengths for rays P(...,3) + t d(b,3): list of (b, ...) tensors."""
    out = []
    for e in phantom:
        M = torch.tensor(e.quad, dtype=P.dtype, device=P.device)
        c = torch.tensor(e.centre, dtype=P.dtype, device=P.device)
        q = P - c
        a = torch.einsum("bi,ij,bj->b", d, M, d)
        qM = q @ M
        b = torch.einsum("...i,bi->b...", qM, d)
        cc = (qM * q).sum(-1) - 1.0
        disc = b * b - a.view(-1, *([1] * cc.dim())) * cc
        out.append(2.0 * torch.sqrt(torch.clamp(disc, min=0.0)) /
                   a.view(-1, *([1] * cc.dim())))
    return out


def simulate_torch(n, nxmax, circle_cx, circle_cy, pix, lam, n0, kx, ky, phantom,
                   seed=0, photons=20000.0, ref_amp=1.0, device="cuda", chunk=8,
                   lowpass=True):
    """Same physics as ``simulate`` with Poisson noise, evaluated on the GPU
    with torch: returns a uint16 (A, N, N) device tensor and the uint8 mask."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    A = len(kx)
    out = torch.empty((A, n, n), dtype=torch.uint16, device=device)
    ar = torch.arange(n, device=device, dtype=torch.float32)
    yy, xx = torch.meshgrid(ar, ar, indexing="ij")
    P = torch.stack([(xx - n / 2) * pix, (yy - n / 2) * pix, torch.zeros_like(xx)], -1)
    f = torch.fft.fftfreq(n, device=device) * n
    pupil = ((f[None, :] ** 2 + f[:, None] ** 2) < nxmax * nxmax).to(torch.complex64)
    fcx, fcy = n // 2 - circle_cx, n // 2 - circle_cy
    full = (ref_amp + 1.0) ** 2
    kvac = 2 * math.pi / lam
    for a0 in range(0, A, chunk):
        a1 = min(A, a0 + chunk)
        kxs = torch.tensor(np.asarray(kx[a0:a1], dtype=np.float64), device=device)
        kys = torch.tensor(np.asarray(ky[a0:a1], dtype=np.float64), device=device)
        d = torch.tensor(_directions(np.asarray(kx[a0:a1], np.float64),
                                     np.asarray(ky[a0:a1], np.float64), n, pix, lam, n0),
                         dtype=torch.float32, device=device)
        phi = torch.zeros((a1 - a0, n, n), dtype=torch.float32, device=device)
        for e, L in zip(phantom, _torch_chords(P, d, phantom, torch)):
            phi += float(kvac * e.dn) * L
        ramp = 2 * math.pi * (kxs.view(-1, 1, 1).float() * xx + kys.view(-1, 1, 1).float() * yy) / n
        u = torch.polar(torch.ones_like(phi), torch.remainder(ramp, 2 * math.pi) - phi)
        if lowpass:
            u = torch.fft.ifft2(torch.fft.fft2(u) * pupil)
        psi = (math.pi * ((kxs - fcx) + (kys - fcy))).float().view(-1, 1, 1)
        rphase = torch.remainder(2 * math.pi * (fcx * xx + fcy * yy) / n, 2 * math.pi)
        r = torch.polar(torch.full_like(phi, ref_amp), rphase + psi)
        inten = (r + u).abs() ** 2
        v = torch.poisson(inten * (photons / full), generator=gen)
        out[a0:a1] = torch.clamp(v, 0, 65535).to(torch.int32).to(torch.uint16)
    # mask on the field grid, union over all illuminations
    C = 2 * nxmax
    arc = torch.arange(C, device=device, dtype=torch.float32)
    cy_, cx_ = torch.meshgrid(arc, arc, indexing="ij")
    Pm = torch.stack([(cx_ * n / C - n / 2) * pix, (cy_ * n / C - n / 2) * pix,
                      torch.zeros_like(cx_)], -1)
    inside = torch.zeros((C, C), dtype=torch.bool, device=device)
    dall = torch.tensor(_directions(np.asarray(kx, np.float64), np.asarray(ky, np.float64),
                                    n, pix, lam, n0), dtype=torch.float32, device=device)
    for b0 in range(0, A, 64):
        for L in _torch_chords(Pm, dall[b0:b0 + 64], phantom, torch):
            inside |= (L > 0).any(0)
    grown = inside.cpu().numpy()
    for _ in range(2):
        g = grown.copy()
        g[1:, :] |= grown[:-1, :]
        g[:-1, :] |= grown[1:, :]
        g[:, 1:] |= grown[:, :-1]
        g[:, :-1] |= grown[:, 1:]
        grown = g
    grown[:2, :] = grown[-2:, :] = True
    grown[:, :2] = grown[:, -2:] = True
    return out, np.where(grown, 0, 255).astype(np.uint8)


def voxelize_torch(phantom, V, voxel, device="cuda"):
    """Index contrast of the phantom on the reconstruction grid (V^3 float32
    torch tensor, [z][y][x]): voxel (x, y, z) sits at ((x - V/2) voxel,
    (y - V/2) voxel, (z - V/2) voxel), the frame of the reconstructed index
    volume (its lateral centre is the hologram centre, its height centre the
    z = 0 plane the fields are recorded in). Overlapping ellipsoids add, as
    in the projected phase. Sampled at voxel centres."""
    import torch
    ar = (torch.arange(V, device=device, dtype=torch.float64) - V // 2) * voxel
    out = torch.zeros((V, V, V), device=device, dtype=torch.float32)
    for z0 in range(0, V, 32):  # slabs bound the temporary memory
        z1 = min(V, z0 + 32)
        Z, Y, X = torch.meshgrid(ar[z0:z1], ar, ar, indexing="ij")
        P = torch.stack([X, Y, Z], -1)
        for e in phantom:
            M = torch.tensor(e.quad, device=device)
            q = P - torch.tensor(e.centre, device=device)
            out[z0:z1] += (((q @ M) * q).sum(-1) <= 1.0).float() * float(e.dn)
    return out
