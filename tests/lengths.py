"""Shared pieces of the transform-length parity tests: seeded acquisitions at
a given hologram side / crop / volume, the oracle's index volume at sampled
voxels, and the FP64 error bound.

The index at voxel x is the centred inverse transform of the count-normalised
spectrum (finalize, Reconstruction/main.cpp:243-352):
index(x) = cte * df^3 * sum_u t(u) exp(+2 pi i (u - V/2).(x - V/2) / V),
so for volumes too large for a dense oracle FFT the oracle evaluates that sum
directly over the hit voxels (ewald_scatter_sparse) at sampled x, in FP64.
"""
import math

import numpy as np

from cases import LAM, N0, NA, rel_l2
from oracle import ref


def acquisition(n, nx, V, A, seed, deroul=0):
    """Seeded holograms (torch, on the GPU) of a small ellipsoid phantom,
    Fermat-spiral illumination, with the geometry of the reference's
    conventions: pix = NXMAX lambda / (NA N), carrier at 0.6 Nyquist.
    deroul 1: Volkov unwrapping (DEROUL 1, VOLKOV 1) with Born fields."""
    import mtd_transmission_amd as m
    from mtd_transmission_amd import synth
    pix = nx * LAM / (NA * n)
    cx, cy = n // 2 - round(0.3 * n), n // 2 + round(0.3 * n)
    ph = synth.scale_phantom(synth.make_phantom(seed, 3), 0.45 * n * pix / 33.8e-6)
    kx, ky = synth.scan_fermat(A, nx, 0.95)
    holos, mask = synth.simulate_torch(n, nx, cx, cy, pix, LAM, N0, kx, ky, ph, seed=seed + 1)
    gl = m.geom_derive(n, nx, V, cx, cy, NA, N0, LAM, pix, 1.0, 1.0, 1.0, 1.0, 1, 1, deroul)
    go = ref.derive_geom(n, nx, V, cx, cy, NA, N0, LAM, pix, 1, 1, 1, deroul=deroul)
    return holos, mask, gl, go


def spectra(fields, pos, g, prec):
    A = fields.shape[0]
    window = ref.tukey2d(g.C, g.alpha_field)
    return np.stack([ref.field_to_spectrum(fields[a], int(pos[a]), int(pos[a + A]), g, prec,
                                           window) for a in range(A)])


def index_at(fields, pos, g, prec, flat_idx, chunk=32):
    """The oracle's index volume (FP64 sum over the hit voxels of the
    spectrum it accumulates in `prec`) at flat voxel indices."""
    V = g.dim_final
    vox, sums, cnt = ref.ewald_scatter_sparse(spectra(fields, pos, g, prec), pos, g, prec)
    rt = np.float32 if prec == "f32" else np.float64
    c = cnt.astype(rt)
    t = (sums.real.astype(rt) / c).astype(np.float64) + 1j * (sums.imag.astype(rt) / c)
    uz, rem = np.divmod(vox, V * V)
    uy, ux = np.divmod(rem, V)
    z, rem = np.divmod(np.asarray(flat_idx), V * V)
    y, x = np.divmod(rem, V)
    cen = np.arange(V) - V // 2
    out = np.empty(len(flat_idx), np.complex128)
    for s0 in range(0, len(flat_idx), chunk):
        sl = slice(s0, s0 + chunk)
        # exp(2 pi i (u - V/2)(x - V/2) / V) per axis: V x chunk tables
        ex = np.exp(2j * np.pi * np.outer(cen, cen[x[sl]]) / V)
        ey = np.exp(2j * np.pi * np.outer(cen, cen[y[sl]]) / V)
        ez = np.exp(2j * np.pi * np.outer(cen, cen[z[sl]]) / V)
        out[sl] = (t[:, None] * ex[ux] * ey[uy] * ez[uz]).sum(axis=0)
    kv = 2 * math.pi / g.lam
    return out * (-1.0 / (2 * g.n0 * kv * kv)) * g.delta_f ** 3


def fp64_bound(o32, o64):
    """The error allowed against the FP64 oracle: the FP32 oracle's own
    distance from it, twice, but at least the 1e-5 parity bar."""
    return max(1e-5, 2 * rel_l2(o32, o64))
