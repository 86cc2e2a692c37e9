"""The reference CPU path timed as a baseline: the oracle's FP64 arithmetic
(the reference computes in FP64) with the reference's stage boundaries and
threading, for bench.py's ``cpu_baseline`` leg.

TEST INFRASTRUCTURE. Only ``bench.py``'s ``cpu_baseline`` leg and ``tests/``
use this module; it is a timed restatement, never the product path.

Stages and the reference lines they time:

* T_pre1 -- hologram window, centred c2c FFT / N^2, crop
  (pretraitement_hors_axe_CPU/main.cpp:122-169, holo2TF_UBorn);
* T_pre2 -- per-angle specular, demodulation, phase, least-squares
  backgrounds and Born field (main.cpp:226-346);
* T_rec1 -- per-angle window + FFT and the Ewald-cap accumulation
  (Reconstruction/main.cpp:200-229, retroPropag_Born);
* T_rec2 -- count normalisation, fftshift3D, 3-D inverse FFT, index
  (Reconstruction/main.cpp:243-352).

As in the reference, the host loops run serially and only the FFTs use
threads (``NB_THREADS``): FFTW when the host has ``libfftw3.so.3`` and
``libfftw3_threads.so.3`` (loaded with ctypes; the reference's planner
flags: MEASURE for the hologram plan, PATIENT for the crop-sized plans,
ESTIMATE for the volume), else scipy.fft (pocketfft) with that many workers.
Every FFT is FP64 complex, as the reference's.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math
import time

import numpy as np

from oracle import ref

FFTW_FORWARD, FFTW_BACKWARD = -1, 1
FFTW_MEASURE, FFTW_PATIENT, FFTW_ESTIMATE = 0, 1 << 5, 1 << 6


class _Fftw:
    """Minimal ctypes binding of FFTW's double-precision c2c plans on numpy
    complex128 arrays (fftw_complex is two doubles)."""

    def __init__(self, threads):
        name = ctypes.util.find_library("fftw3")
        tname = ctypes.util.find_library("fftw3_threads")
        if not name or not tname:
            raise OSError("libfftw3 / libfftw3_threads not found")
        self.lib = ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
        self.tlib = ctypes.CDLL(tname)
        self.tlib.fftw_init_threads.restype = ctypes.c_int
        if not self.tlib.fftw_init_threads():
            raise OSError("fftw_init_threads failed")
        self.tlib.fftw_plan_with_nthreads.argtypes = [ctypes.c_int]
        self.tlib.fftw_plan_with_nthreads(int(threads))
        L = self.lib
        P = ctypes.c_void_p
        L.fftw_plan_dft_2d.restype = P
        L.fftw_plan_dft_2d.argtypes = [ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int,
                                       ctypes.c_uint]
        L.fftw_plan_dft_3d.restype = P
        L.fftw_plan_dft_3d.argtypes = [ctypes.c_int] * 3 + [P, P, ctypes.c_int, ctypes.c_uint]
        L.fftw_execute_dft.argtypes = [P, P, P]
        L.fftw_destroy_plan.argtypes = [P]
        self.plans = {}
        self.name = name

    def run(self, x, sign, flags):
        """Unnormalised c2c transform of a C-contiguous complex128 array
        (2-D or 3-D), out of place, with a cached plan per shape and sign."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        key = (x.shape, sign)
        if key not in self.plans:
            # planning may overwrite its arrays: plan on scratch copies
            a, b = np.empty_like(x), np.empty_like(x)
            pa, pb = a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p)
            if x.ndim == 2:
                p = self.lib.fftw_plan_dft_2d(x.shape[0], x.shape[1], pa, pb, sign, flags)
            else:
                p = self.lib.fftw_plan_dft_3d(*x.shape, pa, pb, sign, flags)
            if not p:
                raise OSError("FFTW planning failed")
            self.plans[key] = (p, a, b)
        p = self.plans[key][0]
        self.lib.fftw_execute_dft(p, x.ctypes.data_as(ctypes.c_void_p),
                                  out.ctypes.data_as(ctypes.c_void_p))
        return out


class Backend:
    """The FFTs of the baseline: FFTW with `threads` threads, else scipy.fft
    with `threads` workers. `name` says which."""

    def __init__(self, threads):
        self.threads = int(threads)
        self.fftw = None
        try:
            self.fftw = _Fftw(self.threads)
            self.name = f"FFTW ({self.fftw.name}), {self.threads} threads"
        except OSError:
            self.name = (f"FFTW unavailable on this host -- scipy.fft (pocketfft) FP64 in "
                         f"its place, {self.threads} workers")

    def fft2(self, x, flags=FFTW_MEASURE):
        if self.fftw:
            return self.fftw.run(x, FFTW_FORWARD, flags)
        import scipy.fft as sf
        return sf.fft2(x, workers=self.threads)

    def ifft2_unnorm(self, x, flags=FFTW_PATIENT):
        if self.fftw:
            return self.fftw.run(x, FFTW_BACKWARD, flags)
        import scipy.fft as sf
        return sf.ifft2(x, norm="forward", workers=self.threads)

    def ifftn_unnorm(self, x, flags=FFTW_ESTIMATE):
        if self.fftw:
            return self.fftw.run(x, FFTW_BACKWARD, flags)
        import scipy.fft as sf
        return sf.ifftn(x, norm="forward", workers=self.threads)


def run(holos, g: "ref.Geom", mask, threads):
    """The plain method (BORN 1 / C_ABER 1 / DEROUL 0) over every hologram,
    FP64. Returns (index volume, posspec, seconds per stage)."""
    be = Backend(threads)
    t = {}
    A, N, C = len(holos), g.n, g.C
    x0, y0 = g.coin
    win = ref.tukey2d(N, g.alpha_holo)
    # T_pre1: holo2TF_UBorn per hologram (FP64, c2c, /N^2, both fftshifts)
    t0 = time.perf_counter()
    crops = np.empty((A, C, C), np.complex128)
    for a in range(A):
        h = np.asarray(holos[a], np.float64) * win
        spec = be.fft2(ref.fftshift2(h).astype(np.complex128)) / float(N * N)
        crops[a] = ref.fftshift2(spec)[y0:y0 + C, x0:x0 + C]
    t["T_pre1"] = time.perf_counter() - t0
    # T_pre2: specular, demodulation, fits, Born field per angle
    t0 = time.perf_counter()
    plan = ref.LsqPlan.build(ref.load_mask(C, mask))
    fields = np.empty((A, C, C), np.complex128)
    pos = np.empty(2 * A, np.int64)
    for a in range(A):
        i = ref.coord_spec(crops[a], "f64")
        px, py = i % C, i // C
        pos[a], pos[a + A] = px, py
        U = ref.decal2d(be.ifft2_unnorm(ref.decal2d(crops[a], -px, -py)), C // 2, C // 2)
        fields[a] = ref.form_field(U, plan, g, "f64", spec_value=crops[a][py, px])
    t["T_pre2"] = time.perf_counter() - t0
    # T_rec1: per-angle window + FFT * pixUBorn^2, +posSpec shift, scatter
    t0 = time.perf_counter()
    w2 = ref.tukey2d(C, g.alpha_field)
    S = np.empty((A, C, C), np.complex128)
    for a in range(A):
        ub = ref.decal2d(fields[a] * w2, g.nxmax, g.nxmax)
        S[a] = ref.decal2d(be.fft2(ub, FFTW_MEASURE) * g.pix_uborn ** 2, int(pos[a]),
                           int(pos[a + A]))
    vol, cnt = ref.ewald_scatter(S, pos, g, "f64")
    t["T_rec1"] = time.perf_counter() - t0
    del S
    # T_rec2: normalise, fftshift3D, 3-D inverse FFT * delta_f^3, index
    t0 = time.perf_counter()
    V = g.dim_final
    c = np.where(cnt == 0, 1, cnt).astype(np.float64)
    tv = ref.fftshift3((vol / c).reshape(V, V, V))
    del vol, c
    p = ref.fftshift3(be.ifftn_unnorm(tv) * g.delta_f ** 3)
    kv = 2 * math.pi / g.lam
    index = p * (-1.0 / (2 * g.n0 * kv * kv))
    t["T_rec2"] = time.perf_counter() - t0
    return index, pos, t, be.name
