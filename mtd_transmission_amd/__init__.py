"""MI355X-native off-axis preprocessing and Ewald-cap Fourier reconstruction.

Python binding of the C ABI in ``include/mtd.h`` (ctypes over the in-tree
``libmtd.so``). Device arrays are torch CUDA tensors (HIP on ROCm); the
library launches on torch's current stream. There is no CPU fallback: every
compute call goes through the gfx950 kernels and fails loudly when the
library or the GPU is missing.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmtd.so")
BIN_DIR = os.path.join(_HERE, "bin")

MTD_U8, MTD_U16, MTD_F32 = 0, 1, 2
MTD_BATCH_GRAPH, MTD_BATCH_AUTOTUNE = 1, 2
MTD_DEROUL_OFF, MTD_DEROUL_VOLKOV, MTD_DEROUL_HERRAEZ = 0, 1, 2


class MTDError(RuntimeError):
    pass


class Geom(ctypes.Structure):
    """mtd_geom (include/mtd.h)."""
    _fields_ = [
        ("n", ctypes.c_int32), ("nxmax", ctypes.c_int32),
        ("dim_final", ctypes.c_int32), ("circle_cx", ctypes.c_int32),
        ("circle_cy", ctypes.c_int32), ("born", ctypes.c_int32),
        ("c_aber", ctypes.c_int32), ("deroul", ctypes.c_int32),
        ("n0", ctypes.c_double), ("na", ctypes.c_double),
        ("lambda_", ctypes.c_double), ("pix_holo", ctypes.c_double),
        ("rayon", ctypes.c_double), ("pix_uborn", ctypes.c_double),
        ("delta_f", ctypes.c_double), ("coef_na", ctypes.c_double),
        ("alpha_holo", ctypes.c_double), ("alpha_field", ctypes.c_double),
        ("alpha_crop", ctypes.c_double),
    ]

    @property
    def C(self):
        return 2 * self.nxmax

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# name -> (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_GP = ctypes.POINTER(Geom)
SIGNATURES = {
    "mtd_version": (ctypes.c_char_p, []),
    "mtd_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "mtd_ctx_destroy": (_I, [_P]),
    "mtd_ctx_set_stream": (_I, [_P, _P]),
    "mtd_ctx_sync": (_I, [_P]),
    "mtd_last_error": (ctypes.c_char_p, [_P]),
    "mtd_ctx_set_timing": (_I, [_P, _I]),
    "mtd_ctx_set_timing_only": (_I, [_P, ctypes.c_char_p]),
    "mtd_ctx_set_schedule": (_I, [_P, _I, _I]),
    "mtd_ctx_set_schedule_param": (_I, [_P, ctypes.c_char_p, _I]),
    "mtd_ctx_get_schedule_param": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_I)]),
    "mtd_ctx_timing_report": (_I, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "mtd_ctx_timing_reset": (_I, [_P]),
    "mtd_config_extract_val": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, _D,
                                    ctypes.POINTER(ctypes.c_float)]),
    "mtd_config_extract_string": (_I, [ctypes.c_char_p, ctypes.c_char_p,
                                       ctypes.c_char_p, ctypes.c_size_t]),
    "mtd_geom_derive": (_I, [_D] * 13 + [_I, _I, _I, _GP]),
    "mtd_tukey1d": (_I, [_I, _D, ctypes.POINTER(_D)]),
    "mtd_write_tiff_f32": (_I, [ctypes.c_char_p, _P, _I, _I, _I, _D, _I, _I]),
    "mtd_holo_to_crop": (_I, [_P, _GP, _P, _I, _I, _P]),
    "mtd_holo_to_crop_noalloc": (_I, [_P, _GP, _P, _I, _I, _P]),
    "mtd_crop_to_field": (_I, [_P, _GP, _P, _I, _P, _P, _P]),
    "mtd_preprocess": (_I, [_P, _GP, _P, _I, _I, _P, _P, _P]),
    "mtd_field_to_spectrum": (_I, [_P, _GP, _P, _P, _I, _P]),
    "mtd_ewald_scatter": (_I, [_P, _GP, _P, _P, _I, _P, _P]),
    "mtd_ewald_accumulate": (_I, [_P, _GP, _P, _P, _I, _P, _P]),
    "mtd_finalize": (_I, [_P, _GP, _P, _P, _P, _P, _P]),
    "mtd_reconstruct_plain": (_I, [_P, _GP, _P, _I, _I, _P, _P, _P, _P]),
    "mtd_reconstruct_batch": (_I, [_P, _I, _GP, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
    "mtd_batch_tuned_chunk": (_I, [_P, _I, _I, _I]),
    "mtd_batch_tuned_window": (_I, [_P, _I, _I, _I]),
    "mtd_batch_tuned_schedule": (_I, [_P, _I, _I, _I, _I, ctypes.POINTER(_I)]),
    "mtd_ctx_set_tune_cache": (_I, [_P, ctypes.c_char_p]),
    "mtd_demodulate": (_I, [_P, _P, _P, _I, _P, _P]),
    "mtd_forward_spectra": (_I, [_P, _GP, _P, _P, _P, _I, _P]),
    "mtd_adjoint_spectra": (_I, [_P, _GP, _P, _P, _I, _P, _P]),
    "mtd_cap_weights": (_I, [_P, _GP, _P, _I, _P]),
    "mtd_born_plan": (_I, [_P, _GP, _P, _I]),
    "mtd_spectra_to_fields": (_I, [_P, _GP, _P, _P, _I, _P]),
    "mtd_tv_fista_step": (_I, [_P, _I, _P, _P, _P, _P, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_float, _I]),
    "mtd_tv_solve": (_I, [_P, _GP, _P, _P, _I, _I, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, _I, _P, _I]),
}


def lib():
    """Load libmtd.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MTDError(f"{LIB_PATH} is missing: build it first "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
        # torch's ROCm wheel carries its own libamdhip64 and asks for it as
        # "libamdhip64.so", while libmtd.so asks for the soname
        # "libamdhip64.so.7". Loading libmtd.so first would bring the system
        # runtime in and torch would then load a second HIP/HSA runtime
        # beside it (hipSetDevice then fails in one of them). Importing torch
        # first makes libmtd.so bind to the runtime torch already holds.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# host-only helpers
# --------------------------------------------------------------------------
def geom_derive(ccd_roix, nxmax, dim_final, circle_cx, circle_cy, na, n0, lam,
                tpcam, f_tube=1.0, f_obj=1.0, rf=1.0, coef_na=1.0, born=1,
                c_aber=1, deroul=0) -> Geom:
    g = Geom()
    rc = lib().mtd_geom_derive(ccd_roix, nxmax, dim_final, circle_cx, circle_cy,
                               na, n0, lam, tpcam, f_tube, f_obj, rf, coef_na,
                               born, c_aber, deroul, ctypes.byref(g))
    if rc:
        raise MTDError(f"mtd_geom_derive failed ({rc})")
    return g


def extract_val(path, token, default=None):
    """Returns (value, found) with the reference's float semantics."""
    out = ctypes.c_float()
    rc = lib().mtd_config_extract_val(str(path).encode(), token.encode(),
                                      int(default is not None),
                                      float(default or 0.0), ctypes.byref(out))
    if rc < 0:
        raise MTDError(f"cannot read {path}")
    return float(out.value), bool(rc)


def extract_string(path, token):
    buf = ctypes.create_string_buffer(4096)
    rc = lib().mtd_config_extract_string(str(path).encode(), token.encode(), buf, 4096)
    if rc < 0:
        raise MTDError(f"cannot read {path}")
    return buf.value.decode()


def tukey1d(n, alpha):
    arr = (ctypes.c_double * n)()
    if lib().mtd_tukey1d(n, alpha, arr):
        raise MTDError("mtd_tukey1d failed")
    return [arr[i] for i in range(n)]


def write_tiff_f32(path, pages, pixel_m=1.0, orientation=4, fmt=0):
    """pages: float32 (pages, h, w) or (h, w) host array -> TIFF file (fmt 0:
    classic, BigTIFF past 4 GiB; 1: classic only; 2: BigTIFF). Returns
    "tiff" or "bigtiff"."""
    import numpy as np
    a = np.ascontiguousarray(pages, dtype=np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise MTDError("pages must be (pages, h, w)")
    rc = lib().mtd_write_tiff_f32(str(path).encode(), a.ctypes.data_as(ctypes.c_void_p),
                                  a.shape[2], a.shape[1], a.shape[0], float(pixel_m),
                                  int(orientation), int(fmt))
    if rc == -2:
        raise MTDError("classic TIFF asked for, but the file would pass 4 GiB")
    if rc < 0:
        raise MTDError(f"cannot write {path}")
    return "bigtiff" if rc == 1 else "tiff"


# --------------------------------------------------------------------------
# device context
# --------------------------------------------------------------------------
def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise MTDError("device arrays must be GPU tensors")
    if not t.is_contiguous():
        raise MTDError("device arrays must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


# element kinds of the C ABI's arrays: accepted torch dtypes and the scalars
# of that dtype per element
_KINDS = {
    "c32": (("complex64", 1), ("float32", 2)),  # interleaved (re, im) pairs
    "f32": (("float32", 1),),
    "i32": (("int32", 1),),
    "u8": (("uint8", 1), ("bool", 1)),
}


def _dev(t, kind, count, name, optional=False):
    """Pointer of a device array after checking its dtype and that it holds
    at least `count` elements of `kind` (the library reads and writes that
    many without further checks)."""
    if t is None:
        if optional:
            return None
        raise MTDError(f"{name}: array required")
    dt = str(t.dtype).replace("torch.", "")
    per = dict(_KINDS[kind]).get(dt)
    if per is None:
        raise MTDError(f"{name}: dtype {dt} not accepted (expects {kind}: "
                       f"{', '.join(d for d, _ in _KINDS[kind])})")
    if t.numel() < per * count:
        raise MTDError(f"{name}: {t.numel()} {dt} values, needs {per * count}")
    return _ptr(t)


def _holo_count(holos, g):
    """Angles of a hologram stack [A][N][N] (numel checked against N)."""
    A = holos.shape[0] if holos.dim() > 0 else 0
    if holos.numel() < A * g.n * g.n:
        raise MTDError(f"holograms: {holos.numel()} values for {A} x {g.n}^2")
    return A


def _ptr_holos(t):
    """Hologram stacks may be GPU tensors or host (pinned or pageable) CPU
    tensors: the library streams host stacks to the GPU itself."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise MTDError("hologram arrays must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _dtype_code(t):
    import torch
    if t.dtype == torch.uint8:
        return MTD_U8
    if t.dtype in (torch.uint16, torch.int16):
        return MTD_U16
    if t.dtype == torch.float32:
        return MTD_F32
    raise MTDError(f"unsupported hologram dtype {t.dtype}")


def _mask_ptr(mask, C):
    if mask is None:
        return None, None
    import numpy as np
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if m.shape != (C, C):
        raise MTDError(f"mask must be {C}x{C}, got {m.shape}")
    return m, m.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One mtd_ctx bound to a device; launches on torch's current stream."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise MTDError("no GPU visible: the MI355X path has no CPU fallback")
        self.device = device
        h = ctypes.c_void_p()
        torch.cuda.set_device(device)
        rc = lib().mtd_ctx_create(device, ctypes.byref(h))
        if rc:
            msg = lib().mtd_last_error(None).decode()
            raise MTDError(f"mtd_ctx_create failed: {msg} (code {rc})")
        self.h = h

    def close(self):
        if self.h:
            lib().mtd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        lib().mtd_ctx_set_stream(self.h, ctypes.c_void_p(s))

    def _check(self, rc, what):
        if rc:
            msg = lib().mtd_last_error(self.h).decode()
            raise MTDError(f"{what}: {msg} (code {rc})")

    def sync(self):
        self._check(lib().mtd_ctx_sync(self.h), "sync")

    def set_timing(self, on=True, only=None):
        """Per-kernel event timing; only: time just the kernel of that report
        name (the others then run without event markers between them)."""
        lib().mtd_ctx_set_timing_only(self.h, (only or "").encode())
        lib().mtd_ctx_set_timing(self.h, int(bool(on)))

    def set_schedule(self, holo_chunk=0, gather_window=0):
        """Launch schedule: holograms per hologram launch pair and the gather's
        window height (0: defaults); results do not depend on it."""
        self._check(lib().mtd_ctx_set_schedule(self.h, int(holo_chunk), int(gather_window)),
                    "set_schedule")

    def set_schedule_param(self, name, value):
        """One named schedule parameter ("holo_chunk", "gather_window",
        "gather_batch", "vol_rows_tile", "vol_cols_tile"; 0: its default);
        results do not depend on it."""
        self._check(lib().mtd_ctx_set_schedule_param(self.h, name.encode(), int(value)),
                    "set_schedule_param")

    def schedule_param(self, name):
        v = _I(0)
        self._check(lib().mtd_ctx_get_schedule_param(self.h, name.encode(), ctypes.byref(v)),
                    "schedule_param")
        return v.value

    def timing_reset(self):
        lib().mtd_ctx_timing_reset(self.h)

    def timing(self):
        """{kernel: (launches, total_ms)} recorded since the last reset."""
        buf = ctypes.create_string_buffer(1 << 16)
        lib().mtd_ctx_timing_report(self.h, buf, 1 << 16)
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    # ---- stages ---------------------------------------------------------
    # Every wrapper checks dtypes and sizes against the geometry before the
    # call: the library trusts the pointers it is given.
    def holo_to_crop(self, g: Geom, holos, crops):
        self._stream()
        A = _holo_count(holos, g)
        self._check(lib().mtd_holo_to_crop(self.h, ctypes.byref(g), _ptr_holos(holos),
                                           _dtype_code(holos), A,
                                           _dev(crops, "c32", A * g.C * g.C, "crops")),
                    "holo_to_crop")

    def holo_to_crop_noalloc(self, g: Geom, holos, crops):
        """holo_to_crop with no device allocation: the library's per-device
        scratch array holds its tables and row intermediate (holograms in
        device memory; calls overlapping on one GPU race)."""
        self._stream()
        A = _holo_count(holos, g)
        self._check(lib().mtd_holo_to_crop_noalloc(self.h, ctypes.byref(g), _ptr(holos),
                                                   _dtype_code(holos), A,
                                                   _dev(crops, "c32", A * g.C * g.C, "crops")),
                    "holo_to_crop_noalloc")

    def crop_to_field(self, g: Geom, crops, mask, posspec, fields):
        self._stream()
        keep, mp = _mask_ptr(mask, g.C)
        A = crops.shape[0]
        n = A * g.C * g.C
        self._check(lib().mtd_crop_to_field(self.h, ctypes.byref(g), _dev(crops, "c32", n, "crops"),
                                            A, mp, _dev(posspec, "i32", 2 * A, "posspec"),
                                            _dev(fields, "c32", n, "fields")), "crop_to_field")

    def demodulate(self, g: Geom, crops, posspec, u):
        """calc_Uborn2: the demodulated field of each crop for its specular
        (U frame), complex64 [A][C][C]."""
        self._stream()
        A = crops.shape[0]
        n = A * g.C * g.C
        self._check(lib().mtd_demodulate(self.h, ctypes.byref(g), _dev(crops, "c32", n, "crops"),
                                         A, _dev(posspec, "i32", 2 * A, "posspec"),
                                         _dev(u, "c32", n, "u")), "demodulate")

    def preprocess(self, g: Geom, holos, mask, posspec, fields):
        self._stream()
        keep, mp = _mask_ptr(mask, g.C)
        A = _holo_count(holos, g)
        self._check(lib().mtd_preprocess(self.h, ctypes.byref(g), _ptr_holos(holos),
                                         _dtype_code(holos), A, mp,
                                         _dev(posspec, "i32", 2 * A, "posspec"),
                                         _dev(fields, "c32", A * g.C * g.C, "fields")),
                    "preprocess")

    def field_to_spectrum(self, g: Geom, fields, posspec, spectra):
        self._stream()
        A = fields.shape[0]
        n = A * g.C * g.C
        self._check(lib().mtd_field_to_spectrum(self.h, ctypes.byref(g),
                                                _dev(fields, "c32", n, "fields"),
                                                _dev(posspec, "i32", 2 * A, "posspec"), A,
                                                _dev(spectra, "c32", n, "spectra")),
                    "field_to_spectrum")

    def ewald_scatter(self, g: Geom, spectra, posspec, vol, counts):
        self._stream()
        A = spectra.shape[0] if spectra is not None else 0
        nv = g.dim_final ** 3
        self._check(lib().mtd_ewald_scatter(
            self.h, ctypes.byref(g), _dev(spectra, "c32", A * g.C * g.C, "spectra", A == 0),
            _dev(posspec, "i32", 2 * A, "posspec", A == 0), A, _dev(vol, "c32", nv, "vol"),
            _dev(counts, "i32", nv, "counts")), "ewald_scatter")

    def ewald_accumulate(self, g: Geom, fields, posspec, vol, counts):
        self._stream()
        A = fields.shape[0]
        nv = g.dim_final ** 3
        self._check(lib().mtd_ewald_accumulate(
            self.h, ctypes.byref(g), _dev(fields, "c32", A * g.C * g.C, "fields"),
            _dev(posspec, "i32", 2 * A, "posspec"), A, _dev(vol, "c32", nv, "vol"),
            _dev(counts, "i32", nv, "counts")), "ewald_accumulate")

    def finalize(self, g: Geom, vol, counts, index_re, index_im, otf=None):
        self._stream()
        nv = g.dim_final ** 3
        self._check(lib().mtd_finalize(self.h, ctypes.byref(g), _dev(vol, "c32", nv, "vol"),
                                       _dev(counts, "i32", nv, "counts"),
                                       _dev(index_re, "f32", nv, "index_re"),
                                       _dev(index_im, "f32", nv, "index_im"),
                                       _dev(otf, "u8", nv, "otf", True)),
                    "finalize")

    def reconstruct_plain(self, g: Geom, holos, mask, index_re, index_im, posspec=None):
        self._stream()
        keep, mp = _mask_ptr(mask, g.C)
        A = _holo_count(holos, g)
        nv = g.dim_final ** 3
        self._check(lib().mtd_reconstruct_plain(self.h, ctypes.byref(g), _ptr_holos(holos),
                                                _dtype_code(holos), A, mp,
                                                _dev(posspec, "i32", 2 * A, "posspec", True),
                                                _dev(index_re, "f32", nv, "index_re"),
                                                _dev(index_im, "f32", nv, "index_im")),
                    "reconstruct_plain")

    def reconstruct_batch(self, geoms, holos, masks, index_re, index_im, posspecs=None,
                          graph=True, autotune=False, times=False):
        """Batched form: acquisition i = (geoms[i], holos[i] (GPU tensor),
        masks[i] (host C x C or None), index_re[i], index_im[i], posspecs[i]
        or None), reconstructed in order; graph: captured on the first call
        and replayed on later calls with the same arguments. times: returns
        each acquisition's device milliseconds (the call then waits)."""
        import numpy as np
        self._stream()
        n = len(geoms)
        G = (Geom * n)(*geoms)
        VP = ctypes.c_void_p * n
        keep = [_mask_ptr(masks[i] if masks else None, geoms[i].C) for i in range(n)]
        cnts = [_holo_count(h, g) for h, g in zip(holos, geoms)]
        hp = VP(*[_ptr_holos(h) for h in holos])
        dt = (ctypes.c_int * n)(*[_dtype_code(h) for h in holos])
        ct = (ctypes.c_int * n)(*cnts)
        mp = VP(*[k[1] for k in keep])
        pp = (VP(*[_dev(p, "i32", 2 * a, "posspec", True) for p, a in zip(posspecs, cnts)])
              if posspecs else None)
        rp = VP(*[_dev(t, "f32", g.dim_final ** 3, "index_re") for t, g in zip(index_re, geoms)])
        ip = VP(*[_dev(t, "f32", g.dim_final ** 3, "index_im") for t, g in zip(index_im, geoms)])
        ms = np.zeros(n, dtype=np.float32) if times else None
        flags = (MTD_BATCH_GRAPH if graph else 0) | (MTD_BATCH_AUTOTUNE if autotune else 0)
        self._check(lib().mtd_reconstruct_batch(
            self.h, n, G, hp, dt, ct, mp, pp, rp, ip, flags,
            None if ms is None else ms.ctypes.data_as(ctypes.c_void_p)), "reconstruct_batch")
        return ms

    def batch_tuned_chunk(self, n, C, dtype=MTD_U16):
        return lib().mtd_batch_tuned_chunk(self.h, n, C, dtype)

    def batch_tuned_window(self, n, C, dtype=MTD_U16):
        return lib().mtd_batch_tuned_window(self.h, n, C, dtype)

    def batch_tuned_schedule(self, n, C, V, dtype=MTD_U16):
        """The autotuner's schedule for (n, C, V, dtype) as a dict, or None
        if that shape is not tuned."""
        out = (_I * 6)()
        found = lib().mtd_batch_tuned_schedule(self.h, n, C, V, dtype, out)
        if found < 0:
            self._check(found, "batch_tuned_schedule")
        if found == 0:
            return None
        keys = ("holo_chunk", "gather_window", "gather_batch", "vol_rows_tile",
                "vol_cols_tile", "from_cache")
        return {k: int(v) for k, v in zip(keys, out)}

    def set_tune_cache(self, path):
        """Load the autotuner's cache file (entries of this library version
        and device) and append what it tunes from now on; None: no file."""
        self._check(lib().mtd_ctx_set_tune_cache(
            self.h, None if path is None else os.fsencode(path)), "set_tune_cache")

    # ---- Born forward model and adjoint ---------------------------------
    def spectra_to_fields(self, g: Geom, spectra, posspec, fields):
        """The simulator's field synthesis: fields[A][C][C] (UBornfinal layout)."""
        self._stream()
        A = spectra.shape[0]
        n = A * g.C * g.C
        self._check(lib().mtd_spectra_to_fields(self.h, ctypes.byref(g),
                                                _dev(spectra, "c32", n, "spectra"),
                                                _dev(posspec, "i32", 2 * A, "posspec"), A,
                                                _dev(fields, "c32", n, "fields")),
                    "spectra_to_fields")

    def born_plan(self, g: Geom, posspec):
        """Geometry plan: later forward/adjoint/weights calls with
        posspec=None reuse its angle geometry, band and tile lists."""
        self._stream()
        if posspec.numel() % 2:
            raise MTDError("posspec: 2*A values (all x, then all y)")
        A = posspec.numel() // 2
        self._check(lib().mtd_born_plan(self.h, ctypes.byref(g),
                                        _dev(posspec, "i32", 2 * A, "posspec"), A), "born_plan")

    @staticmethod
    def _angles(posspec, arr):
        return posspec.numel() // 2 if posspec is not None else arr.shape[0]

    def forward_spectra(self, g: Geom, dn_re, dn_im, posspec, spectra):
        """spectra[A][C][C] c32 = A(dn) (Simul_holo's calcHolo over all angles);
        posspec=None: the geometry plan."""
        self._stream()
        A = self._angles(posspec, spectra)
        nv = g.dim_final ** 3
        self._check(lib().mtd_forward_spectra(self.h, ctypes.byref(g),
                                              _dev(dn_re, "f32", nv, "dn_re"),
                                              _dev(dn_im, "f32", nv, "dn_im", True),
                                              _dev(posspec, "i32", 2 * A, "posspec", True), A,
                                              _dev(spectra, "c32", A * g.C * g.C, "spectra")),
                    "forward_spectra")

    def adjoint_spectra(self, g: Geom, spectra, posspec, out_re, out_im):
        """(out_re, out_im) = A^H spectra; posspec=None: the geometry plan."""
        self._stream()
        A = self._angles(posspec, spectra)
        nv = g.dim_final ** 3
        self._check(lib().mtd_adjoint_spectra(self.h, ctypes.byref(g),
                                              _dev(spectra, "c32", A * g.C * g.C, "spectra"),
                                              _dev(posspec, "i32", 2 * A, "posspec", True), A,
                                              _dev(out_re, "f32", nv, "out_re"),
                                              _dev(out_im, "f32", nv, "out_im", True)),
                    "adjoint_spectra")

    def cap_weights(self, g: Geom, posspec, weights):
        """weights[A][C][C] float = 1 / (q^2 * hits of the sample's voxel);
        posspec=None: the geometry plan."""
        self._stream()
        A = self._angles(posspec, weights)
        self._check(lib().mtd_cap_weights(self.h, ctypes.byref(g),
                                          _dev(posspec, "i32", 2 * A, "posspec", True), A,
                                          _dev(weights, "f32", A * g.C * g.C, "weights")),
                    "cap_weights")

    def tv_fista_step(self, z, grad_data, x, z_next, data_scale, lam, eps, step, beta,
                      positivity=True):
        """x' = P(z - step (data_scale grad_data + lam grad TV_eps(z))),
        z_next = x' + beta (x' - x), x <- x' (float V^3 volumes)."""
        self._stream()
        V = z.shape[-1]
        nv = V ** 3
        if z.numel() != nv:
            raise MTDError(f"z: {z.numel()} values, not a {V}^3 volume")
        self._check(lib().mtd_tv_fista_step(self.h, V, _dev(z, "f32", nv, "z"),
                                            _dev(grad_data, "f32", nv, "grad_data"),
                                            _dev(x, "f32", nv, "x"),
                                            _dev(z_next, "f32", nv, "z_next"), data_scale, lam,
                                            eps, step, beta, int(bool(positivity))),
                    "tv_fista_step")

    def tv_solve(self, g: Geom, spectra, weights, x, n_iter, lam, eps, inv_kappa,
                 positivity=True, graph=False):
        """n_iter FISTA/TV iterations in the library on the geometry plan
        (born_plan first); x (float V^3) is the start and the result.
        graph: the iterations are captured into one HIP graph, launched once."""
        self._stream()
        A = spectra.shape[0]
        n = A * g.C * g.C
        self._check(lib().mtd_tv_solve(self.h, ctypes.byref(g), _dev(spectra, "c32", n, "spectra"),
                                       _dev(weights, "f32", n, "weights", True), A, int(n_iter),
                                       lam, eps, inv_kappa, int(bool(positivity)),
                                       _dev(x, "f32", g.dim_final ** 3, "x"),
                                       1 if graph else 0),
                    "tv_solve")
