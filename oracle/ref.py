"""Oracle: a CPU restatement of the reference's off-axis preprocessing
(``Tomo_prog/pretraitement_hors_axe_CPU``) and Fourier-space reconstruction
(``Tomo_prog/Reconstruction``) for the plain method (BORN 1 / C_ABER 1 /
DEROUL 0, plus the Rytov and C_ABER 0 branches).

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as what
results are compared against (or the timed CPU baseline). The product path
never calls it.

Two precisions are restated:

* ``prec="f64"`` -- the FP64 oracle: every step in float64, as the reference
  computes.
* ``prec="f32"`` -- the reference FP32 path: the same steps with FFTs, field
  and volume arithmetic in float32; the least-squares polynomial fit and the
  background evaluation stay in float64 (the only FP64 stage of the GPU path).

Every function cites the reference lines it follows. Pins (checks that do not
retype these formulas) live in ``tests/test_oracle_*.py``:

* config parsing        -- hand-computed known answers (float rounding, last key
                           wins, comments, defaults);
* tukey windows         -- borne1/borne2 known answers, plateau, end values;
* centred FFT + crop    -- brute-force DFT sum on tiny inputs, pure-carrier
                           closed form, Parseval;
* argmax                -- brute-force scan on tiny inputs with ties and the
                           slot-0 squared quirk;
* demodulation          -- plane wave at a known sideband offset gives a
                           constant field of known amplitude;
* least squares         -- exact recovery of a degree-4 background, residual
                           orthogonality, agreement with numpy.linalg.lstsq;
* Born/Rytov formation  -- empty sample gives U_B == 0, known phase object;
* Ewald scatter         -- closed-form total count, spherical-cap z for a
                           centred specular, single-pixel spectrum gives one
                           voxel with the closed-form value;
* finalize              -- single normalised delta gives the closed-form plane
                           wave; Parseval.

Known, documented deviations from the reference (both unpinnable without
running OpenCV):

* ``compuBackgr2`` reads ``coef[15]`` past the end of the 15 coefficients
  (Correction_aberration.cpp:157). The restatement evaluates the intended
  15-term polynomial.
* ``cv::solve(..., DECOMP_NORMAL)`` rounding is restated as LAPACK LU with
  partial pivoting on the same normal equations (numpy.linalg.solve).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ``#define PI 3.14159`` at Reconstruction/src/fonctions.cpp:2 (used by
# retroPropag_Born); ``3.1415`` is written inline in tukey2D
# (pretraitement_hors_axe_CPU/FFT_fonctions.cpp:75,81); Reconstruction/main.cpp:28
# uses M_PI for the potential-to-index constant.
PI_RETRO = 3.14159
PI_TUKEY = 3.1415


def _real(prec):
    return np.float32 if prec == "f32" else np.float64


def _cplx(prec):
    return np.complex64 if prec == "f32" else np.complex128


def _mk(re, im, ct):
    """Assemble a complex array from separately computed parts (no extra
    rounding: the parts are stored as computed)."""
    out = np.empty(np.shape(re), dtype=ct)
    out.real = re
    out.imag = im
    return out


# --------------------------------------------------------------------------
# A1: config files
# --------------------------------------------------------------------------
def _config_lines(path):
    """Lines of a config file as the reference tokenises them
    (pretraitement_hors_axe_CPU/src/IO_fonctions.cpp:85-95): split on '\\n',
    drop lines whose first character is '#'."""
    with open(path, "rb") as fh:
        text = fh.read().decode("latin-1")
    return [ln for ln in text.split("\n") if not ln.startswith("#")]


def _lookup(path, token):
    """Last value string for ``token`` or None (IO_fonctions.cpp:97-112).

    The key is the text before the first space; the value is the text after
    it. A line without a space is its own key with the whole line as value.
    """
    value = None
    for ln in _config_lines(path):
        if ln == "":
            continue
        pos = ln.find(" ")
        key = ln if pos < 0 else ln[:pos]
        if key == token:
            value = ln if pos < 0 else ln[pos + 1:]
    return value


def _atof(s):
    """C ``atof``: longest leading decimal-float prefix, 0 when none."""
    import re
    m = re.match(r"\s*([+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?))", s)
    return float(m.group(1)) if m else 0.0


def extract_val(path, token, default=None):
    """``extract_val``: value parsed with atof and stored in a C ``float``
    (IO_fonctions.cpp:81-118; the 3-argument default form is
    Reconstruction/src/fonctions.cpp:137-180). Returns the float value
    promoted to a Python float (= C double)."""
    v = _lookup(path, token)
    if v is None or v == "":
        return float(np.float32(default)) if default is not None else 0.0
    return float(np.float32(_atof(v)))


def extract_string(path, token):
    """``extract_string`` (IO_fonctions.cpp:44-79): last value, '' if absent."""
    v = _lookup(path, token)
    return "" if v is None else v


# --------------------------------------------------------------------------
# small helpers
# --------------------------------------------------------------------------
def round_half_away(x):
    """C ``round``: halves away from zero, exact for every double."""
    x = np.asarray(x, dtype=np.float64)
    t = np.trunc(x)
    frac = x - t
    return t + np.where(np.abs(frac) >= 0.5, np.sign(x), 0.0)


def fftshift2(a):
    """fftshift2D for even sides (FFT_fonctions.cpp:89-114,237-264):
    quadrant swap == np.roll by half a side."""
    n0, n1 = a.shape[-2:]
    return np.roll(a, (n0 // 2, n1 // 2), axis=(-2, -1))


def fftshift3(a):
    """fftshift3D for even sides (Reconstruction/src/FFT_fonctions.cpp:78-151)."""
    return np.roll(a, tuple(s // 2 for s in a.shape), axis=(0, 1, 2))


def decal2d(a, dx, dy):
    """decal2DCplxGen / decal2DCplxGen2 (fonctions.cpp:64-85,
    Reconstruction/src/fonctions.cpp:1268-1309): out[(y+dy)%C][(x+dx)%C] =
    in[y][x]."""
    return np.roll(a, (dy, dx), axis=(-2, -1))


# --------------------------------------------------------------------------
# A3: Tukey window
# --------------------------------------------------------------------------
def tukey1d(n, alpha):
    """1-D factor of tukey2D (FFT_fonctions.cpp:61-87; identical copy in
    Reconstruction/src/FFT_fonctions.cpp:224-248). ``alpha`` is a C float and
    ``2*cpt/(alpha*(N-1))`` is a float division; pi is written 3.1415."""
    f = np.float32
    a = f(alpha)
    nm1 = f(n - 1)
    borne1 = int(round_half_away(float(f(a * nm1) / f(2))))
    borne2 = int(round_half_away(float(f(nm1 * f(f(1) - f(a / f(2)))))))
    den = f(a * nm1)
    two_over_a = f(f(2) / a)
    t = np.empty(n, dtype=np.float64)
    for cpt in range(n):
        if cpt < borne1 + 1:
            arg = f(f(f(2 * cpt) / den) - f(1))
            t[cpt] = 0.5 * (1 + math.cos(PI_TUKEY * float(arg)))
        elif cpt < borne2 + 1:
            t[cpt] = 1.0
        else:
            arg = f(f(f(f(2 * cpt) / den) - two_over_a) + f(1))
            t[cpt] = 0.5 * (1 + math.cos(PI_TUKEY * float(arg)))
    return t, borne1, borne2


def tukey2d(n, alpha):
    """tukey2D: tuk2D[y*N+x] = tuk1D[x]*tuk1D[y] in double."""
    t, _, _ = tukey1d(n, alpha)
    return t[None, :] * t[:, None]


def holo_window(n, alpha, prec="f64"):
    """Window of the hologram stage: tukey2D in double (FP64 path); the FP32
    path forms the same separable product from the float-rounded 1-D factors
    (tuk1D[x] * tuk1D[y] in float32)."""
    if prec == "f32":
        t = tukey1d(n, alpha)[0].astype(np.float32)
        return t[None, :] * t[:, None]
    return tukey2d(n, alpha)


# --------------------------------------------------------------------------
# geometry (A1 derivations)
# --------------------------------------------------------------------------
@dataclass
class Geom:
    """Derived acquisition geometry, as the two ``manip`` constructors compute
    it (pretraitement_hors_axe_CPU/manip.cpp:9-157,
    Reconstruction/src/manip.cpp:10-163)."""
    n: int                 # CCD_ROIX (hologram side N)
    nxmax: int             # NXMAX (crop side C = 2*NXMAX)
    dim_final: int         # DIM_FINAL (volume side V)
    circle_cx: int
    circle_cy: int
    n0: float
    na: float
    lam: float
    pix_holo: float        # tailleTheoPixelHolo = TPCAM/Gt
    rayon: float           # round(NXMAX*n0/NA)
    pix_uborn: float       # tailleTheoPixelHolo*N/(2*NXMAX)
    delta_f: float         # 1/(pix_uborn*2*NXMAX)
    coef_na: float = 1.0   # COEF_NA_OBJ_LIMIT (default 1)
    born: int = 1
    c_aber: int = 1
    deroul: int = 0
    alpha_holo: float = 0.05   # main.cpp:105
    alpha_field: float = 0.1   # main.cpp:180, Reconstruction/main.cpp:185
    # the ArrayFire variant's crop window (1+i) tukey2D(C, alpha_crop)
    # (pretraitement_hors_axe_GPU/main.cpp:74-79, src/GPU_fonctions.cpp:122);
    # 0: none (the CPU program)
    alpha_crop: float = 0.0

    @property
    def C(self):
        return 2 * self.nxmax

    @property
    def coin(self):
        """coinHA = (CIRCLE_CX-NXMAX, CIRCLE_CY-NXMAX) (main.cpp:116)."""
        return self.circle_cx - self.nxmax, self.circle_cy - self.nxmax


def derive_geom(n, nxmax, dim_final, circle_cx, circle_cy, na, n0, lam,
                tpcam, f_tube, f_obj, rf, coef_na=1.0, born=1, c_aber=1,
                deroul=0):
    """Apply the reference's config arithmetic: every config value goes through
    a C float (IO_fonctions.cpp:109) before the double derivations of
    manip.cpp:37-57,104 and Reconstruction/src/manip.cpp:52-94."""
    f = lambda v: float(np.float32(v))
    n0, na, lam = f(n0), f(na), f(lam)
    g_ = f(f_tube) / f(f_obj)
    gt = g_ / f(rf)
    pix_holo = f(tpcam) / gt
    nxmax = int(np.float32(nxmax))
    n = int(np.float32(n))
    dim_final = int(np.float32(dim_final))
    rayon = float(round_half_away(float(nxmax) * n0 / na))
    pix_uborn = pix_holo * n / (2 * nxmax)
    delta_f = 1.0 / (pix_uborn * 2 * nxmax)
    return Geom(n=n, nxmax=nxmax, dim_final=dim_final,
                circle_cx=int(np.float32(circle_cx)),
                circle_cy=int(np.float32(circle_cy)), n0=n0, na=na, lam=lam,
                pix_holo=pix_holo, rayon=rayon, pix_uborn=pix_uborn,
                delta_f=delta_f, coef_na=f(coef_na), born=born, c_aber=c_aber,
                deroul=deroul)


# --------------------------------------------------------------------------
# A4: hologram -> centred spectrum -> crop
# --------------------------------------------------------------------------
def holo_to_crop(holo, g: Geom, prec="f64", window=None):
    """holo2TF_UBorn (fonctions.cpp:279-304): H*tukey(0.05); fftshift2D;
    FFTW c2c forward; /N^2 (FFT_fonctions.cpp:316-330); fftshift2D;
    coupeCplx at coinHA (fonctions.cpp:104-124)."""
    rt, ct = _real(prec), _cplx(prec)
    N, C = g.n, g.C
    if window is None:
        window = holo_window(N, g.alpha_holo, prec)
    h = np.asarray(holo).astype(rt) * window.astype(rt)
    spec = np.fft.fft2(fftshift2(h).astype(ct))
    spec = (spec / rt(N * N)).astype(ct)
    spec = fftshift2(spec)
    x0, y0 = g.coin
    crop = spec[y0:y0 + C, x0:x0 + C].copy()
    if g.alpha_crop:
        # crop * af::complex(T, T): (re - im, re + im) * T, T the separable
        # window (products of the float-rounded factors in FP32)
        t = tukey1d(C, g.alpha_crop)[0]
        w = (t.astype(np.float32)[None, :] * t.astype(np.float32)[:, None]).astype(rt) \
            if prec == "f32" else (t[None, :] * t[:, None])
        crop = _mk((crop.real - crop.imag) * w, (crop.real + crop.imag) * w, ct)
    return crop


# --------------------------------------------------------------------------
# A5: specular position
# --------------------------------------------------------------------------
def coord_spec(crop, prec="f64"):
    """coordSpec (fonctions.cpp:88-101): slot 0 holds |TF[0]|^2, every other
    slot |TF|; a strictly-greater scan keeps the first maximum, i.e. the first
    index of the overall maximum."""
    rt = _real(prec)
    flat = crop.ravel()
    mod = np.abs(flat).astype(rt)
    mod[0] = rt(flat[0].real) * rt(flat[0].real) + rt(flat[0].imag) * rt(flat[0].imag)
    return int(np.argmax(mod))


# --------------------------------------------------------------------------
# A6: demodulation
# --------------------------------------------------------------------------
def demod(crop, px, py, prec="f64"):
    """calc_Uborn2 (fonctions.cpp:348-374): circular shift by -posSpec
    (decal2DCplxGen2), unnormalised backward FFT (TF2Dcplx_INV,
    FFT_fonctions.cpp:420-438), circular shift by (NXMAX, NXMAX)."""
    ct = _cplx(prec)
    C = crop.shape[0]
    t = decal2d(crop.astype(ct), -px, -py)
    iu = np.fft.ifft2(t, norm="forward").astype(ct)
    return decal2d(iu, C // 2, C // 2)


# --------------------------------------------------------------------------
# A9-A11: aberration (polynomial background) least squares, always FP64
# --------------------------------------------------------------------------
POLY_DEG = 4


def poly_terms(deg=POLY_DEG):
    """Monomial order of CalcPolyUs_xy / CalcPoly_xy
    (Correction_aberration.cpp:108-149): powX outer, powY inner."""
    return [(i, j) for i in range(deg + 1) for j in range(deg + 1 - i)]


def sample_points(mask):
    """countM + the ==220 selection (Correction_aberration.cpp:86-106,
    116-127, 196-203): pixels with row%5==0 and col%5==0 and mask>45, in
    row-major (y outer, x inner) order. Returns (ys, xs) int arrays."""
    C = mask.shape[0]
    ys, xs = np.mgrid[0:C, 0:C]
    sel = (ys % 5 == 0) & (xs % 5 == 0) & (np.asarray(mask) > 45)
    return ys[sel].astype(np.int64), xs[sel].astype(np.int64)


def poly_matrix(xs, ys, deg=POLY_DEG):
    """Rows x^i*y^j in raw pixel units (exact in double)."""
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    return np.stack([xs ** i * ys ** j for i, j in poly_terms(deg)], axis=1)


@dataclass
class LsqPlan:
    """Angle-independent part of the fit: sample set, normal matrix A^T A and
    the full-image basis."""
    ys: np.ndarray
    xs: np.ndarray
    A: np.ndarray
    G: np.ndarray
    full: np.ndarray

    @staticmethod
    def build(mask):
        C = mask.shape[0]
        ys, xs = sample_points(mask)
        if len(ys) < 15:
            raise ValueError("aberration mask leaves %d sample points; the degree-4 "
                             "fit needs at least 15" % len(ys))
        A = poly_matrix(xs, ys)
        gy, gx = np.mgrid[0:C, 0:C]
        full = poly_matrix(gx.ravel(), gy.ravel())
        return LsqPlan(ys, xs, A, A.T @ A, full)


def lsq_background(values, plan: LsqPlan):
    """compuCoefPoly2 (Correction_aberration.cpp:186-217) with
    cv::solve(DECOMP_NORMAL): LU on A^T A c = A^T b; then compuBackgr2
    (:151-164) evaluating the 15-term polynomial at every pixel (the
    reference's 16th, out-of-bounds term is not restated)."""
    C = values.shape[0]
    b = np.asarray(values, dtype=np.float64)[plan.ys, plan.xs]
    coef = np.linalg.solve(plan.G, plan.A.T @ b)
    return (plan.full @ coef).reshape(C, C), coef


# --------------------------------------------------------------------------
# A7, A10-A13: field formation
# --------------------------------------------------------------------------
def form_field(U, plan: LsqPlan, g: Geom, prec="f64", spec_value=None, angle=0):
    """main.cpp:254-341 for one angle, from the demodulated field U.

    C_ABER 1: phase = atan2 (fonctions.cpp:39-43); DEROUL 0 keeps it wrapped
    (main.cpp:277), DEROUL 1 / VOLKOV 1 unwraps it (main.cpp:271); phase and amplitude backgrounds by aberCorr2 / ampliCorr2
    (Correction_aberration.cpp:166-184); amplitude clamp (main.cpp:301-302);
    Born (main.cpp:307,309) or Rytov (:312-313).
    C_ABER 0: the spectrum is multiplied by conj(s)/|s| before the demodulation
    (main.cpp:332-337); by linearity this equals scaling U."""
    rt, ct = _real(prec), _cplx(prec)
    if not g.c_aber:
        s = complex(spec_value)
        mod = math.sqrt(s.real * s.real + s.imag * s.imag)
        f = ct(complex(s.real / mod, -s.imag / mod))
        return (U.astype(ct) * f).astype(ct)
    phi = np.arctan2(U.imag, U.real).astype(rt)
    if g.deroul == DEROUL_VOLKOV:
        phi = volkov_unwrap(phi, prec)
    elif g.deroul == DEROUL_HERRAEZ:
        # float32 values (the reference's PIXEL.value), used as doubles next
        phi = herraez_unwrap(phi, angle).astype(rt)
    elif g.deroul:
        raise ValueError("deroul must be 0, 1 (Volkov) or 2 (Herraez)")
    amp = np.abs(U).astype(rt)
    bg_phi, _ = lsq_background(phi, plan)
    bg_amp, _ = lsq_background(amp, plan)
    phic = phi.astype(np.float64) - bg_phi
    a = amp.astype(np.float64) / (bg_amp + 0.00000001)
    a = np.where(a > 10, 1.0, a)
    a = np.where(a < -10, 1.0, a)
    if prec == "f32":
        phic = phic.astype(np.float32)
        aa = np.abs(a).astype(np.float32)
        if g.born:
            re = aa * np.cos(phic) - np.float32(1)
            im = aa * np.sin(phic)
        else:
            re = np.log(aa)
            im = phic
        return _mk(re, im, np.complex64)
    aa = np.sqrt(a * a)
    if g.born:
        return _mk(aa * np.cos(phic) - 1, aa * np.sin(phic), np.complex128)
    return _mk(np.log(aa), phic, np.complex128)


# --------------------------------------------------------------------------
# A8: phase unwrapping, DEROUL 1 with VOLKOV 1 (deroul_volkov4_total_sym_paire)
# --------------------------------------------------------------------------
DEROUL_OFF, DEROUL_VOLKOV, DEROUL_HERRAEZ = 0, 1, 2


def kvect_shift(D):
    """init_kvect_shift (FFT_fonctions.cpp:7-17): (i - round(D/2))/D, then
    fftshift2D: the frequency, in cycles per pixel, of FFT index i (the
    Nyquist index D/2 gets -1/2). Same for x and y (square images)."""
    i = np.arange(D)
    return (((i + D // 2) % D) - D // 2).astype(np.float64) / D


def symmetrise_mirror(img):
    """Symetrise_mirror (src/deroulement_volkov4.cpp:221-240): the 2C x 2C
    even extension; the original is the [C,2C)^2 quadrant, the other three are
    its mirror images about the half-pixel lines."""
    top = np.concatenate([img[::-1, ::-1], img[::-1, :]], axis=1)
    bottom = np.concatenate([img[:, ::-1], img], axis=1)
    return np.concatenate([top, bottom], axis=0)


def _fft2_fwd(x, prec):
    """TF2Dcplx (FFT_fonctions.cpp:366-400): forward, divided by the pixel count."""
    return np.fft.fft2(x.astype(_cplx(prec)), norm="forward").astype(_cplx(prec))


def _fft2_bwd(x, prec):
    """TF2Dcplx_INV (FFT_fonctions.cpp:420-438): backward, unnormalised."""
    return np.fft.ifft2(x.astype(_cplx(prec)), norm="forward").astype(_cplx(prec))


def gradient_fft(f, prec="f64"):
    """gradient_fft4 (src/deroulement_volkov4.cpp:98-159), real or complex
    input: spectrum = TF2Dcplx(f); grad = TF2Dcplx_INV(2 pi k * spectrum * i)
    for k = kx and k = ky. Returns (gx, gy), complex."""
    rt = _real(prec)
    D = f.shape[0]
    k = kvect_shift(D).astype(rt)
    two_pi = rt(2 * math.pi)
    S = _fft2_fwd(f, prec)
    jx = (two_pi * k)[None, :].astype(rt)
    jy = (two_pi * k)[:, None].astype(rt)
    gx = _fft2_bwd(_mk(-jx * S.imag, jx * S.real, _cplx(prec)), prec)
    gy = _fft2_bwd(_mk(-jy * S.imag, jy * S.real, _cplx(prec)), prec)
    return gx, gy


def integ_grad(gx, gy, prec="f64"):
    """integ_grad4 (src/deroulement_volkov4.cpp:162-191): least-squares
    integration of a gradient field, tampon = -i (TFx kx + TFy ky) /
    (2 pi |k|^2), then the unnormalised backward FFT.

    Deviation: the reference sets the zero frequency to
    (tampon[nbPix] + tampon[1]) / 2, reading one element past the array
    (:189). It only adds a constant to the result, which the phase fit that
    always follows unwrapping removes (the constant monomial is in the basis);
    here it is 0."""
    rt, ct = _real(prec), _cplx(prec)
    D = gx.shape[0]
    k = kvect_shift(D).astype(rt)
    kx = k[None, :]
    ky = k[:, None]
    Fx = _fft2_fwd(gx, prec)
    Fy = _fft2_fwd(gy, prec)
    num = Fx * kx.astype(ct) + Fy * ky.astype(ct)            # complex
    den = (rt(2 * math.pi) * (kx * kx + ky * ky)).astype(rt)  # real
    den[0, 0] = 1
    t = _mk(num.imag / den, -num.real / den, ct)              # -i num / den
    t[0, 0] = 0
    return _fft2_bwd(t, prec)


def volkov_unwrap(phi, prec="f64"):
    """deroul_volkov4_total_sym_paire (src/deroulement_volkov4.cpp:14-92):
    mirror-extend the wrapped phase, compare the FFT gradient of the phase
    with the FFT gradient of exp(i phase) to get the gradient of the 2 pi
    multiples, integrate it (least squares, FFT Poisson solve), cut the
    original quadrant and add 2 pi times it. The multiples are not rounded."""
    rt, ct = _real(prec), _cplx(prec)
    C = phi.shape[0]
    ps = symmetrise_mirror(np.asarray(phi, dtype=rt))
    gxw, gyw = gradient_fft(ps, prec)
    Z = _mk(np.cos(ps).astype(rt), np.sin(ps).astype(rt), ct)
    gzx, gzy = gradient_fft(Z, prec)
    two_pi = rt(2 * math.pi)
    ax = (-1j * (gzx / Z)).astype(ct)
    ay = (-1j * (gzy / Z)).astype(ct)
    gix = ((ax.real - gxw.real) / two_pi).astype(rt)
    giy = ((ay.real - gyw.real) / two_pi).astype(rt)
    intm = integ_grad(gix, giy, prec)[C:, C:].real.astype(rt)
    return (np.asarray(phi, dtype=rt) + two_pi * intm).astype(rt)


# --------------------------------------------------------------------------
# A8: phase unwrapping, DEROUL 1 with VOLKOV 0 (Herraez: phaseUnwrapping_Mat,
# pretraitement_hors_axe_CPU/deroulement_herraez.cpp:32-75)
# --------------------------------------------------------------------------
HERR_PI = np.float32(3.141592654)      # deroulement_herraez.cpp:9 (float)
HERR_TWOPI = np.float32(6.283185307)   # :10


def glibc_rand(start, count):
    """rand() of glibc with its default seed 1 (the reference never calls
    srand): outputs start .. start+count-1 of the TYPE_3 additive generator,
    restated from glibc's published random_r: a 31-word state seeded by the
    Park-Miller recurrence 16807*x mod (2^31-1), r[i] = r[i-31] + r[i-3]
    (mod 2^32) with the first 310 words discarded, output r >> 1."""
    total = start + count
    r = [0] * (344 + total)
    r[0] = 1
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1], 127773)
        w = 16807 * lo - 2836 * hi
        r[i] = w + 2147483647 if w < 0 else w
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + total):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return np.array([x >> 1 for x in r[344 + start:344 + total]], dtype=np.int64)


def herraez_border_reliability(rand_vals):
    """initialisePIXELs (deroulement_herraez.cpp:335-356): every pixel starts
    with reliability 9999999 + rand() as a C int (wrapping past INT_MAX, as
    the 32-bit addition does), stored in a float; only border pixels keep it."""
    v = (rand_vals + 9999999) & 0xFFFFFFFF
    v = np.where(v >= 2 ** 31, v - 2 ** 32, v)
    return v.astype(np.float32)


def _herr_wrap(x):
    """gamma (deroulement_herraez.cpp:360-367), float arithmetic."""
    x = x.astype(np.float32)
    return np.where(x > HERR_PI, x - HERR_TWOPI,
                    np.where(x < -HERR_PI, x + HERR_TWOPI, x)).astype(np.float32)


def herraez_reliability(W, border):
    """calculate_reliability (:382-404): interior pixels get H^2 + V^2 + D1^2
    + D2^2 in float, each second difference from double differences of the
    wrapped image rounded to float; border pixels keep `border` (row-major
    image of the initial values)."""
    Wd = np.asarray(W, np.float64)
    c = Wd[1:-1, 1:-1]
    d = lambda a, b: _herr_wrap((a - b).astype(np.float32))
    H = d(Wd[1:-1, :-2], c) - d(c, Wd[1:-1, 2:])
    V = d(Wd[:-2, 1:-1], c) - d(c, Wd[2:, 1:-1])
    D1 = d(Wd[:-2, :-2], c) - d(c, Wd[2:, 2:])
    D2 = d(Wd[:-2, 2:], c) - d(c, Wd[2:, :-2])
    rel = np.array(border, dtype=np.float32).copy()
    rel[1:-1, 1:-1] = ((H * H + V * V) + D1 * D1) + D2 * D2
    return rel


def herraez_edges(v, rel):
    """horizentalEDGEs / verticalEDGEs (:409-452): all horizontal edges in
    row-major order, then all vertical ones; reliab = float sum of the two
    pixels' reliabilities; increment = find_wrap (:370-380) of the float
    values. Returns (p1, p2, reliab, increment)."""
    Hh, Ww = v.shape
    idx = np.arange(Hh * Ww).reshape(Hh, Ww)
    p1 = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    p2 = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    vf, rf = v.ravel().astype(np.float32), rel.ravel().astype(np.float32)
    reliab = (rf[p1] + rf[p2]).astype(np.float32)
    diff = (vf[p1] - vf[p2]).astype(np.float32)
    inc = np.where(diff > HERR_PI, -1, np.where(diff < -HERR_PI, 1, 0)).astype(np.int64)
    return p1, p2, reliab, inc


def _herr_mix(keys, run_l, run_r):
    """Mix (:113-157): the two-pointer merge of two runs (index arrays) by
    key, the left element first on ties. The runs need not be sorted (odd
    sizes leave elements out of the recursion, below); such a merge orders
    elements exactly as a stable merge of the runs' running maxima."""
    kl = np.maximum.accumulate(keys[run_l]) if len(run_l) else keys[run_l]
    kr = np.maximum.accumulate(keys[run_r]) if len(run_r) else keys[run_r]
    order = np.argsort(np.concatenate([kl, kr]), kind="stable")
    return np.concatenate([run_l, run_r])[order]


def herraez_sort(keys):
    """quick_sort (:176-195) -> sort (:160-174): a recursive merge sort of
    the edge indices that splits a range of `size` into two halves of
    size // 2, so an odd range leaves its last element where it is, and size
    2 is a compare-and-swap. Returns the processed order (edge indices)."""
    keys = np.asarray(keys)
    idx = np.arange(len(keys))

    def rec(start, size):
        if size == 2:
            if keys[idx[start]] > keys[idx[start + 1]]:
                idx[start], idx[start + 1] = idx[start + 1], idx[start]
        elif size > 2:
            h = size // 2
            rec(start, h)
            rec(start + h, h)
            idx[start:start + 2 * h] = _herr_mix(keys, idx[start:start + h].copy(),
                                                 idx[start + h:start + 2 * h].copy())

    rec(0, len(keys))
    return idx


def herraez_gather(order, p1, p2, inc, npix):
    """gatherPIXELs (:455-552): edges in the sorted order; an edge joining two
    groups merges them and fixes the number of 2 pi between them: a pixel
    alone joins the other's group (PIXEL 2 tested first), else the group
    with more pixels absorbs the other (PIXEL 2's on equal counts), its
    members' increments shifted. Returns the increments."""
    head = list(range(npix))
    members = [[p] for p in range(npix)]
    incr = [0] * npix
    for e in order:
        a, b, w = int(p1[e]), int(p2[e]), int(inc[e])
        ha, hb = head[a], head[b]
        if ha == hb:
            continue
        if len(members[hb]) == 1:                  # PIXEL 2 alone
            members[ha].append(b)
            members[hb] = None
            head[b] = ha
            incr[b] = incr[a] - w
        elif len(members[ha]) == 1:                # PIXEL 1 alone
            members[hb].append(a)
            members[ha] = None
            head[a] = hb
            incr[a] = incr[b] + w
        elif len(members[ha]) > len(members[hb]):  # group 2 into group 1
            d = incr[a] - w - incr[b]
            for q in members[hb]:
                head[q] = ha
                incr[q] += d
            members[ha].extend(members[hb])
            members[hb] = None
        else:                                      # group 1 into group 2
            d = incr[b] + w - incr[a]
            for q in members[ha]:
                head[q] = hb
                incr[q] += d
            members[hb].extend(members[ha])
            members[ha] = None
    return np.asarray(incr, dtype=np.int64)


def herraez_unwrap(phi, angle=0):
    """phaseUnwrapping_Mat (:32-75) of one C x C wrapped phase: reliability,
    edges, the reference's sort, the grouping, then value + TWOPI * increment
    in float (unwrapImage :555-566). `angle`: how many images the program
    unwrapped before this one (each consumed one rand() per pixel). Returns
    float32."""
    Hh, Ww = phi.shape
    npix = Hh * Ww
    border = herraez_border_reliability(glibc_rand(angle * npix, npix)).reshape(Hh, Ww)
    v = np.asarray(phi).astype(np.float32)
    rel = herraez_reliability(phi, border)
    p1, p2, reliab, inc = herraez_edges(v, rel)
    order = herraez_sort(reliab)
    incr = herraez_gather(order, p1, p2, inc, npix)
    out = v.ravel() + HERR_TWOPI * incr.astype(np.float32)
    return out.astype(np.float32).reshape(Hh, Ww)


def load_mask(C, mask=None):
    """init_mask_aber (Correction_aberration.cpp:43-70): absent mask => 255."""
    if mask is None:
        return np.full((C, C), 255, dtype=np.uint8)
    return np.asarray(mask, dtype=np.uint8)


def preprocess(holos, g: Geom, mask=None, prec="f64", return_crops=False):
    """The two hot loops of pretraitement_hors_axe_CPU/main.cpp:130-342 over
    holograms that all exist. Returns (fields[A,C,C], posspec[2A] as ints:
    all x then all y (main.cpp:250-251))."""
    C = g.C
    A = len(holos)
    window = holo_window(g.n, g.alpha_holo, prec)
    plan = LsqPlan.build(load_mask(C, mask))
    fields = np.empty((A, C, C), dtype=_cplx(prec))
    crops = np.empty((A, C, C), dtype=_cplx(prec)) if return_crops else None
    pos = np.empty(2 * A, dtype=np.int64)
    for a in range(A):
        crop = holo_to_crop(holos[a], g, prec, window)
        if return_crops:
            crops[a] = crop
        i = coord_spec(crop, prec)
        px, py = i % C, i // C
        pos[a], pos[a + A] = px, py
        U = demod(crop, px, py, prec)
        fields[a] = form_field(U, plan, g, prec, spec_value=crop[py, px], angle=a)
    if return_crops:
        return fields, pos, crops
    return fields, pos


# --------------------------------------------------------------------------
# A16: per-angle spectrum in the reconstruction
# --------------------------------------------------------------------------
def field_to_spectrum(field_, px, py, g: Geom, prec="f64", window=None):
    """Reconstruction/main.cpp:207-217: U_B * tukey(C, 0.1); decal2DCplxGen by
    (NXMAX, NXMAX); FFTW forward * pixUBorn^2 (src/FFT_fonctions.cpp:333-349);
    decal2DCplxGen by +posSpec. Returns S[C][C] indexed like TF_UBorn_normC."""
    rt, ct = _real(prec), _cplx(prec)
    C = g.C
    if window is None:
        window = tukey2d(C, g.alpha_field)
    w = window.astype(rt)
    ub = _mk(field_.real.astype(rt) * w, field_.imag.astype(rt) * w, ct)
    ub = decal2d(ub, g.nxmax, g.nxmax)
    F = np.fft.fft2(ub).astype(ct)
    F = (F * rt(g.pix_uborn ** 2)).astype(ct)
    return decal2d(F, px, py)


# --------------------------------------------------------------------------
# A17: Ewald-cap scatter with hit counts
# --------------------------------------------------------------------------
def ewald_constants(g: Geom):
    """Constants of retroPropag_Born (Reconstruction/src/fonctions.cpp:875-905):
    float n0, kv = 2*PI/lambda0 (double product stored in a float),
    k0 = kv*n0 in float; cteUb2Pot = i*k0/PI; cteNorm = -2*PI; the per-pixel
    factor beta = ((k0/PI)*cteNorm)*sdz."""
    n0f = np.float32(g.n0)
    kv = np.float32(2 * PI_RETRO / g.lam)
    k0 = np.float32(kv * n0f)
    gamma = (float(k0) / PI_RETRO) * (-2 * PI_RETRO)
    return gamma


def pupil_tables(g: Geom):
    """fd grid, pupil mask (|fd|^2 < coef*NXMAX^2, fonctions.cpp:915), the
    height h = sqrt(R^2 - fdx^2 - fdy^2) and beta per pixel."""
    nx = g.nxmax
    fdy, fdx = np.mgrid[-nx:nx, -nx:nx]
    r2i = fdx * fdx + fdy * fdy
    pupil = r2i.astype(np.float64) < g.coef_na * float(nx * nx)
    R = g.rayon
    h = np.sqrt(np.maximum((R * R - fdx * fdx) - fdy * fdy, 0.0))
    sdz = h * (1.0 / R)
    beta = ewald_constants(g) * sdz
    return fdx, fdy, pupil, h, beta


def ewald_scatter(S, pos, g: Geom, prec="f64"):
    """retroPropag_Born (Reconstruction/src/fonctions.cpp:875-935) over all
    angles in order (main.cpp:203-225). S: [A][C][C] (TF_UBorn_normC frame);
    pos: ints [2A]. Returns (vol[V^3] complex, counts[V^3] int32).

    koz = round(sqrt(R^2-|fd|^2) - fzm0); vox = (fdx-fxm0+V/2) +
    (fdy-fym0+V/2)*V + (koz+V/2)*V^2; the angle is skipped when
    round(R^2-|fm0|^2) < 0. Values (i*beta)*S = (-beta*Si, beta*Sr) summed in
    angle order."""
    V = g.dim_final
    vol = np.zeros(V ** 3, dtype=_cplx(prec))
    cnt = np.zeros(V ** 3, dtype=np.int32)
    for vox, val in _scatter_hits(S, pos, g, prec):
        vol[vox] += val
        cnt[vox] += 1
    return vol, cnt


def _scatter_hits(S, pos, g: Geom, prec):
    """Per non-skipped angle in order: the voxel indices of its pupil pixels
    and their values (i*beta)*S (retroPropag_Born's index and value)."""
    ct = _cplx(prec)
    rt = _real(prec)
    A = S.shape[0]
    nx, V, R = g.nxmax, g.dim_final, g.rayon
    half = V // 2
    fdx, fdy, pupil, h, beta = pupil_tables(g)
    fdx_p, fdy_p, h_p = fdx[pupil], fdy[pupil], h[pupil]
    beta_p = beta[pupil].astype(rt)
    for a in range(A):
        px, py = int(pos[a]), int(pos[a + A])
        fxm0, fym0 = px - nx, py - nx
        fz2 = R * R - fxm0 * fxm0 - fym0 * fym0
        if round_half_away(fz2) < 0:
            continue
        fzm0 = math.sqrt(fz2)
        koz = round_half_away(h_p - fzm0).astype(np.int64)
        vox = (fdx_p - fxm0 + half) + (fdy_p - fym0 + half) * V + (koz + half) * V * V
        s = S[a][pupil]
        yield vox, _mk(-(beta_p * s.imag.astype(rt)), beta_p * s.real.astype(rt), ct)


def ewald_scatter_sparse(S, pos, g: Geom, prec="f64"):
    """ewald_scatter restricted to the voxels that receive hits, for volumes
    too large to hold densely: (sorted voxel indices, sums, counts), each sum
    accumulated in angle order in the working precision, as the dense form
    (within one angle no voxel is hit twice)."""
    vs, xs = [], []
    for vox, val in _scatter_hits(S, pos, g, prec):
        vs.append(vox)
        xs.append(val)
    if not vs:
        return np.zeros(0, np.int64), np.zeros(0, _cplx(prec)), np.zeros(0, np.int32)
    vox = np.concatenate(vs)
    val = np.concatenate(xs)
    uniq, inv = np.unique(vox, return_inverse=True)
    sums = np.zeros(uniq.size, dtype=_cplx(prec))
    cnt = np.zeros(uniq.size, dtype=np.int32)
    np.add.at(sums, inv, val)  # unbuffered, in input (angle) order
    np.add.at(cnt, inv, 1)
    return uniq, sums, cnt


def ewald_bounds_ok(pos, g: Geom):
    """True when every voxel index of every non-skipped angle lies inside the
    V^3 volume (the reference has no check, fonctions.cpp:926-928)."""
    A = len(pos) // 2
    nx, V, R = g.nxmax, g.dim_final, g.rayon
    half = V // 2
    fdx, fdy, pupil, h, _ = pupil_tables(g)
    for a in range(A):
        fxm0, fym0 = int(pos[a]) - nx, int(pos[a + A]) - nx
        fz2 = R * R - fxm0 * fxm0 - fym0 * fym0
        if round_half_away(fz2) < 0:
            continue
        koz = round_half_away(h[pupil] - math.sqrt(fz2))
        for arr in (fdx[pupil] - fxm0 + half, fdy[pupil] - fym0 + half, koz + half):
            if arr.min() < 0 or arr.max() >= V:
                return False
    return True


# --------------------------------------------------------------------------
# A18-A22: normalisation, 3-D inverse FFT, potential -> index
# --------------------------------------------------------------------------
def finalize(vol, cnt, g: Geom, prec="f64"):
    """Reconstruction/main.cpp:243-352: counts 0 -> 1 and divide; OTF mask
    Re != 0 (:262-267); fftshift3D; FFTW backward 3-D * delta_f^3
    (src/FFT_fonctions.cpp:394-412); fftshift3D; * -1/(2 n0 kv^2) with
    kv = 2*M_PI/lambda0 (:346-352). Returns (index complex V^3, otf bool V^3)."""
    rt, ct = _real(prec), _cplx(prec)
    V = g.dim_final
    c = np.where(cnt == 0, 1, cnt).astype(rt)
    t = _mk(vol.real.astype(rt) / c, vol.imag.astype(rt) / c, ct)
    otf = t.real != 0
    t = fftshift3(t.reshape(V, V, V))
    p = np.fft.ifftn(t, norm="forward").astype(ct)
    p = (p * rt(g.delta_f ** 3)).astype(ct)
    p = fftshift3(p)
    kv = 2 * math.pi / g.lam
    cte = -1.0 / (2 * g.n0 * kv * kv)
    idx = (p * rt(cte)).astype(ct)
    return idx, otf.reshape(V, V, V)


# --------------------------------------------------------------------------
# Born forward model (Simul_holo) and its adjoint
# --------------------------------------------------------------------------
# ``#define PI 3.14159`` at Simul_holo/src/fonctions.cpp:10 (calcHolo).
PI_SIMUL = 3.14159


def spectra_to_fields(S, pos, g: Geom):
    """The simulator's field synthesis (Simul_holo/main.cpp:174-180):
    decal2DCplxGen moving the specular to bin 0, TF2Dcplx_INV (FFTW backward
    x Delta_f^2, Simul_holo/src/FFT_fonctions.cpp:280-299), fftshift2D. FP64."""
    A = len(pos) // 2
    out = np.empty((A,) + S.shape[1:], np.complex128)
    for a in range(A):
        t = decal2d(np.asarray(S[a], np.complex128), -int(pos[a]), -int(pos[a + A]))
        out[a] = fftshift2(np.fft.ifft2(t, norm="forward") * g.delta_f ** 2)
    return out


def tomo_pixel(g: Geom):
    """Voxel size 1/(V delta_f) (Simul_holo/src/manip.cpp:59: Tp_Tomo =
    Tp_Uborn * 2 NXMAX / dim_final)."""
    return 1.0 / (g.dim_final * g.delta_f)


def born_weight(g: Geom):
    """calcHolo's factor (Simul_holo/src/fonctions.cpp:115-122,136):
    cteNormalisation * ctePot2UBorn * cteInd2Pot / sdz with sdz =
    sqrt(R^2 - |fd|^2)/R, in FP64. Returns the complex factor per bin of the
    C x C pupil frame (0 outside the pupil)."""
    kv = 2 * PI_SIMUL / g.lam
    k0 = kv * g.n0
    cte = (-1.0 / (2 * PI_SIMUL)) * (-1j * PI_SIMUL / k0) * (-2 * kv * kv * g.n0)
    _, _, pupil, h, _ = pupil_tables(g)
    sdz = np.where(pupil, h / g.rayon, 1.0)
    return np.where(pupil, cte / sdz, 0.0)


def contrast_spectrum(dn, g: Geom, prec="f64"):
    """TF_obj = fftshift3D(TF3Dcplx(fftshift3D(vol_obj))) with the forward
    FFT times Tp^3 (Simul_holo/main.cpp:106-109, src/FFT_fonctions.cpp:301-318)."""
    ct, rt = _cplx(prec), _real(prec)
    V = g.dim_final
    t = fftshift3(np.asarray(dn).reshape(V, V, V).astype(ct))
    F = np.fft.fftn(t).astype(ct)
    F = (F * rt(tomo_pixel(g) ** 3)).astype(ct)
    return fftshift3(F)


def _cap_voxels(pos, a, g: Geom):
    """Voxel index of every pupil pixel of angle a (the map of retroPropag_Born,
    Reconstruction/src/fonctions.cpp:918-926, which is calcHolo's
    fobj = round(fd - fi) in centred coordinates, Simul_holo/src/fonctions.cpp:
    124-136 with Point3D::coordI, src/Point3D.cpp:55-58). None for a skipped
    angle; -1 where the voxel leaves the volume (the height, or laterally for
    a specular outside the crop, where the reference would index out of
    bounds)."""
    A = len(pos) // 2
    nx, V, R = g.nxmax, g.dim_final, g.rayon
    half = V // 2
    fdx, fdy, pupil, h, _ = pupil_tables(g)
    fxm0, fym0 = int(pos[a]) - nx, int(pos[a + A]) - nx
    fz2 = R * R - fxm0 * fxm0 - fym0 * fym0
    if round_half_away(fz2) < 0:
        return None
    koz = round_half_away(h[pupil] - math.sqrt(fz2)).astype(np.int64) + half
    X = fdx[pupil] - fxm0 + half
    Y = fdy[pupil] - fym0 + half
    vox = X + Y * V + koz * V * V
    ok = (koz >= 0) & (koz < V) & (X >= 0) & (X < V) & (Y >= 0) & (Y < V)
    return np.where(ok, vox, -1)


def born_forward(dn, pos, g: Geom, prec="f64"):
    """Simul_holo's per-angle hologram spectrum (calcHolo,
    Simul_holo/src/fonctions.cpp:109-140, loop main.cpp:162-171) in the frame
    of field_to_spectrum: S[a][row][col], fd = (col - NXMAX, row - NXMAX)."""
    ct = _cplx(prec)
    A = len(pos) // 2
    C = g.C
    TF = contrast_spectrum(dn, g, prec).ravel()
    w = born_weight(g)
    _, _, pupil, _, _ = pupil_tables(g)
    S = np.zeros((A, C, C), dtype=ct)
    for a in range(A):
        vox = _cap_voxels(pos, a, g)
        if vox is None:
            continue
        vals = np.where(vox >= 0, TF[np.maximum(vox, 0)], 0)
        S[a][pupil] = (w[pupil] * vals).astype(ct)
    return S


def born_adjoint(S, pos, g: Geom):
    """Adjoint of born_forward in FP64, written as its explicit transpose:
    scatter conj(w) S onto the cap voxels, then the adjoint of the centred
    FFT (shift, unnormalised backward FFT, shift) times Tp^3."""
    A = len(pos) // 2
    V = g.dim_final
    w = born_weight(g)
    _, _, pupil, _, _ = pupil_tables(g)
    vol = np.zeros(V ** 3, dtype=np.complex128)
    for a in range(A):
        vox = _cap_voxels(pos, a, g)
        if vox is None:
            continue
        ok = vox >= 0
        np.add.at(vol, vox[ok], (np.conj(w[pupil]) * np.asarray(S[a])[pupil])[ok])
    t = fftshift3(vol.reshape(V, V, V))
    t = np.fft.ifftn(t, norm="forward") * tomo_pixel(g) ** 3
    return fftshift3(t)


def cap_weights(pos, g: Geom):
    """Data weights of the adjoint-iterative form (no reference counterpart):
    1 / (q^2 * hits of the sample's voxel) per cap sample, with q the float32
    magnitude of calcHolo's factor and the hits of retroPropag_Born's map;
    0 outside the pupil and for skipped angles. FP32 arithmetic as the GPU."""
    A = len(pos) // 2
    C, V = g.C, g.dim_final
    _, cnt = ewald_scatter(np.zeros((A, C, C), np.complex64), pos, g, "f32")
    q = np.abs(born_weight(g)).astype(np.float32)
    _, _, pupil, _, _ = pupil_tables(g)
    w = np.zeros((A, C, C), np.float32)
    for a in range(A):
        vox = _cap_voxels(pos, a, g)
        if vox is None:
            continue
        ok = vox >= 0
        qq = q[pupil][ok]
        c = cnt[vox[ok]].astype(np.float32)
        wa = np.zeros(pupil.sum(), np.float32)
        wa[ok] = np.float32(1) / (qq * qq * c)
        w[a][pupil] = wa
    return w


def reconstruct(fields, pos, g: Geom, prec="f64"):
    """tomo_reconstruction's compute (Reconstruction/main.cpp:203-352) from
    the preprocessing outputs. Returns (index, vol, counts, otf)."""
    A = fields.shape[0]
    window = tukey2d(g.C, g.alpha_field)
    S = np.empty_like(fields, dtype=_cplx(prec))
    for a in range(A):
        S[a] = field_to_spectrum(fields[a], int(pos[a]), int(pos[a + A]), g,
                                 prec, window)
    vol, cnt = ewald_scatter(S, pos, g, prec)
    idx, otf = finalize(vol, cnt, g, prec)
    return idx, vol, cnt, otf
