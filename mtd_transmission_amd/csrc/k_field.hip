// Field stage on the C x C crops (C = 2*NXMAX), batched over angles.
//
//   argmax     coordSpec (pretraitement_hors_axe_CPU/fonctions.cpp:88-101)
//   demod      calc_Uborn2 (fonctions.cpp:348-374): the -posSpec shift is an
//              index remap on load, then an unnormalised 2-D backward FFT.
//              The result IU is kept in the frame before the (NXMAX, NXMAX)
//              shift: U[y][x] = IU[(y+NXMAX)%C][(x+NXMAX)%C].
//   lsq        aberCorr2 / ampliCorr2 right-hand sides and solves in FP64
//              (Correction_aberration.cpp:166-217) with the LU factors of the
//              angle-independent normal matrix.
//   form       main.cpp:289-316 (Born / Rytov) or main.cpp:332-337 (C_ABER 0).
//   spectrum   Reconstruction/main.cpp:207-217: window, shift by NXMAX,
//              forward 2-D FFT * pixUBorn^2, optional shift by +posSpec.
#include "argmax.hpp"
#include "dft.hpp"
#include "gen_fft.hpp"
#include "internal.hpp"
#include "passes.hpp"
#include "trig.hpp"

namespace mtd {

// ---------------------------------------------------------------------------
// argmax with the slot-0 squared quirk; first index of the maximum.
// ---------------------------------------------------------------------------
constexpr int ARGMAX_SEG = 16;  // blocks per crop in the first stage

template <int CT>
__global__ void __launch_bounds__(256)
k_argmax_part(const float2* __restrict__ crop, float2* __restrict__ part, Dim<CT> C) {
  const int a = blockIdx.y, seg = blockIdx.x;
  const int LEN = (C * C + ARGMAX_SEG - 1) / ARGMAX_SEG;
  const float2* p = crop + (size_t)a * C * C;
  float best = -1.f;
  int bi = 0x7fffffff;
  const int i0 = seg * LEN, i1 = min(i0 + LEN, C * C);
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    const float m = argmax_key(p[i], i == 0);
    if (m > best) { best = m; bi = i; }
  }
  __shared__ float red[8];
  argmax_block(best, bi, red);
  if (threadIdx.x == 0) part[(size_t)a * ARGMAX_SEG + seg] = make_float2(best, __int_as_float(bi));
}

// second stage: one thread per angle over its S partials
template <int CT>
__global__ void k_argmax_fin(const float2* __restrict__ part, int S, int A, int* __restrict__ posx,
                             int* __restrict__ posy, Dim<CT> C) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A) return;
  float best = -1.f;
  int bi = 0x7fffffff;
  for (int s = 0; s < S; ++s) {
    const float2 p = part[(size_t)a * S + s];
    const int i = __float_as_int(p.y);
    if (argmax_better(p.x, i, best, bi)) { best = p.x; bi = i; }
  }
  posx[a] = bi % C;
  posy[a] = bi / C;
}

// ---------------------------------------------------------------------------
// demodulation: rows then columns of the backward FFT
// ---------------------------------------------------------------------------
template <int CT>
struct DemodRowOp {
  const float2* crop;
  const int* posx;
  const int* posy;
  float2* out;
  const float2* tw;
  int px, py;
  __device__ void begin(int a) { px = posx[a]; py = posy[a]; }
  __device__ float2 load(int a, int v, int u) const {
    return crop[((size_t)a * C + (v + py) % C) * C + (u + px) % C];
  }
  __device__ void store(int a, int v, int u, float2 x) const {
    out[((size_t)a * C + v) * C + u] = x;
  }
  Dim<CT> C{};
};

template <int CT>
struct ColInPlaceOp {  // column pass reading and writing one stack
  const float2* in;
  float2* out;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int col, int row) const {
    return in[((size_t)a * C + row) * C + col];
  }
  __device__ void store(int a, int col, int row, float2 x) const {
    out[((size_t)a * C + row) * C + col] = x;
  }
  Dim<CT> C{};
};

// Column pass of the simulator's field synthesis: the IU frame moved to the
// centred field frame (fftshift2D, a shift by C/2) and scaled by delta_f^2.
template <int CT>
struct ColCenterScaleOp {
  const float2* in;
  float2* out;
  const float2* tw;
  float scale;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int col, int row) const {
    return in[((size_t)a * C + row) * C + col];
  }
  __device__ void store(int a, int col, int row, float2 x) const {
    out[((size_t)a * C + (row + C / 2) % C) * C + (col + C / 2) % C] =
        make_float2(x.x * scale, x.y * scale);
  }
  Dim<CT> C{};
};

// ---------------------------------------------------------------------------
// least squares: per angle, FP64 right-hand sides A^T b for the phase and
// amplitude fits over the sample set, then the LU solve.
// ---------------------------------------------------------------------------
constexpr int NCOEF = 15;

// rows of the U frame that hold least-squares samples (y = 0, 5, 10, ...)
__host__ __device__ inline int lsq_sample_rows(int C) { return (C + 4) / 5; }

__device__ __forceinline__ void monomials(double X, double Y, double* mono) {
  double xp[5], yp[5];
  xp[0] = 1.0; yp[0] = 1.0;
#pragma unroll
  for (int i = 1; i < 5; ++i) { xp[i] = xp[i - 1] * X; yp[i] = yp[i - 1] * Y; }
  int k = 0;
#pragma unroll
  for (int i = 0; i <= 4; ++i)
#pragma unroll
    for (int j = 0; j <= 4 - i; ++j) mono[k++] = xp[i] * yp[j];
}

// Solves (LU) x = P b in place; b holds P b on entry. Fully unrolled so that
// b stays in registers; lu is block-uniform (scalar loads).
__device__ __forceinline__ void lu_solve15(const double* __restrict__ lu, double (&b)[NCOEF]) {
#pragma unroll
  for (int i = 1; i < NCOEF; ++i) {
    double s = b[i];
#pragma unroll
    for (int j = 0; j < i; ++j) s -= lu[i * NCOEF + j] * b[j];
    b[i] = s;
  }
#pragma unroll
  for (int i = NCOEF - 1; i >= 0; --i) {
    double s = b[i];
#pragma unroll
    for (int j = i + 1; j < NCOEF; ++j) s -= lu[i * NCOEF + j] * b[j];
    b[i] = s / lu[i * NCOEF + i];
  }
}

template <int CT>
__global__ void __launch_bounds__(256)
k_lsq(const float2* __restrict__ iu, const float* __restrict__ phu,
      const int2* __restrict__ samples, int M, const double* __restrict__ lu,
      const int* __restrict__ perm, double* __restrict__ coef, Dim<CT> C, int sample_rows) {
  const int a = blockIdx.x;
  const int Ci = C;
  const float2* u = iu + (size_t)a * (sample_rows ? lsq_sample_rows(Ci) : Ci) * Ci;
  double rp[NCOEF], ra[NCOEF];
#pragma unroll
  for (int k = 0; k < NCOEF; ++k) { rp[k] = 0.0; ra[k] = 0.0; }
  // unrolled so that several (sample, IU) load chains are in flight
#pragma unroll 4
  for (int m = threadIdx.x; m < M; m += 256) {
    const int2 s = samples[m];
    // sample_rows: iu holds only the IU rows of the sample rows (U-frame y a
    // multiple of 5), row y / 5
    const int row = sample_rows ? s.y / 5 : (s.y + C / 2) % C;
    const float2 v = u[row * C + (s.x + C / 2) % C];
    // phase: wrapped (atan2) or, with unwrapping, the unwrapped U-frame phase
    const double ph = phu ? (double)phu[((size_t)a * C + s.y) * C + s.x] : (double)atan2f(v.y, v.x);
    const double am = (double)hypotf(v.x, v.y);
    double mono[NCOEF];
    monomials((double)s.x, (double)s.y, mono);
#pragma unroll
    for (int k = 0; k < NCOEF; ++k) {
      rp[k] += ph * mono[k];
      ra[k] += am * mono[k];
    }
  }
  // deterministic reduction: butterfly within the wave, then waves in order
  __shared__ double red[4][2 * NCOEF];
  __shared__ double rhs[2 * NCOEF];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NCOEF; ++k) {
    double x = rp[k], y = ra[k];
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, 64);
      y += __shfl_xor(y, off, 64);
    }
    if (lane == 0) { red[wave][k] = x; red[wave][NCOEF + k] = y; }
  }
  __syncthreads();
  if (threadIdx.x < 2 * NCOEF) {
    const int idx = threadIdx.x;
    rhs[idx] = ((red[0][idx] + red[1][idx]) + red[2][idx]) + red[3][idx];
  }
  __syncthreads();
  if (threadIdx.x < 2) {  // thread 0: phase fit, thread 1: amplitude fit
    const double* r = rhs + threadIdx.x * NCOEF;
    double b[NCOEF];
#pragma unroll
    for (int k = 0; k < NCOEF; ++k) b[k] = r[perm[k]];
    lu_solve15(lu, b);
#pragma unroll
    for (int k = 0; k < NCOEF; ++k) coef[((size_t)a * 2 + threadIdx.x) * NCOEF + k] = b[k];
  }
}

// ---------------------------------------------------------------------------
// field formation for one pixel of the U frame at (X, Y)
// ---------------------------------------------------------------------------
struct FieldMode {
  int born;
  int c_aber;
};

// The fitted background sum_k c_k x^i y^j (compuBackgr2,
// Correction_aberration.cpp:151-164; k runs i-outer over i + j <= 4) in
// Horner form: sum_i x^i P_i(y), P_i(y) = sum_j c_(i,j) y^j. Along a row y is
// fixed: the P_i(y) are computed once per row (RowPoly) and each pixel costs
// four FP64 multiply-adds.
struct RowPoly {
  double p[5];
};
__device__ __forceinline__ RowPoly row_poly(const double* __restrict__ c, double y) {
  RowPoly r;
  r.p[0] = (((c[4] * y + c[3]) * y + c[2]) * y + c[1]) * y + c[0];
  r.p[1] = ((c[8] * y + c[7]) * y + c[6]) * y + c[5];
  r.p[2] = (c[11] * y + c[10]) * y + c[9];
  r.p[3] = c[13] * y + c[12];
  r.p[4] = c[14];
  return r;
}
__device__ __forceinline__ double row_poly_at(const RowPoly& r, double x) {
  return (((r.p[4] * x + r.p[3]) * x + r.p[2]) * x + r.p[1]) * x + r.p[0];
}

// Field of one pixel from its demodulated value u (main.cpp:289-316), given
// the phase (wrapped atan2 of u, or unwrapped) and the two fitted
// backgrounds at that pixel; C_ABER 0 (main.cpp:332-337): U * conj(s)/|s|.
__device__ __forceinline__ float2 form_from(float2 u, float phf, double bgp, double bga,
                                            float2 spec_norm, int mode) {
  if (!(mode & 2)) {
    return make_float2(u.x * spec_norm.x - u.y * spec_norm.y,
                       u.x * spec_norm.y + u.y * spec_norm.x);
  }
  const float amf = hypotf(u.x, u.y);
  const double phic = (double)phf - bgp;
  double A = (double)amf / (bga + 0.00000001);
  if (A > 10) A = 1;
  if (A < -10) A = 1;
  const float pc = (float)phic;
  const float aa = (float)fabs(A);
  if (mode & 1) {
    float sp, cp;
    sincos_cw(pc, &sp, &cp);
    return make_float2(aa * cp - 1.0f, aa * sp);
  }
  return make_float2(logf(aa), pc);
}

// Born mode with the wrapped phase (no unwrapping): the phase enters only
// through e^{i pc}, pc = atan2(u) - bgp, so
//   aa e^{i pc} - 1 = (aa/|u|) u e^{-i bgp} - 1
// gives the same field without the arctangent (an FP32 rounding apart:
// rel-L2 about 2e-7 against form_from on a 512^3 reconstruction); u = 0 keeps
// the reference's -1 (atan2f(0, 0) = 0, aa = 0). |u| by the hardware square
// root of x^2 + y^2 (within about 2 ulp of hypotf; field values here stay far
// from the float range's ends, and |u| cancels in aa / |u| unless clamped).
__device__ __forceinline__ float2 form_born_wrapped(float2 u, double bgp, double bga) {
  const float amf = __builtin_amdgcn_sqrtf(u.x * u.x + u.y * u.y);
  double A = (double)amf / (bga + 0.00000001);
  if (A > 10) A = 1;
  if (A < -10) A = 1;
  const float aa = (float)fabs(A);
  float sb, cb;
  sincos_cw((float)bgp, &sb, &cb);
  const float k = amf > 0.f ? aa * __builtin_amdgcn_rcpf(amf) : 0.f;
  return make_float2(k * (u.x * cb + u.y * sb) - 1.0f, k * (u.y * cb - u.x * sb));
}

// conj(s)/|s| of the specular value, in double then rounded
__device__ __forceinline__ float2 spec_norm_of(const float2* crop_a, int C,
                                               int px, int py) {
  const float2 s = crop_a[py * C + px];
  const double re = s.x, im = s.y;
  const double mod = sqrt(re * re + im * im);
  return make_float2((float)(re / mod), (float)(-im / mod));
}

template <int CT>
__global__ void __launch_bounds__(256)
k_form_field(const float2* __restrict__ iu, const float2* __restrict__ crop,
             const int* __restrict__ posx, const int* __restrict__ posy,
             const double* __restrict__ coef, int mode, const float* __restrict__ phu,
             float2* __restrict__ fields, Dim<CT> C) {
  const int a = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= C * C) return;
  const int X = i % C, Y = i / C;
  const double* cph = coef + (size_t)a * 2 * NCOEF;
  const float2 sn = (mode & 2) ? make_float2(0.f, 0.f)
                               : spec_norm_of(crop + (size_t)a * C * C, C, posx[a], posy[a]);
  const float2 u = iu[((size_t)a * C + (Y + C / 2) % C) * C + (X + C / 2) % C];
  float2 f;
  if (mode & 2) {
    const double bgp = row_poly_at(row_poly(cph, (double)Y), (double)X);
    const double bga = row_poly_at(row_poly(cph + NCOEF, (double)Y), (double)X);
    if (!phu && (mode & 1)) {
      f = form_born_wrapped(u, bgp, bga);
    } else {
      const float phf = phu ? phu[((size_t)a * C + Y) * C + X] : atan2f(u.y, u.x);
      f = form_from(u, phf, bgp, bga, sn, mode);
    }
  } else {
    f = form_from(u, 0.f, 0.0, 0.0, sn, mode);
  }
  fields[(size_t)a * C * C + i] = f;
}

// ---------------------------------------------------------------------------
// per-angle forward spectrum
// G[y][x] = field_U[(y+NX)%C][(x+NX)%C] * w[(y+NX)%C] * w[(x+NX)%C]
// ---------------------------------------------------------------------------
template <int CT>
struct FwdRowOp {
  const float2* src;      // fields (U frame) or IU
  const float2* crop;
  const double* coef;
  const int* posx;
  const int* posy;
  const double* win;      // tukey1d(C, 0.1), double
  float2* out;
  const float2* tw;
  int from_iu, mode;
  const float* phu;  // unwrapped phase (U frame) or nullptr
  float2 sn;
  const double* cph;
  const float* phu_a;
  // per row (line): its U-frame index, window factor and background polys
  int Y;
  double wy;
  RowPoly pp, pa;
  __device__ void begin(int a) {
    cph = coef + (size_t)a * 2 * NCOEF;
    phu_a = phu ? phu + (size_t)a * C * C : nullptr;
    if (from_iu && !(mode & 2))
      sn = spec_norm_of(crop + (size_t)a * C * C, C, posx[a], posy[a]);
  }
  __device__ void line(int, int y) {
    Y = (y + C / 2) % C;
    wy = win[Y];
    if (from_iu && (mode & 2)) {
      pp = row_poly(cph, (double)Y);
      pa = row_poly(cph + NCOEF, (double)Y);
    }
  }
  // windowed field of the IU value u at column x of the line set up by line()
  __device__ float2 form_iu(int x, float2 u) const {
    const int X = (x + C / 2) % C;
    const float w = (float)(win[X] * wy);
    float2 f;
    if (mode & 2) {
      if (!phu_a && (mode & 1)) {
        f = form_born_wrapped(u, row_poly_at(pp, (double)X), row_poly_at(pa, (double)X));
      } else {
        const float phf = phu_a ? phu_a[Y * C + X] : atan2f(u.y, u.x);
        f = form_from(u, phf, row_poly_at(pp, (double)X), row_poly_at(pa, (double)X), sn, mode);
      }
    } else {
      f = form_from(u, 0.f, 0.0, 0.0, sn, mode);
    }
    return make_float2(f.x * w, f.y * w);
  }
  __device__ float2 load(int a, int y, int x) const {
    if (from_iu) return form_iu(x, src[((size_t)a * C + y) * C + x]);
    const int X = (x + C / 2) % C;
    const float w = (float)(win[X] * wy);
    const float2 f = src[((size_t)a * C + Y) * C + X];
    return make_float2(f.x * w, f.y * w);
  }
  __device__ void store(int a, int y, int x, float2 v) const {
    out[((size_t)a * C + y) * C + x] = v;
  }
  Dim<CT> C{};
};

template <int CT>
struct FwdColOp {
  const float2* in;
  float2* out;
  const int* posx;
  const int* posy;
  const float2* tw;
  float dx2;
  int shift_pos;
  // beta2 != nullptr (unshifted output only): each bin is also multiplied
  // by the Ewald factor beta(|fd|^2) of its pupil pixel, as the accumulation
  // would (retroPropag_Born, Reconstruction/src/fonctions.cpp:927: one FP32
  // product, so the weighted value is bit-identical to the one the gather
  // would form); bins outside the pupil, which no hit reads, become 0.
  // beta2 is the C x C table of those factors in the pupil frame.
  const float* beta2;
  int px, py, bx, by;
  __device__ void begin(int a) {
    if (shift_pos) { px = posx[a]; py = posy[a]; } else { px = 0; py = 0; }
    if (beta2) { bx = posx[a]; by = posy[a]; }
  }
  __device__ float2 load(int a, int col, int row) const {
    return in[((size_t)a * C + row) * C + col];
  }
  __device__ void store(int a, int col, int row, float2 v) const {
    float2 s = make_float2(v.x * dx2, v.y * dx2);
    if (beta2) {
      // unshifted bin (row, col) holds pupil pixel fd = wrap((col, row) + fm0),
      // i.e. pupil-frame index ((row + posy) % C, (col + posx) % C)
      const float b = beta2[((row + by) % C) * C + (col + bx) % C];
      s = make_float2(b * s.x, b * s.y);
    }
    out[((size_t)a * C + (row + py) % C) * C + (col + px) % C] = s;
  }
  Dim<CT> C{};
};

// ---------------------------------------------------------------------------
// Fused field stage from Q, the crops' column inverse DFTs (k_holo_cols's
// QOUT form). The demodulation's -posSpec shift (calc_Uborn2) is a shift by
// px along each row of Q on load and, by the shift theorem, the phase ramp
// exp(-2 pi i py k / C) on row k; the row inverse DFT then gives row k of IU.
// QRowOp goes on through the field formation, the alpha = 0.1 window and
// the forward row FFT in the same kernel (k_rows2); SampleRowOp only writes
// the IU rows that hold least-squares samples (U-frame y = 5 j), for the fit
// that the field formation needs first.
// ---------------------------------------------------------------------------
template <int CT>
struct QLoad {
  const float2* q;
  const int* posx;
  const int* posy;
  const float2* tw;
  int px;
  float2 ramp;
  __device__ void begin(int a, int C) {
    (void)C;
    px = posx[a];
  }
  __device__ void set_row(int a, int k, int C) {
    const float2 w = tw[(posy[a] * k) % C];  // exp(-2 pi i py k / C)
    ramp = w;
  }
  __device__ float2 at(int a, int k, int u, int C) const {
    return cmul(q[((size_t)a * C + k) * C + (u + px) % C], ramp);
  }
};

template <int CT>
struct QRowOp {
  QLoad<CT> ql;
  FwdRowOp<CT> f;  // field formation, window and output (from_iu = 1)
  const float2* tw;
  __device__ void begin(int a) {
    ql.begin(a, f.C);
    f.begin(a);
  }
  __device__ void line(int a, int k) { ql.set_row(a, k, f.C); }
  // the field formation's row state (window factor, background polynomials)
  // once the row's inverse transform is done
  __device__ void mid_line(int a, int k) { f.line(a, k); }
  __device__ float2 load(int a, int k, int u) const { return ql.at(a, k, u, f.C); }
  __device__ float2 mid(int, int, int l, float2 iu) const { return f.form_iu(l, iu); }
  __device__ void store(int a, int k, int j, float2 v) const { f.store(a, k, j, v); }
};

template <int CT>
struct SampleRowOp {
  QLoad<CT> ql;
  float2* ius;  // [A][lsq_sample_rows(C)][C]
  const float2* tw;
  int ns;
  Dim<CT> C{};
  int j;
  __device__ void begin(int a) { ql.begin(a, C); }
  __device__ void line(int a, int jj) {
    j = jj;
    ql.set_row(a, (5 * (jj < ns ? jj : 0) + C / 2) % C, C);
  }
  __device__ float2 load(int a, int, int u) const {
    if (j >= ns) return make_float2(0.f, 0.f);
    return ql.at(a, (5 * j + C / 2) % C, u, C);
  }
  __device__ void store(int a, int, int l, float2 v) const {
    if (j < ns) ius[((size_t)a * ns + j) * C + l] = v;
  }
};

// ---------------------------------------------------------------------------
// launchers: compiled crop sides (CT > 0, register FFTs) or any other side
// (CT == 0, the runtime-length FFTs of gen_fft.hpp)
// ---------------------------------------------------------------------------
template <int CT>
static hipError_t argmax_c(Dim<CT> C, const float2* crop, int A, int* posx, int* posy,
                           float2* part, hipStream_t s) {
  {
    Timed tm("argmax", s);
    hipLaunchKernelGGL((k_argmax_part<CT>), dim3(ARGMAX_SEG, A), dim3(256), 0, s, crop, part, C);
  }
  Timed tm("argmax_fin", s);
  hipLaunchKernelGGL((k_argmax_fin<CT>), dim3((A + 63) / 64), dim3(64), 0, s, part, ARGMAX_SEG, A,
                     posx, posy, C);
  return hipGetLastError();
}

template <int CT>
static hipError_t argmax_fin_c(Dim<CT> C, const float2* part, int S, int A, int* posx, int* posy,
                               hipStream_t s) {
  Timed tm("argmax_fin", s);
  hipLaunchKernelGGL((k_argmax_fin<CT>), dim3((A + 63) / 64), dim3(64), 0, s, part, S, A, posx,
                     posy, C);
  return hipGetLastError();
}

size_t argmax_scratch_bytes(int A) { return (size_t)(A > 0 ? A : 1) * ARGMAX_SEG * sizeof(float2); }

// one row pass and one column pass of op types R and K over A batches
template <int CT, bool INV, class R, class K>
static hipError_t rows_cols(Dim<CT> C, const R& r, const K& c, int A, const FftTab& ft,
                            const char* rname, const char* cname, hipStream_t s) {
  if constexpr (CT > 0) {
    {
      Timed tm(rname, s);
      hipLaunchKernelGGL((k_rows<CT, INV, R>), rows_grid<CT>(A), rows_block<CT>(), 0, s, r);
    }
    Timed tm(cname, s);
    hipLaunchKernelGGL((k_cols<CT, INV, K>), (cols_grid<CT, K>(A)), (cols_block<CT, K>()), 0, s,
                       c);
    return hipGetLastError();
  } else {
    if (!ft.gen) return hipErrorInvalidValue;
    hipError_t e;
    {
      Timed tm(rname, s);
      e = launch_gen<INV, true>(r, *ft.gen, C, A, s);
    }
    if (e != hipSuccess) return e;
    Timed tm(cname, s);
    return launch_gen<INV, false>(c, *ft.gen, C, A, s);
  }
}

template <int CT>
static hipError_t demod_c(Dim<CT> C, const float2* crop, const int* posx, const int* posy,
                          int A, const FftTab& ft, float2* work, float2* iu,
                          hipStream_t s) {
  DemodRowOp<CT> r{crop, posx, posy, work, ft.tw, 0, 0, C};
  ColInPlaceOp<CT> c{work, iu, ft.tw, C};
  return rows_cols<CT, true>(C, r, c, A, ft, "demod_rows", "demod_cols", s);
}

template <int CT>
static hipError_t lsq_c(Dim<CT> C, const float2* iu, const float* phu, int A,
                        const LsqPlan& plan, double* coef, hipStream_t s, int sample_rows = 0) {
  Timed tm("lsq", s);
  hipLaunchKernelGGL((k_lsq<CT>), dim3(A), dim3(256), 0, s, iu, phu, plan.d_samples,
                     plan.M, plan.d_lu, plan.d_perm, coef, C, sample_rows);
  return hipGetLastError();
}

template <int CT>
static hipError_t form_c(Dim<CT> C, const float2* iu, const float2* crop, const int* posx,
                         const int* posy, int A, const double* coef, int mode,
                         const float* phu, float2* fields, hipStream_t s) {
  Timed tm("form_field", s);
  hipLaunchKernelGGL((k_form_field<CT>), dim3((C * C + 255) / 256, A), dim3(256),
                     0, s, iu, crop, posx, posy, coef, mode, phu, fields, C);
  return hipGetLastError();
}

template <int CT>
static hipError_t fwd_c(Dim<CT> C, const float2* src, int from_iu, const float2* crop,
                        const double* coef, int mode, const float* phu, const int* posx,
                        const int* posy, int A, const double* win,
                        const FftTab& ft, float dx2, int shift_pos, const float* beta2,
                        float2* work, float2* out, hipStream_t s) {
  FwdRowOp<CT> r{src, crop, coef, posx, posy, win, work, ft.tw, from_iu, mode, phu,
                 make_float2(0.f, 0.f), nullptr, nullptr};
  r.C = C;
  FwdColOp<CT> c{work, out, posx, posy, ft.tw, dx2, shift_pos, beta2, 0, 0, 0, 0, C};
  return rows_cols<CT, false>(C, r, c, A, ft, from_iu ? "fwd_rows_field" : "fwd_rows",
                              "fwd_cols", s);
}

template <int CT>
static hipError_t synth_fields_c(Dim<CT> C, const float2* spec, const int* posx, const int* posy,
                                 int A, const FftTab& ft, float2* work, float scale,
                                 float2* fields, hipStream_t s) {
  DemodRowOp<CT> r{spec, posx, posy, work, ft.tw, 0, 0, C};
  ColCenterScaleOp<CT> c{work, fields, ft.tw, scale, C};
  return rows_cols<CT, true>(C, r, c, A, ft, "synth_rows", "synth_cols", s);
}

#define MTD_CROP_SIZES(X) \
  X(16) X(32) X(64) X(128) X(160) X(192) X(224) X(250) X(256) X(288) X(320) X(352) X(384) \
  X(512)

hipError_t launch_argmax(int C, const float2* crop, int A, int* posx,
                         int* posy, float2* part, hipStream_t s) {
  switch (C) {
#define X(n) case n: return argmax_c<n>(Dim<n>{}, crop, A, posx, posy, part, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return argmax_c<0>(Dim<0>(C), crop, A, posx, posy, part, s);
}

hipError_t launch_argmax_fin(int C, const float2* part, int S, int A, int* posx, int* posy,
                             hipStream_t s) {
  switch (C) {
#define X(n) case n: return argmax_fin_c<n>(Dim<n>{}, part, S, A, posx, posy, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return argmax_fin_c<0>(Dim<0>(C), part, S, A, posx, posy, s);
}

hipError_t launch_spectra_to_fields(int C, const float2* spec, const int* posx, const int* posy,
                                    int A, const FftTab& ft, float2* work, float scale,
                                    float2* fields, hipStream_t s) {
  switch (C) {
#define X(n) \
  case n: return synth_fields_c<n>(Dim<n>{}, spec, posx, posy, A, ft, work, scale, fields, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return synth_fields_c<0>(Dim<0>(C), spec, posx, posy, A, ft, work, scale, fields, s);
}

hipError_t launch_demod(int C, const float2* crop, const int* posx,
                        const int* posy, int A, const FftTab& ft,
                        float2* work, float2* iu, hipStream_t s) {
  switch (C) {
#define X(n) case n: return demod_c<n>(Dim<n>{}, crop, posx, posy, A, ft, work, iu, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return demod_c<0>(Dim<0>(C), crop, posx, posy, A, ft, work, iu, s);
}

// IU frame to the U frame of calc_Uborn2's output: U[y][x] =
// IU[(y + C/2) % C][(x + C/2) % C]
__global__ void __launch_bounds__(256)
k_iu_to_u(const float2* __restrict__ iu, int C, float2* __restrict__ u) {
  const int a = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= C * C) return;
  const int x = i % C, y = i / C, h = C / 2;
  u[(size_t)a * C * C + i] = iu[((size_t)a * C + (y + h) % C) * C + (x + h) % C];
}

hipError_t launch_iu_to_u(int C, const float2* iu, int A, float2* u, hipStream_t s) {
  if (A <= 0) return hipSuccess;
  Timed tm("iu_to_u", s);
  hipLaunchKernelGGL(k_iu_to_u, dim3((C * C + 255) / 256, A), dim3(256), 0, s, iu, C, u);
  return hipGetLastError();
}

hipError_t launch_lsq(int C, const float2* iu, const float* phu, int A, const LsqPlan& plan,
                      double* coef, hipStream_t s) {
  switch (C) {
#define X(n) case n: return lsq_c<n>(Dim<n>{}, iu, phu, A, plan, coef, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return lsq_c<0>(Dim<0>(C), iu, phu, A, plan, coef, s);
}

hipError_t launch_form_field(int C, const float2* iu, const float2* crop,
                             const int* posx, const int* posy, int A,
                             const double* coef, int mode, const float* phu, float2* fields,
                             hipStream_t s) {
  switch (C) {
#define X(n) \
  case n: return form_c<n>(Dim<n>{}, iu, crop, posx, posy, A, coef, mode, phu, fields, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return form_c<0>(Dim<0>(C), iu, crop, posx, posy, A, coef, mode, phu, fields, s);
}

hipError_t launch_fwd_spectrum(int C, const float2* src, int from_iu,
                               const float2* crop, const double* coef,
                               int mode, const float* phu, const int* posx, const int* posy,
                               int A, const double* win, const FftTab& ft,
                               float dx2, int shift_pos, const float* beta2,
                               float2* work, float2* out, hipStream_t s) {
  switch (C) {
#define X(n)                                                                          \
  case n:                                                                             \
    return fwd_c<n>(Dim<n>{}, src, from_iu, crop, coef, mode, phu, posx, posy, A, win, \
                    ft, dx2, shift_pos, beta2, work, out, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return fwd_c<0>(Dim<0>(C), src, from_iu, crop, coef, mode, phu, posx, posy, A, win, ft, dx2,
                  shift_pos, beta2, work, out, s);
}

// the fused field stage from Q (compiled crop sides): the sample rows, the
// fit, then one row kernel to the forward row FFT and the column pass
template <int C>
static hipError_t fused_c(const float2* q, const int* posx, const int* posy, int A,
                          const LsqPlan& plan, double* coef, float2* ius, int mode,
                          const double* win, const float2* tw, float dx2, const float* beta2,
                          float2* work, float2* out, hipStream_t s) {
  const int ns = lsq_sample_rows(C);
  SampleRowOp<C> sr{QLoad<C>{q, posx, posy, tw, 0, make_float2(0.f, 0.f)}, ius, tw, ns};
  {
    Timed tm("sample_rows", s);
    constexpr int NT = rows_nt<C>();
    hipLaunchKernelGGL((k_rows<C, true, SampleRowOp<C>>), dim3((ns + NT - 1) / NT, A),
                       rows_block<C>(), 0, s, sr);
  }
  hipError_t e = lsq_c<C>(Dim<C>{}, ius, nullptr, A, plan, coef, s, 1);
  if (e != hipSuccess) return e;
  FwdRowOp<C> f{nullptr, nullptr, coef, posx, posy, win, work, tw, 1, mode, nullptr,
                make_float2(0.f, 0.f), nullptr, nullptr};
  QRowOp<C> r{QLoad<C>{q, posx, posy, tw, 0, make_float2(0.f, 0.f)}, f, tw};
  {
    Timed tm("fwd_rows_field", s);
    hipLaunchKernelGGL((k_rows2<C, true, false, QRowOp<C>>), rows_grid<C>(A), rows_block<C>(), 0,
                       s, r);
  }
  FwdColOp<C> c{work, out, posx, posy, tw, dx2, 0, beta2, 0, 0, 0, 0};
  Timed tm("fwd_cols", s);
  hipLaunchKernelGGL((k_cols<C, false, FwdColOp<C>>), (cols_grid<C, FwdColOp<C>>(A)),
                     (cols_block<C, FwdColOp<C>>()), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_fused_field(int C, const float2* q, const int* posx, const int* posy, int A,
                              const LsqPlan& plan, double* coef, float2* ius, int mode,
                              const double* win, const FftTab& ft, float dx2,
                              const float* beta2, float2* work, float2* out, hipStream_t s) {
  switch (C) {
#define X(n)                                                                                \
  case n:                                                                                   \
    return fused_c<n>(q, posx, posy, A, plan, coef, ius, mode, win, ft.tw, dx2, beta2, work, \
                      out, s);
    MTD_CROP_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

bool crop_len_supported(int C) {
  switch (C) {
#define X(n) case n: return true;
    MTD_CROP_SIZES(X)
#undef X
  }
  return false;
}

}  // namespace mtd
