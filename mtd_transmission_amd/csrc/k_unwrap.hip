// Phase unwrapping, DEROUL 1 with VOLKOV 1 (deroul_volkov4_total_sym_paire,
// pretraitement_hors_axe_CPU/src/deroulement_volkov4.cpp:14-92).
//
// For each angle the wrapped phase phi = atan2(U) of the C x C demodulated
// field is mirror-extended to D x D (D = 2C, Symetrise_mirror :221-240), the
// FFT gradients of phi_s and of Z = exp(i phi_s) give the gradient of the
// field of 2 pi multiples, g = (Re(-i grad Z / Z) - Re(grad phi_s)) / 2 pi,
// which is integrated by the FFT least-squares (Poisson) solve of integ_grad4
// (:162-191); the original quadrant of the result times 2 pi is added to phi.
//
// GPU restructuring (same mathematics):
//  - a gradient along x only involves x frequencies: the reference's 2-D
//    forward FFT (scaled 1/D^2), multiplication by 2 pi i kx and unnormalised
//    2-D backward FFT equal a 1-D forward FFT along each row, 2 pi i kx / D,
//    and a 1-D backward FFT. k_grad_rows / k_grad_cols do both transforms in
//    one kernel per line, for phi_s and Z together;
//  - the two real gradient fields are packed as gx + i gy into one 2-D
//    forward FFT and separated with the Hermitian symmetry inside the
//    integration's first inverse pass;
//  - only the [C,2C)^2 quadrant of the integral is needed: the inverse row
//    pass keeps the right half, the inverse column pass the lower half.
// The mirror images are index arithmetic on load (no D x D phase is stored).
// The reference's zero frequency reads one element past its array (:189);
// it only adds a constant that the phase fit always removes; here it is 0.
#include "dft.hpp"
#include "gen_fft.hpp"
#include "internal.hpp"
#include "passes.hpp"

namespace mtd {

// source index in [0, C) of position i in [0, 2C) of the even extension
__device__ __forceinline__ int mirror_src(int i, int C) {
  return i < C ? C - 1 - i : i - C;
}
template <int C>
__device__ __forceinline__ int mirror_src(int i) {
  return mirror_src(i, C);
}

// wrapped phase of U[y][x], U held in the IU frame
__device__ __forceinline__ float wrapped_phase(const float2* __restrict__ iu_a, int y, int x,
                                               int C) {
  const float2 u = iu_a[((y + C / 2) % C) * C + (x + C / 2) % C];
  return atan2f(u.y, u.x);
}
template <int C>
__device__ __forceinline__ float wrapped_phase(const float2* __restrict__ iu_a, int y, int x) {
  return wrapped_phase(iu_a, y, x, C);
}

// frequency in cycles per pixel of FFT index k (init_kvect_shift,
// FFT_fonctions.cpp:7-17; the Nyquist index gets -1/2)
__device__ __forceinline__ float kfreq(int k, int D) {
  return (float)(((k + D / 2) % D) - D / 2) / (float)D;
}
template <int D>
__device__ __forceinline__ float kfreq(int k) {
  return kfreq(k, D);
}

// d/dx of one line held in the two-step layout: v (input layout) in, u
// (threads n2 < L1 hold positions n2 + L1*k2) out. Forward FFT, times
// 2 pi i k / D, back to the input layout through the exchange storage x,
// backward FFT.
template <int D, class X>
__device__ __forceinline__ void line_derivative(X x, const float2* __restrict__ tw, int n2,
                                                float2 (&v)[Fft2<D, false>::L1],
                                                float2 (&u)[Fft2<D, false>::L2]) {
  using F = Fft2<D, false>;
  using B = Fft2<D, true>;
  const float two_pi = (float)(2.0 * 3.14159265358979323846);
  F::run(v, u, n2, x, tw);
  if (n2 < F::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) {
      const float f = two_pi * kfreq<D>(n2 + F::L1 * k2) / (float)D;
      u[k2] = make_float2(-f * u[k2].y, f * u[k2].x);
    }
  }
  fft2_relayout<D>(u, v, n2, x);
  B::run(v, u, n2, x, tw);
}

// Gradient along one line of the extension: at the outputs held by threads
// n2 < L1 (positions n2 + L1*k2), g = (Im(dZ / Z) - Re(dphi_s)) / (2 pi),
// Z = exp(i phi_s). `pos(i)` maps line position i to the (y, x) of the U
// image; x is this line's exchange storage. The two derivatives run one
// after the other so that only the real part of the first stays live.
template <int D, class Pos, class X>
__device__ __forceinline__ void line_gradient(const float2* __restrict__ iu_a, Pos pos, int n2,
                                              X x, const float2* __restrict__ tw,
                                              float (&g)[Fft2<D, false>::L2]) {
  using F = Fft2<D, false>;
  constexpr int C = D / 2;
  const float two_pi = (float)(2.0 * 3.14159265358979323846);
  float dphi[F::L2];
  {
    float2 v[F::L1], u[F::L2];
#pragma unroll
    for (int n1 = 0; n1 < F::L1; ++n1) {
      int y, xx;
      pos(F::L2 * n1 + n2, y, xx);
      v[n1] = make_float2(wrapped_phase<C>(iu_a, y, xx), 0.f);
    }
    line_derivative<D>(x, tw, n2, v, u);
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) dphi[k2] = u[k2].x;
  }
  __syncthreads();  // the exchange storage is reused
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1) {
    int y, xx;
    pos(F::L2 * n1 + n2, y, xx);
    float zs, zc;
    sincosf(wrapped_phase<C>(iu_a, y, xx), &zs, &zc);
    v[n1] = make_float2(zc, zs);
  }
  line_derivative<D>(x, tw, n2, v, u);
  if (n2 < F::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) {
      int y, xx;
      pos(n2 + F::L1 * k2, y, xx);
      float zi, zr;
      sincosf(wrapped_phase<C>(iu_a, y, xx), &zi, &zr);
      // -i (dZ / Z): its real part is Im(dZ / Z)
      const float2 dz = u[k2];
      const float im_q = (dz.y * zr - dz.x * zi) / (zr * zr + zi * zi);
      g[k2] = (im_q - dphi[k2]) / two_pi;
    }
  }
}

// rows of the extension: gx[a][r][c]
template <int C>
__global__ void __launch_bounds__(256)
k_grad_rows(const float2* __restrict__ iu, const float2* __restrict__ tw, float* __restrict__ gx) {
  constexpr int D = 2 * C;
  using F = Fft2<D, false>;
  constexpr int NT = rows_nt<D>();
  __shared__ float lds[NT * F::LDS_ROW];
  const int2 ab = xcd_block();
  const int a = ab.x;
  const int t = threadIdx.x / F::T, n2 = threadIdx.x % F::T;
  const int r = ab.y * NT + t;
  const float2* iu_a = iu + (size_t)a * C * C;
  const int ys = mirror_src<C>(r);
  auto pos = [ys](int i, int& y, int& x) { y = ys; x = mirror_src<C>(i); };
  float g[F::L2];
  line_gradient<D>(iu_a, pos, n2, typename F::Row{lds + t * F::LDS_ROW}, tw, g);
  if (n2 < F::L1) {
    float* o = gx + ((size_t)a * D + r) * D;
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) o[n2 + F::L1 * k2] = g[k2];
  }
}

// columns per block of the gradient kernels: two transforms' registers
// are live, so at most 256 threads
template <int D>
constexpr int grad_nc() {
  constexpr int by_thr = 256 / Fft2<D, false>::T;
  constexpr int by_lds = kLdsBlockMax / (Fft2<D, false>::LDS_COL * 4);
  return divisor_le(D, by_thr < by_lds ? by_thr : by_lds);
}

// columns of the extension: gy[a][r][c]; NC adjacent columns per block
template <int C>
__global__ void __launch_bounds__(256)
k_grad_cols(const float2* __restrict__ iu, const float2* __restrict__ tw, float* __restrict__ gy) {
  constexpr int D = 2 * C;
  using F = Fft2<D, false>;
  constexpr int NC = grad_nc<D>();
  __shared__ float lds[NC * F::LDS_COL];
  const int2 ab = xcd_block();
  const int a = ab.x;
  const int cl = threadIdx.x % NC, n2 = threadIdx.x / NC;
  const int c = ab.y * NC + cl;
  const float2* iu_a = iu + (size_t)a * C * C;
  const int xs = mirror_src<C>(c);
  auto pos = [xs](int i, int& y, int& x) { y = mirror_src<C>(i); x = xs; };
  float g[F::L2];
  line_gradient<D>(iu_a, pos, n2, typename F::template Col<NC>{lds + cl}, tw, g);
  if (n2 < F::L1) {
    float* o = gy + (size_t)a * D * D + c;
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) o[(size_t)(n2 + F::L1 * k2) * D] = g[k2];
  }
}

// integration pass 1: rows of gx + i gy, forward
template <int DT>
struct PackRowOp {
  const float* gx;
  const float* gy;
  float2* out;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int r, int c) const {
    const size_t i = ((size_t)a * D + r) * D + c;
    return make_float2(gx[i], gy[i]);
  }
  __device__ void store(int a, int r, int k, float2 v) const { out[((size_t)a * D + r) * D + k] = v; }
  Dim<DT> D{};
};

// integration pass 2: columns, forward, in place
template <int DT>
struct InPlaceColOp {
  float2* buf;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int c, int r) const { return buf[((size_t)a * D + r) * D + c]; }
  __device__ void store(int a, int c, int r, float2 v) const { buf[((size_t)a * D + r) * D + c] = v; }
  Dim<DT> D{};
};

// integration pass 3: P = FFT2(gx + i gy) (unnormalised). Per frequency,
// TFx = (P[k] + conj(P[-k])) / (2 D^2), TFy = (P[k] - conj(P[-k])) / (2i D^2),
// T = -i (TFx kx + TFy ky) / (2 pi |k|^2), T[0] = 0; backward along rows,
// keeping x in [C, 2C): Q[a][ky][x - C].
template <int DT>
struct PoissonRowOp {
  const float2* P;
  float2* Q;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int ky, int kx) const {
    if (ky == 0 && kx == 0) return make_float2(0.f, 0.f);
    const float2* Pa = P + (size_t)a * D * D;
    const float2 p = Pa[(size_t)ky * D + kx];
    const float2 m = Pa[(size_t)((D - ky) % D) * D + (D - kx) % D];
    const float inv = 1.0f / (2.0f * (float)D * (float)D);
    // TFx = (p + conj(m)) * inv, TFy = (p - conj(m)) * inv / i
    const float2 fx = make_float2((p.x + m.x) * inv, (p.y - m.y) * inv);
    const float2 fy = make_float2((p.y + m.y) * inv, -(p.x - m.x) * inv);
    const float kxf = kfreq(kx, D), kyf = kfreq(ky, D);
    const float2 num = make_float2(fx.x * kxf + fy.x * kyf, fx.y * kxf + fy.y * kyf);
    const float den = (float)(2.0 * 3.14159265358979323846) * (kxf * kxf + kyf * kyf);
    return make_float2(num.y / den, -num.x / den);  // -i num / den
  }
  __device__ void store(int a, int ky, int x, float2 v) const {
    const int C = D / 2;
    if (x >= C) Q[((size_t)a * D + ky) * C + (x - C)] = v;
  }
  Dim<DT> D{};
};

// integration pass 4: backward along columns of Q, keeping y in [C, 2C);
// unwrapped phase = phi + 2 pi Re(.), in the U frame: phu[a][y][x]
template <int DT>
struct UnwrapColOp {
  const float2* Q;
  const float2* iu;
  float* phu;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int c, int ky) const {
    const int C = D / 2;
    if (c >= C) return make_float2(0.f, 0.f);
    return Q[((size_t)a * D + ky) * C + c];
  }
  __device__ void store(int a, int c, int y, float2 v) const {
    const int C = D / 2;
    if (c >= C || y < C) return;
    const int yy = y - C;
    const float ph = wrapped_phase(iu + (size_t)a * C * C, yy, c, C);
    phu[((size_t)a * C + yy) * C + c] = ph + (float)(2.0 * 3.14159265358979323846) * v.x;
  }
  Dim<DT> D{};
};

template <int C>
static hipError_t unwrap_c(const float2* iu, int A, const float2* tw, float* gx, float* gy,
                           float2* P, float2* Q, float* phu, hipStream_t s) {
  constexpr int D = 2 * C;
  {
    Timed tm("unwrap_grad", s);
    hipLaunchKernelGGL((k_grad_rows<C>), rows_grid<D>(A), rows_block<D>(), 0, s, iu, tw, gx);
    constexpr int GNC = grad_nc<D>();
    hipLaunchKernelGGL((k_grad_cols<C>), dim3(D / GNC, A), dim3(GNC * Fft2<D, false>::T), 0, s,
                       iu, tw, gy);
  }
  {
    Timed tm("unwrap_fwd", s);
    PackRowOp<D> r{gx, gy, P, tw};
    hipLaunchKernelGGL((k_rows<D, false, PackRowOp<D>>), rows_grid<D>(A), rows_block<D>(), 0, s, r);
    InPlaceColOp<D> c{P, tw};
    hipLaunchKernelGGL((k_cols<D, false, InPlaceColOp<D>>), cols_grid<D>(A), cols_block<D>(), 0,
                       s, c);
  }
  Timed tm("unwrap_inv", s);
  PoissonRowOp<D> pr{P, Q, tw};
  hipLaunchKernelGGL((k_rows<D, true, PoissonRowOp<D>>), rows_grid<D>(A), rows_block<D>(), 0, s,
                     pr);
  // only the C columns of Q: ceil(C / NC) column groups
  constexpr int NC = cols_nc<D>();
  UnwrapColOp<D> uc{Q, iu, phu, tw};
  hipLaunchKernelGGL((k_cols<D, true, UnwrapColOp<D>>), dim3((C + NC - 1) / NC, A),
                     cols_block<D>(), 0, s, uc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Crop sides without compiled 2C-point FFTs: the same four steps with the
// runtime-length FFTs. The gradient kernel holds nl lines of the extension
// in LDS and runs both derivatives there (forward transform, times
// 2 pi i k / D^2, backward transform); the first derivative's real part goes
// to the output array and is read back by the same thread for the second.
// ---------------------------------------------------------------------------
template <bool ROWS>
__global__ void __launch_bounds__(kGenThreads)
k_grad_gen(const float2* __restrict__ iu, GenPlan p, int C, float* __restrict__ g, int nl) {
  extern __shared__ float2 gen_lds[];
  const int D = 2 * C;
  const int a = blockIdx.y, line0 = blockIdx.x * nl;
  const int nact = min(nl, D - line0);
  float2* A = gen_lds;
  float2* B = gen_lds + (size_t)nl * p.W;
  const float2* iu_a = iu + (size_t)a * C * C;
  const float two_pi = (float)(2.0 * 3.14159265358979323846);
  auto at = [&](int t, int& l, int& i) {
    if (ROWS) { l = t / D; i = t - l * D; } else { i = t / nl; l = t - i * nl; }
  };
  auto phase = [&](int l, int i) {
    const int line = line0 + l;
    return ROWS ? wrapped_phase(iu_a, mirror_src(line, C), mirror_src(i, C), C)
                : wrapped_phase(iu_a, mirror_src(i, C), mirror_src(line, C), C);
  };
  auto out = [&](int l, int i) -> size_t {
    return ROWS ? ((size_t)a * D + line0 + l) * D + i : ((size_t)a * D + i) * D + line0 + l;
  };
  for (int pass = 0; pass < 2; ++pass) {  // 0: phi_s, 1: Z = exp(i phi_s)
    for (int t = threadIdx.x; t < nl * D; t += blockDim.x) {
      int l, i;
      at(t, l, i);
      float2 v = make_float2(0.f, 0.f);
      if (l < nact) {
        const float ph = phase(l, i);
        if (pass == 0) {
          v.x = ph;
        } else {
          sincosf(ph, &v.y, &v.x);
        }
      }
      A[l * p.W + i] = v;
    }
    __syncthreads();
    float2* R = gen_fft<false>(A, B, nl, p);
    for (int t = threadIdx.x; t < nl * D; t += blockDim.x) {
      const int l = t / D, k = t - l * D;
      const float f = two_pi * kfreq(k, D) / (float)D;
      const float2 u = R[l * p.W + k];
      R[l * p.W + k] = make_float2(-f * u.y, f * u.x);
    }
    __syncthreads();
    const float2* R2 = gen_fft<true>(R, R == A ? B : A, nl, p);
    for (int t = threadIdx.x; t < nl * D; t += blockDim.x) {
      int l, i;
      at(t, l, i);
      if (l >= nact) continue;
      if (pass == 0) {
        g[out(l, i)] = R2[l * p.W + i].x;  // d phi_s
      } else {
        float zi, zr;
        sincosf(phase(l, i), &zi, &zr);
        const float2 dz = R2[l * p.W + i];
        const float im_q = (dz.y * zr - dz.x * zi) / (zr * zr + zi * zi);
        const size_t o = out(l, i);
        g[o] = (im_q - g[o]) / two_pi;
      }
    }
    __syncthreads();  // A and B are reused
  }
}

static hipError_t unwrap_gen(int C, const float2* iu, int A, const GenPlan& p, float* gx,
                             float* gy, float2* P, float2* Q, float* phu, hipStream_t s) {
  const int D = 2 * C;
  const Dim<0> d(D);
  {
    Timed tm("unwrap_grad", s);
    const int nl = gen_lines_per_block(p.W);
    const size_t lds = (size_t)2 * nl * p.W * sizeof(float2);
    const dim3 grid((D + nl - 1) / nl, A);
    hipLaunchKernelGGL((k_grad_gen<true>), grid, dim3(kGenThreads), lds, s, iu, p, C, gx, nl);
    hipLaunchKernelGGL((k_grad_gen<false>), grid, dim3(kGenThreads), lds, s, iu, p, C, gy, nl);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipError_t e;
  {
    Timed tm("unwrap_fwd", s);
    e = launch_gen<false, true>(PackRowOp<0>{gx, gy, P, nullptr, d}, p, D, A, s);
    if (e == hipSuccess) e = launch_gen<false, false>(InPlaceColOp<0>{P, nullptr, d}, p, D, A, s);
  }
  if (e != hipSuccess) return e;
  Timed tm("unwrap_inv", s);
  e = launch_gen<true, true>(PoissonRowOp<0>{P, Q, nullptr, d}, p, D, A, s);
  if (e != hipSuccess) return e;
  return launch_gen<true, false>(UnwrapColOp<0>{Q, iu, phu, nullptr, d}, p, C, A, s);
}

#define MTD_UNWRAP_SIZES(X) \
  X(16) X(32) X(64) X(128) X(160) X(192) X(224) X(250) X(256) X(288) X(320) X(352) X(384) \
  X(512)

hipError_t launch_unwrap(int C, const float2* iu, int A, const FftTab& ft2C, float* gx,
                         float* gy, float2* P, float2* Q, float* phu, hipStream_t s) {
  switch (C) {
#define X(n) case n: return unwrap_c<n>(iu, A, ft2C.tw, gx, gy, P, Q, phu, s);
    MTD_UNWRAP_SIZES(X)
#undef X
  }
  if (!ft2C.gen) return hipErrorInvalidValue;
  return unwrap_gen(C, iu, A, *ft2C.gen, gx, gy, P, Q, phu, s);
}

bool unwrap_len_supported(int C) {
  switch (C) {
#define X(n) case n: return true;
    MTD_UNWRAP_SIZES(X)
#undef X
  }
  return false;
}

size_t unwrap_scratch_bytes(int C, int A) {
  const size_t D = 2 * (size_t)C, n = (size_t)(A > 0 ? A : 1);
  // gx, gy (float D^2), P (float2 D^2), Q (float2 D*C)
  return n * (2 * D * D * sizeof(float) + D * D * sizeof(float2) + D * C * sizeof(float2));
}

}  // namespace mtd
