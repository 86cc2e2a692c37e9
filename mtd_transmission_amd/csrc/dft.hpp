// Complex helpers and small in-register DFT kernels (radix 2, 3, 4, 5, 7, 8,
// 11, 13) used by the register FFTs of fft_reg.hpp and the runtime-length
// FFTs of gen_fft.hpp.
//
// Sign convention (FFTW): forward  X[k] = sum x[n] exp(-2 pi i n k / L),
//                         backward X[k] = sum x[n] exp(+2 pi i n k / L),
// both unnormalised.
//
// Twiddle tables tw[k] = exp(-2 pi i k / L), k < L, are computed on the host
// in double precision and rounded once to float.
#pragma once
#include <hip/hip_runtime.h>

namespace mtd {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by -i (forward) or +i (backward)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Radix-R DFT in registers.
// ---------------------------------------------------------------------------
template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]), t3 = mul_mi<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV>
struct Dft<8, INV> {
  __device__ __forceinline__ static void run(float2* a) {
    constexpr float h = 0.70710678118654752440f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, INV>::run(e);
    Dft<4, INV>::run(o);
    // o[k] *= W8^k (forward W8 = exp(-i pi/4))
    float2 o1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                    : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    float2 o2 = mul_mi<INV>(o[2]);
    float2 o3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                    : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  }
};

// Odd prime radices: y_p = a0 + sum_q (a_q + a_{R-q}) cos(2 pi p q / R)
//                          -/+ i sum_q (a_q - a_{R-q}) sin(2 pi p q / R).
template <int R>
struct OddConst;
template <>
struct OddConst<3> {
  __device__ static constexpr float c(int k) { return k == 1 ? -0.5f : 0.f; }
  __device__ static constexpr float s(int k) { return k == 1 ? 0.86602540378443865f : 0.f; }
};
template <>
struct OddConst<5> {
  __device__ static constexpr float c(int k) {
    return k == 1 ? 0.30901699437494742f : (k == 2 ? -0.80901699437494742f : 0.f);
  }
  __device__ static constexpr float s(int k) {
    return k == 1 ? 0.95105651629515357f : (k == 2 ? 0.58778525229247313f : 0.f);
  }
};
template <>
struct OddConst<7> {
  __device__ static constexpr float c(int k) {
    return k == 1 ? 0.62348980185873353f
                  : (k == 2 ? -0.22252093395631440f : (k == 3 ? -0.90096886790241913f : 0.f));
  }
  __device__ static constexpr float s(int k) {
    return k == 1 ? 0.78183148246802981f
                  : (k == 2 ? 0.97492791218182361f : (k == 3 ? 0.43388373911755812f : 0.f));
  }
};
template <>
struct OddConst<11> {
  __device__ static constexpr float c(int k) {
    return k == 1 ? 0.84125353283118117f
         : k == 2 ? 0.41541501300188643f
         : k == 3 ? -0.14231483827328514f
         : k == 4 ? -0.65486073394528506f
         : k == 5 ? -0.95949297361449739f : 0.f;
  }
  __device__ static constexpr float s(int k) {
    return k == 1 ? 0.54064081745559756f
         : k == 2 ? 0.90963199535451837f
         : k == 3 ? 0.98982144188093273f
         : k == 4 ? 0.75574957435425828f
         : k == 5 ? 0.28173255684142969f : 0.f;
  }
};

template <>
struct OddConst<13> {
  __device__ static constexpr float c(int k) {
    return k == 1 ? 0.88545602565320990f
         : k == 2 ? 0.56806474673115581f
         : k == 3 ? 0.12053668025532300f
         : k == 4 ? -0.35460488704253545f
         : k == 5 ? -0.74851074817110120f
         : k == 6 ? -0.97094181742605200f : 0.f;
  }
  __device__ static constexpr float s(int k) {
    return k == 1 ? 0.46472317204376850f
         : k == 2 ? 0.82298386589365640f
         : k == 3 ? 0.99270887409805400f
         : k == 4 ? 0.93501624268541480f
         : k == 5 ? 0.66312265824079520f
         : k == 6 ? 0.23931566428755768f : 0.f;
  }
};

template <int R, bool INV>
struct DftOdd {
  __device__ __forceinline__ static void run(float2* a) {
    constexpr int H = (R - 1) / 2;
    float2 sp[H], sm[H];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      sp[q - 1] = cadd(a[q], a[R - q]);
      sm[q - 1] = csub(a[q], a[R - q]);
    }
    float2 y0 = a[0];
#pragma unroll
    for (int q = 0; q < H; ++q) y0 = cadd(y0, sp[q]);
    float2 out[R];
    out[0] = y0;
#pragma unroll
    for (int p = 1; p <= H; ++p) {
      float2 re = a[0];
      float2 im = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        int k = (p * q) % R;
        float c = OddConst<R>::c(k <= H ? k : R - k);
        float s = k <= H ? OddConst<R>::s(k) : -OddConst<R>::s(R - k);
        re.x += sp[q - 1].x * c;
        re.y += sp[q - 1].y * c;
        im.x += sm[q - 1].x * s;
        im.y += sm[q - 1].y * s;
      }
      // forward: y_p = re - i*im, y_{R-p} = re + i*im
      float2 mi = mul_mi<INV>(im);
      out[p] = cadd(re, mi);
      out[R - p] = csub(re, mi);
    }
#pragma unroll
    for (int p = 0; p < R; ++p) a[p] = out[p];
  }
};
template <bool INV> struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};
template <bool INV> struct Dft<11, INV> : DftOdd<11, INV> {};
template <bool INV> struct Dft<13, INV> : DftOdd<13, INV> {};

// Largest d <= cap dividing n (at least 1).
__host__ __device__ constexpr int divisor_le(int n, int cap) {
  for (int d = cap; d > 1; --d)
    if (n % d == 0) return d;
  return 1;
}

}  // namespace mtd
