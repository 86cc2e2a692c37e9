// Register-resident FFTs for gfx950.
//
// A length-L transform is split as L = L1*L2 (L1 <= L2, L2 the smallest
// divisor >= sqrt L) and run by L2 threads:
//   step 1  thread n2 holds x[L2*n1 + n2], n1 < L1 (a strided subsequence,
//           loaded straight from global memory: coalesced across threads),
//           and does an L1-point DFT in registers;
//   twiddle Y[n2][k1] *= W_L^(n2*k1) (table);
//   exchange one padded LDS transpose;
//   step 2  thread k1 < L1 does the L2-point DFT over n2 in registers and
//           holds X[k1 + L1*k2], k2 < L2 (stored straight to global memory).
// The in-register DFTs are mixed-radix decimation in time, fully unrolled at
// compile time with twiddle constants evaluated by constexpr trigonometry.
//
// Sign convention: forward exp(-2 pi i n k / L), backward exp(+...), both
// unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dft.hpp"

namespace mtd {

// ---------------------------------------------------------------------------
// constexpr cos/sin of 2*pi*k/n (exact octant reduction, Taylor series)
// ---------------------------------------------------------------------------
constexpr double kQuarterPi = 0.785398163397448309615660845819875721;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 12; ++i) {
    term *= -x * x / ((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 12; ++i) {
    term *= -x * x / ((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}
// returns cos (which=0) or sin (which=1) of 2*pi*k/n
constexpr double cis2pi(long k, long n, int which) {
  k %= n;
  if (k < 0) k += n;
  const long m = 8 * k;
  const long o = m / n;            // octant 0..7
  const long rem = m - o * n;      // angle within octant = (pi/4) * rem / n
  const double x = kQuarterPi * (double)rem / (double)n;
  const double s = taylor_sin(x), c = taylor_cos(x);
  const double sc = taylor_sin(kQuarterPi - x), cc = taylor_cos(kQuarterPi - x);
  double co = 0, si = 0;
  switch (o) {
    case 0: co = c; si = s; break;
    case 1: co = sc; si = cc; break;
    case 2: co = -s; si = c; break;
    case 3: co = -cc; si = sc; break;
    case 4: co = -c; si = -s; break;
    case 5: co = -sc; si = -cc; break;
    case 6: co = s; si = -c; break;
    default: co = cc; si = -sc; break;
  }
  return which ? si : co;
}

// W_P^M with the transform's sign
template <int P, int M, bool INV>
__device__ __forceinline__ float2 twc() {
  constexpr float c = (float)cis2pi(M, P, 0);
  constexpr float s = (float)cis2pi(M, P, 1);
  return make_float2(c, INV ? s : -s);
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

__host__ __device__ constexpr int reg_radix(int P) {
  return (P % 4 == 0 && P != 8) ? 4
       : (P % 8 == 0) ? 8
       : (P % 2 == 0) ? 2
       : (P % 3 == 0) ? 3
       : (P % 5 == 0) ? 5
       : (P % 7 == 0) ? 7
       : (P % 11 == 0) ? 11 : 0;
}

// In-register DFT of size P (in place, natural order in and out).
template <int P, bool INV>
struct RegFFT {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int R = reg_radix(P);
    static_assert(R != 0, "unsupported register FFT size");
    constexpr int Q = P / R;
    if constexpr (Q == 1) {
      Dft<R, INV>::run(v);
    } else {
      float2 sub[R][Q];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int q = 0; q < Q; ++q) sub[r][q] = v[q * R + r];
#pragma unroll
      for (int r = 0; r < R; ++r) RegFFT<Q, INV>::run(sub[r]);
      static_for<0, Q>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        float2 t[R];
        t[0] = sub[0][k];
        static_for<1, R>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          if constexpr (k == 0) {
            t[r] = sub[r][0];
          } else {
            t[r] = cmul(sub[r][k], twc<P, (r * k) % P, INV>());
          }
        });
        Dft<R, INV>::run(t);
#pragma unroll
        for (int s = 0; s < R; ++s) v[k + Q * s] = t[s];
      });
    }
  }
};

// ---------------------------------------------------------------------------
// two-step transform
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int fft2_l2(int L) {
  for (int d = 1; d <= L; ++d)
    if (L % d == 0 && d * d >= L) return d;
  return L;
}
__host__ __device__ constexpr int fft2_l1(int L) { return L / fft2_l2(L); }

// static LDS per workgroup stays within 64 KiB
constexpr int kLdsBlockMax = 65536;

// LDS floats2 needed per transform (padded transpose)
__host__ __device__ constexpr int fft2_lds(int L) { return fft2_l1(L) * (fft2_l2(L) + 1); }

template <int L, bool INV>
struct Fft2 {
  static constexpr int L1 = fft2_l1(L);
  static constexpr int L2 = fft2_l2(L);
  static constexpr int T = L2;  // threads per transform
  static constexpr int LDS = fft2_lds(L);

  // v[n1] = x[L2*n1 + n2] on entry (thread n2). On exit, threads n2 < L1
  // hold u[k2] = X[n2 + L1*k2]. lds: LDS floats2 of this transform.
  // Contains one barrier: every thread of the workgroup must call it.
  __device__ __forceinline__ static void run(float2 (&v)[L1], float2 (&u)[L2], int n2,
                                             float2* lds, const float2* __restrict__ tw) {
    if constexpr (L1 > 1) RegFFT<L1, INV>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < L1; ++k1) {
      float2 w = tw[n2 * k1];
      if (INV) w.y = -w.y;
      v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1) lds[k1 * (L2 + 1) + n2] = v[k1];
    __syncthreads();
    if (n2 < L1) {
#pragma unroll
      for (int n = 0; n < L2; ++n) u[n] = lds[n2 * (L2 + 1) + n];
      RegFFT<L2, INV>::run(u);
    }
  }
};

}  // namespace mtd
