// Register-resident FFTs for gfx950.
//
// A length-L transform is split as L = L1*L2 (L1 <= L2, L2 the smallest
// divisor >= sqrt L) and run by L2 threads:
//   step 1  thread n2 holds x[L2*n1 + n2], n1 < L1 (a strided subsequence,
//           loaded straight from global memory: coalesced across threads),
//           and does an L1-point DFT in registers;
//   twiddle Y[n2][k1] *= W_L^(n2*k1) (table);
//   exchange one LDS transpose (real parts, then imaginary parts);
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
       : (P % 11 == 0) ? 11
       : (P % 13 == 0) ? 13 : 0;
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

// LDS floats per transform of the two exchange layouts (see Fft2)
__host__ __device__ constexpr int fft2_lds_row(int L) { return fft2_l1(L) * (fft2_l2(L) + 1); }
__host__ __device__ constexpr int fft2_lds_col(int L) { return L; }

// Exchange storage of one transform, seen as an L1 x L2 matrix (at) or as a
// line of L values (lin). Rows: a padded L1 x (L2+1) block per transform,
// transforms one after the other (the lanes of a transform are adjacent).
template <int L1, int L2>
struct RowX {
  float* p;
  __device__ __forceinline__ float* at(int i, int j) const { return p + i * (L2 + 1) + j; }
  __device__ __forceinline__ float* lin(int k) const { return p + k; }
};
// Columns: the NC transforms of a block interleaved (column fastest, as the
// lanes are). A wave spans G = 64/NC consecutive matrix rows (or threads of a
// transform); XOR-swizzling the column index with the row index mod G puts
// those G rows on distinct groups of NC banks, for the transposing writes
// and reads alike, while keeping only G distinct runtime addresses (the rest
// are immediate offsets). No swizzle where G does not fit (correct, slower).
template <int L2, int NC>
constexpr int colx_group() {
  return (NC <= 64 && 64 % NC == 0 && L2 % (64 / NC) == 0) ? 64 / NC : 1;
}
template <int L1, int L2, int NC>
struct ColX {
  static constexpr int G = colx_group<L2, NC>();
  float* p;  // base + column
  __device__ __forceinline__ float* at(int i, int j) const {
    return p + (i * L2 + (j ^ (i & (G - 1)))) * NC;
  }
  __device__ __forceinline__ float* lin(int k) const { return p + k * NC; }
};

template <int L, bool INV>
struct Fft2 {
  static constexpr int L1 = fft2_l1(L);
  static constexpr int L2 = fft2_l2(L);
  static constexpr int T = L2;  // threads per transform
  static constexpr int LDS_ROW = fft2_lds_row(L);
  static constexpr int LDS_COL = fft2_lds_col(L);
  template <int NC>
  using Col = ColX<L1, L2, NC>;
  using Row = RowX<L1, L2>;

  // v[n1] = x[L2*n1 + n2] on entry (thread n2). On exit, threads n2 < L1
  // hold u[k2] = X[n2 + L1*k2]. x: this transform's exchange storage. The
  // transpose moves the real parts, then the imaginary parts, through a
  // float buffer (half the LDS of a complex one). Contains three barriers:
  // every thread of the workgroup must call it. tw: the L-point twiddle
  // table, or with TS > 1 the (TS*L)-point one (entry TS*j is entry j of the
  // L-point table).
  template <int TS = 1, class X>
  __device__ __forceinline__ static void run(float2 (&v)[L1], float2 (&u)[L2], int n2, X x,
                                             const float2* __restrict__ tw) {
    if constexpr (L1 > 1) RegFFT<L1, INV>::run(v);
#pragma unroll
    for (int k1 = 1; k1 < L1; ++k1) {
      float2 w = tw[TS * n2 * k1];
      if (INV) w.y = -w.y;
      v[k1] = cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1) *x.at(k1, n2) = v[k1].x;
    __syncthreads();
    if (n2 < L1) {
#pragma unroll
      for (int n = 0; n < L2; ++n) u[n].x = *x.at(n2, n);
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1) *x.at(k1, n2) = v[k1].y;
    __syncthreads();
    if (n2 < L1) {
#pragma unroll
      for (int n = 0; n < L2; ++n) u[n].y = *x.at(n2, n);
      RegFFT<L2, INV>::run(u);
    }
  }

  // run with the transform's work confined to the threads where act holds
  // (the barriers are still every thread's): for a block whose lines have
  // more threads than this length's transform needs
  template <class X>
  __device__ __forceinline__ static void run_masked(float2 (&v)[L1], float2 (&u)[L2], int n2,
                                                    X x, const float2* __restrict__ tw,
                                                    bool act) {
    if (act) {
      if constexpr (L1 > 1) RegFFT<L1, INV>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < L1; ++k1) {
        float2 w = tw[n2 * k1];
        if (INV) w.y = -w.y;
        v[k1] = cmul(v[k1], w);
      }
#pragma unroll
      for (int k1 = 0; k1 < L1; ++k1) *x.at(k1, n2) = v[k1].x;
    }
    __syncthreads();
    if (act && n2 < L1) {
#pragma unroll
      for (int n = 0; n < L2; ++n) u[n].x = *x.at(n2, n);
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int k1 = 0; k1 < L1; ++k1) *x.at(k1, n2) = v[k1].y;
    }
    __syncthreads();
    if (act && n2 < L1) {
#pragma unroll
      for (int n = 0; n < L2; ++n) u[n].y = *x.at(n2, n);
      RegFFT<L2, INV>::run(u);
    }
  }
};

// ---------------------------------------------------------------------------
// Balanced two-step transform, for L = L1*L2 with L1 | L2 (512, 128, ...):
// Fft2's second step runs on L1 of its L2 threads (when L1 < L2); here both
// steps run on all T = L1 threads of a line, R = L2/L1 transforms each in the
// second, and the exchange reuses the registers of the values it has just
// written out (one array of L2 values per thread):
//   step 1  thread t holds x[L1*m + t], m < L2, and does the L2-point DFT
//           over m: Y[t][k2];
//   twiddle Y[t][k2] *= W_L^(t*k2) (table);
//   exchange (real parts, then imaginary parts) through the storage seen as
//           an L2 x L1 matrix (k2, t);
//   step 2  thread s does the L1-point DFTs over t of rows k2 = R*s + j,
//           j < R, and holds X[k2 + L2*k1] at v[j*L1 + k1].
// ---------------------------------------------------------------------------
// Exchange storage of a block's columns for Fft2B: NC columns interleaved
// (column fastest, as the lanes are). A wave holds G = 64/NC consecutive
// threads t of each column; in the second step they read rows R*s + j of G
// consecutive s, so t is XORed with (k2/R) mod G, which puts the G reads on
// distinct groups of NC banks (the writes, one row k2 per instruction, stay
// on distinct groups too).
template <int L1, int L2, int NC>
struct ColB {
  static constexpr int R = L2 / L1;
  static constexpr int G = 64 / NC;
  static_assert(NC <= 64 && 64 % NC == 0 && G <= L1, "swizzle within a row");
  float* p;  // base + column
  __device__ __forceinline__ float* at(int k2, int t) const {
    return p + (k2 * L1 + (t ^ ((k2 / R) & (G - 1)))) * NC;
  }
};

template <int L, bool INV>
struct Fft2B {
  static constexpr int L1 = fft2_l1(L);
  static constexpr int L2 = fft2_l2(L);
  static constexpr int T = L1;  // threads per transform
  static constexpr int R = L2 / L1;
  static_assert(L1 > 1 && L2 % L1 == 0, "balanced split needs L1 | L2");
  template <int NC>
  using Col = ColB<L1, L2, NC>;

  // v[m] = x[L1*m + t] on entry (thread t); on exit v[j*L1 + k1] =
  // X[R*t + j + L2*k1]. Three barriers: every thread of the workgroup calls
  // it. The exchange reuses v: each component is read back into the slots
  // whose values were just written out.
  template <class X>
  __device__ __forceinline__ static void run(float2 (&v)[L2], int t, X x,
                                             const float2* __restrict__ tw) {
    RegFFT<L2, INV>::run(v);
#pragma unroll
    for (int k2 = 1; k2 < L2; ++k2) {
      float2 w = tw[t * k2];
      if (INV) w.y = -w.y;
      v[k2] = cmul(v[k2], w);
    }
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) *x.at(k2, t) = v[k2].x;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int i = 0; i < L1; ++i) v[j * L1 + i].x = *x.at(R * t + j, i);
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) *x.at(k2, t) = v[k2].y;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int i = 0; i < L1; ++i) v[j * L1 + i].y = *x.at(R * t + j, i);
#pragma unroll
    for (int j = 0; j < R; ++j) RegFFT<L1, INV>::run(v + j * L1);
  }
};

// Moves a line from the output layout of Fft2::run (threads k1 < L1 hold
// u[k2] at position k1 + L1*k2) to its input layout (thread n2 holds v[n1]
// at position L2*n1 + n2), one component at a time. Starts and ends with a
// barrier, so the storage is free on entry and on exit.
template <int L, class X>
__device__ __forceinline__ void fft2_relayout(const float2 (&u)[Fft2<L, false>::L2],
                                              float2 (&v)[Fft2<L, false>::L1], int n2, X x) {
  constexpr int L1 = Fft2<L, false>::L1, L2 = Fft2<L, false>::L2;
  __syncthreads();
  if (n2 < L1) {
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) *x.lin(n2 + L1 * k2) = u[k2].x;
  }
  __syncthreads();
#pragma unroll
  for (int n1 = 0; n1 < L1; ++n1) v[n1].x = *x.lin(L2 * n1 + n2);
  __syncthreads();
  if (n2 < L1) {
#pragma unroll
    for (int k2 = 0; k2 < L2; ++k2) *x.lin(n2 + L1 * k2) = u[k2].y;
  }
  __syncthreads();
#pragma unroll
  for (int n1 = 0; n1 < L1; ++n1) v[n1].y = *x.lin(L2 * n1 + n2);
  __syncthreads();
}

}  // namespace mtd
