// Runtime-length FFTs for the transform lengths that have no compiled
// register FFT (fft_reg.hpp): any hologram side, crop side or volume side the
// configuration files can hold, as FFTW plans any size
// (pretraitement_hors_axe_CPU/main.cpp:103, src/FFTW_init.cpp:70-71,
// Reconstruction/src/FFT_encaps.cpp:33-53).
//
// A block transforms nl lines held in LDS (two buffers of nl x W complex):
//  - direct: L factors into radices 2, 3, 4, 5, 7, 8, 11, 13; Stockham
//    autosort passes (each a full out-of-place pass between the two LDS
//    buffers, one barrier per pass), twiddles from the length-L table;
//  - Bluestein (chirp-z) for lengths with a larger prime factor, written
//    from the published method: X_k = w_k sum_n (x_n w_n) conj(w_(k-n)) with
//    w_m = exp(-i pi m^2 / L), the convolution done as a cyclic one of length
//    W = a power of two >= 2L - 1 with two forward length-W Stockham FFTs
//    (the inverse through conjugation) and the host-computed spectrum of the
//    chirp (scaled 1/W). The backward transform is conj(forward(conj(x))).
// Operations load and store through the same op interface as the compiled
// passes (passes.hpp): load(batch, line, element), store(batch, line,
// element, value), optional begin/skip/line.
#pragma once
#include <hip/hip_runtime.h>

#include "dft.hpp"
#include "passes.hpp"

namespace mtd {

constexpr int kGenMaxPasses = 24;
constexpr int kGenThreads = 256;
// LDS of a block: two buffers of nl x W complex values
constexpr int kGenLdsBytes = 65536;

struct GenPlan {
  int L;                   // transform length
  int W;                   // working length: L (direct) or the Bluestein length
  int nf;                  // radix passes over W
  int f[kGenMaxPasses];    // radices
  const float2* tw;        // exp(-2 pi i k / W), k < W
  const float2* chirp;     // Bluestein: exp(-i pi n^2 / L), n < L; nullptr: direct
  const float2* bk;        // Bluestein: FFT_W of the wrapped conj(chirp), times 1/W
};

// Lines per block: as many as the two LDS buffers hold (at most 64).
__host__ __device__ inline int gen_lines_per_block(int W) {
  int nl = kGenLdsBytes / (2 * W * (int)sizeof(float2));
  if (nl > 64) nl = 64;
  return nl < 1 ? 1 : nl;
}

template <int R, bool INV>
__device__ __forceinline__ void gen_butterfly(const float2* __restrict__ src,
                                              float2* __restrict__ dst, int j, int nb, int Ns,
                                              int step, const float2* __restrict__ tw) {
  float2 v[R];
  const int k = j % Ns;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v[r] = src[j + r * nb];
    if (r > 0 && k > 0) {
      float2 w = tw[r * k * step];  // W_(Ns R)^(r k) = W_W^(r k step), r k step < W
      if (INV) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
  }
  Dft<R, INV>::run(v);
  const int d = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[d + r * Ns] = v[r];
}

// Stockham passes over nl lines of length W (line l at in + l * W); returns
// the buffer holding the result (in or out). All threads of the block call
// it; it begins with the caller's data in place (barrier done by the caller)
// and ends with a barrier.
template <bool INV>
__device__ float2* gen_stockham(float2* in, float2* out, int nl, int W, const int* f, int nf,
                                const float2* __restrict__ tw) {
  int Ns = 1;
  for (int p = 0; p < nf; ++p) {
    const int R = f[p];
    const int nb = W / R;
    const int step = W / (Ns * R);
    for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
      const int l = t / nb, j = t - l * nb;
      const float2* s = in + l * W;
      float2* d = out + l * W;
      switch (R) {
        case 2: gen_butterfly<2, INV>(s, d, j, nb, Ns, step, tw); break;
        case 3: gen_butterfly<3, INV>(s, d, j, nb, Ns, step, tw); break;
        case 4: gen_butterfly<4, INV>(s, d, j, nb, Ns, step, tw); break;
        case 5: gen_butterfly<5, INV>(s, d, j, nb, Ns, step, tw); break;
        case 7: gen_butterfly<7, INV>(s, d, j, nb, Ns, step, tw); break;
        case 8: gen_butterfly<8, INV>(s, d, j, nb, Ns, step, tw); break;
        case 11: gen_butterfly<11, INV>(s, d, j, nb, Ns, step, tw); break;
        default: gen_butterfly<13, INV>(s, d, j, nb, Ns, step, tw); break;
      }
    }
    __syncthreads();
    float2* t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

// Transforms nl lines of length p.L held at a + l * p.W (elements [0, L) of
// each line; b is the second buffer). Returns the buffer whose lines hold
// the result in [0, L). Begins after the caller's barrier, ends with one.
template <bool INV>
__device__ float2* gen_fft(float2* a, float2* b, int nl, const GenPlan& p) {
  if (!p.chirp) return gen_stockham<INV>(a, b, nl, p.W, p.f, p.nf, p.tw);
  const int L = p.L, W = p.W;
  // x_n w_n (conj(x) for the backward transform), zero-padded to W
  for (int t = threadIdx.x; t < nl * W; t += blockDim.x) {
    const int l = t / W, n = t - l * W;
    float2 v = make_float2(0.f, 0.f);
    if (n < L) {
      v = a[l * W + n];
      if (INV) v.y = -v.y;
      v = cmul(v, p.chirp[n]);
    }
    a[l * W + n] = v;
  }
  __syncthreads();
  float2* r = gen_stockham<false>(a, b, nl, W, p.f, p.nf, p.tw);
  float2* o = r == a ? b : a;
  // times the chirp's spectrum, conjugated: the inverse length-W FFT is
  // conj(FFT(conj(.)))
  for (int t = threadIdx.x; t < nl * W; t += blockDim.x) {
    const int k = t % W;
    const float2 z = cmul(r[t], p.bk[k]);
    r[t] = make_float2(z.x, -z.y);
  }
  __syncthreads();
  r = gen_stockham<false>(r, o, nl, W, p.f, p.nf, p.tw);
  for (int t = threadIdx.x; t < nl * W; t += blockDim.x) {
    const int l = t / W, k = t - l * W;
    if (k < L) {
      const float2 c = r[l * W + k];
      float2 y = cmul(make_float2(c.x, -c.y), p.chirp[k]);
      if (INV) y.y = -y.y;
      r[l * W + k] = y;
    }
  }
  __syncthreads();
  return r;
}

// Batched 1-D transforms of nlines lines per batch (grid: line groups x
// batches). ROWS: a block's lines are consecutive rows and element loads run
// along each line; otherwise its lines are adjacent columns and the loads of
// one element index run across them (coalesced either way).
template <bool INV, bool ROWS, class Op>
__global__ void __launch_bounds__(kGenThreads) k_gen(Op op, GenPlan p, int nlines, int nl) {
  extern __shared__ float2 gen_lds[];
  const int a = blockIdx.y, line0 = blockIdx.x * nl;
  if (skip_batch(op, a)) return;  // block-uniform, before any barrier
  op.begin(a);
  float2* A = gen_lds;
  float2* B = gen_lds + (size_t)nl * p.W;
  const int nact = min(nl, nlines - line0);
  const int L = p.L;
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    int l, n;
    if (ROWS) { l = t / L; n = t - l * L; } else { n = t / nl; l = t - n * nl; }
    float2 v = make_float2(0.f, 0.f);
    if (l < nact) {
      if constexpr (has_line<Op>::value) {
        Op o = op;
        o.line(a, line0 + l);
        v = o.load(a, line0 + l, n);
      } else {
        v = op.load(a, line0 + l, n);
      }
    }
    A[l * p.W + n] = v;
  }
  __syncthreads();
  const float2* R = gen_fft<INV>(A, B, nl, p);
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    int l, n;
    if (ROWS) { l = t / L; n = t - l * L; } else { n = t / nl; l = t - n * nl; }
    if (l < nact) op.store(a, line0 + l, n, R[l * p.W + n]);
  }
}

template <bool INV, bool ROWS, class Op>
inline hipError_t launch_gen(const Op& op, const GenPlan& p, int nlines, int batches,
                             hipStream_t s) {
  if (nlines <= 0 || batches <= 0) return hipSuccess;
  const int nl = gen_lines_per_block(p.W);
  const size_t lds = (size_t)2 * nl * p.W * sizeof(float2);
  hipLaunchKernelGGL((k_gen<INV, ROWS, Op>), dim3((nlines + nl - 1) / nl, batches),
                     dim3(kGenThreads), lds, s, op, p, nlines, nl);
  return hipGetLastError();
}

}  // namespace mtd
