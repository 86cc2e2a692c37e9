// Hologram stage: windowed, centred 2-D FFT of each N x N hologram, keeping
// only the C x C window around the +1 order (holo2TF_UBorn,
// pretraitement_hors_axe_CPU/fonctions.cpp:279-304).
//
// The centred transform fftshift(FFT(fftshift(h)))/N^2 equals
// (-1)^(kx+ky) * FFT(h * (-1)^(x+y))[ky][kx] / N^2 for even N, so both shifts
// become sign flips applied on load and on store.
//
// Pass 1 (rows): two real rows are packed into one complex FFT of length N
// (two-step register FFT; from N = 2048, each row's even and odd samples into
// one of length N/2); only the C bins kx in [x0, x0+C) are separated and
// written, to t1 (N x C per hologram) in column tiles of W bins (W = 16 when
// it divides C; from N = 2048 the column block width):
// t1[a][d / W][y][d % W]. A row pass stores runs of W bins (128 bytes at
// W = 16), and the blocks of the column pass that share a tile read it as
// one contiguous N x W slab.
// Pass 2 (columns): each of the C columns of t1 is transformed along y and
// only the C bins ky in [y0, y0+C) are kept, scaled by 1/N^2 and signed.
#include "argmax.hpp"
#include "fft_reg.hpp"
#include "gen_fft.hpp"
#include "internal.hpp"
#include "passes.hpp"

namespace mtd {

template <int N>
constexpr int holo_cols_nc() {  // a power of two <= 16: divides every crop side
  constexpr int by_thr = (N > 512 ? 256 : 1024) / Fft2<N, false>::T;
  constexpr int by_lds = kLdsBlockMax / (Fft2<N, false>::LDS_COL * 4);
  return divisor_le(16, by_thr < by_lds ? by_thr : by_lds);
}

// width of the column tiles of t1: the largest power of two from 16 down to
// the column block width nc (a power of two dividing C) that divides C
static int holo_tile_w(int C, int nc) {
  for (int w = 16; w > nc; w /= 2)
    if (C % w == 0) return w;
  return nc;
}

// the t1 tile width for hologram side N: from N = 2048 the column block
// width itself (4 columns there), so that each column block reads one
// contiguous N x 4 slab instead of 32-byte pieces of 256-byte rows (column
// pass 4.32 -> 3.76 ms, row pass 5.24 -> 5.37 ms per 1000 holograms; 8-bin
// tiles 3.90 + 5.29 ms); below, 16 bins where they divide C: the row pass
// stores whole 128-byte lines and two column blocks share a slab (at 1024,
// against 32-bin tiles: column pass -12 us per volume, row pass unchanged;
// 8-bin tiles: -28 us and +52 us)
template <int N>
static int holo_w(int C) {
  if constexpr (N >= 2048) return holo_cols_nc<N>();
  return holo_tile_w(C, holo_cols_nc<N>());
}

template <int N>
constexpr int holo_pairs() {  // row pairs per block: <= 256 threads, <= 64 KiB
  constexpr int by_lds = kLdsBlockMax / (Fft2<N, false>::LDS_ROW * 4);
  constexpr int cap = 256 / Fft2<N, false>::T;
  return divisor_le(N / 2, by_lds < cap ? (by_lds > 0 ? by_lds : 1) : cap);
}

// waves per SIMD the two-row kernel is compiled for (blocks of at most 256
// threads; four such blocks fit in the LDS of a CU; sides up to 1024)
template <int N>
constexpr int holo_rows_waves() {
  return N == 768 ? 3 : 4;
}

template <int N, typename TIn>
__global__ void __launch_bounds__(256, holo_rows_waves<N>())
k_holo_rows(const TIn* __restrict__ holos, const float* __restrict__ ws,
            const float2* __restrict__ tw, int x0, int C, int cp, int lw,
            float2* __restrict__ t1) {
  using F = Fft2<N, false>;
  constexpr int NP = holo_pairs<N>();
  static_assert(F::LDS_ROW >= N, "exchange storage must hold a full row");
  __shared__ float lds[NP * F::LDS_ROW];
  const int2 ab = xcd_block();
  const int a = ab.x;
  const int p = threadIdx.x / F::T, n2 = threadIdx.x % F::T;
  const int y0 = 2 * (ab.y * NP + p), y1 = y0 + 1;
  const TIn* h0 = holos + ((size_t)a * N + y0) * N;
  const TIn* h1 = h0 + N;
  // window tuk(x) tuk(y) with the centring sign (-1)^(x+y): ws holds
  // (-1)^i (float)tuk(i), so one product per sample carries both
  const float wy0 = ws[y0], wy1 = ws[y1];
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1) {
    const int x = F::L2 * n1 + n2;
    const float wx = ws[x];
    v[n1] = make_float2((float)h0[x] * (wx * wy0), (float)h1[x] * (wx * wy1));
  }
  float* buf = lds + p * F::LDS_ROW;
  F::run(v, u, n2, typename F::Row{buf}, tw);
  // Separation of the two rows, for k = x0 + d, m = N - k:
  //   X0 = Z[k] + conj(Z[m]) = (Zk.x + Zm.x, Zk.y - Zm.y)
  //   X1 = -i (Z[k] - conj(Z[m])) = (Zk.y + Zm.y, Zm.x - Zk.x)
  // (the 1/2 is folded into the column pass scale). The row goes through the
  // float storage twice: real parts give X0.x and X1.y, imaginary parts
  // X0.y and X1.x.
  // The values of the first pass stay in registers (crops up to N/2 wide)
  // so that each output is stored once, as a whole complex value.
  // Bin d of row y0 sits at q0[tile(d)], of row y1 at q1[tile(d)].
  // W = 2^lw: shifts and masks, no integer division in the store loop
  const int W = 1 << lw;
  float2* q0 = t1 + (size_t)a * N * cp + (size_t)y0 * W;
  float2* q1 = q0 + W;
  auto tile = [lw, W](int d) { return (((d >> lw) * N) << lw) + (d & (W - 1)); };
  constexpr int DK = (N / 2 + F::T - 1) / F::T;
  float r0[DK], r1[DK];
  const bool keep = C <= N / 2;  // uniform
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    __syncthreads();  // the storage is free
    if (n2 < F::L1) {
#pragma unroll
      for (int k2 = 0; k2 < F::L2; ++k2) buf[n2 + F::L1 * k2] = part ? u[k2].y : u[k2].x;
    }
    __syncthreads();
    if (keep) {
#pragma unroll
      for (int j = 0; j < DK; ++j) {
        const int d = n2 + F::T * j;
        if (d < C) {
          const int k = x0 + d;
          const int m = (N - k) % N;
          const float zk = buf[k], zm = buf[m];
          if (part == 0) {
            r0[j] = zk + zm;
            r1[j] = zm - zk;
          } else {
            q0[tile(d)] = make_float2(r0[j], zk - zm);
            q1[tile(d)] = make_float2(zk + zm, r1[j]);
          }
        }
      }
    } else {
      float* o0 = (float*)q0;
      float* o1 = (float*)q1;
      for (int d = n2; d < C; d += F::T) {
        const int k = x0 + d;
        const int m = (N - k) % N;
        const float zk = buf[k], zm = buf[m];
        const int e = 2 * tile(d);
        if (part == 0) {
          o0[e] = zk + zm;
          o1[e + 1] = zm - zk;
        } else {
          o0[e + 1] = zk - zm;
          o1[e] = zk + zm;
        }
      }
    }
  }
}

// Hologram sides from 2048: each real row is one complex transform of half
// the length, z[n] = x[2n] + i x[2n+1], n < H = N/2. At N = 2048 that is the
// balanced 32 x 32 register FFT (the two-row packing needs the 32 x 64
// split: 64 values per thread, half the threads idle in its second step, two
// waves per SIMD). With Z = FFT_H(z) and w = exp(-2 pi i / N), the row's bins
// follow from the even and odd samples' spectra:
//   X[k] = A[k] + w^k B[k],  A = (Z[k] + conj Z[m]) / 2,
//   B = (Z[k] - conj Z[m]) / 2i,  m = H - k,  Z's indices mod H.
// With E = Z[k] + conj Z[m] and D = Z[k] - conj Z[m], the stored value is
//   2 X[k] = (E.x + w.x D.y + w.y D.x, E.y - w.x D.x + w.y D.y)
// (2 X: as in the two-row pass, the 1/2 is in the column pass's scale).
template <int N>
constexpr bool holo_rows_half() {
  return N >= 2048;
}

__device__ __forceinline__ float2 load_pair(const uint8_t* p) {
  const uchar2 q = *(const uchar2*)p;
  return make_float2((float)q.x, (float)q.y);
}
__device__ __forceinline__ float2 load_pair(const uint16_t* p) {
  const ushort2 q = *(const ushort2*)p;
  return make_float2((float)q.x, (float)q.y);
}
__device__ __forceinline__ float2 load_pair(const float* p) { return *(const float2*)p; }

template <int N, typename TIn>
__global__ void __launch_bounds__(256, 4)
k_holo_rows_half(const TIn* __restrict__ holos, const float* __restrict__ ws,
                 const float2* __restrict__ tw, int x0, int C, int cp, int lw,
                 float2* __restrict__ t1) {
  constexpr int H = N / 2;
  using F = Fft2<H, false>;
  constexpr int NR = 256 / F::T;  // rows per block
  static_assert(F::LDS_ROW >= H, "exchange storage must hold a full line");
  __shared__ float lds[NR * F::LDS_ROW];
  const int2 ab = xcd_block();
  const int a = ab.x;
  const int r = threadIdx.x / F::T, n2 = threadIdx.x % F::T;
  const int y = ab.y * NR + r;
  const TIn* h = holos + ((size_t)a * N + y) * N;
  // ws holds (-1)^i (float)tuk(i): window and centring sign in one product
  const float wy = ws[y];
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1) {
    const int n = F::L2 * n1 + n2;
    const float2 s = load_pair(h + 2 * n);
    const float2 wx = *(const float2*)(ws + 2 * n);
    v[n1] = make_float2(s.x * (wx.x * wy), s.y * (wx.y * wy));
  }
  float* buf = lds + r * F::LDS_ROW;
  // tw is the N-point table: stride 2 gives the H-point one
  F::template run<2>(v, u, n2, typename F::Row{buf}, tw);
  // Z goes through the float storage twice, real parts (E.x, D.x) then
  // imaginary parts (E.y, D.y); crops up to N/4 keep the first pass's values
  // in registers, wider ones stage them in their own output slots.
  const int W = 1 << lw;
  float2* q = t1 + (size_t)a * N * cp + (size_t)y * W;
  auto tile = [lw, W](int d) { return (((d >> lw) * N) << lw) + (d & (W - 1)); };
  constexpr int DK = (N / 4 + F::T - 1) / F::T;
  float ex[DK], dx[DK];
  const bool keep = C <= N / 4;  // uniform
#pragma unroll
  for (int part = 0; part < 2; ++part) {
    __syncthreads();  // the storage is free
    if (n2 < F::L1) {
#pragma unroll
      for (int k2 = 0; k2 < F::L2; ++k2) buf[n2 + F::L1 * k2] = part ? u[k2].y : u[k2].x;
    }
    __syncthreads();
    if (keep) {
#pragma unroll
      for (int j = 0; j < DK; ++j) {
        const int d = n2 + F::T * j;
        if (d < C) {
          const int k = x0 + d;
          const int kk = k & (H - 1);
          const float zk = buf[kk], zm = buf[(H - kk) & (H - 1)];
          if (part == 0) {
            ex[j] = zk + zm;
            dx[j] = zk - zm;
          } else {
            const float2 w = tw[k];
            const float ey = zk - zm, dy = zk + zm;
            q[tile(d)] = make_float2(ex[j] + w.x * dy + w.y * dx[j], ey - w.x * dx[j] + w.y * dy);
          }
        }
      }
    } else {
      for (int d = n2; d < C; d += F::T) {
        const int k = x0 + d;
        const int kk = k & (H - 1);
        const float zk = buf[kk], zm = buf[(H - kk) & (H - 1)];
        float2* o = q + tile(d);
        if (part == 0) {
          *o = make_float2(zk + zm, zk - zm);
        } else {
          const float2 e = *o;  // this thread's own store of the first pass
          const float2 w = tw[k];
          const float ey = zk - zm, dy = zk + zm;
          *o = make_float2(e.x + w.x * dy + w.y * e.y, ey - w.x * e.y + w.y * dy);
        }
      }
    }
  }
}

template <int N>
constexpr int holo_cols_threads() {
  return holo_cols_nc<N>() * Fft2<N, false>::T;
}

// waves per SIMD the column kernel is compiled for: from N = 1024 its
// register FFT plus the argmax take more than a third of the VGPRs
template <int N>
constexpr int holo_cols_waves() {
  return N >= 1024 ? 2 : col_waves_per_simd<N>();
}

// Crop sides whose column inverse DFT the column pass can produce (QOUT):
// compiled register transforms, for hologram sides of at least twice the
// crop (the kept column is staged in the exchange storage as complex values)
#define MTD_QCOL_SIZES(X) X(128) X(192) X(250) X(256) X(384) X(512)

template <int N, int CQ>
constexpr bool qcol_fits() {
  return N >= 512 && 2 * CQ <= N && Fft2<CQ, true>::T <= Fft2<N, false>::T;
}

// Column c's inverse DFT from the staging st ([CQ][NC], column fastest)
// through the exchange storage x (which aliases st: read first) to Q's
// rows; threads n2 < Fft2<CQ>::T of the column work, every thread passes the
// barriers.
template <int CQ, int NC>
__device__ __forceinline__ void q_column(const float2* st, float* x, int c, int n2,
                                         const float2* __restrict__ tq, float2* __restrict__ q,
                                         bool keep) {
  using G = Fft2<CQ, true>;
  const bool act = n2 < G::T;
  float2 v[G::L1], u[G::L2];
  if (act) {
#pragma unroll
    for (int n1 = 0; n1 < G::L1; ++n1) v[n1] = st[(G::L2 * n1 + n2) * NC + c];
  }
  __syncthreads();
  G::run_masked(v, u, n2, typename G::template Col<NC>{x + c}, tq, act);
  if (act && keep && n2 < G::L1) {
#pragma unroll
    for (int k2 = 0; k2 < G::L2; ++k2) q[(size_t)(n2 + G::L1 * k2) * CQ] = u[k2];
  }
}

// QOUT: instead of the crop, each column's unnormalised C-point inverse DFT
// along y, Q[a][k][x] = sum_e crop[a][e][x] exp(+2 pi i e k / C) (the column
// half of the demodulation's 2-D inverse transform, calc_Uborn2,
// fonctions.cpp:348-374; its -posSpec shift becomes a phase ramp on Q's
// rows by the shift theorem, applied by the row pass once the specular is
// known). tq: the C-point twiddle table; C one of MTD_QCOL_SIZES with
// qcol_fits<N, C>.
template <int N, bool QOUT = false>
__global__ void __launch_bounds__(holo_cols_threads<N>(), holo_cols_waves<N>())
k_holo_cols(const float2* __restrict__ t1, const float2* __restrict__ tw,
            int C, int cp, int W, int x0, int y0, float scale, float2* __restrict__ crop,
            float2* __restrict__ part, const float2* __restrict__ tq) {
  using F = Fft2<N, false>;
  constexpr int NC = holo_cols_nc<N>();
  __shared__ float lds[NC * F::LDS_COL];
  const int2 ab = xcd_block();
  const int a = ab.x, d0 = ab.y * NC;
  const int c = threadIdx.x % NC, n2 = threadIdx.x / NC;
  // this block's contiguous N x NC tile through a buffer view: one lane
  // offset, uniform row offsets
  const BufView src(t1 + (size_t)a * N * cp + (size_t)(d0 / W) * N * W + d0 % W,
                    (unsigned)((N * W - d0 % W) * sizeof(float2)));
  const unsigned lane = (unsigned)((n2 * W + c) * sizeof(float2));
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1)
    v[n1] = src.f2(lane, (unsigned)(F::L2 * n1 * W * sizeof(float2)));
  F::run(v, u, n2, typename F::template Col<NC>{lds + c}, tw);
  // argmax partial of the crop values this block writes (coordSpec's
  // first stage, fused: the crop is not read again for it)
  // lanes past the crop's last column (a crop side that is not a multiple
  // of the block width) transformed padding: nothing of theirs is kept
  float best = -1.f;
  int bi = 0x7fffffff;
  // QOUT: the kept column is staged as complex values in the exchange
  // storage, free once every thread is past the transform
  float2* st = (float2*)lds;
  if constexpr (QOUT) __syncthreads();
  if (n2 < F::L1 && d0 + c < C) {
    float2* dst = crop + (size_t)a * C * C + d0 + c;
    const int kx = x0 + d0 + c;
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) {
      const int ky = n2 + F::L1 * k2;
      const int e = ky - y0;
      if (e >= 0 && e < C) {
        const float s = ((kx + ky) & 1) ? -scale : scale;
        const float2 val = make_float2(u[k2].x * s, u[k2].y * s);
        if constexpr (QOUT) {
          st[e * NC + c] = val;
        } else {
          dst[(size_t)e * C] = val;
        }
        const int i = e * C + d0 + c;
        const float m = argmax_key(val, i == 0);
        if (argmax_better(m, i, best, bi)) { best = m; bi = i; }
      }
    }
  }
  if (part) {  // uniform
    __shared__ float red[2 * (holo_cols_threads<N>() / 64)];
    argmax_block(best, bi, red);
    if (threadIdx.x == 0)
      part[(size_t)a * ((C + NC - 1) / NC) + ab.y] = make_float2(best, __int_as_float(bi));
  }
  if constexpr (QOUT) {
    __syncthreads();
    float2* q = crop + (size_t)a * C * C + d0 + c;
    const bool keep = d0 + c < C;
    switch (C) {  // uniform
#define X(n)                                                         \
  case n:                                                            \
    if constexpr (qcol_fits<N, n>()) q_column<n, NC>(st, lds, c, n2, tq, q, keep); \
    break;
      MTD_QCOL_SIZES(X)
#undef X
    }
  }
}

// columns of t1 per hologram row: C rounded up to whole tiles of W (a
// multiple of the column block width)
template <int N>
static int t1_cols_n(int C) {
  const int W = holo_w<N>(C);
  return (C + W - 1) / W * W;
}

template <int N, typename TIn>
static void rows_launch(const void* holos, int A, const float* win, const float2* tw, int x0,
                        int C, int cp, int lw, float2* t1, hipStream_t s) {
  if constexpr (holo_rows_half<N>()) {
    constexpr int NR = 256 / Fft2<N / 2, false>::T;
    hipLaunchKernelGGL((k_holo_rows_half<N, TIn>), dim3(N / NR, A), dim3(NR * Fft2<N / 2, false>::T),
                       0, s, (const TIn*)holos, win, tw, x0, C, cp, lw, t1);
  } else {
    constexpr int NP = holo_pairs<N>();
    hipLaunchKernelGGL((k_holo_rows<N, TIn>), dim3(N / (2 * NP), A), dim3(NP * Fft2<N, false>::T),
                       0, s, (const TIn*)holos, win, tw, x0, C, cp, lw, t1);
  }
}

template <int N>
static hipError_t rows_n(int dtype, const void* holos, int A, const float* win,
                         const float2* tw, int x0, int C, float2* t1,
                         hipStream_t s) {
  const int W = holo_w<N>(C);
  const int cp = t1_cols_n<N>(C);
  int lw = 0;
  while ((1 << lw) < W) ++lw;  // W is a power of two
  Timed tm("holo_rows", s);
  switch (dtype) {
    case MTD_U8:
      rows_launch<N, uint8_t>(holos, A, win, tw, x0, C, cp, lw, t1, s);
      break;
    case MTD_U16:
      rows_launch<N, uint16_t>(holos, A, win, tw, x0, C, cp, lw, t1, s);
      break;
    default:
      rows_launch<N, float>(holos, A, win, tw, x0, C, cp, lw, t1, s);
      break;
  }
  return hipGetLastError();
}

template <int N>
static bool qcol_n(int C) {
  switch (C) {
#define X(n) case n: return qcol_fits<N, n>();
    MTD_QCOL_SIZES(X)
#undef X
  }
  return false;
}

template <int N>
static hipError_t cols_n(const float2* t1, int A, const float2* tw, int C,
                         int x0, int y0, float2* crop, float2* part, const float2* tq,
                         hipStream_t s) {
  constexpr int NC = holo_cols_nc<N>();
  const float scale = (float)(0.5 / ((double)N * (double)N));
  if (tq && !qcol_n<N>(C)) return hipErrorInvalidValue;
  Timed tm("holo_cols", s);
  const dim3 grid((C + NC - 1) / NC, A), block(holo_cols_threads<N>());
  if (tq) {
    hipLaunchKernelGGL((k_holo_cols<N, true>), grid, block, 0, s, t1, tw, C, t1_cols_n<N>(C),
                       holo_w<N>(C), x0, y0, scale, crop, part, tq);
  } else {
    hipLaunchKernelGGL((k_holo_cols<N, false>), grid, block, 0, s, t1, tw, C, t1_cols_n<N>(C),
                       holo_w<N>(C), x0, y0, scale, crop, part, nullptr);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Hologram sides without a compiled FFT: one complex transform per row of
// the windowed, sign-modulated hologram (its C kept bins into t1[a][y][d],
// untiled), then one per kept column (its C kept bins, / N^2 and signed).
// The specular is then found from the crops (no argmax partials).
// ---------------------------------------------------------------------------
template <typename TIn>
struct HoloRowGenOp {
  const TIn* holos;
  const float* ws;  // (float)tukey1d(N)[i] * (-1)^i
  float2* t1;
  int N, C, x0;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int y, int x) const {
    return make_float2((float)holos[((size_t)a * N + y) * N + x] * (ws[x] * ws[y]), 0.f);
  }
  __device__ void store(int a, int y, int k, float2 v) const {
    const int d = k - x0;
    if (d >= 0 && d < C) t1[((size_t)a * N + y) * C + d] = v;
  }
};

struct HoloColGenOp {
  const float2* t1;
  float2* crop;
  int N, C, x0, y0;
  float scale;
  __device__ void begin(int) {}
  __device__ float2 load(int a, int d, int y) const { return t1[((size_t)a * N + y) * C + d]; }
  __device__ void store(int a, int d, int ky, float2 v) const {
    const int e = ky - y0;
    if (e < 0 || e >= C) return;
    const float sc = ((x0 + d + ky) & 1) ? -scale : scale;
    crop[((size_t)a * C + e) * C + d] = make_float2(v.x * sc, v.y * sc);
  }
};

static hipError_t rows_gen(int N, int dtype, const void* holos, int A, const float* win,
                           const GenPlan& p, int x0, int C, float2* t1, hipStream_t s) {
  Timed tm("holo_rows", s);
  switch (dtype) {
    case MTD_U8:
      return launch_gen<false, true>(HoloRowGenOp<uint8_t>{(const uint8_t*)holos, win, t1, N, C, x0},
                                     p, N, A, s);
    case MTD_U16:
      return launch_gen<false, true>(
          HoloRowGenOp<uint16_t>{(const uint16_t*)holos, win, t1, N, C, x0}, p, N, A, s);
    default:
      return launch_gen<false, true>(HoloRowGenOp<float>{(const float*)holos, win, t1, N, C, x0},
                                     p, N, A, s);
  }
}

// The ArrayFire variant's crop window (pretraitement_hors_axe_GPU
// src/GPU_fonctions.cpp:122): crop * af::complex(T, T), T(x, y) the product
// of the float-rounded 1-D factors, i.e. (re - im, re + im) * T.
__global__ void __launch_bounds__(256)
k_crop_window(float2* __restrict__ crop, int C, const double* __restrict__ win) {
  const int a = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= C * C) return;
  const int y = i / C, x = i - y * C;
  const float w = (float)win[x] * (float)win[y];
  float2* c = crop + (size_t)a * C * C + i;
  const float2 v = *c;
  *c = make_float2((v.x - v.y) * w, (v.x + v.y) * w);
}

hipError_t launch_crop_window(int C, float2* crop, int A, const double* win, hipStream_t s) {
  if (A <= 0) return hipSuccess;
  Timed tm("crop_window", s);
  hipLaunchKernelGGL(k_crop_window, dim3((C * C + 255) / 256, A), dim3(256), 0, s, crop, C, win);
  return hipGetLastError();
}

#define MTD_HOLO_SIZES(X) X(64) X(128) X(256) X(512) X(768) X(1024) X(2048)

hipError_t launch_holo_rows(int N, int dtype, const void* holos, int A,
                            const float* win, const FftTab& ft, int x0,
                            int C, float2* t1, hipStream_t s) {
  switch (N) {
#define X(n) case n: return rows_n<n>(dtype, holos, A, win, ft.tw, x0, C, t1, s);
    MTD_HOLO_SIZES(X)
#undef X
  }
  if (!ft.gen) return hipErrorInvalidValue;
  return rows_gen(N, dtype, holos, A, win, *ft.gen, x0, C, t1, s);
}

hipError_t launch_holo_cols(int N, const float2* t1, int A, const FftTab& ft,
                            int C, int x0, int y0, float2* crop, float2* part,
                            hipStream_t s, const float2* qtw) {
  switch (N) {
#define X(n) case n: return cols_n<n>(t1, A, ft.tw, C, x0, y0, crop, part, qtw, s);
    MTD_HOLO_SIZES(X)
#undef X
  }
  if (qtw) return hipErrorInvalidValue;  // the fused column IDFT is for compiled sides
  if (!ft.gen) return hipErrorInvalidValue;
  Timed tm("holo_cols", s);
  const float scale = (float)(1.0 / ((double)N * (double)N));
  return launch_gen<false, false>(HoloColGenOp{t1, crop, N, C, x0, y0, scale}, *ft.gen, C, A, s);
}

int holo_argmax_parts(int N, int C) {
  switch (N) {
#define X(n) case n: return (C + holo_cols_nc<n>() - 1) / holo_cols_nc<n>();
    MTD_HOLO_SIZES(X)
#undef X
  }
  return 0;
}

int holo_t1_cols(int N, int C) {
  switch (N) {
#define X(n) case n: return t1_cols_n<n>(C);
    MTD_HOLO_SIZES(X)
#undef X
  }
  return C;
}

bool holo_q_supported(int N, int C) {
  switch (N) {
#define X(n) case n: return qcol_n<n>(C);
    MTD_HOLO_SIZES(X)
#undef X
  }
  return false;
}

bool holo_len_supported(int N) {
  switch (N) {
#define X(n) case n: return true;
    MTD_HOLO_SIZES(X)
#undef X
  }
  return false;
}

}  // namespace mtd
