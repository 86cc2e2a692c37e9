// Generic batched 1-D FFT passes over stacks of lines, parameterised by an
// operation object that maps (batch, line, element) to memory on load and on
// store. Transforms use the two-step register FFT (fft_reg.hpp): inputs are
// loaded straight from global memory into registers and outputs stored
// straight from registers, with one LDS transpose in between.
//   rows:    each transform is one contiguous line; NT lines per block.
//   columns: NC adjacent columns per block; lane = column fastest, so every
//            load and store touches NC contiguous elements of a row.
#pragma once
#include <type_traits>
#include <utility>

#include "fft_reg.hpp"

namespace mtd {

// A side length fixed at compile time (N > 0: folds into the index
// arithmetic of the compiled lengths) or held at run time (N == 0: the
// runtime-length path). Converts to int.
template <int N>
struct Dim {
  __host__ __device__ constexpr Dim(int = N) {}
  __host__ __device__ constexpr operator int() const { return N; }
};
template <>
struct Dim<0> {
  int v = 0;
  __host__ __device__ constexpr Dim(int x = 0) : v(x) {}
  __host__ __device__ constexpr operator int() const { return v; }
};

// Raw buffer view of one device array (at most 2 GiB). A load takes a 32-bit
// per-lane byte offset and a uniform byte offset held in a scalar register,
// so a run of loads at uniform strides keeps one address register in place of
// a 64-bit address per load. Offsets at or beyond `bytes` read as zero (the
// buffer range check).
struct BufView {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ BufView(const void* p, unsigned bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                            0x00020000)) {}
  __device__ __forceinline__ float2 f2(unsigned lane_bytes, unsigned uni_bytes) const {
    return __builtin_bit_cast(float2,
                              __builtin_amdgcn_raw_buffer_load_b64(r, lane_bytes, uni_bytes, 0));
  }
};

// Row passes: NT transforms per block (at most 256 threads, 64 KiB LDS).
template <int L>
constexpr int rows_nt() {
  constexpr int by_thr = 256 / Fft2<L, false>::T;
  constexpr int by_lds = kLdsBlockMax / (Fft2<L, false>::LDS_ROW * 4);
  return divisor_le(L, by_thr < by_lds ? by_thr : by_lds);
}
// Column passes: NC adjacent columns per block, up to 32 (128-byte segments
// even for float outputs), at most 1024 threads and 64 KiB LDS; lines of
// 513-1023 keep 256 threads (their register FFTs allow 3 waves per SIMD,
// which 4-wave blocks fill), lines from 1024 take 512 (2 waves per SIMD:
// one 8-wave block per CU either way, and twice the row segment).
// MAXT lowers the thread cap (see op_col_threads).
template <int L, int MAXT = 1024>
constexpr int cols_nc() {
  constexpr int cap = L >= 1024 ? 512 : L > 512 ? 256 : 1024;
  constexpr int by_thr = (MAXT < cap ? MAXT : cap) / Fft2<L, false>::T;
  constexpr int by_lds = kLdsBlockMax / (Fft2<L, false>::LDS_COL * 4);
  constexpr int m = by_thr < by_lds ? by_thr : by_lds;
  return divisor_le(L, m < 32 ? m : 32);
}
template <int L, int MAXT = 1024>
constexpr int cols_threads() {
  return cols_nc<L, MAXT>() * Fft2<L, false>::T;
}

// An operation may cap the threads of its column blocks with a static
// kColThreads: two half-size blocks per CU overlap one block's barriers with
// the other's memory traffic, which pays where the stores are whole 128-byte
// lines either way (complex outputs) but not for float outputs.
template <class Op, class = void>
struct op_col_threads {
  static constexpr int value = 1024;
};
template <class Op>
struct op_col_threads<Op, std::void_t<decltype(Op::kColThreads)>> {
  static constexpr int value = Op::kColThreads;
};
// Column blocks are up to 1024 threads; ask the compiler for at least 4
// waves per SIMD (at most 128 VGPRs) so that a CU keeps 16 waves in flight,
// 3 for lines longer than 512 and 2 from 1024 (a 32 + 32 point register FFT
// needs about 170 VGPRs: 3 waves would spill).
// (HIP's second __launch_bounds__ argument is amdgpu_waves_per_eu.)
template <int L>
constexpr int col_waves_per_simd() {
  return L >= 1024 ? 2 : L > 512 ? 3 : 4;
}

// Tile variants TV of the passes (a schedule the autotuner picks; every
// variant does the same arithmetic in the same order, so results are
// identical): TV 0 the shapes above; TV 1 half the lines (rows) or columns
// per block, twice the blocks; TV 2 twice the columns per block where the
// block still fits one CU at the pass's waves per SIMD, its LDS within
// 64 KiB and the width within 32 and dividing L (else TV 0's width).
template <int L, int TV>
constexpr int rows_nt_tv() {
  constexpr int nt = rows_nt<L>();
  return TV == 1 && nt % 2 == 0 ? nt / 2 : nt;
}
template <int L, int MAXT, int TV>
constexpr int cols_nc_tv() {
  constexpr int nc = cols_nc<L, MAXT>();
  constexpr int T = Fft2<L, false>::T;
  constexpr int waves = col_waves_per_simd<L>();
  if constexpr (TV == 1) {
    return nc % 2 == 0 ? nc / 2 : nc;
  } else if constexpr (TV == 2) {
    constexpr int w = 2 * nc;
    constexpr bool fits = w <= 32 && L % w == 0 && w * T <= 1024 && w * T <= 256 * waves &&
                          w * Fft2<L, false>::LDS_COL * 4 <= kLdsBlockMax;
    return fits ? w : nc;
  }
  return nc;
}

// Workgroups are dealt round-robin to the 8 XCDs in dispatch order (linear
// id, x fastest). Renumber them so that XCD k runs the k-th eighth of the
// line/column groups of every batch, in order: neighbouring column groups
// share that XCD's L2 for the 128-byte lines they split, and every batch
// (e.g. a z plane, some of which may be skipped) is spread over all XCDs.
// A bijection when gridDim.x is a multiple of 8; the identity otherwise.
__device__ __forceinline__ int2 xcd_block() {
  const int gx = gridDim.x;
  if (gx % 8) return make_int2(blockIdx.y, blockIdx.x);
  const int id = blockIdx.x + blockIdx.y * gx;
  const int per = gx / 8, n = id / 8;
  return make_int2(n / per, (id % 8) * per + n % per);
}

// An operation may define skip(batch): the whole batch is then left alone
// (neither read nor written) when it returns true.
template <class Op, class = void>
struct has_skip : std::false_type {};
template <class Op>
struct has_skip<Op, std::void_t<decltype(std::declval<const Op&>().skip(0))>>
    : std::true_type {};

template <class Op>
__device__ __forceinline__ bool skip_batch(const Op& op, int a) {
  if constexpr (has_skip<Op>::value) return op.skip(a);
  return false;
}

// An operation may define line(batch, line), called once per transform
// before its loads, to hoist per-line work out of load().
template <class Op, class = void>
struct has_line : std::false_type {};
template <class Op>
struct has_line<Op, std::void_t<decltype(std::declval<Op&>().line(0, 0))>> : std::true_type {};

template <int L, bool INV, class Op, int TV = 0>
__global__ void __launch_bounds__(256) k_rows(Op op) {
  using F = Fft2<L, INV>;
  constexpr int NT = rows_nt_tv<L, TV>();
  __shared__ float lds[NT * F::LDS_ROW];
  const int2 ab = xcd_block();
  const int a = ab.x;
  if (skip_batch(op, a)) return;  // block-uniform, before the barrier
  const int r = threadIdx.x / F::T, n2 = threadIdx.x % F::T;
  const int line = ab.y * NT + r;
  op.begin(a);
  if constexpr (has_line<Op>::value) op.line(a, line);
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1) v[n1] = op.load(a, line, F::L2 * n1 + n2);
  F::run(v, u, n2, typename F::Row{lds + r * F::LDS_ROW}, op.tw);
  if (n2 < F::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) op.store(a, line, n2 + F::L1 * k2, u[k2]);
  }
}

template <int L, bool INV, class Op, int TV = 0>
__global__ void __launch_bounds__((cols_nc_tv<L, op_col_threads<Op>::value, TV>() *
                                   Fft2<L, false>::T),
                                  col_waves_per_simd<L>()) k_cols(Op op) {
  using F = Fft2<L, INV>;
  constexpr int NC = cols_nc_tv<L, op_col_threads<Op>::value, TV>();
  __shared__ float lds[NC * F::LDS_COL];
  const int2 ab = xcd_block();
  const int a = ab.x, c0 = ab.y * NC;
  if (skip_batch(op, a)) return;  // block-uniform, before the barrier
  const int c = threadIdx.x % NC, n2 = threadIdx.x / NC;
  op.begin(a);
  float2 v[F::L1], u[F::L2];
#pragma unroll
  for (int n1 = 0; n1 < F::L1; ++n1) v[n1] = op.load(a, c0 + c, F::L2 * n1 + n2);
  F::run(v, u, n2, typename F::template Col<NC>{lds + c}, op.tw);
  if (n2 < F::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F::L2; ++k2) op.store(a, c0 + c, n2 + F::L1 * k2, u[k2]);
  }
}

// The column pass on the balanced transform (Fft2B): NC adjacent columns per
// block (32: 128-byte rows even for float outputs), T = L1 threads per
// column, compiled for 4 waves per SIMD: two blocks per CU at L = 512 (512
// threads, 64 KiB), where k_cols runs one 1024-thread block and leaves half
// of each column's threads idle in its second step. An operation opts in with
// a static constexpr bool kBalancedCols = true.
template <class Op, class = void>
struct op_cols_balanced : std::false_type {};
template <class Op>
struct op_cols_balanced<Op, std::void_t<decltype(Op::kBalancedCols)>>
    : std::integral_constant<bool, Op::kBalancedCols> {};

template <int L>
constexpr int cols_b_nc() {
  return 32 * L * 4 <= kLdsBlockMax ? 32 : 16;
}
// lengths that take the balanced pass: L1 | L2 with L1 < L2, where Fft2
// idles half a line's threads in its second step. At L = 1024 (32 x 32, 16
// columns per block) it took 128 VGPRs with 12-19 spilled and was slower
// than k_cols (1024^3 z pass 4.4-5.05 -> 5.4 ms, y pass 1.08 -> 1.21 ms).
template <int L>
constexpr bool cols_b_fits() {
  constexpr int l1 = fft2_l1(L), l2 = fft2_l2(L), nc = cols_b_nc<L>();
  return l1 > 1 && l2 % l1 == 0 && l1 < l2 && L % nc == 0 && nc * L * 4 <= kLdsBlockMax &&
         64 / nc <= l1;
}
template <int L, bool INV, class Op>
__global__ void __launch_bounds__((cols_b_nc<L>() * Fft2B<L, false>::T), 4) k_cols_b(Op op) {
  using F = Fft2B<L, INV>;
  constexpr int NC = cols_b_nc<L>();
  __shared__ float lds[NC * L];
  const int2 ab = xcd_block();
  const int a = ab.x, c0 = ab.y * NC;
  if (skip_batch(op, a)) return;  // block-uniform, before the barrier
  const int c = threadIdx.x % NC, t = threadIdx.x / NC;
  op.begin(a);
  float2 v[F::L2];
#pragma unroll
  for (int m = 0; m < F::L2; ++m) v[m] = op.load(a, c0 + c, F::L1 * m + t);
  F::run(v, t, typename F::template Col<NC>{lds + c}, op.tw);
#pragma unroll
  for (int j = 0; j < F::R; ++j)
#pragma unroll
    for (int k1 = 0; k1 < F::L1; ++k1)
      op.store(a, c0 + c, F::R * t + j + F::L2 * k1, v[j * F::L1 + k1]);
}
template <int L>
inline dim3 cols_b_grid(int A) {
  return dim3(L / cols_b_nc<L>(), A);
}
template <int L>
inline dim3 cols_b_block() {
  return dim3(cols_b_nc<L>() * Fft2B<L, false>::T);
}

// An operation may define mid_line(batch, line), called once per transform
// between the first transform and its mid() calls: per-line state that only
// mid() reads is then not held in registers through the first transform.
template <class Op, class = void>
struct has_mid_line : std::false_type {};
template <class Op>
struct has_mid_line<Op, std::void_t<decltype(std::declval<Op&>().mid_line(0, 0))>>
    : std::true_type {};

// Two chained transforms per row: the INV1 transform of the loaded line,
// op.mid(batch, line, element, value) on every output element, the values
// moved back to the input layout through the exchange storage, the INV2
// transform, op.store. The line set up by op.line(batch, line) is kept for
// the whole kernel (one line per NT-th of the block).
// waves per SIMD asked of the chained row pass: 4 (at most 128 VGPRs) for
// 128- and 256-point lines, which then fit without spilling (other lengths
// of that size would spill); the rest leave the register budget to the
// compiler
template <int L>
constexpr int rows2_waves_per_simd() {
  return (L == 128 || L == 256) ? 4 : 1;
}
template <int L, bool INV1, bool INV2, class Op>
__global__ void __launch_bounds__(256, rows2_waves_per_simd<L>()) k_rows2(Op op) {
  using F1 = Fft2<L, INV1>;
  using F2 = Fft2<L, INV2>;
  constexpr int NT = rows_nt<L>();
  __shared__ float lds[NT * F1::LDS_ROW];
  const int2 ab = xcd_block();
  const int a = ab.x;
  const int r = threadIdx.x / F1::T, n2 = threadIdx.x % F1::T;
  const int line = ab.y * NT + r;
  op.begin(a);
  op.line(a, line);
  float2 v[F1::L1], u[F1::L2];
#pragma unroll
  for (int n1 = 0; n1 < F1::L1; ++n1) v[n1] = op.load(a, line, F1::L2 * n1 + n2);
  const typename F1::Row x{lds + r * F1::LDS_ROW};
  F1::run(v, u, n2, x, op.tw);
  if constexpr (has_mid_line<Op>::value) op.mid_line(a, line);
  if (n2 < F1::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F1::L2; ++k2) u[k2] = op.mid(a, line, n2 + F1::L1 * k2, u[k2]);
  }
  fft2_relayout<L>(u, v, n2, x);
  F2::run(v, u, n2, x, op.tw);
  if (n2 < F2::L1) {
#pragma unroll
    for (int k2 = 0; k2 < F2::L2; ++k2) op.store(a, line, n2 + F2::L1 * k2, u[k2]);
  }
}

template <int L, int TV = 0>
inline dim3 rows_grid(int A) {
  return dim3(L / rows_nt_tv<L, TV>(), A);
}
template <int L, int TV = 0>
inline dim3 rows_block() {
  return dim3(rows_nt_tv<L, TV>() * Fft2<L, false>::T);
}
template <int L, class Op = void, int TV = 0>
inline dim3 cols_grid(int A) {
  return dim3(L / cols_nc_tv<L, op_col_threads<Op>::value, TV>(), A);
}
template <int L, class Op = void, int TV = 0>
inline dim3 cols_block() {
  return dim3(cols_nc_tv<L, op_col_threads<Op>::value, TV>() * Fft2<L, false>::T);
}

}  // namespace mtd
