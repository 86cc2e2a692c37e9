// Generic batched 1-D FFT passes over stacks of lines, parameterised by an
// operation object that maps (batch, line, element) to memory on load and on
// store. Rows: each transform is one contiguous line. Columns: NC adjacent
// columns are loaded as a tile (coalesced along the row) and transposed into
// LDS.
#pragma once
#include "fft_lds.hpp"

namespace mtd {

template <int C, bool INV, class Op>
__global__ void __launch_bounds__(256) k_rows(Op op) {
  constexpr int T = fft_threads(C);
  constexpr int NT = divisor_le(C, fft_per_block(C));
  __shared__ float2 lds[NT][C];
  const int a = blockIdx.y;
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int line = blockIdx.x * NT + r;
  op.begin(a);
  for (int i = t; i < C; i += T) lds[r][i] = op.load(a, line, i);
  __syncthreads();
  fft_lds<C, T, INV>(lds[r], t, op.tw);
  for (int i = t; i < C; i += T) op.store(a, line, i, lds[r][i]);
}

template <int C, bool INV, int NC, class Op>
__global__ void __launch_bounds__(256) k_cols(Op op) {
  constexpr int T = 256 / NC;
  __shared__ float2 lds[NC][C + 1];
  const int a = blockIdx.y, c0 = blockIdx.x * NC;
  op.begin(a);
  for (int e = threadIdx.x; e < C * NC; e += 256) {
    const int y = e / NC, c = e % NC;
    lds[c][y] = op.load(a, c0 + c, y);
  }
  __syncthreads();
  const int c = threadIdx.x / T, t = threadIdx.x % T;
  fft_lds<C, T, INV>(lds[c], t, op.tw);
  for (int e = threadIdx.x; e < C * NC; e += 256) {
    const int y = e / NC, cc = e % NC;
    op.store(a, c0 + cc, y, lds[cc][y]);
  }
}

template <int C>
inline dim3 rows_grid(int A) {
  return dim3(C / divisor_le(C, fft_per_block(C)), A);
}
template <int C>
inline dim3 rows_block() {
  return dim3(divisor_le(C, fft_per_block(C)) * fft_threads(C));
}
template <int C>
constexpr int cols_nc() {
  return C >= 256 ? 16 : (C >= 64 ? 8 : 4);
}

}  // namespace mtd
