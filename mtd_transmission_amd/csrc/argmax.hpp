// Argmax of coordSpec (pretraitement_hors_axe_CPU/fonctions.cpp:88-101):
// slot 0 compares |TF[0]|^2, every other slot |TF|; strict '>' keeps the
// first maximum. Shared by the standalone argmax and the hologram column
// pass, which computes the partials of the crops it writes.
#pragma once
#include <hip/hip_runtime.h>

namespace mtd {

// key of crop element i (slot 0 squared, computed unfused as the reference)
__device__ __forceinline__ float argmax_key(float2 v, bool slot0) {
#pragma clang fp contract(off)
  return slot0 ? v.x * v.x + v.y * v.y : hypotf(v.x, v.y);
}

// (v2, i2) beats (v1, i1): larger value, then smaller index
__device__ __forceinline__ bool argmax_better(float v2, int i2, float v1, int i1) {
  return v2 > v1 || (v2 == v1 && i2 < i1);
}

// Reduces (best, bi) over the workgroup; thread 0 gets the result. red must
// hold 2 * (blockDim.x / 64) words. Contains barriers.
__device__ __forceinline__ void argmax_block(float& best, int& bi, float* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    const float v2 = __shfl_xor(best, o, 64);
    const int i2 = __shfl_xor(bi, o, 64);
    if (argmax_better(v2, i2, best, bi)) { best = v2; bi = i2; }
  }
  __syncthreads();
  if (lane == 0) { red[2 * wave] = best; red[2 * wave + 1] = __int_as_float(bi); }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      const float v2 = red[2 * w];
      const int i2 = __float_as_int(red[2 * w + 1]);
      if (argmax_better(v2, i2, best, bi)) { best = v2; bi = i2; }
    }
  }
}

}  // namespace mtd
