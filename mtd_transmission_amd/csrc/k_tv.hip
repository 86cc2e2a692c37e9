// One step of the adjoint-iterative form's solver, fused into a single pass
// over the volume: the gradient of the smoothed isotropic total variation
//   TV(z) = sum_i sqrt(dx_i^2 + dy_i^2 + dz_i^2 + eps^2) - eps,
//   d*_i = z[i + e*] - z[i] (0 at the far face of each axis),
// the gradient step with the data gradient, the projection on z >= 0 and the
// FISTA momentum:
//   x' = P(z - step * (sd * gd + lam * gTV(z))),  z' = x' + beta * (x' - x).
// dTV/dz_i collects the terms of voxel i and of its three lower neighbours:
//   -(dx_i + dy_i + dz_i)/s_i + dx_{i-ex}/s_{i-ex} + dy_{i-ey}/s_{i-ey}
//   + dz_{i-ez}/s_{i-ez},  s_j = sqrt(dx_j^2 + dy_j^2 + dz_j^2 + eps^2).
// z and z' must not overlap (neighbours of z are read); x is updated in place.
#include "internal.hpp"

namespace mtd {

// 2.5-D blocking: a block owns a TVX x TVY (64 x 4) column tile and marches
// a segment of TVSEG (32) heights. Planes k-1, k, k+1 of the tile with a one-voxel halo in x
// and y sit in LDS (four rotating slots: plane k+2 is loaded into the slot
// plane k-2 used while step k reads the other three), so z is read from HBM
// about once (the halo and the segment ends re-read a few percent, mostly
// from L2) instead of through thirteen cached loads per voxel.
constexpr int TVX = 64, TVY = 4, TVSEG = 32;
constexpr int THX = TVX + 2, THY = TVY + 2, THP = THX * THY;

__global__ void __launch_bounds__(256)
k_tv_fista(const float* __restrict__ z, const float* __restrict__ gd, float* __restrict__ x,
           float* __restrict__ zn, int V, float sd, float lam, float eps, float step,
           float beta, int positivity) {
  __shared__ float pl[4][THP];
  const int nbx = (V + TVX - 1) / TVX, nby = (V + TVY - 1) / TVY;
  // Blocks are dealt round-robin to the 8 XCDs; renumbered so that XCD k
  // runs the k-th eighth of the tiles in order: the tiles above and below a
  // tile, whose rows form its y halo, then sit in the same XCD's L2
  // (a bijection when the block count is a multiple of 8)
  const unsigned nb = gridDim.x;
  int b = blockIdx.x;
  if (nb % 8 == 0) b = (int)((blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8);
  const int bx = b % nbx;
  b /= nbx;
  const int by = b % nby, seg = b / nby;
  const int X0 = bx * TVX, Y0 = by * TVY;
  const int k0 = seg * TVSEG, k1 = min(k0 + TVSEG, V);
  const int tid = threadIdx.x, tx = tid % TVX, ty = tid / TVX;
  const int X = X0 + tx, Y = Y0 + ty;
  const size_t plane = (size_t)V * V;
  // The tile + halo of a plane is THP < 512 values: elements tid and
  // tid + 256 of each thread. Voxels outside the volume hold 0 (no term
  // reads them: the far-face conditions below exclude them).
  auto fetch = [&](int k, int e) {
    const int hy = e / THX, hx = e - hy * THX;
    const int gx = X0 + hx - 1, gy = Y0 + hy - 1;
    float v = 0.f;
    if (e < THP && k >= 0 && k < V && gx >= 0 && gx < V && gy >= 0 && gy < V)
      v = z[(size_t)k * plane + (size_t)gy * V + gx];
    return v;
  };
  auto put = [&](float* dst, float v0, float v1) {
    dst[tid] = v0;
    if (tid + 256 < THP) dst[tid + 256] = v1;
  };
  // plane k lives in slot (k - k0 + 1) & 3
  put(pl[0], fetch(k0 - 1, tid), fetch(k0 - 1, tid + 256));
  put(pl[1], fetch(k0, tid), fetch(k0, tid + 256));
  put(pl[2], fetch(k0 + 1, tid), fetch(k0 + 1, tid + 256));
  __syncthreads();
  const float eps2 = eps * eps;
  const bool live = X < V && Y < V;
  const size_t col = (size_t)Y * V + X;
  // this column's data gradient and x at the current height, loaded one
  // step ahead (each thread owns its column: no other thread writes them)
  float gdc = 0.f, xc = 0.f;
  if (live) {
    gdc = gd[(size_t)k0 * plane + col];
    xc = x[(size_t)k0 * plane + col];
  }
  for (int k = k0; k < k1; ++k) {
    // issue the next plane's and the next height's loads first; their
    // latency hides behind this height's arithmetic
    const bool more = k + 1 < k1;
    float n0 = 0.f, n1 = 0.f, gdn = 0.f, xn1 = 0.f;
    if (more) {
      n0 = fetch(k + 2, tid);
      n1 = fetch(k + 2, tid + 256);
      if (live) {
        gdn = gd[(size_t)(k + 1) * plane + col];
        xn1 = x[(size_t)(k + 1) * plane + col];
      }
    }
    const float* Pm = pl[(k - k0) & 3];
    const float* Pc = pl[(k - k0 + 1) & 3];
    const float* Pp = pl[(k - k0 + 2) & 3];
    auto at = [&](const float* P, int dx, int dy) {
      return P[(ty + 1 + dy) * THX + (tx + 1 + dx)];
    };
    if (live) {
      // forward differences (0 across the far faces) and the reciprocal of
      // their smoothed norm, of this voxel and of its three lower neighbours
      const float c = at(Pc, 0, 0);
      const float dx = X + 1 < V ? at(Pc, 1, 0) - c : 0.f;
      const float dy = Y + 1 < V ? at(Pc, 0, 1) - c : 0.f;
      const float dz = k + 1 < V ? at(Pp, 0, 0) - c : 0.f;
      const float rs = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz + eps2);
      float gtv = -(dx + dy + dz) * rs;
      if (X > 0) {
        const float cl = at(Pc, -1, 0);
        const float ax = c - cl;
        const float ay = Y + 1 < V ? at(Pc, -1, 1) - cl : 0.f;
        const float az = k + 1 < V ? at(Pp, -1, 0) - cl : 0.f;
        gtv += ax * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az + eps2);
      }
      if (Y > 0) {
        const float cl = at(Pc, 0, -1);
        const float ax = X + 1 < V ? at(Pc, 1, -1) - cl : 0.f;
        const float ay = c - cl;
        const float az = k + 1 < V ? at(Pp, 0, -1) - cl : 0.f;
        gtv += ay * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az + eps2);
      }
      if (k > 0) {
        const float cl = at(Pm, 0, 0);
        const float ax = X + 1 < V ? at(Pm, 1, 0) - cl : 0.f;
        const float ay = Y + 1 < V ? at(Pm, 0, 1) - cl : 0.f;
        const float az = c - cl;
        gtv += az * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az + eps2);
      }
      const size_t i = (size_t)k * plane + col;
      float xn = c - step * (sd * gdc + lam * gtv);
      if (positivity) xn = fmaxf(xn, 0.f);
      x[i] = xn;
      zn[i] = xn + beta * (xn - xc);
    }
    // plane k+2 into the slot plane k-2 used (no thread reads it this step)
    if (more) put(pl[(k - k0 + 3) & 3], n0, n1);
    gdc = gdn;
    xc = xn1;
    __syncthreads();  // plane k+2 is complete; slot of plane k-1 is free
  }
}

// Data-term gradient of the weighted least squares on the spectra, in place:
// S = (S - y) * w (w NULL: 1), the product taken per component as a complex
// times a real weight; q2 (C x C pupil frame, or NULL): then also times the
// forward factor q of the bin, the factor the adjoint's gather would apply
// per hit (the same FP32 product, so the adjoint can skip it).
__global__ void __launch_bounds__(256)
k_residual(float2* __restrict__ S, const float2* __restrict__ y, const float* __restrict__ w,
           const float* __restrict__ q2, int C2) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= C2) return;
  const size_t i = (size_t)blockIdx.y * C2 + j;
  const float2 s = S[i], m = y[i];
  float2 r = make_float2(s.x - m.x, s.y - m.y);
  if (w) {
    const float k = w[i];
    r = make_float2(r.x * k, r.y * k);
  }
  if (q2) {
    const float k = q2[j];
    r = make_float2(r.x * k, r.y * k);
  }
  S[i] = r;
}

// Solver start: x = max(x, 0) when positivity holds, z = x.
__global__ void __launch_bounds__(256)
k_tv_start(float* __restrict__ x, float* __restrict__ z, size_t n, int positivity) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float v = x[i];
  if (positivity) v = fmaxf(v, 0.f);
  x[i] = v;
  z[i] = v;
}

hipError_t launch_residual(float2* S, const float2* y, const float* w, const float* q2, int A,
                           int C2, hipStream_t s) {
  if (A <= 0 || C2 <= 0) return hipSuccess;
  Timed tm("residual", s);
  hipLaunchKernelGGL(k_residual, dim3((C2 + 255) / 256, A), dim3(256), 0, s, S, y, w, q2, C2);
  return hipGetLastError();
}

hipError_t launch_tv_start(float* x, float* z, size_t n, int positivity, hipStream_t s) {
  hipLaunchKernelGGL(k_tv_start, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, z, n,
                     positivity);
  return hipGetLastError();
}

hipError_t launch_tv_fista(int V, const float* z, const float* gd, float* x, float* zn,
                           float sd, float lam, float eps, float step, float beta,
                           int positivity, hipStream_t s) {
  Timed tm("tv_fista", s);
  const unsigned nb = (unsigned)((V + TVX - 1) / TVX) * (unsigned)((V + TVY - 1) / TVY) *
                      (unsigned)((V + TVSEG - 1) / TVSEG);
  hipLaunchKernelGGL(k_tv_fista, dim3(nb), dim3(256), 0, s, z, gd, x, zn, V, sd, lam, eps, step,
                     beta, positivity);
  return hipGetLastError();
}

}  // namespace mtd
