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

namespace {

struct Vol {
  const float* z;
  int V;
  __device__ __forceinline__ float at(int x, int y, int k) const {
    return z[((size_t)k * V + y) * V + x];
  }
  // forward differences of voxel (x, y, k) and the reciprocal of their
  // smoothed norm (one hardware reciprocal square root, about 1 ulp)
  __device__ __forceinline__ void diffs(int x, int y, int k, float c, float eps2, float& dx,
                                        float& dy, float& dz, float& rs) const {
    dx = x + 1 < V ? at(x + 1, y, k) - c : 0.f;
    dy = y + 1 < V ? at(x, y + 1, k) - c : 0.f;
    dz = k + 1 < V ? at(x, y, k + 1) - c : 0.f;
    rs = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz + eps2);
  }
};

}  // namespace

__global__ void __launch_bounds__(256)
k_tv_fista(const float* __restrict__ z, const float* __restrict__ gd, float* __restrict__ x,
           float* __restrict__ zn, int V, float sd, float lam, float eps, float step,
           float beta, int positivity) {
  // 1-D grid of V*V rows x cx = ceil(V/256) segments. Blocks are dealt
  // round-robin to the 8 XCDs; renumbered so that XCD k sweeps the k-th
  // eighth of the rows in order, the rows y +- 1 and planes k +- 1 the
  // stencil reads were just read by the same XCD (its L2).
  const int cx = (V + 255) / 256;
  const unsigned nb = gridDim.x;
  unsigned r = blockIdx.x;
  if (nb % 8 == 0) r = (r % 8) * (nb / 8) + r / 8;
  const int seg = (int)(r % cx);
  const unsigned row = r / cx;
  const int xi = seg * 256 + threadIdx.x, yi = (int)(row % V), ki = (int)(row / V);
  if (xi >= V) return;
  const size_t i = ((size_t)ki * V + yi) * V + xi;
  const Vol vz{z, V};
  const float eps2 = eps * eps;
  const float c = z[i];
  float dx, dy, dz, rs;
  vz.diffs(xi, yi, ki, c, eps2, dx, dy, dz, rs);
  float gtv = -(dx + dy + dz) * rs;
  if (xi > 0) {
    const float cl = vz.at(xi - 1, yi, ki);
    float ax, ay, az, ar;
    vz.diffs(xi - 1, yi, ki, cl, eps2, ax, ay, az, ar);
    gtv += ax * ar;
  }
  if (yi > 0) {
    const float cl = vz.at(xi, yi - 1, ki);
    float ax, ay, az, ar;
    vz.diffs(xi, yi - 1, ki, cl, eps2, ax, ay, az, ar);
    gtv += ay * ar;
  }
  if (ki > 0) {
    const float cl = vz.at(xi, yi, ki - 1);
    float ax, ay, az, ar;
    vz.diffs(xi, yi, ki - 1, cl, eps2, ax, ay, az, ar);
    gtv += az * ar;
  }
  float xn = c - step * (sd * gd[i] + lam * gtv);
  if (positivity) xn = fmaxf(xn, 0.f);
  const float xo = x[i];
  x[i] = xn;
  zn[i] = xn + beta * (xn - xo);
}

// Data-term gradient of the weighted least squares on the spectra, in place:
// S = (S - y) * w (w NULL: 1), the product taken per component as a complex
// times a real weight.
__global__ void __launch_bounds__(256)
k_residual(float2* __restrict__ S, const float2* __restrict__ y, const float* __restrict__ w,
           size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float2 s = S[i], m = y[i];
  float2 r = make_float2(s.x - m.x, s.y - m.y);
  if (w) {
    const float k = w[i];
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

hipError_t launch_residual(float2* S, const float2* y, const float* w, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  Timed tm("residual", s);
  hipLaunchKernelGGL(k_residual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, y, w, n);
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
  const unsigned nb = (unsigned)((V + 255) / 256) * (unsigned)V * (unsigned)V;
  hipLaunchKernelGGL(k_tv_fista, dim3(nb), dim3(256), 0, s, z, gd, x, zn, V, sd, lam, eps, step,
                     beta, positivity);
  return hipGetLastError();
}

}  // namespace mtd
