// Volume stage: deterministic Ewald-cap accumulation and the finalisation to
// a refractive-index volume.
//
// Accumulation (retroPropag_Born, Reconstruction/src/fonctions.cpp:875-935)
// is restated as a gather. For a lateral voxel column q = (X - V/2, Y - V/2),
// angle a contributes at most one hit: at pupil pixel fd = q + fm0(a), at
// height z = round(h(|fd|^2) - fzm0(a)) + V/2 with h(r2) = sqrt(R^2 - r2),
// provided fd lies in the [-NXMAX, NXMAX)^2 crop, |fd|^2 < coef*NXMAX^2 and the
// angle is not skipped. Each thread owns one column and walks the angles in
// order, so every voxel receives its hits in the reference's angle order:
// this is the sort-by-voxel of the accumulation done analytically (the
// scatter map is invertible per angle), and the sum is deterministic and
// equal to the sequential FP32 sum. Partial columns live in LDS windows of
// ZW heights; every voxel of the V^3 spectrum and count volume is written
// exactly once (zeros included), so no clearing pass is needed.
//
// Finalisation (Reconstruction/main.cpp:243-352): divide by counts (0 -> 1),
// backward 3-D FFT with both fftshift3D folded into (-1)^(x+y+z) sign flips,
// times delta_f^3 and -1/(2 n0 kv^2), split into index and absorption.
#include "fft_lds.hpp"
#include "internal.hpp"
#include "passes.hpp"

namespace mtd {

// C round(): halves away from zero, exact.
__device__ __forceinline__ double round_away(double x) {
  const double t = trunc(x);
  return (fabs(x - t) >= 0.5) ? t + copysign(1.0, x) : t;
}

__global__ void k_angle_geo(const int* __restrict__ posx,
                            const int* __restrict__ posy, int A, int nx,
                            double R2, const double* __restrict__ htab,
                            int r2tab, AngleGeo* __restrict__ ageo) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A) return;
  AngleGeo g;
  g.fx = posx[a] - nx;
  g.fy = posy[a] - nx;
  const int r2 = g.fx * g.fx + g.fy * g.fy;
  const double fz2 = (R2 - (double)(g.fx * g.fx)) - (double)(g.fy * g.fy);
  g.skip = round_away(fz2) < 0.0;
  g.pad = 0;
  g.fz = (!g.skip && r2 < r2tab) ? htab[r2] : 0.0;
  ageo[a] = g;
}

struct GatherArgs {
  const float2* spec;   // [A][C][C]
  int frame_pos;        // 1: spectra in the +posSpec frame, 0: unshifted
  const AngleGeo* ageo;
  int A, C, nx, V;
  const double* htab;   // h(r2) = sqrt(R^2 - r2)
  const float* beta;    // per r2: ((k0/PI)*(-2PI)) * h/R
  int r2max;            // largest r2 inside the pupil
  float2* vol;
  int32_t* cnt;
};

constexpr int GT = 16;  // tile side (columns per block = 256)

// One hit test; returns true and fills z/val when angle a hits this column.
__device__ __forceinline__ bool gather_hit(const GatherArgs& g, const AngleGeo& ag,
                                           int a, int qx, int qy, int& z,
                                           float2& val) {
  const int fdx = qx + ag.fx, fdy = qy + ag.fy;
  if (fdx < -g.nx || fdx >= g.nx || fdy < -g.nx || fdy >= g.nx) return false;
  const int r2 = fdx * fdx + fdy * fdy;
  if (r2 > g.r2max) return false;
  z = (int)round_away(g.htab[r2] - ag.fz) + g.V / 2;
  int row, col;
  if (g.frame_pos) {
    row = fdy + g.nx;
    col = fdx + g.nx;
  } else {
    row = ((qy % g.C) + g.C) % g.C;
    col = ((qx % g.C) + g.C) % g.C;
  }
  const float2 s = g.spec[((size_t)a * g.C + row) * g.C + col];
  const float b = g.beta[r2];
  val = make_float2(-__fmul_rn(b, s.y), __fmul_rn(b, s.x));
  return true;
}

// Uniform (whole-tile) rejection of angle a.
__device__ __forceinline__ bool tile_misses(const GatherArgs& g, const AngleGeo& ag,
                                            int qx0, int qy0) {
  if (ag.skip) return true;
  const int lx = qx0 + ag.fx, hx = lx + GT - 1;
  const int ly = qy0 + ag.fy, hy = ly + GT - 1;
  if (hx < -g.nx || lx >= g.nx || hy < -g.nx || ly >= g.nx) return true;
  const int dx = lx > 0 ? lx : (hx < 0 ? -hx : 0);
  const int dy = ly > 0 ? ly : (hy < 0 ? -hy : 0);
  return dx * dx + dy * dy > g.r2max;
}

template <int ZW>
__global__ void __launch_bounds__(256) k_gather(GatherArgs g) {
  __shared__ float2 acc[ZW][256];
  __shared__ int cn[ZW][256];
  __shared__ int zred[2][4];
  const int tid = threadIdx.x;
  const int tx = tid % GT, ty = tid / GT;
  const int X = blockIdx.x * GT + tx, Y = blockIdx.y * GT + ty;
  const int qx0 = blockIdx.x * GT - g.V / 2, qy0 = blockIdx.y * GT - g.V / 2;
  const int qx = X - g.V / 2, qy = Y - g.V / 2;
  const size_t plane = (size_t)g.V * g.V;
  const size_t col = (size_t)Y * g.V + X;

  // height range of this tile
  int zlo = 0x7fffffff, zhi = -0x7fffffff;
  for (int a = 0; a < g.A; ++a) {
    const AngleGeo ag = g.ageo[a];
    if (tile_misses(g, ag, qx0, qy0)) continue;
    int z;
    float2 v;
    if (gather_hit(g, ag, a, qx, qy, z, v)) {
      zlo = min(zlo, z);
      zhi = max(zhi, z);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    zlo = min(zlo, __shfl_xor(zlo, off, 64));
    zhi = max(zhi, __shfl_xor(zhi, off, 64));
  }
  if ((tid & 63) == 0) { zred[0][tid >> 6] = zlo; zred[1][tid >> 6] = zhi; }
  __syncthreads();
  zlo = min(min(zred[0][0], zred[0][1]), min(zred[0][2], zred[0][3]));
  zhi = max(max(zred[1][0], zred[1][1]), max(zred[1][2], zred[1][3]));
  if (zlo < 0) zlo = 0;
  if (zhi > g.V - 1) zhi = g.V - 1;

  // zero heights outside [zlo, zhi]
  const float2 zero2 = make_float2(0.f, 0.f);
  for (int z = 0; z < g.V; ++z) {
    if (z >= zlo && z <= zhi) { z = zhi; continue; }
    g.vol[z * plane + col] = zero2;
    g.cnt[z * plane + col] = 0;
  }
  // windows of ZW heights
  for (int w0 = zlo; w0 <= zhi; w0 += ZW) {
#pragma unroll
    for (int k = 0; k < ZW; ++k) { acc[k][tid] = zero2; cn[k][tid] = 0; }
    for (int a = 0; a < g.A; ++a) {
      const AngleGeo ag = g.ageo[a];
      if (tile_misses(g, ag, qx0, qy0)) continue;
      int z;
      float2 v;
      if (gather_hit(g, ag, a, qx, qy, z, v)) {
        const int k = z - w0;
        if (k >= 0 && k < ZW) {
          float2 s = acc[k][tid];
          s.x = __fadd_rn(s.x, v.x);
          s.y = __fadd_rn(s.y, v.y);
          acc[k][tid] = s;
          cn[k][tid] += 1;
        }
      }
    }
    const int kend = min(ZW, zhi - w0 + 1);
    for (int k = 0; k < kend; ++k) {
      g.vol[(size_t)(w0 + k) * plane + col] = acc[k][tid];
      g.cnt[(size_t)(w0 + k) * plane + col] = cn[k][tid];
    }
  }
}

hipError_t launch_angle_geo(const int* posx, const int* posy, int A, int nxmax,
                            double R2, const double* htab, int r2tab,
                            AngleGeo* ageo, hipStream_t s) {
  Timed tm("angle_geo", s);
  hipLaunchKernelGGL(k_angle_geo, dim3((A + 127) / 128), dim3(128), 0, s, posx,
                     posy, A, nxmax, R2, htab, r2tab, ageo);
  return hipGetLastError();
}

hipError_t launch_gather(const float2* spec, int frame_pos, const int* posx,
                         const int* posy, const AngleGeo* ageo, int A, int C,
                         int nxmax, int V, const double* htab,
                         const float* beta, int r2max, float2* vol,
                         int32_t* cnt, hipStream_t s) {
  (void)posx;
  (void)posy;
  GatherArgs g{spec, frame_pos, ageo, A, C, nxmax, V, htab, beta, r2max, vol, cnt};
  Timed tm("ewald_gather", s);
  hipLaunchKernelGGL((k_gather<16>), dim3(V / GT, V / GT), dim3(256), 0, s, g);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// finalisation passes
// ---------------------------------------------------------------------------
template <int V>
struct NormXOp {  // rows along x; a = z, line = y
  const float2* vol;
  const int32_t* cnt;
  float2* out;
  uint8_t* otf;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int z, int y, int x) const {
    const size_t i = ((size_t)z * V + y) * V + x;
    const float2 v = vol[i];
    const int c0 = cnt[i];
    const float c = (float)(c0 == 0 ? 1 : c0);
    float2 t = make_float2(v.x / c, v.y / c);
    if (otf) otf[i] = (t.x != 0.f) ? 1 : 0;
    if ((x + y + z) & 1) { t.x = -t.x; t.y = -t.y; }
    return t;
  }
  __device__ void store(int z, int y, int x, float2 v) const {
    out[((size_t)z * V + y) * V + x] = v;
  }
};

template <int V>
struct PlaneYOp {  // columns of each z plane; a = z, col = x, row = y
  float2* buf;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int z, int x, int y) const { return buf[((size_t)z * V + y) * V + x]; }
  __device__ void store(int z, int x, int y, float2 v) const { buf[((size_t)z * V + y) * V + x] = v; }
};

template <int V>
struct ZOutOp {  // lines along z; a = y, col = x, row = z
  const float2* buf;
  float* re;
  float* im;
  const float2* tw;
  float df3, cte, gsign;
  __device__ void begin(int) {}
  __device__ float2 load(int y, int x, int z) const { return buf[((size_t)z * V + y) * V + x]; }
  __device__ void store(int y, int x, int z, float2 v) const {
    const float sg = ((x + y + z) & 1) ? -gsign : gsign;
    const size_t i = ((size_t)z * V + y) * V + x;
    re[i] = (v.x * sg * df3) * cte;
    im[i] = (v.y * sg * df3) * cte;
  }
};

template <int V>
static hipError_t finalize_v(const float2* vol, const int32_t* cnt, float2* work,
                             float df3, float cte, float gsign, float* re,
                             float* im, uint8_t* otf, const float2* tw,
                             hipStream_t s) {
  NormXOp<V> x{vol, cnt, work, otf, tw};
  {
    Timed tm("vol_x", s);
    hipLaunchKernelGGL((k_rows<V, true, NormXOp<V>>), rows_grid<V>(V), rows_block<V>(),
                       0, s, x);
  }
  constexpr int NC = cols_nc<V>();
  PlaneYOp<V> y{work, tw};
  {
    Timed tm("vol_y", s);
    hipLaunchKernelGGL((k_cols<V, true, NC, PlaneYOp<V>>), dim3(V / NC, V), dim3(256),
                       0, s, y);
  }
  ZOutOp<V> z{work, re, im, tw, df3, cte, gsign};
  Timed tm("vol_z", s);
  hipLaunchKernelGGL((k_cols<V, true, NC, ZOutOp<V>>), dim3(V / NC, V), dim3(256),
                     0, s, z);
  return hipGetLastError();
}

#define MTD_VOL_SIZES(X) \
  X(32) X(64) X(128) X(256) X(320) X(384) X(448) X(512) X(576) X(640) X(704) X(768) X(1024)

hipError_t launch_finalize(int V, const float2* vol, const int32_t* cnt,
                           float2* work, float df3, float cte, float gsign,
                           float* out_re, float* out_im, uint8_t* otf,
                           const float2* tw, hipStream_t s) {
  switch (V) {
#define X(n) \
  case n: return finalize_v<n>(vol, cnt, work, df3, cte, gsign, out_re, out_im, otf, tw, s);
    MTD_VOL_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

bool vol_len_supported(int V) {
  switch (V) {
#define X(n) case n: return true;
    MTD_VOL_SIZES(X)
#undef X
  }
  return false;
}

}  // namespace mtd
