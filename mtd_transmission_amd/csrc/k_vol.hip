// Volume stage: deterministic Ewald-cap accumulation and the finalisation to
// a refractive-index volume.
//
// Accumulation (retroPropag_Born, Reconstruction/src/fonctions.cpp:875-935)
// is restated as a gather. For a lateral voxel column q = (X - V/2, Y - V/2),
// angle a contributes at most one hit: at pupil pixel fd = q + fm0(a), at
// height z = round(h(|fd|^2) - fzm0(a)) + V/2 with h(r2) = sqrt(R^2 - r2),
// provided fd lies in the [-NXMAX, NXMAX)^2 crop, |fd|^2 < coef*NXMAX^2 and the
// angle is not skipped. Each thread owns one column of a (column tile, height
// window) work item and walks the angles in order, so every voxel receives
// its hits in the reference's angle order: this is the sort-by-voxel of the
// accumulation done analytically (the scatter map is invertible per angle),
// and the sum is deterministic and equal to the sequential FP32 sum. Partial
// columns live in LDS; every voxel of the V^3 spectrum and count volume is
// written exactly once (zeros included), so no clearing pass is needed.
// Banded mode (the fused plain method): hits of all angles lie in a height
// band about V/2 of half-width about R - sqrt(R^2 - NXMAX^2) (a third of the
// volume for the usual geometry); only the windows meeting it are written and
// the inverse transform reads nothing else.
//
// Finalisation (Reconstruction/main.cpp:243-352): divide by counts (0 -> 1),
// backward 3-D FFT with both fftshift3D folded into (-1)^(x+y+z) sign flips,
// times delta_f^3 and -1/(2 n0 kv^2), split into index and absorption.
#include <algorithm>

#include "dft.hpp"
#include "gen_fft.hpp"
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

// ---------------------------------------------------------------------------
// Windowed gather. A prepass lists, per 16x16 column tile and in angle order,
// the angles whose pupil disc touches the tile, with a conservative height
// range derived from the tile box's nearest and farthest pupil radii (h is
// decreasing in r^2 and round is monotonic). The main kernel gives each block
// one (tile, ZW-height window); a wave tests 64 list entries against the
// window at once and takes the overlapping ones in angle order, KB at a
// time. Each column's exact hit height is computed in registers, and only
// the hits' spectrum values are loaded (a miss's buffer-load offset is out
// of range, which fetches nothing). Near the lateral origin nearly every
// angle hits the same few voxels, and those sums are serial by definition
// (angle order): the loads of KB entries are issued together so that only
// the additions stay serial, and those tiles are dispatched first (tiles in
// order of distance from the lateral origin).
// ---------------------------------------------------------------------------
struct TileEntry {
  int a;
  int zpack;  // zlo | zhi << 16, absolute heights
  int fxy;    // fxm0 & 0xffff | fym0 << 16 (signed 16-bit each)
  float fz;   // fzm0 rounded to float (the registers' height path)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// h(r2) = sqrt(R^2 - r2) for 0 <= r2 <= R^2: IEEE sqrt of the same exact
// double, so the host height table's value bit for bit
__device__ __forceinline__ double height(double R2, int r2) { return sqrt(R2 - (double)r2); }

// Height offset round(h(r2) - fz) of a hit (before + V/2), as the reference
// computes it: FP64, round half away from zero. First in FP32: with R^2 <
// 2^24, R^2 - r2 is an exact float; the hardware square root (v_sqrt_f32,
// within 1 ulp, not correctly rounded) of it is within 1 ulp(R), fz32 is fz
// rounded once and the difference is rounded once, each within half an ulp
// of R (h, fz and |h - fz| are at most R): the FP32 difference is within
// 2 ulp(R) of the FP64 one. Wherever the fractional part of |difference| is
// farther than thr = 8 ulp(R) from 1/2 both round alike; near 1/2 (*near is
// set) the caller takes the FP64 path.
__device__ __forceinline__ int height_off_f32(float R2f, int r2, float fz32, float thr,
                                              bool* near) {
  const float d = __builtin_amdgcn_sqrtf(R2f - (float)r2) - fz32;
  const float ad = fabsf(d);
  const float fl = floorf(ad);
  const float fr = ad - fl;  // exact (Sterbenz)
  *near = fabsf(fr - 0.5f) < thr;
  const int zr = (int)fl + (fr >= 0.5f ? 1 : 0);
  return d < 0.f ? -zr : zr;
}

// Does angle a's pupil disc touch the 16x16 column tile at (qx0, qy0)? If so,
// the nearest and farthest pupil radii^2 of the tile box (the farthest
// clamped to the pupil edge).
__device__ __forceinline__ bool tile_touch(const GatherArgs& g, const AngleGeo& ag, int qx0,
                                           int qy0, int& dmin, int& dmax) {
  if (ag.skip) return false;
  const int cxl = max(qx0 + ag.fx, -g.nx), cxh = min(qx0 + GT - 1 + ag.fx, g.nx - 1);
  const int cyl = max(qy0 + ag.fy, -g.nx), cyh = min(qy0 + GT - 1 + ag.fy, g.nx - 1);
  if (cxl > cxh || cyl > cyh) return false;
  const int mx = cxl > 0 ? cxl : (cxh < 0 ? -cxh : 0);
  const int my = cyl > 0 ? cyl : (cyh < 0 ? -cyh : 0);
  const int Mx = max(abs(cxl), abs(cxh)), My = max(abs(cyl), abs(cyh));
  dmin = mx * mx + my * my;
  dmax = min(Mx * Mx + My * My, g.r2max);
  return dmin <= g.r2max;
}

// Prepass: one wave per 16x16 column tile lists, in angle order, the angles
// whose pupil disc touches the tile, 64 angles per ballot, with a
// conservative height range from the tile box's nearest and farthest pupil
// radii (h decreases with r^2, round is monotonic; FP64 heights and
// round-half-away, as the reference), and writes the tile's entry count and
// height range.
__global__ void __launch_bounds__(64)
k_tile_angles(GatherArgs g, int ntx, TileEntry* __restrict__ list, int4* __restrict__ tinfo) {
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int qx0 = (tile % ntx) * GT - g.V / 2, qy0 = (tile / ntx) * GT - g.V / 2;
  const double R2 = g.htab[0] * g.htab[0];  // h(0) = R, an integer: exact
  TileEntry* lst = list + (size_t)tile * g.A;
  int off = 0, zmin = 0x7fffffff, zmax = -1;
  for (int a0 = 0; a0 < g.A; a0 += 64) {
    const int a = a0 + lane;
    bool act = false;
    int zl = 0, zh = 0;
    AngleGeo ag{};
    if (a < g.A) {
      ag = g.ageo[a];
      int dmin, dmax;
      if (tile_touch(g, ag, qx0, qy0, dmin, dmax)) {
        zl = clampi((int)round_away(height(R2, dmax) - ag.fz) + g.V / 2, 0, g.V - 1);
        zh = clampi((int)round_away(height(R2, dmin) - ag.fz) + g.V / 2, 0, g.V - 1);
        act = true;
        zmin = min(zmin, zl);
        zmax = max(zmax, zh);
      }
    }
    const unsigned long long bal = __ballot(act);
    if (act) {
      const int p = __popcll(bal & ((1ull << lane) - 1ull));
      lst[off + p] = TileEntry{a, zl | (zh << 16), (ag.fx & 0xffff) | (ag.fy << 16), (float)ag.fz};
    }
    off += __popcll(bal);
  }
  for (int o = 32; o > 0; o >>= 1) {
    zmin = min(zmin, __shfl_xor(zmin, o, 64));
    zmax = max(zmax, __shfl_xor(zmax, o, 64));
  }
  if (lane == 0) tinfo[tile] = make_int4(off, zmin, zmax, 0);
}

// 8 bytes of a spectrum at byte offset off from base; an offset past nbytes
// returns zeros without a memory access (a buffer load's range check)
__device__ __forceinline__ float2 load_or_zero(const float2* base, uint32_t nbytes, uint32_t off) {
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)nbytes, 0x00020000);
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
constexpr uint32_t kMissOffset = 0x80000000u;  // past every spectrum (< 2^31 bytes)

// one column's fixed data for the gather's hit tests
struct GatherCol {
  int qx, qy;       // lateral column (voxel - V/2)
  int zbase;        // the window's first height - V/2
  uint32_t fixed;   // byte offset of the column's element in the unshifted frame
  float R2f, thr;   // R^2 in FP32 and the FP32 height path's margin
  double R2;
};

// A batch of KB list entries: per entry the window slot of this column's hit
// (-1: none), the angle, the spectrum element's byte offset (kMissOffset for
// a miss) and, after issue(), the loaded value (and beta when !PRE).
template <int KB>
struct GBatch {
  int kv[KB], ea[KB];
  uint32_t el[KB];
  float2 sp[KB];
  float b[KB];
  __device__ __forceinline__ void issue(const GatherArgs& g, size_t C2, uint32_t C2b) {
#pragma unroll
    for (int t = 0; t < KB; ++t) sp[t] = load_or_zero(g.spec + (size_t)ea[t] * C2, C2b, el[t]);
    // keep the loads ahead of what follows (the other batch's sums): the
    // scheduler would otherwise sink them below those independent LDS ops
    __builtin_amdgcn_sched_barrier(0);
  }
  // Only hits update their LDS slot (a lane-masked read-modify-write: the
  // misses of a wave move no LDS data); the loads were issued before the
  // scheduling barrier in issue(), so the branch cannot sink them.
  template <bool PRE, int ZW>
  __device__ __forceinline__ void apply(float2 (&acc)[ZW][256], unsigned short (&cn)[ZW][256],
                                        int tid) const {
#pragma clang fp contract(off)
#pragma unroll
    for (int t = 0; t < KB; ++t) {
      if (kv[t] >= 0) {
        const int kc = kv[t];
        // value (-b Si, b Sr)
        const float pim = PRE ? sp[t].y : b[t] * sp[t].y;
        const float pre = PRE ? sp[t].x : b[t] * sp[t].x;
        float2 s = acc[kc][tid];
        s.x = s.x + -pim;
        s.y = s.y + pre;
        acc[kc][tid] = s;
        cn[kc][tid] = (unsigned short)(cn[kc][tid] + 1);
      }
    }
  }
};

// Batch bi of a segment's compacted entries (seg[0 .. total), angle order):
// entries i0 .. i0 + KB - 1, those past total miss (angle 0, no load). The
// entries are read from LDS at a uniform address (a broadcast).
template <int ZW, bool PRE, int KB>
__device__ __forceinline__ void gather_fill(GBatch<KB>& B, const int4* seg, int i0, int total,
                                            const GatherArgs& g, const GatherCol& c) {
#pragma unroll
  for (int t = 0; t < KB; ++t) {
    const bool valid = i0 + t < total;  // uniform
    const int4 en = seg[valid ? i0 + t : 0];
    const int ea = en.x, fxy = en.y;
    const float fz32 = __int_as_float(en.z);
    const int fdx = c.qx + (int)(short)(fxy & 0xffff), fdy = c.qy + (fxy >> 16);
    const int r2raw = fdx * fdx + fdy * fdy;
    const bool inp = valid && fdx >= -g.nx && fdx < g.nx && fdy >= -g.nx && fdy < g.nx &&
                     r2raw <= g.r2max;
    const int r2 = inp ? r2raw : 0;
    bool near;
    int dz = height_off_f32(c.R2f, r2, fz32, c.thr, &near);
    if (near && inp) dz = (int)round_away(height(c.R2, r2) - g.ageo[ea].fz);  // rare
    const int k = dz - c.zbase;
    const bool hit = inp && k >= 0 && k < ZW;
    B.kv[t] = hit ? k : -1;
    B.ea[t] = valid ? ea : 0;
    B.el[t] = !hit ? kMissOffset
                   : (g.frame_pos ? (uint32_t)((fdy + g.nx) * g.C + (fdx + g.nx)) * 8u : c.fixed);
    if constexpr (!PRE) B.b[t] = g.beta[r2];
  }
}

constexpr int kGatherSeg = 256;  // list entries tested (and compacted) per segment

// flags: bit0 = normalise (divide by counts, 0 -> 1) and apply the centring
// sign (-1)^(x+y+z) of the 3-D inverse FFT; counts are then not written.
// bit2 = the plain sums with that centring sign, no counts (the adjoint of
// the Born forward model).
// PRE: the spectra are already multiplied by the Ewald factor beta of their
// pupil pixel (the same FP32 product), so beta is not looked up here.
// KB: list entries whose loads are issued together.
template <int ZW, bool PRE, int KB>
__global__ void __launch_bounds__(256)
k_gather_win(GatherArgs g, int ntx, int w0, int nwin, int items, const int* __restrict__ order,
             const TileEntry* __restrict__ list, const int4* __restrict__ tinfo, int flags,
             const int2* __restrict__ band) {
  // The sums must be the sequential FP32 sums of the reference: no fused
  // multiply-add. The pragma is lexical, so the products and sums are written
  // here as plain operators (HIP's __fmul_rn/__fadd_rn are header functions
  // outside its scope and would still be contracted).
#pragma clang fp contract(off)
  // Blocks are dealt round-robin to the 8 XCDs. Tiles go to the XCDs in
  // order of distance from the lateral origin (order[]: the heaviest first,
  // spread over the XCDs), and the windows of one tile stay on one XCD, in
  // order (they read the same spectrum patches). A bijection when the tile
  // count is a multiple of 8. Windows w0 .. w0 + nwin - 1 are launched.
  int tile, w;
  {
    const int b = blockIdx.x, tiles = items / nwin;
    int rank;
    if (tiles % 8 == 0) {
      const int n = b / 8;
      rank = (n / nwin) * 8 + b % 8;
      w = w0 + n % nwin;
    } else {
      rank = b / nwin;
      w = w0 + b % nwin;
    }
    tile = order[rank];
  }
  const int tid = threadIdx.x;
  const int X = (tile % ntx) * GT + tid % GT, Y = (tile / ntx) * GT + tid / GT;
  const int qx = X - g.V / 2, qy = Y - g.V / 2;
  const size_t plane = (size_t)g.V * g.V;
  const size_t col = (size_t)Y * g.V + X;
  // the last tiles of a volume side that is not a multiple of the tile side
  // hold columns past the volume: they walk the lists like the others (the
  // ballots are per wave) but write nothing
  const bool inb = X < g.V && Y < g.V;
  const int z0 = w * ZW, z1 = min(z0 + ZW, g.V) - 1;
  // heights this block writes: the window, cut to the hit band when the
  // output is banded (no pass reads a plane outside the band)
  int zs0 = z0, zs1 = z1;
  if (band) {
    const int2 b = *band;
    if (z1 < b.x || z0 > b.y) return;
    zs0 = max(z0, b.x);
    zs1 = min(z1, b.y);
  }
  const int4 info = tinfo[tile];
  const bool norm = flags & 1, sgn = flags & 5;
  if (info.x == 0 || z1 < info.y || z0 > info.z) {
    if (inb) {
      for (int z = zs0; z <= zs1; ++z) {
        g.vol[z * plane + col] = make_float2(0.f, 0.f);
        if (!sgn) g.cnt[z * plane + col] = 0;
      }
    }
    return;
  }
  // this column's ZW partial sums and hit counts (< 2^16: at most one hit
  // per angle) live in LDS; each thread only touches its own column
  __shared__ float2 acc[ZW][256];
  __shared__ unsigned short cn[ZW][256];
#pragma unroll
  for (int k = 0; k < ZW; ++k) { acc[k][tid] = make_float2(0.f, 0.f); cn[k][tid] = 0; }
  const TileEntry* lst = list + (size_t)tile * g.A;
  // spectrum element read by this column: the unshifted frame holds it at
  // (qy mod C, qx mod C) for every angle; the +posSpec frame at fd + NXMAX
  const int frow = ((qy % g.C) + g.C) % g.C, fcol = ((qx % g.C) + g.C) % g.C;
  const uint32_t fixed = (uint32_t)(frow * g.C + fcol) * 8u;
  const uint32_t C2b = (uint32_t)g.C * (uint32_t)g.C * 8u;  // bytes per angle
  const size_t C2 = (size_t)g.C * g.C;
  const double R2 = g.htab[0] * g.htab[0];
  const float R2f = (float)R2;  // an integer: exact below 2^24
  // the FP32 height path's margin, 8 ulp(R); everything takes the FP64 path
  // where R^2 - r2 would not be exact in FP32
  const float thr = R2 < 16777216.0 ? ldexpf(1.f, ilogbf((float)g.htab[0]) - 20) : 2.f;
  const int zbase = z0 - g.V / 2;
  // The tile's list in segments of kGatherSeg entries: each thread tests one
  // entry against the window, the overlapping ones are compacted into LDS in
  // list (angle) order, and every wave takes them KB at a time,
  // software-pipelined two deep: the loads of batch i + 1 are in flight
  // while the sums of batch i are applied (each voxel's sum still takes its
  // hits in angle order). Both batches always issue their KB loads (misses
  // fetch nothing), so the outstanding-load count is the same on every path.
  __shared__ int4 seg[kGatherSeg];
  __shared__ int wcount[4];
  const int lane = tid & 63, wave = tid >> 6;
  const GatherCol cg{qx, qy, zbase, fixed, R2f, thr, R2};
  for (int s0 = 0; s0 < info.x; s0 += kGatherSeg) {  // uniform
    __syncthreads();  // the previous segment is consumed
    const int j = s0 + tid;
    bool ov = false;
    TileEntry e{0, 0, 0, 0.f};
    if (j < info.x) {
      e = lst[j];
      const int zl = e.zpack & 0xffff, zh = e.zpack >> 16;
      ov = !(zh < z0 || zl > z1);
    }
    const unsigned long long bal = __ballot(ov);
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    int base = 0;
#pragma unroll
    for (int w2 = 0; w2 < 4; ++w2) base += w2 < wave ? wcount[w2] : 0;
    const int total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (ov)
      seg[base + __popcll(bal & ((1ull << lane) - 1ull))] =
          make_int4(e.a, e.fxy, __float_as_int(e.fz), 0);
    __syncthreads();
    const int nb = (total + KB - 1) / KB;
    GBatch<KB> A, B;
    gather_fill<ZW, PRE, KB>(A, seg, 0, total, g, cg);
    A.issue(g, C2, C2b);
    for (int bi = 0; bi < nb; bi += 2) {
      gather_fill<ZW, PRE, KB>(B, seg, (bi + 1) * KB, total, g, cg);
      B.issue(g, C2, C2b);
      A.template apply<PRE>(acc, cn, tid);
      if (bi + 1 >= nb) break;
      gather_fill<ZW, PRE, KB>(A, seg, (bi + 2) * KB, total, g, cg);
      A.issue(g, C2, C2b);
      B.template apply<PRE>(acc, cn, tid);
    }
  }
  if (!inb) return;
#pragma unroll
  for (int k = 0; k < ZW; ++k) {
    const int z = z0 + k;
    if (z > zs1) break;
    if (z < zs0) continue;
    const float2 s = acc[k][tid];
    const int c = cn[k][tid];
    if (sgn) {
      float2 t = s;
      if (norm) {
        const float d = (float)(c == 0 ? 1 : c);
        t = make_float2(s.x / d, s.y / d);
      }
      if ((X + Y + z) & 1) { t.x = -t.x; t.y = -t.y; }
      g.vol[z * plane + col] = t;
    } else {
      g.vol[z * plane + col] = s;
      g.cnt[z * plane + col] = c;
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

__global__ void __launch_bounds__(256)
k_band(const AngleGeo* __restrict__ ageo, int A, int V, const double* __restrict__ htab,
       int r2max, int2* __restrict__ band) {
  __shared__ int red[2][4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int lo = 0x7fffffff, hi = -0x7fffffff;
  for (int a = tid; a < A; a += 256) {
    const AngleGeo ag = ageo[a];
    if (ag.skip) continue;
    // hits of angle a lie between the heights at the pupil's rim and centre
    lo = min(lo, (int)round_away(htab[r2max] - ag.fz) + V / 2);
    hi = max(hi, (int)round_away(htab[0] - ag.fz) + V / 2);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  if (lane == 0) { red[0][wave] = lo; red[1][wave] = hi; }
  __syncthreads();
  if (tid == 0) {
    lo = min(min(red[0][0], red[0][1]), min(red[0][2], red[0][3]));
    hi = max(max(red[1][0], red[1][1]), max(red[1][2], red[1][3]));
    lo = max(lo, 0);
    hi = min(hi, V - 1);
    band[0] = lo <= hi ? make_int2(lo, hi) : make_int2(1, 0);
  }
}

hipError_t launch_band(const AngleGeo* ageo, int A, int V, const double* htab, int r2max,
                       int2* band, hipStream_t s) {
  Timed tm("ewald_band", s);
  hipLaunchKernelGGL(k_band, dim3(1), dim3(256), 0, s, ageo, A, V, htab, r2max, band);
  return hipGetLastError();
}

// the windowed gather with height windows of ZW voxels (LDS: 10 B per
// voxel of the 256 x ZW block) and batches of KB list entries
template <int ZW, int KB>
static hipError_t gather_windows(const GatherArgs& g, int ntx, int tiles, const int* order,
                                 const TileEntry* list, const int4* tinfo, int flags,
                                 const int2* band, int zlo, int zhi, hipStream_t s) {
  const int V = g.V;
  // banded output: only the windows that can meet the band are launched
  // (the exact band is only known on the device)
  int w0 = 0, w1 = (V + ZW - 1) / ZW - 1;
  if (band) {
    w0 = std::max(zlo, 0) / ZW;
    w1 = std::min(zhi, V - 1) / ZW;
  }
  const int nwin = w1 - w0 + 1, items = tiles * nwin;
  Timed tm("ewald_gather", s);
  if (flags & 2) {
    hipLaunchKernelGGL((k_gather_win<ZW, true, KB>), dim3(items), dim3(256), 0, s, g, ntx, w0,
                       nwin, items, order, list, tinfo, flags, band);
  } else {
    hipLaunchKernelGGL((k_gather_win<ZW, false, KB>), dim3(items), dim3(256), 0, s, g, ntx, w0,
                       nwin, items, order, list, tinfo, flags, band);
  }
  return hipGetLastError();
}

bool gather_window_supported(int zw) { return zw == 8 || zw == 16 || zw == 24; }

// scratch: [tiles][A] entries | [tiles] info
static void gather_scratch_split(void* scratch, int tiles, int A, TileEntry** list, int4** tinfo) {
  const size_t nA = (size_t)(A > 0 ? A : 1);
  *list = (TileEntry*)scratch;
  *tinfo = (int4*)((char*)scratch + (size_t)tiles * nA * sizeof(TileEntry));
}

static hipError_t tile_prepass(const GatherArgs& g, int ntx, int tiles, TileEntry* list,
                               int4* tinfo, hipStream_t s) {
  Timed tm("ewald_tiles", s);
  hipLaunchKernelGGL(k_tile_angles, dim3(tiles), dim3(64), 0, s, g, ntx, list, tinfo);
  return hipGetLastError();
}

hipError_t launch_gather(const float2* spec, int frame_pos, const AngleGeo* ageo, int A, int C,
                         int nxmax, int V, const double* htab, const float* beta, int r2max,
                         float2* vol, int32_t* cnt, void* scratch, const int* order, int flags,
                         const int2* band, int zlo, int zhi, bool prepass, int zw, int kb,
                         hipStream_t s) {
  if ((size_t)C * C * sizeof(float2) >= ((size_t)1 << 31)) return hipErrorInvalidValue;
  GatherArgs g{spec, frame_pos, ageo, A, C, nxmax, V, htab, beta, r2max, vol, cnt};
  const int ntx = (V + GT - 1) / GT, tiles = ntx * ntx;
  TileEntry* list;
  int4* tinfo;
  gather_scratch_split(scratch, tiles, A, &list, &tinfo);
  if (prepass) {
    const hipError_t e = tile_prepass(g, ntx, tiles, list, tinfo, s);
    if (e != hipSuccess) return e;
  }
#define MTD_GW(ZW, KB) \
  return gather_windows<ZW, KB>(g, ntx, tiles, order, list, tinfo, flags, band, zlo, zhi, s)
  if (kb == 16) {
    switch (zw) {
      case 8: MTD_GW(8, 16);
      case 24: MTD_GW(24, 16);
      default: MTD_GW(16, 16);
    }
  }
  switch (zw) {
    case 8: MTD_GW(8, 8);
    case 24: MTD_GW(24, 8);
    default: MTD_GW(16, 8);
  }
#undef MTD_GW
}

hipError_t launch_gather_prepass(const AngleGeo* ageo, int A, int C, int nxmax, int V,
                                 const double* htab, int r2max, void* scratch, hipStream_t s) {
  GatherArgs g{nullptr, 0, ageo, A, C, nxmax, V, htab, nullptr, r2max, nullptr, nullptr};
  const int ntx = (V + GT - 1) / GT, tiles = ntx * ntx;
  TileEntry* list;
  int4* tinfo;
  gather_scratch_split(scratch, tiles, A, &list, &tinfo);
  return tile_prepass(g, ntx, tiles, list, tinfo, s);
}

size_t gather_scratch_bytes(int V, int A) {
  const size_t tiles = (size_t)((V + GT - 1) / GT) * ((V + GT - 1) / GT);
  const size_t nA = (size_t)(A > 0 ? A : 1);
  return tiles * nA * sizeof(TileEntry) + tiles * sizeof(int4);  // 16 B per entry
}

// The gather's dispatch order of the tiles of a V^3 volume: by distance of
// the tile centre from the lateral origin, nearest first (the tiles there
// hold the longest hit lists), ties by index.
void gather_tile_order_host(int V, int* order) {
  const int ntx = (V + GT - 1) / GT, tiles = ntx * ntx;
  std::vector<std::pair<long, int>> key(tiles);
  for (int t = 0; t < tiles; ++t) {
    const long dx = 2L * ((t % ntx) * GT) + GT - V, dy = 2L * ((t / ntx) * GT) + GT - V;
    key[t] = {dx * dx + dy * dy, t};
  }
  std::sort(key.begin(), key.end());
  for (int t = 0; t < tiles; ++t) order[t] = key[t].second;
}

// ---------------------------------------------------------------------------
// finalisation passes
// ---------------------------------------------------------------------------
// heights outside [band.x, band.y] hold zeros; band == nullptr: none
__device__ __forceinline__ bool outside_band(const int2* band, int z) {
  if (!band) return false;
  const int2 b = *band;
  return z < b.x || z > b.y;
}

template <int VT>
struct NormXOp {  // rows along x; a = z, line = y
  const float2* vol;
  const int32_t* cnt;
  float2* out;
  uint8_t* otf;
  const float2* tw;
  int prenorm;
  const int2* band;
  __device__ bool skip(int z) const { return outside_band(band, z); }
  __device__ void begin(int) {}
  __device__ float2 load(int z, int y, int x) const {
    const size_t i = ((size_t)z * V + y) * V + x;
    if (prenorm) {  // already t * (-1)^(x+y+z)
      const float2 t = vol[i];
      if (otf) otf[i] = (t.x != 0.f) ? 1 : 0;
      return t;
    }
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
  Dim<VT> V{};
};

template <int VT>
struct PlaneYOp {  // columns of each z plane; a = z, col = x, row = y
  static constexpr int kColThreads = 512;  // complex in and out
  static constexpr bool kBalancedCols = true;
  float2* buf;
  const float2* tw;
  const int2* band;
  __device__ bool skip(int z) const { return outside_band(band, z); }
  __device__ void begin(int) {}
  __device__ float2 load(int z, int x, int y) const { return buf[((size_t)z * V + y) * V + x]; }
  __device__ void store(int z, int x, int y, float2 v) const { buf[((size_t)z * V + y) * V + x] = v; }
  Dim<VT> V{};
};

template <int VT>
struct ZOutOp {  // lines along z; a = y, col = x, row = z
  static constexpr bool kBalancedCols = true;
  const float2* buf;
  float* re;
  float* im;
  const float2* tw;
  float df3, cte, gsign;
  const int2* band;
  int lo, hi;
  __device__ void begin(int) {
    if (band) {
      const int2 b = *band;
      lo = b.x;
      hi = b.y;
    } else {
      lo = 0;
      hi = V - 1;
    }
  }
  __device__ float2 load(int y, int x, int z) const {
    if (z < lo || z > hi) return make_float2(0.f, 0.f);
    return buf[((size_t)z * V + y) * V + x];
  }
  __device__ void store(int y, int x, int z, float2 v) const {
    const float sg = ((x + y + z) & 1) ? -gsign : gsign;
    const size_t i = ((size_t)z * V + y) * V + x;
    re[i] = (v.x * sg * df3) * cte;
    if (im) im[i] = (v.y * sg * df3) * cte;  // uniform: NULL skips the imaginary part
  }
  Dim<VT> V{};
};

// the compiled row / column pass of tile variant TV (passes.hpp)
template <int VT, bool INV, int TV, class Op>
static void vol_rows_tv(const Op& op, hipStream_t s) {
  hipLaunchKernelGGL((k_rows<VT, INV, Op, TV>), (rows_grid<VT, TV>(VT)), (rows_block<VT, TV>()),
                     0, s, op);
}
template <int VT, bool INV, int TV, class Op>
static void vol_cols_tv(const Op& op, hipStream_t s) {
  hipLaunchKernelGGL((k_cols<VT, INV, Op, TV>), (cols_grid<VT, Op, TV>(VT)),
                     (cols_block<VT, Op, TV>()), 0, s, op);
}

// one row pass and two column passes over V batches: compiled register FFTs
// (VT > 0) or the runtime-length FFTs (VT == 0). TILES: the tile variants
// tiles.x (rows: 0, 1) and tiles.y (columns: 0, 1, 2) are compiled and
// picked at run time; otherwise variant 0.
template <int VT, bool INV, bool TILES, class RX, class CY, class CZ>
static hipError_t vol_passes(Dim<VT> V, const RX& x, const CY& y, const CZ& z, const FftTab& ft,
                             const char* nx, const char* ny, const char* nz, int2 tiles,
                             hipStream_t s) {
  if constexpr (VT > 0) {
    const int rt = TILES ? tiles.x : 0, ct = TILES ? tiles.y : 0;
    {
      Timed tm(nx, s);
      if constexpr (TILES) {
        if (rt == 1) vol_rows_tv<VT, INV, 1>(x, s);
        else vol_rows_tv<VT, INV, 0>(x, s);
      } else {
        vol_rows_tv<VT, INV, 0>(x, s);
      }
    }
    auto cols = [&](const auto& op) {
      using Op = std::decay_t<decltype(op)>;
      if constexpr (op_cols_balanced<Op>::value && cols_b_fits<VT>()) {
        hipLaunchKernelGGL((k_cols_b<VT, INV, Op>), cols_b_grid<VT>(VT), cols_b_block<VT>(), 0,
                           s, op);
      } else if constexpr (TILES) {
        if (ct == 1) vol_cols_tv<VT, INV, 1>(op, s);
        else if (ct == 2) vol_cols_tv<VT, INV, 2>(op, s);
        else vol_cols_tv<VT, INV, 0>(op, s);
      } else {
        vol_cols_tv<VT, INV, 0>(op, s);
      }
    };
    {
      Timed tm(ny, s);
      cols(y);
    }
    Timed tm(nz, s);
    cols(z);
    return hipGetLastError();
  } else {
    if (!ft.gen) return hipErrorInvalidValue;
    hipError_t e;
    {
      Timed tm(nx, s);
      e = launch_gen<INV, true>(x, *ft.gen, V, V, s);
    }
    if (e != hipSuccess) return e;
    {
      Timed tm(ny, s);
      e = launch_gen<INV, false>(y, *ft.gen, V, V, s);
    }
    if (e != hipSuccess) return e;
    Timed tm(nz, s);
    return launch_gen<INV, false>(z, *ft.gen, V, V, s);
  }
}

template <int VT>
static hipError_t finalize_v(Dim<VT> V, const float2* vol, const int32_t* cnt, int prenorm,
                             const int2* band, float2* work,
                             float df3, float cte, float gsign, float* re,
                             float* im, uint8_t* otf, const FftTab& ft, int2 tiles,
                             hipStream_t s) {
  NormXOp<VT> x{vol, cnt, work, otf, ft.tw, prenorm, band, V};
  PlaneYOp<VT> y{work, ft.tw, band, V};
  ZOutOp<VT> z{work, re, im, ft.tw, df3, cte, gsign, band, 0, V - 1, V};
  return vol_passes<VT, true, true>(V, x, y, z, ft, "vol_x", "vol_y", "vol_z", tiles, s);
}

// ---------------------------------------------------------------------------
// Born forward model (Simul_holo/main.cpp:106-109 and calcHolo,
// Simul_holo/src/fonctions.cpp:109-140): the centred 3-D spectrum of the
// index contrast, TF3D = fftshift3D(FFT3D(fftshift3D(dn))) * Tp^3
// (TF3Dcplx, Simul_holo/src/FFT_fonctions.cpp:301-318), sampled on each
// angle's Ewald cap. Both shifts are (-1)^(x+y+z) sign flips (even V), the
// output ones times (-1)^(3V/2) = gsign. Only the heights of the hit band are
// stored, the only ones the caps read.
// ---------------------------------------------------------------------------
template <int VT>
struct ContrastXOp {  // rows along x of the contrast volume; a = z, line = y
  const float* re;
  const float* im;  // nullptr: real contrast
  float2* out;
  const float2* tw;
  __device__ void begin(int) {}
  __device__ float2 load(int z, int y, int x) const {
    const size_t i = ((size_t)z * V + y) * V + x;
    float2 t = make_float2(re[i], im ? im[i] : 0.f);
    if ((x + y + z) & 1) { t.x = -t.x; t.y = -t.y; }
    return t;
  }
  __device__ void store(int z, int y, int x, float2 v) const {
    out[((size_t)z * V + y) * V + x] = v;
  }
  Dim<VT> V{};
};

template <int VT>
struct SpectrumZOp {  // lines along z, in place; heights outside the band are not stored
  static constexpr int kColThreads = 512;  // complex in and out
  float2* buf;
  const float2* tw;
  float scale;  // Tp^3 * gsign
  const int2* band;
  int lo, hi;
  __device__ void begin(int) {
    const int2 b = *band;
    lo = b.x;
    hi = b.y;
  }
  __device__ float2 load(int y, int x, int z) const { return buf[((size_t)z * V + y) * V + x]; }
  __device__ void store(int y, int x, int z, float2 v) const {
    if (z < lo || z > hi) return;
    const float s = ((x + y + z) & 1) ? -scale : scale;
    buf[((size_t)z * V + y) * V + x] = make_float2(v.x * s, v.y * s);
  }
  Dim<VT> V{};
};

template <int VT>
static hipError_t contrast_spectrum_v(Dim<VT> V, const float* re, const float* im, float2* work,
                                      float scale, const int2* band, const FftTab& ft,
                                      hipStream_t s) {
  ContrastXOp<VT> x{re, im, work, ft.tw, V};
  PlaneYOp<VT> y{work, ft.tw, nullptr, V};
  SpectrumZOp<VT> z{work, ft.tw, scale, band, 0, V - 1, V};
  return vol_passes<VT, false, false>(V, x, y, z, ft, "fwd_vol_x", "fwd_vol_y", "fwd_vol_z",
                                      make_int2(0, 0), s);
}

// calcHolo's sampling, in the frame of the per-angle spectra (+posSpec frame:
// bin (row, col) is pupil pixel fd = (col - NXMAX, row - NXMAX)): the voxel is
// retroPropag_Born's (the simulator's fobj = round(fd - fi) in centred
// coordinates is the same integer map), and the value is
// -i * q(|fd|^2) * TF3D[voxel] with q = kv / (sqrt(R^2 - |fd|^2) / R). Bins
// outside the pupil, of skipped angles, or whose height leaves the volume
// are 0, and so are bins whose voxel would leave the volume laterally (a
// caller's specular outside the crop; the accumulation has no hit there).
__global__ void __launch_bounds__(256)
k_cap_sample(const float2* __restrict__ vol, const AngleGeo* __restrict__ ageo,
             const double* __restrict__ htab, const float* __restrict__ q2, int r2max, int nx,
             int C, int V, float2* __restrict__ spec) {
  // grid (column blocks, rows, angles): no division by the runtime C; the
  // forward factor is read from the C x C pupil-frame table (coalesced)
  const int a = blockIdx.z, row = blockIdx.y;
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= C) return;
  const int i = row * C + col;
  const int fdx = col - nx, fdy = row - nx;
  const int r2 = fdx * fdx + fdy * fdy;
  const AngleGeo ag = ageo[a];
  float2 s = make_float2(0.f, 0.f);
  if (!ag.skip && r2 <= r2max) {
    const int z = (int)round_away(htab[r2] - ag.fz) + V / 2;
    const int X = fdx - ag.fx + V / 2, Y = fdy - ag.fy + V / 2;
    if (z >= 0 && z < V && X >= 0 && X < V && Y >= 0 && Y < V) {
      const float2 t = vol[((size_t)z * V + Y) * V + X];
      const float w = q2[i];
      s = make_float2(w * t.y, -(w * t.x));
    }
  }
  spec[(size_t)a * C * C + i] = s;
}

// Data weights of the adjoint-iterative form: w = 1 / (q(|fd|^2)^2 * hits of
// the sample's voxel) on the cap samples, 0 elsewhere. With them A^H W A is
// Tp^6 times the centred 3-D DFT's projection on the hit voxels: the least
// squares are then as well conditioned as the plain method's count
// normalisation.
__global__ void __launch_bounds__(256)
k_cap_weights(const int32_t* __restrict__ cnt, const AngleGeo* __restrict__ ageo,
              const double* __restrict__ htab, const float* __restrict__ q, int r2max, int nx,
              int C, int V, float* __restrict__ w) {
  const int a = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= C * C) return;
  const int fdx = i % C - nx, fdy = i / C - nx;
  const int r2 = fdx * fdx + fdy * fdy;
  const AngleGeo ag = ageo[a];
  float out = 0.f;
  if (!ag.skip && r2 <= r2max) {
    const int z = (int)round_away(htab[r2] - ag.fz) + V / 2;
    const int X = fdx - ag.fx + V / 2, Y = fdy - ag.fy + V / 2;
    if (z >= 0 && z < V && X >= 0 && X < V && Y >= 0 && Y < V) {
      const int c = cnt[((size_t)z * V + Y) * V + X];  // >= 1: this sample hits it
      const float qq = q[r2];
      out = 1.f / (qq * qq * (float)c);
    }
  }
  w[(size_t)a * C * C + i] = out;
}

#define MTD_VOL_SIZES(X)                                                                   \
  X(32) X(64) X(128) X(256) X(320) X(384) X(416) X(448) X(512) X(576) X(640) X(704) X(768) \
  X(1024)

hipError_t launch_finalize(int V, const float2* vol, const int32_t* cnt,
                           int prenorm, const int2* band, float2* work, float df3, float cte,
                           float gsign, float* out_re, float* out_im,
                           uint8_t* otf, const FftTab& ft, int2 tiles, hipStream_t s) {
  switch (V) {
#define X(n)                                                                                   \
  case n:                                                                                      \
    return finalize_v<n>(Dim<n>{}, vol, cnt, prenorm, band, work, df3, cte, gsign, out_re, \
                         out_im, otf, ft, tiles, s);
    MTD_VOL_SIZES(X)
#undef X
  }
  return finalize_v<0>(Dim<0>(V), vol, cnt, prenorm, band, work, df3, cte, gsign, out_re, out_im,
                       otf, ft, tiles, s);
}

hipError_t launch_contrast_spectrum(int V, const float* re, const float* im, float2* work,
                                   float scale, const int2* band, const FftTab& ft,
                                   hipStream_t s) {
  switch (V) {
#define X(n) \
  case n: return contrast_spectrum_v<n>(Dim<n>{}, re, im, work, scale, band, ft, s);
    MTD_VOL_SIZES(X)
#undef X
  }
  return contrast_spectrum_v<0>(Dim<0>(V), re, im, work, scale, band, ft, s);
}

hipError_t launch_cap_sample(const float2* vol, const AngleGeo* ageo, int A, const double* htab,
                             const float* q2, int r2max, int nx, int C, int V, float2* spec,
                             hipStream_t s) {
  Timed tm("cap_sample", s);
  if (A <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cap_sample, dim3((C + 255) / 256, C, A), dim3(256), 0, s, vol, ageo,
                     htab, q2, r2max, nx, C, V, spec);
  return hipGetLastError();
}

hipError_t launch_cap_weights(const int32_t* cnt, const AngleGeo* ageo, int A,
                              const double* htab, const float* q, int r2max, int nx, int C,
                              int V, float* w, hipStream_t s) {
  Timed tm("cap_weights", s);
  hipLaunchKernelGGL(k_cap_weights, dim3((C * C + 255) / 256, A), dim3(256), 0, s, cnt, ageo,
                     htab, q, r2max, nx, C, V, w);
  return hipGetLastError();
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
