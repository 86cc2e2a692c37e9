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
// decreasing in r^2 and round is monotonic), and each column's exact hit
// height per listed angle as one byte. The main kernel gives each block one
// (tile, ZW-height window); a wave tests 64 list entries against the window
// at once, so each hit is visited in about 1 + span/ZW windows. Near the
// lateral origin nearly every angle hits the same few voxels, and those
// sums are serial by definition (angle order): the loads of kBatch hits are
// issued together so that only the additions stay serial.
// ---------------------------------------------------------------------------
struct TileEntry {
  int a;
  int zpack;  // zlo | zhi << 16, absolute heights
  int fxy;    // fxm0 & 0xffff | fym0 << 16 (signed 16-bit each)
  int pad;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

constexpr int kBatch = 8;  // hits whose loads are issued together
constexpr int kPrepBatch = 8;  // prepass: entries whose heights are computed together
constexpr int kTileChunk = 64;  // prepass: angles per block

// h(r2) = sqrt(R^2 - r2) for 0 <= r2 <= R^2: IEEE sqrt of the same exact
// double, so the host height table's value bit for bit
__device__ __forceinline__ double height(double R2, int r2) { return sqrt(R2 - (double)r2); }

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

__device__ __forceinline__ int block_sum256(int v, int* wsum) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  const int t = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  __syncthreads();  // wsum is reused
  return t;
}

// Prepass: the angles whose pupil disc touches each 16x16 column tile, in
// angle order, with a conservative height range from the tile box's nearest
// and farthest pupil radii (h decreases with r^2, round is monotonic), and
// every column's exact hit height per listed angle as a byte offset from the
// entry's lower bound (255 = no hit): FP64 heights and round-half-away, as
// the reference. One block per (tile, chunk of kTileChunk angles), so that
// the tiles near the lateral origin, which nearly every angle touches, are
// spread over many blocks: a block first counts the tile's entries of the
// earlier chunks (its offset in the tile's list; the chunk-0 block counts all
// of them and writes the tile's total and height range), then lists its own
// chunk and computes the heights of those entries, one column per thread.
__global__ void __launch_bounds__(256)
k_tile_angles(GatherArgs g, int ntx, int nchunk, TileEntry* __restrict__ list,
              int4* __restrict__ tinfo, uint8_t* __restrict__ zoff) {
  __shared__ int wsum[4];
  __shared__ int zr[2][4];
  __shared__ int nown;
  __shared__ int4 sent[kTileChunk];  // a, zl, fx, fy of this chunk's entries
  __shared__ double sfz[kTileChunk];
  const int tile = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
  const int tid = threadIdx.x, lane = tid & 63;
  const int qx0 = (tile % ntx) * GT - g.V / 2, qy0 = (tile / ntx) * GT - g.V / 2;
  const int qx = qx0 + tid % GT, qy = qy0 + tid / GT;  // this thread's column
  const double R2 = g.htab[0] * g.htab[0];  // h(0) = R, an integer: exact
  const int a_own = chunk * kTileChunk;
  // entries of the earlier chunks; chunk 0: all entries and the height range
  const int lim = chunk == 0 ? g.A : a_own;
  int cnt = 0, zmin = 0x7fffffff, zmax = -1;
  for (int a = tid; a < lim; a += 256) {
    const AngleGeo ag = g.ageo[a];
    int dmin, dmax;
    if (tile_touch(g, ag, qx0, qy0, dmin, dmax)) {
      ++cnt;
      if (chunk == 0) {
        zmin = min(zmin, clampi((int)round_away(height(R2, dmax) - ag.fz) + g.V / 2, 0, g.V - 1));
        zmax = max(zmax, clampi((int)round_away(height(R2, dmin) - ag.fz) + g.V / 2, 0, g.V - 1));
      }
    }
  }
  const int total = block_sum256(cnt, wsum);
  if (chunk == 0) {
    for (int o = 32; o > 0; o >>= 1) {
      zmin = min(zmin, __shfl_xor(zmin, o, 64));
      zmax = max(zmax, __shfl_xor(zmax, o, 64));
    }
    if (lane == 0) { zr[0][tid >> 6] = zmin; zr[1][tid >> 6] = zmax; }
    __syncthreads();
    if (tid == 0) {
      const int lo = min(min(zr[0][0], zr[0][1]), min(zr[0][2], zr[0][3]));
      const int hi = max(max(zr[1][0], zr[1][1]), max(zr[1][2], zr[1][3]));
      tinfo[tile] = make_int4(total, lo, hi, 0);
    }
  }
  const int off = chunk == 0 ? 0 : total;
  // this chunk's entries, in angle order (one wave)
  if (tid < 64) {
    const int a = a_own + lane;
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
      }
    }
    const unsigned long long bal = __ballot(act);
    if (act) {
      const int p = __popcll(bal & ((1ull << lane) - 1ull));
      list[(size_t)tile * g.A + off + p] =
          TileEntry{a, zl | (zh << 16), (ag.fx & 0xffff) | (ag.fy << 16), 0};
      sent[p] = make_int4(a, zl, ag.fx, ag.fy);
      sfz[p] = ag.fz;
    }
    if (lane == 0) nown = __popcll(bal);
  }
  __syncthreads();
  const int n = nown;
  // exact hit heights of this thread's column for the chunk's entries
  uint8_t* zo = zoff + ((size_t)tile * g.A + off) * 256 + tid;
  for (int i0 = 0; i0 < n; i0 += kPrepBatch) {
    // branch-free batch (slots past n repeat the last entry and are not stored)
    double hb[kPrepBatch];
    bool inp[kPrepBatch];
#pragma unroll
    for (int t = 0; t < kPrepBatch; ++t) {
      const int4 e = sent[min(i0 + t, n - 1)];
      const int fdx = qx + e.z, fdy = qy + e.w;
      const int r2 = fdx * fdx + fdy * fdy;
      inp[t] = fdx >= -g.nx && fdx < g.nx && fdy >= -g.nx && fdy < g.nx && r2 <= g.r2max;
      hb[t] = height(R2, inp[t] ? r2 : 0);
    }
#pragma unroll
    for (int t = 0; t < kPrepBatch; ++t) {
      const int it = min(i0 + t, n - 1);
      const int z = (int)round_away(hb[t] - sfz[it]) + g.V / 2;
      const uint8_t o = inp[t] ? (uint8_t)(z - sent[it].y) : (uint8_t)255;
      if (i0 + t < n) zo[(size_t)(i0 + t) * 256] = o;  // uniform
    }
  }
}

// flags: bit0 = normalise (divide by counts, 0 -> 1) and apply the centring
// sign (-1)^(x+y+z) of the 3-D inverse FFT; counts are then not written.
// bit2 = the plain sums with that centring sign, no counts (the adjoint of
// the Born forward model).
// PRE: the spectra are already multiplied by the Ewald factor beta of their
// pupil pixel (the same FP32 product), so beta is not looked up here.
template <int ZW, bool PRE>
__global__ void __launch_bounds__(256)
k_gather_win(GatherArgs g, int ntx, int w0, int nwin, int items,
             const TileEntry* __restrict__ list, const int4* __restrict__ tinfo,
             const uint8_t* __restrict__ zoff, int flags, const int2* __restrict__ band) {
  // The sums must be the sequential FP32 sums of the reference: no fused
  // multiply-add. The pragma is lexical, so the products and sums are written
  // here as plain operators (HIP's __fmul_rn/__fadd_rn are header functions
  // outside its scope and would still be contracted).
#pragma clang fp contract(off)
  // Blocks are dealt round-robin to the 8 XCDs. Tiles go round-robin too
  // (the heavy tiles near the lateral origin are spread), and the windows
  // of one tile stay on one XCD, in order (they read the same spectrum
  // patches). A bijection when the tile count is a multiple of 8. Windows
  // w0 .. w0 + nwin - 1 are launched.
  int tile, w;
  {
    const int b = blockIdx.x, tiles = items / nwin;
    if (tiles % 8 == 0) {
      const int n = b / 8;
      tile = (n / nwin) * 8 + b % 8;
      w = w0 + n % nwin;
    } else {
      tile = b / nwin;
      w = w0 + b % nwin;
    }
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
  if (band) {  // banded output: windows outside the hit band are not written
    const int2 b = *band;
    if (z1 < b.x || z0 > b.y) return;
  }
  const int4 info = tinfo[tile];
  const bool norm = flags & 1, sgn = flags & 5;
  if (info.x == 0 || z1 < info.y || z0 > info.z) {
    if (inb) {
      for (int z = z0; z <= z1; ++z) {
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
  const uint8_t* zo = zoff + (size_t)tile * g.A * 256 + tid;
  // spectrum element read by this column: the unshifted frame holds it at
  // (qy mod C, qx mod C) for every angle; the +posSpec frame at fd + NXMAX
  const int frow = ((qy % g.C) + g.C) % g.C, fcol = ((qx % g.C) + g.C) % g.C;
  const size_t fixed = (size_t)frow * g.C + fcol;
  const size_t C2 = (size_t)g.C * g.C;
  // Each wave tests 64 entries against the window at once, then takes the
  // overlapping ones lowest index first (angle order) kBatch at a time: the
  // loads of a batch are issued together, the sums applied in order.
  const int lane = tid & 63;
  for (int j0 = 0; j0 < info.x; j0 += 64) {
    const int j = j0 + lane;
    TileEntry e{0, 0, 0, 0};
    bool ov = false;
    if (j < info.x) {
      e = lst[j];
      const int zl = e.zpack & 0xffff, zh = e.zpack >> 16;
      ov = !(zh < z0 || zl > z1);
    }
    unsigned long long m = __ballot(ov);  // wave-uniform
    while (m) {
      // Straight-line batch (no branch between the loads, so all of them are
      // in flight together): unused slots repeat the first entry and are
      // masked out afterwards.
      int bitv[kBatch];
      bool valid[kBatch];
#pragma unroll
      for (int t = 0; t < kBatch; ++t) {
        valid[t] = m != 0;  // uniform; true for t = 0
        bitv[t] = valid[t] ? __ffsll((unsigned long long)m) - 1 : bitv[0];
        m &= m - 1;  // m == 0 stays 0
      }
      int offv[kBatch], zlv[kBatch];
      float2 spv[kBatch];
      float bv[kBatch];
#pragma unroll
      for (int t = 0; t < kBatch; ++t) {
        const int i = j0 + bitv[t];
        const int ea = __builtin_amdgcn_readlane(e.a, bitv[t]);
        zlv[t] = __builtin_amdgcn_readlane(e.zpack, bitv[t]) & 0xffff;
        const int fxy = __builtin_amdgcn_readlane(e.fxy, bitv[t]);
        const int fdx = qx + (int)(short)(fxy & 0xffff), fdy = qy + (fxy >> 16);
        offv[t] = zo[(size_t)i * 256];
        // addresses are clamped so that non-hits read in bounds; a hit has
        // |fd| inside the crop and r2 <= r2max, where clamping is the identity
        const int r2 = min(fdx * fdx + fdy * fdy, g.r2max);
        const size_t sp = (size_t)(clampi(fdy, -g.nx, g.nx - 1) + g.nx) * g.C +
                          (clampi(fdx, -g.nx, g.nx - 1) + g.nx);
        spv[t] = g.spec[(size_t)ea * C2 + (g.frame_pos ? sp : fixed)];
        if constexpr (!PRE) bv[t] = g.beta[r2];
      }
#pragma unroll
      for (int t = 0; t < kBatch; ++t) {
        if (!valid[t]) offv[t] = 255;
      }
      // Unconditional updates (a conditional one lets the compiler sink its
      // loads behind the branch): a miss adds -0.0 to slot 0, which leaves
      // every value bit-identical (x + -0.0 == x, also for x = +-0.0), and
      // does not count.
#pragma unroll
      for (int t = 0; t < kBatch; ++t) {
        const int k = offv[t] == 255 ? -1 : zlv[t] + offv[t] - z0;
        const bool hit = k >= 0 && k < ZW;
        const int kc = hit ? k : 0;
        // value (-b Si, b Sr)
        const float pim = PRE ? spv[t].y : bv[t] * spv[t].y;
        const float pre = PRE ? spv[t].x : bv[t] * spv[t].x;
        float2 s = acc[kc][tid];
        s.x = s.x + (hit ? -pim : -0.0f);
        s.y = s.y + (hit ? pre : -0.0f);
        acc[kc][tid] = s;
        cn[kc][tid] = (unsigned short)(cn[kc][tid] + (hit ? 1 : 0));
      }
    }
  }
  if (!inb) return;
#pragma unroll
  for (int k = 0; k < ZW; ++k) {
    const int z = z0 + k;
    if (z > z1) break;
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

static int prep_chunks(int A) { return std::max(1, (A + kTileChunk - 1) / kTileChunk); }

// the windowed gather with height windows of ZW voxels (LDS: 10 B per
// voxel of the 256 x ZW block)
template <int ZW>
static hipError_t gather_windows(const GatherArgs& g, int ntx, int tiles, const TileEntry* list,
                                 const int4* tinfo, const uint8_t* zoff, int flags,
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
    hipLaunchKernelGGL((k_gather_win<ZW, true>), dim3(items), dim3(256), 0, s, g, ntx, w0, nwin,
                       items, list, tinfo, zoff, flags, band);
  } else {
    hipLaunchKernelGGL((k_gather_win<ZW, false>), dim3(items), dim3(256), 0, s, g, ntx, w0,
                       nwin, items, list, tinfo, zoff, flags, band);
  }
  return hipGetLastError();
}

bool gather_window_supported(int zw) { return zw == 8 || zw == 16 || zw == 24; }

hipError_t launch_gather(const float2* spec, int frame_pos, const int* posx,
                         const int* posy, const AngleGeo* ageo, int A, int C,
                         int nxmax, int V, const double* htab,
                         const float* beta, int r2max, float2* vol,
                         int32_t* cnt, void* scratch, int flags, const int2* band,
                         int zlo, int zhi, bool prepass, int zw, hipStream_t s) {
  (void)posx;
  (void)posy;
  GatherArgs g{spec, frame_pos, ageo, A, C, nxmax, V, htab, beta, r2max, vol, cnt};
  const int ntx = (V + GT - 1) / GT, tiles = ntx * ntx;
  // scratch: [tiles][A] entries | [tiles][A][256] height offsets | [tiles] info
  const size_t nA = (size_t)(A > 0 ? A : 1);
  TileEntry* list = (TileEntry*)scratch;
  uint8_t* zoff = (uint8_t*)scratch + (size_t)tiles * nA * sizeof(TileEntry);
  int4* tinfo = (int4*)(zoff + (size_t)tiles * nA * 256);
  if (prepass) {
    Timed tm("ewald_tiles", s);
    hipLaunchKernelGGL(k_tile_angles, dim3(tiles * prep_chunks(A)), dim3(256), 0, s, g, ntx,
                       prep_chunks(A), list, tinfo, zoff);
  }
  switch (zw) {
    case 8:
      return gather_windows<8>(g, ntx, tiles, list, tinfo, zoff, flags, band, zlo, zhi, s);
    case 24:
      return gather_windows<24>(g, ntx, tiles, list, tinfo, zoff, flags, band, zlo, zhi, s);
    default:
      return gather_windows<16>(g, ntx, tiles, list, tinfo, zoff, flags, band, zlo, zhi, s);
  }
}

hipError_t launch_gather_prepass(const AngleGeo* ageo, int A, int C, int nxmax, int V,
                                 const double* htab, int r2max, void* scratch, hipStream_t s) {
  GatherArgs g{nullptr, 0, ageo, A, C, nxmax, V, htab, nullptr, r2max, nullptr, nullptr};
  const int ntx = (V + GT - 1) / GT, tiles = ntx * ntx;
  const size_t nA = (size_t)(A > 0 ? A : 1);
  TileEntry* list = (TileEntry*)scratch;
  uint8_t* zoff = (uint8_t*)scratch + (size_t)tiles * nA * sizeof(TileEntry);
  int4* tinfo = (int4*)(zoff + (size_t)tiles * nA * 256);
  Timed tm("ewald_tiles", s);
  hipLaunchKernelGGL(k_tile_angles, dim3(tiles * prep_chunks(A)), dim3(256), 0, s, g, ntx,
                     prep_chunks(A), list, tinfo, zoff);
  return hipGetLastError();
}

size_t gather_scratch_bytes(int V, int A) {
  const size_t tiles = (size_t)((V + GT - 1) / GT) * ((V + GT - 1) / GT);
  const size_t nA = (size_t)(A > 0 ? A : 1);
  return tiles * nA * (sizeof(TileEntry) + 256) + tiles * sizeof(int4);  // 16 + 256 B per entry
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

// one row pass and two column passes over V batches: compiled register FFTs
// (VT > 0) or the runtime-length FFTs (VT == 0)
template <int VT, bool INV, class RX, class CY, class CZ>
static hipError_t vol_passes(Dim<VT> V, const RX& x, const CY& y, const CZ& z, const FftTab& ft,
                             const char* nx, const char* ny, const char* nz, hipStream_t s) {
  if constexpr (VT > 0) {
    {
      Timed tm(nx, s);
      hipLaunchKernelGGL((k_rows<VT, INV, RX>), rows_grid<VT>(VT), rows_block<VT>(), 0, s, x);
    }
    {
      Timed tm(ny, s);
      hipLaunchKernelGGL((k_cols<VT, INV, CY>), (cols_grid<VT, CY>(VT)), (cols_block<VT, CY>()),
                         0, s, y);
    }
    Timed tm(nz, s);
    hipLaunchKernelGGL((k_cols<VT, INV, CZ>), (cols_grid<VT, CZ>(VT)), (cols_block<VT, CZ>()), 0,
                       s, z);
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
                             float* im, uint8_t* otf, const FftTab& ft,
                             hipStream_t s) {
  NormXOp<VT> x{vol, cnt, work, otf, ft.tw, prenorm, band, V};
  PlaneYOp<VT> y{work, ft.tw, band, V};
  ZOutOp<VT> z{work, re, im, ft.tw, df3, cte, gsign, band, 0, V - 1, V};
  return vol_passes<VT, true>(V, x, y, z, ft, "vol_x", "vol_y", "vol_z", s);
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
  return vol_passes<VT, false>(V, x, y, z, ft, "fwd_vol_x", "fwd_vol_y", "fwd_vol_z", s);
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
                           uint8_t* otf, const FftTab& ft, hipStream_t s) {
  switch (V) {
#define X(n)                                                                                   \
  case n:                                                                                      \
    return finalize_v<n>(Dim<n>{}, vol, cnt, prenorm, band, work, df3, cte, gsign, out_re, \
                         out_im, otf, ft, s);
    MTD_VOL_SIZES(X)
#undef X
  }
  return finalize_v<0>(Dim<0>(V), vol, cnt, prenorm, band, work, df3, cte, gsign, out_re, out_im,
                       otf, ft, s);
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
