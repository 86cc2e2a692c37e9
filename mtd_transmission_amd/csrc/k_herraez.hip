// Phase unwrapping, DEROUL 1 with VOLKOV 0 (Herraez: phaseUnwrapping_Mat,
// pretraitement_hors_axe_CPU/deroulement_herraez.cpp:32-75).
//
// The reference computes a reliability per pixel (second differences of the
// wrapped phase, :382-404; border pixels keep 9999999 + rand(), :348), one
// per edge (the sum of its two pixels', :409-452), sorts the edges with its
// own merge sort (:160-195) and then merges pixel groups edge by edge in that
// order (:455-552): Kruskal's algorithm, so the edges that merge groups form
// the minimum spanning tree of the grid for the edge weights "position in
// the sorted order" (all distinct). The unwrapped phase is the wrapped one
// plus 2 pi times an integer increment that makes every tree edge step the
// wrapped difference (find_wrap, :370-380).
//
// GPU restructuring (same tree, same increments up to one global 2 pi k):
//  - k_herr_rel: wrapped phase (atan2 of the demodulated field) and the
//    reliabilities, one thread per pixel; border values from the host (the
//    reference's unseeded rand() stream, restated on the host);
//  - k_herr_sort: one block per angle replays the reference's sort level by
//    level: the recursion halves a range of s into two runs of s / 2 (an odd
//    range keeps its last element in place) and merges them with a
//    two-pointer merge, left first on ties. Such a merge of runs that need
//    not be sorted orders elements exactly as a stable merge of the runs'
//    running maxima, so each level is a segmented running max and a merge by
//    binary search (every element finds its place independently);
//  - Boruvka on the sort positions: every component takes its lowest-rank
//    edge (atomicMin of the rank), hooks along it unless both picked the
//    same edge and it has the lower root, and the hooked roots' links are
//    flattened by concurrent pointer jumping. Each pixel keeps its increment
//    relative to its component's root; a hook stores the root's increment
//    relative to the new root, fixed by the edge's wrap count. The spanning
//    tree is the unique minimum one, the reference's; only the root whose
//    increment is 0 may differ, a constant 2 pi k that the polynomial fit
//    which always follows unwrapping (its constant monomial) removes.
#include <hip/hip_runtime.h>

#include <climits>

#include "internal.hpp"

namespace mtd {

constexpr float kHerrPi = 3.141592654f;     // deroulement_herraez.cpp:9
constexpr float kHerrTwoPi = 6.283185307f;  // :10

__device__ __forceinline__ float herr_wrap(float x) {
  return x > kHerrPi ? x - kHerrTwoPi : (x < -kHerrPi ? x + kHerrTwoPi : x);
}

// edge e of the C x C grid: horizontal edges row-major first, then vertical
__device__ __forceinline__ void herr_edge(int e, int C, int& p1, int& p2) {
  const int nh = C * (C - 1);
  if (e < nh) {
    const int y = e / (C - 1), x = e - y * (C - 1);
    p1 = y * C + x;
    p2 = p1 + 1;
  } else {
    p1 = e - nh;
    p2 = p1 + C;
  }
}

// number of 2 pi between the two pixels of an edge (find_wrap)
__device__ __forceinline__ int herr_wrapcount(float v1, float v2) {
  const float d = v1 - v2;
  return d > kHerrPi ? -1 : (d < -kHerrPi ? 1 : 0);
}

// index of border pixel (y, x) in the packed border table: row 0, row C-1,
// then columns 0 and C-1 of rows 1 .. C-2
__device__ __forceinline__ int herr_border_slot(int y, int x, int C) {
  if (y == 0) return x;
  if (y == C - 1) return C + x;
  return 2 * C + 2 * (y - 1) + (x == 0 ? 0 : 1);
}

// wrapped phase of the U frame (y, x) and reliability per pixel
__global__ void __launch_bounds__(256)
k_herr_rel(const float2* __restrict__ iu, int C, const float* __restrict__ border,
           float* __restrict__ ph, float* __restrict__ rel) {
  // the reliability sums are the reference's float products and sums: no
  // fused multiply-add
#pragma clang fp contract(off)
  const int a = blockIdx.y;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int n = C * C;
  if (p >= n) return;
  const int y = p / C, x = p - y * C;
  const float2* iu_a = iu + (size_t)a * n;
  auto phase = [&](int yy, int xx) {
    const float2 u = iu_a[((yy + C / 2) % C) * C + (xx + C / 2) % C];
    return atan2f(u.y, u.x);
  };
  const float c = phase(y, x);
  ph[(size_t)a * n + p] = c;
  float r;
  if (y == 0 || y == C - 1 || x == 0 || x == C - 1) {
    r = border[(size_t)a * (4 * C - 4) + herr_border_slot(y, x, C)];
  } else {
    // calculate_reliability (:392-398): second differences of wrapped
    // first differences, summed in float in the reference's order
    const float H = herr_wrap(phase(y, x - 1) - c) - herr_wrap(c - phase(y, x + 1));
    const float V = herr_wrap(phase(y - 1, x) - c) - herr_wrap(c - phase(y + 1, x));
    const float D1 = herr_wrap(phase(y - 1, x - 1) - c) - herr_wrap(c - phase(y + 1, x + 1));
    const float D2 = herr_wrap(phase(y - 1, x + 1) - c) - herr_wrap(c - phase(y + 1, x - 1));
    r = ((H * H + V * V) + D1 * D1) + D2 * D2;
  }
  rel[(size_t)a * n + p] = r;
}

// ---------------------------------------------------------------------------
// the reference's edge sort, one block per angle
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 1024;
constexpr int kMaxDepth = 40;

struct SortLevels {
  int s[kMaxDepth];  // node size at each depth (s[0] = edge count)
  int deepest;       // last depth whose nodes (size >= 2) merge two runs
};

// The node containing position p at depth d: its start, or -1 if p was left
// out (the last element of an odd range) at a shallower depth.
__device__ __forceinline__ int herr_node_start(int p, int d, const SortLevels& L) {
  int start = 0;
  for (int i = 1; i <= d; ++i) {
    const int c = L.s[i], o = p - start;
    if (o < c) continue;
    if (o < 2 * c) {
      start += c;
      continue;
    }
    return -1;
  }
  return start;
}

// run of position p at depth d: (run start, index in run); run start -1 for
// a position that does not move at this depth
__device__ __forceinline__ int2 herr_run(int p, int d, const SortLevels& L) {
  const int S = herr_node_start(p, d, L);
  if (S < 0) return make_int2(-1, 0);
  const int h = L.s[d + 1], o = p - S;
  if (o < h) return make_int2(S, o);
  if (o < 2 * h) return make_int2(S + h, o - h);
  return make_int2(-1, 0);
}

__global__ void __launch_bounds__(kSortThreads)
k_herr_sort(const float* __restrict__ rel, int C, SortLevels L, float* __restrict__ kA_all,
            int* __restrict__ iA_all, float* __restrict__ kB_all, int* __restrict__ iB_all,
            float* __restrict__ rm_all) {
  const int a = blockIdx.x;
  const int m = 2 * C * (C - 1), n = C * C;
  const size_t off = (size_t)a * m;
  float* kA = kA_all + off;
  int* iA = iA_all + off;
  float* kB = kB_all + off;
  int* iB = iB_all + off;
  float* rm = rm_all + off;
  const float* rl = rel + (size_t)a * n;
  const int tid = threadIdx.x;
  // edge reliabilities (:418, :443): float sum of the two pixels'
  for (int e = tid; e < m; e += kSortThreads) {
    int p1, p2;
    herr_edge(e, C, p1, p2);
    kA[e] = rl[p1] + rl[p2];
    iA[e] = e;
  }
  __shared__ int head_run[kSortThreads], tail_run[kSortThreads];
  __shared__ float tail_max[kSortThreads], carry[kSortThreads];
  const int q = (m + kSortThreads - 1) / kSortThreads;
  const int c0 = min(tid * q, m), c1 = min(c0 + q, m);
  for (int d = L.deepest; d >= 0; --d) {
    __syncthreads();
    // 1. running max of the keys within each run: chunk-local pass
    int cur = INT_MIN, first = INT_MIN;
    float mx = -INFINITY;
    for (int p = c0; p < c1; ++p) {
      const int2 r = herr_run(p, d, L);
      const int id = r.x < 0 ? -2 - p : r.x;  // a fixed position is its own run
      if (p == c0) first = id;
      if (id != cur) {
        cur = id;
        mx = -INFINITY;
      }
      mx = fmaxf(mx, kA[p]);
      rm[p] = mx;
    }
    head_run[tid] = first;
    tail_run[tid] = cur;
    tail_max[tid] = mx;
    __syncthreads();
    // 2. the running max carried into each chunk from the chunks before it
    if (tid == 0) {
      float full = -INFINITY;  // max of the run at the end of chunk t-1
      carry[0] = -INFINITY;
      for (int t = 1; t < kSortThreads; ++t) {
        const int pt = t - 1;
        const bool empty = head_run[pt] == INT_MIN;
        if (!empty) {
          const float prev = (head_run[pt] == tail_run[pt]) ? carry[pt] : -INFINITY;
          full = fmaxf(prev, tail_max[pt]);
        }
        carry[t] = (!empty && head_run[t] == tail_run[pt] && head_run[t] != INT_MIN)
                       ? full
                       : (empty ? carry[pt] : -INFINITY);
      }
    }
    __syncthreads();
    // 3. apply the carry to the chunk's leading run
    {
      const float cin = carry[tid];
      if (cin > -INFINITY) {
        for (int p = c0; p < c1; ++p) {
          const int2 r = herr_run(p, d, L);
          const int id = r.x < 0 ? -2 - p : r.x;
          if (id != head_run[tid]) break;
          rm[p] = fmaxf(rm[p], cin);
        }
      }
    }
    __syncthreads();
    // 4. merge: each element's place from a binary search of the other run's
    //    running maxima (left first on ties)
    const int h = L.s[d + 1];
    for (int p = tid; p < m; p += kSortThreads) {
      const int S = herr_node_start(p, d, L);
      int dst = p;
      if (S >= 0 && p - S < 2 * h) {
        const int o = p - S;
        const float key = rm[p];
        if (o < h) {  // left run: count right-run maxima < key
          const float* R = rm + S + h;
          int lo = 0, hi = h;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (R[mid] < key) lo = mid + 1; else hi = mid;
          }
          dst = S + o + lo;
        } else {  // right run: count left-run maxima <= key
          const float* Lr = rm + S;
          int lo = 0, hi = h;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (Lr[mid] <= key) lo = mid + 1; else hi = mid;
          }
          dst = S + (o - h) + lo;
        }
      }
      kB[dst] = kA[p];
      iB[dst] = iA[p];
    }
    // the merged order becomes the current one
    float* tk = kA;
    kA = kB;
    kB = tk;
    int* ti = iA;
    iA = iB;
    iB = ti;
  }
  __syncthreads();
  // the sorted edge order ends in the A buffers of the caller when the
  // number of levels is even; leave it in iA_all either way
  if (iA != iA_all + off) {
    for (int e = tid; e < m; e += kSortThreads) iA_all[off + e] = iA[e];
  }
}

// ---------------------------------------------------------------------------
// Boruvka on the sort positions, batched over angles (grid.y = angle)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_bor_init(const int* __restrict__ sorted, int C, int* __restrict__ rank_of,
           int* __restrict__ comp, int* __restrict__ inc) {
  const int a = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  const int m = 2 * C * (C - 1), n = C * C;
  if (i < m) rank_of[(size_t)a * m + sorted[(size_t)a * m + i]] = i;
  if (i < n) {
    comp[(size_t)a * n + i] = i;
    inc[(size_t)a * n + i] = 0;
  }
}

__global__ void __launch_bounds__(256)
k_bor_reset(int C, unsigned* __restrict__ best) {
  const int a = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int n = C * C;
  if (p < n) best[(size_t)a * n + p] = 0xFFFFFFFFu;
}

__global__ void __launch_bounds__(256)
k_bor_select(int C, const int* __restrict__ rank_of, const int* __restrict__ comp,
             unsigned* __restrict__ best) {
  const int a = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
  const int m = 2 * C * (C - 1), n = C * C;
  if (e >= m) return;
  int p1, p2;
  herr_edge(e, C, p1, p2);
  const int* cp = comp + (size_t)a * n;
  const int r1 = cp[p1], r2 = cp[p2];
  if (r1 == r2) return;
  const unsigned rk = (unsigned)rank_of[(size_t)a * m + e];
  atomicMin(best + (size_t)a * n + r1, rk);
  atomicMin(best + (size_t)a * n + r2, rk);
}

// link of a root: (root it hooks to, its increment relative to that root)
__device__ __forceinline__ unsigned long long bor_pack(int up, int off) {
  return (unsigned long long)(unsigned)up | ((unsigned long long)(unsigned)off << 32);
}
__device__ __forceinline__ int bor_up(unsigned long long v) { return (int)(unsigned)v; }
__device__ __forceinline__ int bor_off(unsigned long long v) { return (int)(unsigned)(v >> 32); }

__global__ void __launch_bounds__(256)
k_bor_hook(int C, const int* __restrict__ sorted, const int* __restrict__ comp,
           const int* __restrict__ inc, const unsigned* __restrict__ best,
           const float* __restrict__ ph, unsigned long long* __restrict__ link) {
  const int a = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int m = 2 * C * (C - 1), n = C * C;
  if (p >= n) return;
  const int* cp = comp + (size_t)a * n;
  if (cp[p] != p) return;  // not a root
  const unsigned* bs = best + (size_t)a * n;
  const unsigned b = bs[p];
  unsigned long long v = bor_pack(p, 0);
  if (b != 0xFFFFFFFFu) {
    const int e = sorted[(size_t)a * m + b];
    int p1, p2;
    herr_edge(e, C, p1, p2);
    const int r1 = cp[p1], r2 = cp[p2];
    const int other = r1 == p ? r2 : r1;
    // both ends picked this edge: the lower root stays
    if (!(bs[other] == b && p < other)) {
      const int* ic = inc + (size_t)a * n;
      const float* pa = ph + (size_t)a * n;
      const int w = herr_wrapcount(pa[p1], pa[p2]);
      // inc(p2) = inc(p1) - w across the edge (gatherPIXELs, :482, :495)
      const int d = r1 == p ? ic[p2] - ic[p1] + w : ic[p1] - w - ic[p2];
      v = bor_pack(other, d);
    }
  }
  link[(size_t)a * n + p] = v;
}

// follows each root's links to the root of its new component, accumulating
// the increments; only thread p writes link[p] (a 64-bit store), so every
// link read is some valid ancestor of its root with a consistent increment
__global__ void __launch_bounds__(256)
k_bor_jump(int C, const int* __restrict__ comp, unsigned long long* link) {
  const int a = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int n = C * C;
  if (p >= n) return;
  if (comp[(size_t)a * n + p] != p) return;
  unsigned long long* lk = link + (size_t)a * n;
  unsigned long long v = __hip_atomic_load(lk + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const int q = bor_up(v);
    if (q == p) break;
    const unsigned long long w =
        __hip_atomic_load(lk + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int r = bor_up(w);
    if (r == q) break;  // q is the new root
    v = bor_pack(r, bor_off(v) + bor_off(w));
    __hip_atomic_store(lk + p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void __launch_bounds__(256)
k_bor_apply(int C, int* __restrict__ comp, int* __restrict__ inc,
            const unsigned long long* __restrict__ link) {
  const int a = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int n = C * C;
  if (p >= n) return;
  int* cp = comp + (size_t)a * n;
  const unsigned long long v = link[(size_t)a * n + cp[p]];
  cp[p] = bor_up(v);
  inc[(size_t)a * n + p] += bor_off(v);
}

// unwrapImage (:555-566): value + TWOPI * increment in float
__global__ void __launch_bounds__(256)
k_herr_out(int C, const float* __restrict__ ph, const int* __restrict__ inc,
           float* __restrict__ phu) {
  const int a = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
  const int n = C * C;
  if (p >= n) return;
  const size_t i = (size_t)a * n + p;
  phu[i] = ph[i] + kHerrTwoPi * (float)inc[i];
}

size_t herraez_scratch_bytes(int C, int A) {
  const size_t n = (size_t)C * C, m = 2 * (size_t)C * (C - 1), na = (size_t)(A > 0 ? A : 1);
  // ph, rel, comp, inc, best (4 B each per pixel), link (8 B per pixel);
  // kA, iA, kB, iB, rm, rank (4 B each per edge); the border table
  return na * (n * 28 + m * 24 + (4 * (size_t)C) * sizeof(float)) + 256 * 8;
}

hipError_t launch_herraez(int C, const float2* iu, int A, const float* border, void* scratch,
                          float* phu, hipStream_t s) {
  if (A <= 0) return hipSuccess;
  if (C < 3) return hipErrorInvalidValue;
  const size_t n = (size_t)C * C, m = 2 * (size_t)C * (C - 1);
  char* w = (char*)scratch;
  auto take = [&](size_t bytes) {
    char* r = w;
    w += (bytes + 255) / 256 * 256;
    return (void*)r;
  };
  float* ph = (float*)take(A * n * 4);
  float* rel = (float*)take(A * n * 4);
  int* comp = (int*)take(A * n * 4);
  int* inc = (int*)take(A * n * 4);
  unsigned* best = (unsigned*)take(A * n * 4);
  unsigned long long* link = (unsigned long long*)take(A * n * 8);
  float* kA = (float*)take(A * m * 4);
  int* iA = (int*)take(A * m * 4);
  float* kB = (float*)take(A * m * 4);
  int* iB = (int*)take(A * m * 4);
  float* rm = (float*)take(A * m * 4);
  int* rank = (int*)take(A * m * 4);
  SortLevels L{};
  L.s[0] = (int)m;
  int depth = 0;
  while (depth + 1 < kMaxDepth && L.s[depth] >= 2) {
    L.s[depth + 1] = L.s[depth] / 2;
    ++depth;
  }
  L.deepest = -1;
  for (int d = 0; d < depth; ++d)
    if (L.s[d] >= 2) L.deepest = d;
  const dim3 gp((unsigned)((n + 255) / 256), (unsigned)A), ge((unsigned)((m + 255) / 256), (unsigned)A);
  {
    Timed tm("herraez_rel", s);
    hipLaunchKernelGGL(k_herr_rel, gp, dim3(256), 0, s, iu, C, border, ph, rel);
  }
  {
    Timed tm("herraez_sort", s);
    hipLaunchKernelGGL(k_herr_sort, dim3(A), dim3(kSortThreads), 0, s, rel, C, L, kA, iA, kB, iB,
                       rm);
  }
  Timed tm("herraez_tree", s);
  hipLaunchKernelGGL(k_bor_init, ge, dim3(256), 0, s, iA, C, rank, comp, inc);
  // every round at least halves the number of components
  int rounds = 1;
  while ((1ull << rounds) < n) ++rounds;
  for (int r = 0; r <= rounds; ++r) {
    hipLaunchKernelGGL(k_bor_reset, gp, dim3(256), 0, s, C, best);
    hipLaunchKernelGGL(k_bor_select, ge, dim3(256), 0, s, C, rank, comp, best);
    hipLaunchKernelGGL(k_bor_hook, gp, dim3(256), 0, s, C, iA, comp, inc, best, ph, link);
    hipLaunchKernelGGL(k_bor_jump, gp, dim3(256), 0, s, C, comp, link);
    hipLaunchKernelGGL(k_bor_apply, gp, dim3(256), 0, s, C, comp, inc, link);
  }
  hipLaunchKernelGGL(k_herr_out, gp, dim3(256), 0, s, C, ph, inc, phu);
  return hipGetLastError();
}

}  // namespace mtd
