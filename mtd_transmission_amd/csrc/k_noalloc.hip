// Allocation-free hologram stage (mtd_holo_to_crop_noalloc): the same
// kernels as mtd_holo_to_crop, with every buffer beyond the caller's
// holograms and crops -- the FFT twiddles, the hologram window and the N x C
// row-pass intermediate -- in one file-scope device array, `scratch`, of
// kNoAllocScratch bytes. Each GPU the library runs on holds its own copy of
// the array until the process ends. Its device address is looked up once per
// device when a context is created there (noalloc_init), and host code only
// ever uses that address, never the array's name.
//
// Every call on a GPU shares that GPU's array and finds what the last call
// left in it: the tables are rewritten whenever the hologram side or window
// changes, and the row intermediate is fully written by each chunk's row
// pass before its column pass reads it, so nothing needs clearing. Calls that
// overlap on one GPU (other streams, other threads, other contexts) race on
// the array.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace mtd {

// tables (twiddles N float2, window N float) first, then the row intermediate
constexpr size_t kNoAllocScratch = 64ull << 20;
constexpr size_t kNoAllocTables = 1ull << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
// what the table region of each device's scratch holds: hologram side and
// the bits of (float)alpha; 0 = nothing yet
uint64_t g_key[kMaxDevices] = {};

// Tables reach the device as kernel arguments (no host buffer, no copy):
// kTableChunk floats per launch keeps the argument block under 4 KiB.
constexpr int kTableChunk = 960;
struct TableChunk {
  float v[kTableChunk];
};
}  // namespace

__global__ void k_fill_table(float* __restrict__ dst, TableChunk c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = c.v[i];
}

// Writes n floats of a host-generated table, kTableChunk at a time; gen(i)
// gives value i.
template <class Gen>
static hipError_t fill_table(float* dst, int n, Gen gen, hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += kTableChunk) {
    TableChunk c;
    const int m = n - i0 < kTableChunk ? n - i0 : kTableChunk;
    for (int i = 0; i < m; ++i) c.v[i] = gen(i0 + i);
    hipLaunchKernelGGL(k_fill_table, dim3((m + 255) / 256), dim3(256), 0, s, dst + i0, c, m);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int noalloc_init(int device, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    if (err) *err = "device ordinal beyond " + std::to_string(kMaxDevices);
    return MTD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return MTD_OK;
  // the caller made `device` current
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    if (err) *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return MTD_ERR_HIP;
  }
  g_ptr[device] = p;
  return MTD_OK;
}

}  // namespace mtd

using namespace mtd;

extern "C" int mtd_holo_to_crop_noalloc(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                                        int dtype, int A, float* crops) {
  if (!ctx) return MTD_ERR_ARG;
  int rc = validate(ctx, g, A, NEED_HOLO);
  if (rc != MTD_OK) return rc;
  if (A == 0) return MTD_OK;
  if (!holos || !crops) return set_err(ctx, MTD_ERR_ARG, "null array");
  if (dtype != MTD_U8 && dtype != MTD_U16 && dtype != MTD_F32)
    return set_err(ctx, MTD_ERR_ARG, "unknown hologram dtype");
  {
    hipPointerAttribute_t at;
    const hipError_t e = hipPointerGetAttributes(&at, holos);
    if (e != hipSuccess) (void)hipGetLastError();
    if (e != hipSuccess || (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged))
      return set_err(ctx, MTD_ERR_ARG,
                     "allocation-free hologram stage: the stack must be in device memory");
  }
  const int N = g->n;
  if (g->alpha_crop > 0)
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   "allocation-free hologram stage: no crop window (alpha_crop 0)");
  if (!holo_len_supported(N))
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   "allocation-free hologram stage: hologram side " + std::to_string(N) +
                       " has no compiled FFT (its runtime-length tables are not kept in the "
                       "scratch array)");
  const size_t t1_per = holo_t1_bytes(g);
  const size_t tables = (size_t)N * (sizeof(float2) + sizeof(float));
  if (tables > kNoAllocTables || t1_per > kNoAllocScratch - kNoAllocTables)
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   "allocation-free hologram stage: one hologram needs " +
                       std::to_string(tables + t1_per) + " bytes of scratch, the maximum is " +
                       std::to_string(kNoAllocScratch));
  std::lock_guard<std::mutex> lk(g_mu);
  if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_ptr[ctx->device])
    return set_err(ctx, MTD_ERR_HIP, "no scratch address for this device (context not created)");
  char* base = (char*)g_ptr[ctx->device];
  float2* tw = (float2*)base;
  float* win = (float*)(base + (size_t)N * sizeof(float2));
  float2* t1 = (float2*)(base + kNoAllocTables);
  const hipStream_t s = ctx->stream;
  const float alpha = (float)g->alpha_holo;
  uint32_t abits;
  std::memcpy(&abits, &alpha, sizeof abits);
  const uint64_t key = (uint64_t)N << 32 | abits;
  if (g_key[ctx->device] != key) {
    // host tables as the allocating path builds them, into the scratch on
    // the stream (ordered after every earlier call on it)
    auto tw_at = [N](int i) {  // float i of the interleaved (re, im) table
      const float2 w = twiddle_at(N, i / 2);
      return (i & 1) ? w.y : w.x;
    };
    hipError_t e = fill_table((float*)tw, 2 * N, tw_at, s);
    if (e == hipSuccess)
      e = fill_table(win, N, [&](int i) { return tukey_signed_at(N, g->alpha_holo, i); }, s);
    if (e != hipSuccess) {
      g_key[ctx->device] = 0;
      return hip_check(ctx, e, "allocation-free tables");
    }
    g_key[ctx->device] = key;
  }
  const int chunk = (int)((kNoAllocScratch - kNoAllocTables) / t1_per);
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = A - a0 < chunk ? A - a0 : chunk;
    FftTab ft;
    ft.tw = tw;
    rc = holo_chunk_with(ctx, g, holos, dtype, a0, n,
                         (float2*)crops + (size_t)a0 * 4 * g->nxmax * g->nxmax, nullptr, win, ft,
                         t1);
    if (rc != MTD_OK) return rc;
  }
  return MTD_OK;
}
