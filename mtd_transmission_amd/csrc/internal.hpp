// Internal declarations shared by the C-ABI layer and the kernel units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mtd.h"

namespace mtd {

// Per-angle Ewald geometry (computed on the device from posspec, FP64 heights
// from a host table): lateral specular offset fm0 = posSpec - NXMAX, its cap
// height fzm0 = sqrt(R^2 - |fm0|^2) and the skip flag of
// Reconstruction/src/fonctions.cpp:897.
struct AngleGeo {
  int fx, fy;
  int skip;
  int pad;
  double fz;
};

// Angle-independent least-squares plan (Correction_aberration.cpp:108-217):
// sample list, LU factors of the normal matrix A^T A with its row permutation.
struct LsqPlan {
  int M = 0;
  int C = 0;
  std::vector<uint8_t> mask;  // host copy used to detect changes
  int2* d_samples = nullptr;  // (x, y) field-frame coordinates
  double* d_lu = nullptr;     // 15x15 row-major, L unit-lower + U
  int* d_perm = nullptr;      // row permutation of the pivoting: (Pb)[i] = b[perm[i]]
};

struct GenPlan;  // gen_fft.hpp

// Tables of one transform length: the twiddles exp(-2 pi i k / L) of the
// compiled register FFTs, and for a length a stage has no compiled FFT for,
// the runtime-length plan (its device tables inside; held by the context).
struct FftTab {
  const float2* tw = nullptr;
  const GenPlan* gen = nullptr;
};

struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct KTime {
  double ms = 0;
  long n = 0;
};

struct EvPair {
  hipEvent_t a, b;
  const char* name;
};

}  // namespace mtd

// A schedule of the plain path picked by the batched form's autotuner (every
// schedule gives the same results): holograms per hologram-stage chunk, the
// gather's height window and load batch (0: defaults), the volume passes'
// tile variants, the time it measured (ms) and whether it came from the
// cache file.
struct TunedSched {
  int chunk, zw, kb, rows_tile, cols_tile;
  float ms;
  bool from_cache;
};
// (hologram side N, crop side C, volume side V, hologram dtype)
using TuneKey = std::tuple<int, int, int, int>;

struct mtd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // host-to-device hologram streaming: copy stream and per-slot events
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_ready[2] = {nullptr, nullptr};
  hipEvent_t ev_free[2] = {nullptr, nullptr};
  // slot k of the staging buffer has been written by an upload since the
  // context was created: the next upload into it (in this call or a later
  // one) must wait for ev_free[k], the last pass that read it
  bool slot_used[2] = {false, false};
  std::string err;
  std::map<int, float2*> twiddles;                       // per FFT length
  // runtime-length FFT plans per length (gen_fft.hpp) and their Bluestein
  // tables (chirp, chirp spectrum)
  std::map<int, mtd::GenPlan*> gen_plans;
  std::map<int, std::pair<float2*, float2*>> bluestein;
  // Herraez border reliabilities per (crop side, first angle, angles)
  std::map<std::tuple<int, int, int>, float*> herraez_border;
  std::map<std::pair<int, uint32_t>, double*> tukey;      // (n, bits of (float)alpha)
  std::map<std::pair<int, uint32_t>, float*> tukey_signed;
  std::map<std::pair<uint64_t, int>, double*> htab;       // heights per (bits of R, table size)
  std::map<std::tuple<uint64_t, int, uint32_t>, float*> betatab;  // (R, r2max, k0) bits
  std::map<std::tuple<uint64_t, int, uint32_t, int>, float*> beta2tab;  // + NXMAX
  std::map<std::tuple<uint64_t, int, uint64_t, uint64_t>, float*> qtab;  // (R, r2max, lambda, n0)
  std::map<std::tuple<uint64_t, int, uint64_t, uint64_t, int>, float*> q2tab;  // + NXMAX
  std::map<int, mtd::Buffer> ws;                          // scratch slots
  // least-squares plans per (crop side, mask); lsq points at the current one
  std::list<mtd::LsqPlan> lsq_plans;
  mtd::LsqPlan* lsq = nullptr;
  // batched form (mtd_reconstruct_batch): the captured graph of the last
  // batch, the arguments it was captured for, its capture stream, the events
  // between its acquisitions, and the hologram chunk picked per shape
  hipStream_t graph_stream = nullptr;
  hipGraph_t batch_graph = nullptr;
  hipGraphExec_t batch_exec = nullptr;
  // the last graph-captured solver loop (mtd_tv_solve), freed at the next one
  hipGraph_t tv_graph = nullptr;
  hipGraphExec_t tv_exec = nullptr;
  std::vector<uint8_t> batch_key;
  // bumped whenever a workspace or fit plan is freed or moved: a graph
  // captured at another generation holds stale addresses and is recaptured
  uint64_t ws_gen = 0;
  uint64_t batch_gen = 0;
  // (N, C, V, dtype) -> the autotuner's schedule; the cache file it is
  // persisted to ("": none)
  std::map<TuneKey, TunedSched> batch_tuned;
  std::string tune_cache;
  int holo_chunk_override = 0;  // > 0: holograms per hologram-stage chunk
  int gather_zw_override = 0;   // > 0: gather height window (8, 16, 24)
  int gather_kb_override = 0;   // > 0: gather entries per load batch (8, 16)
  int vol_rows_tile = 0;        // tile variant of the volume row pass (0, 1)
  int vol_cols_tile = 0;        // tile variant of the volume column passes (0, 1, 2)
  std::map<int, int*> tile_orders;  // the gather's tile dispatch order per volume side
  // geometry plan of the forward model / adjoint (mtd_born_plan): speculars,
  // angle geometry, hit band and the gather's tile lists, in WS_PLAN* slots
  bool plan_valid = false;
  int plan_A = 0;
  mtd_geom plan_g{};
  // per-kernel event timing (mtd_ctx_set_timing)
  bool timing = false;
  std::string timing_only;  // non-empty: only the kernel of that name is timed
  std::vector<mtd::EvPair> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, mtd::KTime> ktimes;
};

namespace mtd {

// Context of the entry point running on this thread (kernel timing hook).
extern thread_local mtd_ctx* tl_ctx;
hipEvent_t ev_take(mtd_ctx* c);

// Brackets one kernel launch with events when timing is on.
struct Timed {
  mtd_ctx* c = nullptr;
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr;
  Timed(const char* n, hipStream_t st) : name(n), s(st) {
    if (tl_ctx && tl_ctx->timing &&
        (tl_ctx->timing_only.empty() || tl_ctx->timing_only == n)) {
      c = tl_ctx;
      a = ev_take(c);
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~Timed() {
    if (!a) return;
    hipEvent_t b = ev_take(c);
    if (!b) return;
    (void)hipEventRecord(b, s);
    c->pending.push_back(EvPair{a, b, name});
  }
};

// scratch slots
enum WsSlot {
  WS_T1 = 0, WS_CROP, WS_WORK, WS_IU, WS_POS, WS_COEF, WS_AGEO, WS_F,
  WS_VOL, WS_CNT, WS_VWORK, WS_HOLOTMP, WS_GATHER, WS_ARGMAX, WS_BAND, WS_PHU, WS_UNWRAP, WS_HOLOPART,
  WS_PLANPOS, WS_PLANAGEO, WS_PLANBAND, WS_PLANGATHER, WS_HSTAGE, WS_BATCHTS,
  WS_TVZ, WS_TVZN, WS_TVGD, WS_TVS
};

int set_err(mtd_ctx* ctx, int code, const std::string& msg);
// geometry and argument checks of every entry point; need: NEED_HOLO
// (hologram side and crop window), NEED_VOL (volume side and optics)
enum { NEED_HOLO = 1, NEED_VOL = 2 };
int validate(mtd_ctx* ctx, const mtd_geom* g, int A, int need);
size_t dtype_bytes(int dtype);
// host copies of the device tables: twiddles exp(-2 pi i k / L), and the
// hologram window (float)tukey1d(n, alpha)[i] * (-1)^i
float2 twiddle_at(int L, int k);
void twiddle_host(int L, float2* out);
float tukey_signed_at(int n, double alpha, int i);
void tukey_signed_host(int n, double alpha, float* out);
// Hologram stage of holograms [a0, a0 + n) with the given tables and a row
// intermediate t1 of n * holo_t1_bytes(g) bytes -> crops (and, when part_out
// is set, the argmax partials).
size_t holo_t1_bytes(const mtd_geom* g);
int holo_chunk_with(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int a0, int n,
                    float2* crop_out, float2* part_out, const float* win, const FftTab& ftN,
                    float2* t1, const float2* qtw = nullptr);
// allocation-free entry points: the per-device scratch address, looked up
// once per device when a context is created on it
int noalloc_init(int device, std::string* err);
int hip_check(mtd_ctx* ctx, hipError_t e, const char* what);
void* ws_get(mtd_ctx* ctx, int slot, size_t bytes);
const float2* twiddle(mtd_ctx* ctx, int L);
// FFT tables of length L; compiled: the stage has a register FFT of that
// length (twiddles only), else the runtime-length plan is built too.
// Returns false when out of memory or L is beyond the runtime-length FFT.
bool fft_tab(mtd_ctx* ctx, int L, bool compiled, FftTab* out);
bool holo_len_supported(int N);
bool crop_len_supported(int C);
bool unwrap_len_supported(int C);
bool vol_len_supported(int V);
// largest length the runtime-length FFT takes (direct: 4096, Bluestein: 2048)
bool gen_len_ok(int L);
const double* tukey_table(mtd_ctx* ctx, int n, double alpha);
// (float)tukey1d[i] * (-1)^i
const float* tukey_signed_table(mtd_ctx* ctx, int n, double alpha);
void tukey1d_host(int n, double alpha, double* out);
int lsq_prepare(mtd_ctx* ctx, int C, const uint8_t* mask);

bool fft_len_supported(int L);

// ---- kernel launchers (return hipError_t of the launch) -------------------
// ws: float tukey1d(N, alpha_holo) with the centring sign (-1)^i folded in
hipError_t launch_holo_rows(int N, int dtype, const void* holos, int A,
                            const float* ws, const FftTab& ft, int x0,
                            int C, float2* t1, hipStream_t s);
// part != nullptr: also writes per-block argmax partials of the crops
// (holo_argmax_parts(N, C) per angle) for launch_argmax_fin
// qtw != nullptr (the C-point twiddle table; holo_q_supported(N, C)): crop
// receives Q, the crops' column inverse DFTs (see k_holo_cols)
hipError_t launch_holo_cols(int N, const float2* t1, int A, const FftTab& ft,
                            int C, int x0, int y0, float2* crop, float2* part,
                            hipStream_t s, const float2* qtw = nullptr);
bool holo_q_supported(int N, int C);
// the ArrayFire variant's crop window (1+i) win[x] win[y] on A crops
hipError_t launch_crop_window(int C, float2* crop, int A, const double* win, hipStream_t s);
// argmax partials per angle the column pass writes (0: none, the field
// stage then finds the specular from the crops)
int holo_argmax_parts(int N, int C);
// row intermediate columns per hologram row (C rounded up to whole tiles)
int holo_t1_cols(int N, int C);

hipError_t launch_argmax(int C, const float2* crop, int A, int* posx,
                         int* posy, float2* part, hipStream_t s);
// second stage only, over S partials per angle
hipError_t launch_argmax_fin(int C, const float2* part, int S, int A, int* posx, int* posy,
                             hipStream_t s);
size_t argmax_scratch_bytes(int A);
hipError_t launch_demod(int C, const float2* crop, const int* posx,
                        const int* posy, int A, const FftTab& ft,
                        float2* work, float2* iu, hipStream_t s);
// the demodulated fields moved from the IU frame to the U frame (a shift by
// C/2 in both axes), A crops of C x C
hipError_t launch_iu_to_u(int C, const float2* iu, int A, float2* u, hipStream_t s);
// The simulator's field synthesis: shift by -posSpec, unnormalised inverse
// FFT, shift by C/2, times scale (delta_f^2): fields[A][C][C].
hipError_t launch_spectra_to_fields(int C, const float2* spec, const int* posx, const int* posy,
                                    int A, const FftTab& ft, float2* work, float scale,
                                    float2* fields, hipStream_t s);
// phu: unwrapped phase [A][C][C] in the U frame, or nullptr for the wrapped
// atan2 of IU (DEROUL 0)
hipError_t launch_lsq(int C, const float2* iu, const float* phu, int A, const LsqPlan& plan,
                      double* coef, hipStream_t s);
// The field stage from Q, the crops' column inverse DFTs (compiled crop
// sides): IU's least-squares sample rows into ius, the fit into coef, then
// one row kernel (row inverse DFT with the specular's shift, field, window,
// forward row FFT) into work and the forward column pass into out, the
// spectra pre-weighted by beta2 (as launch_fwd_spectrum with from_iu = 1).
hipError_t launch_fused_field(int C, const float2* q, const int* posx, const int* posy, int A,
                              const LsqPlan& plan, double* coef, float2* ius, int mode,
                              const double* win, const FftTab& ft, float dx2,
                              const float* beta2, float2* work, float2* out, hipStream_t s);
// mode: bit0 born, bit1 c_aber
hipError_t launch_form_field(int C, const float2* iu, const float2* crop,
                             const int* posx, const int* posy, int A,
                             const double* coef, int mode, const float* phu, float2* fields,
                             hipStream_t s);
// Volkov unwrapping of the wrapped phase of IU into phu; scratch buffers:
// gx, gy [A][2C][2C] float, P [A][2C][2C] float2, Q [A][2C][C] float2;
// tw2C: twiddles of length 2C.
hipError_t launch_unwrap(int C, const float2* iu, int A, const FftTab& ft2C, float* gx,
                         float* gy, float2* P, float2* Q, float* phu, hipStream_t s);
size_t unwrap_scratch_bytes(int C, int A);
// Herraez unwrapping of the wrapped phase of IU into phu (U frame);
// border: the angles' border reliabilities (herraez_border_table), scratch
// of herraez_scratch_bytes(C, A)
hipError_t launch_herraez(int C, const float2* iu, int A, const float* border, void* scratch,
                          float* phu, hipStream_t s);
size_t herraez_scratch_bytes(int C, int A);
const float* herraez_border_table(mtd_ctx* ctx, int C, int a0, int n);
// Per-angle forward spectrum. from_iu=0: src are fields (U frame);
// from_iu=1: src is IU and the field formation runs inside the load.
// Output frame: shift_pos=0 -> unshifted F, 1 -> shifted by +posSpec.
// beta2 != nullptr (shift_pos=0 only): bins pre-multiplied by the Ewald
// factor of their pupil pixel, beta2 being the C x C pupil-frame table
// (gather flag 2).
hipError_t launch_fwd_spectrum(int C, const float2* src, int from_iu,
                               const float2* crop, const double* coef,
                               int mode, const float* phu, const int* posx, const int* posy,
                               int A, const double* win, const FftTab& ft,
                               float dx2, int shift_pos, const float* beta2,
                               float2* work, float2* out, hipStream_t s);

hipError_t launch_angle_geo(const int* posx, const int* posy, int A, int nxmax,
                            double R2, const double* htab, int r2tab,
                            AngleGeo* ageo, hipStream_t s);
// Height band [band.x, band.y] that holds every hit of the non-skipped angles
// (empty: x > y), from the pupil's extreme radii.
hipError_t launch_band(const AngleGeo* ageo, int A, int V, const double* htab, int r2max,
                       int2* band, hipStream_t s);
// band == nullptr: every voxel of the volume (and counts) is written;
// otherwise only the planes of the band are, and [zlo, zhi] is a host-side
// bound of the band (only height windows meeting it are launched).
// order: the tile dispatch order (gather_tile_order); zw: height window
// (8, 16, 24; others fall back to 16); kb: list entries per load batch (8,
// 16). The crop side must keep one spectrum below 2 GiB.
hipError_t launch_gather(const float2* spec, int frame_pos, const AngleGeo* ageo, int A, int C,
                         int nxmax, int V, const double* htab, const float* beta, int r2max,
                         float2* vol, int32_t* cnt, void* scratch, const int* order, int flags,
                         const int2* band, int zlo, int zhi, bool prepass, int zw, int kb,
                         hipStream_t s);
bool gather_window_supported(int zw);
// the gather's prepass alone (per-tile angle lists into scratch)
hipError_t launch_gather_prepass(const AngleGeo* ageo, int A, int C, int nxmax, int V,
                                 const double* htab, int r2max, void* scratch, hipStream_t s);
// the gather's tile dispatch order for a V^3 volume ((V/16)^2 tiles, rounded up)
void gather_tile_order_host(int V, int* order);
size_t gather_scratch_bytes(int V, int A);
// Born forward model: work = centred 3-D spectrum of the contrast (re, im;
// im may be nullptr) times scale (Tp^3 * gsign), stored only on the heights
// of the device band.
hipError_t launch_contrast_spectrum(int V, const float* re, const float* im, float2* work,
                                   float scale, const int2* band, const FftTab& ft,
                                   hipStream_t s);
// One fused FISTA step with the smoothed TV gradient (k_tv.hip).
hipError_t launch_stamp(unsigned long long* out, hipStream_t s);
hipError_t launch_residual(float2* S, const float2* y, const float* w, const float* q2, int A,
                           int C2, hipStream_t s);
hipError_t launch_tv_start(float* x, float* z, size_t n, int positivity, hipStream_t s);
hipError_t launch_tv_fista(int V, const float* z, const float* gd, float* x, float* zn,
                           float sd, float lam, float eps, float step, float beta,
                           int positivity, hipStream_t s);
// Per-sample data weights 1 / (q^2 * hits of the voxel) from the hit counts.
hipError_t launch_cap_weights(const int32_t* cnt, const AngleGeo* ageo, int A,
                              const double* htab, const float* q, int r2max, int nx, int C,
                              int V, float* w, hipStream_t s);
// Per-angle cap samples -i q(|fd|^2) vol[voxel] into spec[A][C][C] (+posSpec frame).
hipError_t launch_cap_sample(const float2* vol, const AngleGeo* ageo, int A, const double* htab,
                             const float* q2, int r2max, int nx, int C, int V, float2* spec,
                             hipStream_t s);
// prenorm=1: vol already divided by counts and centring-signed (gather flag 1).
// band != nullptr: heights outside the band are zero and are not read.
hipError_t launch_finalize(int V, const float2* vol, const int32_t* cnt,
                           int prenorm, const int2* band, float2* work, float df3, float cte,
                           float gsign, float* out_re, float* out_im,
                           uint8_t* otf, const FftTab& ft, int2 tiles, hipStream_t s);

}  // namespace mtd
