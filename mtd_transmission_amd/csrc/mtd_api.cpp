// C-ABI entry points (include/mtd.h): validation, workspaces, host-side
// tables and the stage pipelines over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <complex>

#include "gen_fft.hpp"
#include "internal.hpp"

namespace mtd {

thread_local mtd_ctx* tl_ctx = nullptr;

hipEvent_t ev_take(mtd_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Sets the thread's current context for the kernel-timing hook.
struct CtxScope {
  mtd_ctx* prev;
  explicit CtxScope(mtd_ctx* c) : prev(tl_ctx) { tl_ctx = c; }
  ~CtxScope() { tl_ctx = prev; }
};

static void timing_collect(mtd_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (EvPair& p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      KTime& k = c->ktimes[p.name];
      k.ms += ms;
      k.n += 1;
    }
    c->pool.push_back(p.a);
    c->pool.push_back(p.b);
  }
  c->pending.clear();
}

int set_err(mtd_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_check(mtd_ctx* ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return MTD_OK;
  return set_err(ctx, MTD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void* ws_get(mtd_ctx* ctx, int slot, size_t bytes) {
  Buffer& b = ctx->ws[slot];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    ++ctx->ws_gen;
  }
  if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
    b.ptr = nullptr;
    return nullptr;
  }
  b.bytes = bytes;
  return b.ptr;
}

// device copy of a host table; nullptr when allocation or copy fails
template <class T>
static T* upload(const T* h, size_t n) {
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * n) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

static uint64_t bits_of(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof u);
  return u;
}
static uint32_t bits_of(float v) {
  uint32_t u;
  std::memcpy(&u, &v, sizeof u);
  return u;
}

float2 twiddle_at(int L, int k) {
  // exp(-2 pi i k / L) from an exactly reduced angle
  const long double ang =
      2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)L;
  return make_float2((float)std::cos(ang), (float)-std::sin(ang));
}

void twiddle_host(int L, float2* out) {
  for (int k = 0; k < L; ++k) out[k] = twiddle_at(L, k);
}

const float2* twiddle(mtd_ctx* ctx, int L) {
  auto it = ctx->twiddles.find(L);
  if (it != ctx->twiddles.end()) return it->second;
  std::vector<float2> h(L);
  twiddle_host(L, h.data());
  float2* d = upload(h.data(), h.size());
  if (!d) return nullptr;
  ctx->twiddles[L] = d;
  return d;
}

// the gather's tile dispatch order for a V^3 volume
static const int* tile_order(mtd_ctx* ctx, int V) {
  auto it = ctx->tile_orders.find(V);
  if (it != ctx->tile_orders.end()) return it->second;
  const int ntx = (V + 15) / 16;
  std::vector<int> h((size_t)ntx * ntx);
  gather_tile_order_host(V, h.data());
  int* d = upload(h.data(), h.size());
  if (!d) return nullptr;
  ctx->tile_orders[V] = d;
  return d;
}

// ---------------------------------------------------------------------------
// runtime-length FFT plans (gen_fft.hpp)
// ---------------------------------------------------------------------------
// radices of L from {8, 4, 2, 3, 5, 7, 11, 13}; false if another prime divides L
static bool gen_factor(int L, std::vector<int>& f) {
  f.clear();
  int n = L;
  while (n % 8 == 0) { f.push_back(8); n /= 8; }
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int r : {3, 5, 7, 11, 13})
    while (n % r == 0) { f.push_back(r); n /= r; }
  return n == 1 && (int)f.size() <= kGenMaxPasses;
}

static int bluestein_len(int L) {
  int W = 1;
  while (W < 2 * L - 1) W *= 2;
  return W;
}

bool gen_len_ok(int L) {
  if (L < 2) return false;
  std::vector<int> f;
  const int W = gen_factor(L, f) ? L : bluestein_len(L);
  return gen_lines_per_block(W) >= 1 && (size_t)2 * W * sizeof(float2) <= (size_t)kGenLdsBytes;
}

// in-place radix-2 FFT (forward, unnormalised) of a power-of-two length, in
// long double: the host side of the Bluestein tables
static void fft_pow2_host(std::vector<std::complex<long double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  const long double pi = 3.14159265358979323846264338327950288L;
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const long double ang = -2 * pi * (long double)k / (long double)len;
        const std::complex<long double> w(std::cos(ang), std::sin(ang));
        const std::complex<long double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

static const GenPlan* gen_plan(mtd_ctx* ctx, int L) {
  auto it = ctx->gen_plans.find(L);
  if (it != ctx->gen_plans.end()) return it->second;
  if (!gen_len_ok(L)) return nullptr;
  GenPlan p{};
  p.L = L;
  std::vector<int> f;
  const bool direct = gen_factor(L, f);
  p.W = direct ? L : bluestein_len(L);
  if (!direct) gen_factor(p.W, f);
  p.nf = (int)f.size();
  for (int i = 0; i < p.nf; ++i) p.f[i] = f[i];
  p.tw = twiddle(ctx, p.W);
  if (!p.tw) return nullptr;
  if (!direct) {
    // chirp w_n = exp(-i pi n^2 / L) with n^2 reduced mod 2L exactly; the
    // convolution kernel b = conj(w) wrapped into length W, its spectrum / W
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<float2> ch(L);
    std::vector<std::complex<long double>> b(p.W, 0.0L);
    for (int n = 0; n < L; ++n) {
      const long long m = ((long long)n * n) % (2LL * L);
      const long double ang = -pi * (long double)m / (long double)L;
      const std::complex<long double> w(std::cos(ang), std::sin(ang));
      ch[n] = make_float2((float)w.real(), (float)w.imag());
      b[n] = std::conj(w);
      if (n > 0) b[p.W - n] = std::conj(w);
    }
    fft_pow2_host(b);
    std::vector<float2> bk(p.W);
    for (int k = 0; k < p.W; ++k)
      bk[k] = make_float2((float)(b[k].real() / p.W), (float)(b[k].imag() / p.W));
    float2* dch = upload(ch.data(), ch.size());
    float2* dbk = upload(bk.data(), bk.size());
    if (!dch || !dbk) {
      if (dch) (void)hipFree(dch);
      if (dbk) (void)hipFree(dbk);
      return nullptr;
    }
    ctx->bluestein[L] = std::make_pair(dch, dbk);
    p.chirp = dch;
    p.bk = dbk;
  }
  GenPlan* h = new GenPlan(p);
  ctx->gen_plans[L] = h;
  return h;
}

bool fft_tab(mtd_ctx* ctx, int L, bool compiled, FftTab* out) {
  out->tw = twiddle(ctx, L);
  out->gen = nullptr;
  if (!out->tw) return false;
  if (compiled) return true;
  out->gen = gen_plan(ctx, L);
  return out->gen != nullptr;
}

// tukey2D 1-D factor (pretraitement_hors_axe_CPU/FFT_fonctions.cpp:61-87):
// alpha is a C float, 2*cpt/(alpha*(N-1)) a float division, pi = 3.1415.
static double tukey1d_at(int N, float alpha, int cpt) {
  const int borne1 = (int)std::round(alpha * (float)(N - 1) / 2.0f);
  const int borne2 = (int)std::round((float)(N - 1) * (1.0f - alpha / 2.0f));
  if (cpt < borne1 + 1) {
    const float arg = (float)(2 * cpt) / (alpha * (float)(N - 1)) - 1.0f;
    return 0.5 * (1 + std::cos(3.1415 * (double)arg));
  }
  if (cpt < borne2 + 1) return 1.0;
  const float arg = (float)(2 * cpt) / (alpha * (float)(N - 1)) - 2.0f / alpha + 1.0f;
  return 0.5 * (1 + std::cos(3.1415 * (double)arg));
}

void tukey1d_host(int N, double alpha_d, double* out) {
  for (int cpt = 0; cpt < N; ++cpt) out[cpt] = tukey1d_at(N, (float)alpha_d, cpt);
}

float tukey_signed_at(int n, double alpha, int i) {
  const float w = (float)tukey1d_at(n, (float)alpha, i);
  return (i & 1) ? -w : w;
}

void tukey_signed_host(int n, double alpha, float* out) {
  for (int i = 0; i < n; ++i) out[i] = tukey_signed_at(n, alpha, i);
}

const double* tukey_table(mtd_ctx* ctx, int n, double alpha) {
  // the window depends on alpha only through (float)alpha
  const auto key = std::make_pair(n, bits_of((float)alpha));
  auto it = ctx->tukey.find(key);
  if (it != ctx->tukey.end()) return it->second;
  std::vector<double> h(n);
  tukey1d_host(n, alpha, h.data());
  double* d = upload(h.data(), h.size());
  if (!d) return nullptr;
  ctx->tukey[key] = d;
  return d;
}

const float* tukey_signed_table(mtd_ctx* ctx, int n, double alpha) {
  const auto key = std::make_pair(n, bits_of((float)alpha));
  auto it = ctx->tukey_signed.find(key);
  if (it != ctx->tukey_signed.end()) return it->second;
  std::vector<float> f(n);
  tukey_signed_host(n, alpha, f.data());
  float* d = upload(f.data(), f.size());
  if (!d) return nullptr;
  ctx->tukey_signed[key] = d;
  return d;
}

// ---------------------------------------------------------------------------
// least-squares plan (Correction_aberration.cpp:86-149,186-217)
// ---------------------------------------------------------------------------
static const size_t kLsqPlans = 64;

int lsq_prepare(mtd_ctx* ctx, int C, const uint8_t* mask) {
  std::vector<uint8_t> m(C * C, 255);
  if (mask) std::memcpy(m.data(), mask, C * C);
  for (LsqPlan& q : ctx->lsq_plans)
    if (q.C == C && q.mask == m) {
      ctx->lsq = &q;
      return MTD_OK;
    }
  std::vector<int2> s;
  for (int y = 0; y < C; y += 5)
    for (int x = 0; x < C; x += 5)
      if (m[y * C + x] > 45) s.push_back(make_int2(x, y));
  const int M = (int)s.size();
  if (M < 15)
    return set_err(ctx, MTD_ERR_ARG,
                   "aberration mask leaves " + std::to_string(M) +
                       " sample points; the degree-4 fit needs at least 15");
  // normal matrix A^T A with monomials x^i y^j (i outer, j inner)
  double G[15][15] = {};
  for (const int2& q : s) {
    double xp[5] = {1, 0, 0, 0, 0}, yp[5] = {1, 0, 0, 0, 0}, mono[15];
    for (int i = 1; i < 5; ++i) { xp[i] = xp[i - 1] * q.x; yp[i] = yp[i - 1] * q.y; }
    int k = 0;
    for (int i = 0; i <= 4; ++i)
      for (int j = 0; j <= 4 - i; ++j) mono[k++] = xp[i] * yp[j];
    for (int a = 0; a < 15; ++a)
      for (int b = 0; b < 15; ++b) G[a][b] += mono[a] * mono[b];
  }
  // LU with partial pivoting, sequential row swaps
  int piv[15];
  for (int k = 0; k < 15; ++k) {
    int pr = k;
    for (int i = k + 1; i < 15; ++i)
      if (std::fabs(G[i][k]) > std::fabs(G[pr][k])) pr = i;
    piv[k] = pr;
    if (G[pr][k] == 0.0)
      return set_err(ctx, MTD_ERR_ARG, "aberration fit: singular normal matrix");
    if (pr != k)
      for (int j = 0; j < 15; ++j) std::swap(G[k][j], G[pr][j]);
    for (int i = k + 1; i < 15; ++i) {
      G[i][k] /= G[k][k];
      for (int j = k + 1; j < 15; ++j) G[i][j] -= G[i][k] * G[k][j];
    }
  }
  // a new plan; beyond kLsqPlans the oldest is freed, after the queued fits
  // that may still read it
  if (ctx->lsq_plans.size() >= kLsqPlans) {
    (void)hipStreamSynchronize(ctx->stream);
    LsqPlan& o = ctx->lsq_plans.front();
    (void)hipFree(o.d_samples);
    (void)hipFree(o.d_lu);
    (void)hipFree(o.d_perm);
    ctx->lsq_plans.pop_front();
    ++ctx->ws_gen;
  }
  ctx->lsq_plans.emplace_back();
  LsqPlan& p = ctx->lsq_plans.back();
  ctx->lsq = nullptr;
  p.d_samples = upload(s.data(), s.size());
  p.d_lu = upload(&G[0][0], 225);
  // the sequential swaps as one permutation: (Pb)[i] = b[perm[i]]
  int perm[15];
  for (int k = 0; k < 15; ++k) perm[k] = k;
  for (int k = 0; k < 15; ++k) std::swap(perm[k], perm[piv[k]]);
  p.d_perm = upload(perm, 15);
  p.C = 0;
  if (!p.d_samples || !p.d_lu || !p.d_perm)
    return set_err(ctx, MTD_ERR_HIP, "out of device memory (fit plan)");
  p.M = M;
  p.C = C;
  p.mask = m;
  ctx->lsq = &p;
  return MTD_OK;
}

// ---------------------------------------------------------------------------
// Ewald tables (Reconstruction/src/fonctions.cpp:875-935)
// ---------------------------------------------------------------------------
struct EwaldTables {
  const double* htab;
  const float* beta;   // per r2 <= r2max
  const float* beta2;  // C x C, pupil frame: beta at fd = (c - NXMAX, r - NXMAX), 0 outside
  const float* q;      // Born forward weights per r2 <= r2max (born_weights)
  const float* q2;     // the same in the C x C pupil frame, 0 outside the pupil
  int r2tab;
  int r2max;
  double R2;
};

static int ewald_tables(mtd_ctx* ctx, const mtd_geom* g, EwaldTables* t) {
  const int nx = g->nxmax;
  const double R = g->rayon;
  const double R2 = R * R;
  // pupil: (double)(fdx^2+fdy^2) < coef*NXMAX^2 (fonctions.cpp:915)
  const double lim = g->coef_na * (double)(nx * nx);
  int r2max = -1;
  for (int r = 2 * nx * nx; r >= 0; --r)
    if ((double)r < lim) { r2max = r; break; }
  if (r2max < 0) return set_err(ctx, MTD_ERR_ARG, "empty pupil (COEF_NA_OBJ_LIMIT <= 0)");
  if ((double)r2max > R2)
    return set_err(ctx, MTD_ERR_ARG,
                   "COEF_NA_OBJ_LIMIT*NXMAX^2 exceeds rayon^2: the reference would take "
                   "the square root of a negative height");
  const int r2tab = 2 * nx * nx + 1;
  const auto key = std::make_pair(bits_of(R), r2tab);
  double* d_h;
  auto it = ctx->htab.find(key);
  std::vector<double> h(r2tab);
  for (int r = 0; r < r2tab; ++r) {
    const double v = R2 - (double)r;
    h[r] = v >= 0 ? std::sqrt(v) : 0.0;
  }
  if (it == ctx->htab.end()) {
    d_h = upload(h.data(), h.size());
    if (!d_h) return set_err(ctx, MTD_ERR_HIP, "out of device memory (height table)");
    ctx->htab[key] = d_h;
  } else {
    d_h = it->second;
  }
  // beta = ((k0/PI) * (-2 PI)) * (h * (1/R)); PI = 3.14159, float n0/kv/k0
  const double PI = 3.14159;
  const float n0f = (float)g->n0;
  const float kv = (float)(2 * PI / g->lambda);
  const float k0 = kv * n0f;
  const double gamma = ((double)k0 / PI) * (-2 * PI);
  const double inv_r = 1.0 / R;
  std::vector<float> b(r2max + 1);
  for (int r = 0; r <= r2max; ++r) b[r] = (float)(gamma * (h[r] * inv_r));
  const auto bkey = std::make_tuple(bits_of(R), r2max, bits_of(k0));
  float* d_b;
  auto bt = ctx->betatab.find(bkey);
  if (bt == ctx->betatab.end()) {
    d_b = upload(b.data(), b.size());
    if (!d_b) return set_err(ctx, MTD_ERR_HIP, "out of device memory (beta table)");
    ctx->betatab[bkey] = d_b;
  } else {
    d_b = bt->second;
  }
  // the same factors laid out in the pupil frame, for the forward column
  // pass (coalesced reads instead of a per-bin lookup by r2)
  const int C = 2 * nx;
  const auto b2key = std::make_tuple(bits_of(R), r2max, bits_of(k0), nx);
  float* d_b2;
  auto b2t = ctx->beta2tab.find(b2key);
  if (b2t == ctx->beta2tab.end()) {
    std::vector<float> b2((size_t)C * C);
    for (int r = 0; r < C; ++r)
      for (int c = 0; c < C; ++c) {
        const int fdx = c - nx, fdy = r - nx, r2 = fdx * fdx + fdy * fdy;
        b2[(size_t)r * C + c] = r2 <= r2max ? b[r2] : 0.f;
      }
    d_b2 = upload(b2.data(), b2.size());
    if (!d_b2) return set_err(ctx, MTD_ERR_HIP, "out of device memory (beta table)");
    ctx->beta2tab[b2key] = d_b2;
  } else {
    d_b2 = b2t->second;
  }
  t->htab = d_h;
  t->beta = d_b;
  t->beta2 = d_b2;
  t->q = nullptr;
  t->r2tab = r2tab;
  t->r2max = r2max;
  t->R2 = R2;
  return MTD_OK;
}

// ---------------------------------------------------------------------------
// validation
// ---------------------------------------------------------------------------
int validate(mtd_ctx* ctx, const mtd_geom* g, int A, int need) {
  if (!ctx) return MTD_ERR_ARG;
  if (!g) return set_err(ctx, MTD_ERR_ARG, "null geometry");
  if (A < 0) return set_err(ctx, MTD_ERR_ARG, "negative angle count");
  const int C = 2 * g->nxmax;
  if (g->nxmax <= 0) return set_err(ctx, MTD_ERR_ARG, "NXMAX must be positive");
  // compiled register FFT lengths run fast; any other length the
  // runtime-length FFT takes (direct up to 4096, Bluestein up to 2048) works
  if (!crop_len_supported(C) && !gen_len_ok(C))
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   "crop side 2*NXMAX=" + std::to_string(C) + " beyond the FFT lengths built");
  if (g->c_aber && g->deroul == MTD_DEROUL_VOLKOV && !unwrap_len_supported(C) &&
      !gen_len_ok(2 * C))
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   "unwrapping length 4*NXMAX=" + std::to_string(2 * C) +
                       " beyond the FFT lengths built");
  if (!(g->alpha_crop >= 0) || g->alpha_crop > 1)
    return set_err(ctx, MTD_ERR_ARG, "alpha_crop must be in [0, 1]");
  if (g->deroul != MTD_DEROUL_OFF && g->deroul != MTD_DEROUL_VOLKOV &&
      g->deroul != MTD_DEROUL_HERRAEZ)
    return set_err(ctx, MTD_ERR_ARG, "deroul must be MTD_DEROUL_OFF, _VOLKOV or _HERRAEZ");
  if (need & NEED_HOLO) {
    if (g->n <= 0 || g->n % 2)
      return set_err(ctx, MTD_ERR_ARG,
                     "hologram side CCD_ROIX=" + std::to_string(g->n) +
                         " must be even and positive (the centred transform's fftshift)");
    if (!holo_len_supported(g->n) && !gen_len_ok(g->n))
      return set_err(ctx, MTD_ERR_UNSUPPORTED,
                     "hologram side " + std::to_string(g->n) + " beyond the FFT lengths built");
    const int x0 = g->circle_cx - g->nxmax, y0 = g->circle_cy - g->nxmax;
    if (x0 < 0 || y0 < 0 || x0 + C > g->n || y0 + C > g->n)
      return set_err(ctx, MTD_ERR_ARG,
                     "crop window (CIRCLE_CX/CY +- NXMAX) leaves the hologram");
  }
  if (need & NEED_VOL) {
    const int V = g->dim_final;
    if (V <= 0 || V % 2)
      return set_err(ctx, MTD_ERR_ARG,
                     "DIM_FINAL " + std::to_string(V) + " must be even and positive (fftshift3D)");
    if (!vol_len_supported(V) && !gen_len_ok(V))
      return set_err(ctx, MTD_ERR_UNSUPPORTED,
                     "DIM_FINAL " + std::to_string(V) + " beyond the FFT lengths built");
    if (!(g->rayon > 0) || g->rayon != std::floor(g->rayon))
      return set_err(ctx, MTD_ERR_ARG, "rayon must be a positive integer value");
    // every voxel index must stay inside V^3 (the reference does not check,
    // Reconstruction/src/fonctions.cpp:926-928)
    if (V < 4 * g->nxmax - 2)
      return set_err(ctx, MTD_ERR_ARG, "DIM_FINAL < 4*NXMAX-2: lateral voxel index out of range");
    if (g->rayon > V / 2 - 1)
      return set_err(ctx, MTD_ERR_ARG, "rayon > DIM_FINAL/2-1: axial voxel index out of range");
    if (!(g->delta_f > 0) || !(g->pix_uborn > 0) || !(g->lambda > 0) || !(g->n0 > 0))
      return set_err(ctx, MTD_ERR_ARG, "non-positive optical constants");
  }
  return MTD_OK;
}

#define MTD_TRY(expr)              \
  do {                             \
    int _rc = (expr);              \
    if (_rc != MTD_OK) return _rc; \
  } while (0)

static int chunk_of(size_t bytes_per_angle, size_t budget, int A) {
  size_t c = budget / std::max<size_t>(bytes_per_angle, 1);
  if (c < 1) c = 1;
  if ((int)c > A) c = A;
  return (int)c;
}

size_t dtype_bytes(int dtype) {
  return dtype == MTD_U8 ? 1 : dtype == MTD_U16 ? 2 : 4;
}

// holograms [a0, a0+n) -> crops
// part_out (nullable): argmax partials of these crops, holo_argmax_parts per angle
size_t holo_t1_bytes(const mtd_geom* g) {
  return (size_t)g->n * holo_t1_cols(g->n, 2 * g->nxmax) * sizeof(float2);
}

int holo_chunk_with(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int a0, int n,
                    float2* crop_out, float2* part_out, const float* win, const FftTab& ftN,
                    float2* t1, const float2* qtw) {
  const int N = g->n, C = 2 * g->nxmax;
  const char* src = (const char*)holos + (size_t)a0 * N * N * dtype_bytes(dtype);
  MTD_TRY(hip_check(ctx, launch_holo_rows(N, dtype, src, n, win, ftN, g->circle_cx - g->nxmax,
                                          C, t1, ctx->stream), "hologram row pass"));
  MTD_TRY(hip_check(ctx, launch_holo_cols(N, t1, n, ftN, C, g->circle_cx - g->nxmax,
                                          g->circle_cy - g->nxmax, crop_out, part_out,
                                          ctx->stream, qtw),
                    "hologram column pass"));
  if (g->alpha_crop > 0) {
    const double* wc = tukey_table(ctx, C, g->alpha_crop);
    if (!wc) return set_err(ctx, MTD_ERR_HIP, "out of device memory (crop window)");
    MTD_TRY(hip_check(ctx, launch_crop_window(C, crop_out, n, wc, ctx->stream), "crop window"));
  }
  return MTD_OK;
}

// argmax partials the hologram stage can hand to the field stage: none when
// a crop window follows the column pass (the specular is then found on the
// windowed crops, as the ArrayFire variant's coordSpec sees them)
static int fused_argmax_parts(const mtd_geom* g) {
  return g->alpha_crop > 0 ? 0 : holo_argmax_parts(g->n, 2 * g->nxmax);
}

static int holo_chunk(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                      int dtype, int a0, int n, float2* crop_out, float2* part_out = nullptr,
                      const float2* qtw = nullptr) {
  const float* win = tukey_signed_table(ctx, g->n, g->alpha_holo);
  FftTab ftN;
  const bool tabs = fft_tab(ctx, g->n, holo_len_supported(g->n), &ftN);
  float2* t1 = (float2*)ws_get(ctx, WS_T1, (size_t)n * holo_t1_bytes(g));
  if (!win || !tabs || !t1) return set_err(ctx, MTD_ERR_HIP, "out of device memory (hologram stage)");
  return holo_chunk_with(ctx, g, holos, dtype, a0, n, crop_out, part_out, win, ftN, t1, qtw);
}

static int field_mode(const mtd_geom* g) { return (g->born ? 1 : 0) | (g->c_aber ? 2 : 0); }

// crops [n] (angles a0..) -> fields (out_fields) or unshifted spectra F (out_F)
// part_in (nullable): argmax partials already computed by the hologram
// column pass, S per angle; otherwise the argmax reads the crops
// rand() of glibc with its default seed (the reference never seeds it),
// restated from glibc's published random_r, TYPE_3: a 31-word state seeded by
// the Park-Miller recurrence 16807 x mod (2^31 - 1), r[i] = r[i-31] + r[i-3]
// (mod 2^32), the first 310 words discarded, output r >> 1.
struct GlibcRand {
  uint32_t ring[34];
  uint64_t i = 0;  // index of the next word
  GlibcRand() {
    int32_t r0 = 1;
    ring[0] = 1;
    for (int k = 1; k < 31; ++k) {
      const int32_t hi = r0 / 127773, lo = r0 % 127773;
      int32_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      ring[k] = (uint32_t)w;
      r0 = w;
    }
    for (int k = 31; k < 34; ++k) ring[k] = ring[k - 31];
    i = 34;
    for (int k = 0; k < 310; ++k) word();
  }
  uint32_t word() {
    const uint32_t v = ring[(i - 31) % 34] + ring[(i - 3) % 34];
    ring[i % 34] = v;
    ++i;
    return v;
  }
  int32_t next() { return (int32_t)(word() >> 1); }
};

// Herraez border reliabilities of angles [a0, a0 + n) of a C x C stack:
// initialisePIXELs (deroulement_herraez.cpp:348) gives every pixel
// 9999999 + rand() (a 32-bit int sum, wrapping past INT_MAX) and only the
// border pixels keep it; the program unwrapped a0 images before, each
// drawing C*C numbers. Packed per angle as row 0, row C-1, then columns 0
// and C-1 of rows 1 .. C-2. Cached on the device per (C, a0, n).
const float* herraez_border_table(mtd_ctx* ctx, int C, int a0, int n) {
  const auto key = std::make_tuple(C, a0, n);
  auto it = ctx->herraez_border.find(key);
  if (it != ctx->herraez_border.end()) return it->second;
  const size_t per = 4 * (size_t)C - 4, npix = (size_t)C * C;
  std::vector<float> t((size_t)n * per);
  GlibcRand rng;
  for (size_t k = 0; k < (size_t)a0 * npix; ++k) rng.word();
  for (int a = 0; a < n; ++a) {
    float* o = t.data() + (size_t)a * per;
    for (int y = 0; y < C; ++y)
      for (int x = 0; x < C; ++x) {
        const int32_t r = rng.next();
        const float v = (float)(int32_t)((uint32_t)r + 9999999u);
        if (y == 0) o[x] = v;
        else if (y == C - 1) o[C + x] = v;
        else if (x == 0) o[2 * C + 2 * (y - 1)] = v;
        else if (x == C - 1) o[2 * C + 2 * (y - 1) + 1] = v;
      }
  }
  float* d = upload(t.data(), t.size());
  if (!d) return nullptr;
  if (ctx->herraez_border.size() >= 64) {  // bounded cache: drop all, after queued work
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->herraez_border) (void)hipFree(kv.second);
    ctx->herraez_border.clear();
    ++ctx->ws_gen;
  }
  ctx->herraez_border[key] = d;
  return d;
}

// a0: index of the chunk's first angle in the caller's stack (Herraez
// unwrapping's border reliabilities depend on it)
static int front_chunk(mtd_ctx* ctx, const mtd_geom* g, const float2* crop, int n, int a0,
                       int* posx, int* posy, float2* out_fields, float2* out_F,
                       const float2* part_in = nullptr, int S = 0) {
  const int C = 2 * g->nxmax;
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, crop_len_supported(C), &twC);
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)n * C * C * sizeof(float2));
  float2* iu = (float2*)ws_get(ctx, WS_IU, (size_t)n * C * C * sizeof(float2));
  double* coef = (double*)ws_get(ctx, WS_COEF, (size_t)n * 30 * sizeof(double));
  if (!tabs || !work || !iu || !coef)
    return set_err(ctx, MTD_ERR_HIP, "out of device memory (field stage)");
  if (part_in && S > 0) {
    MTD_TRY(hip_check(ctx, launch_argmax_fin(C, part_in, S, n, posx, posy, ctx->stream), "argmax"));
  } else {
    float2* part = (float2*)ws_get(ctx, WS_ARGMAX, argmax_scratch_bytes(n));
    if (!part) return set_err(ctx, MTD_ERR_HIP, "out of device memory (argmax)");
    MTD_TRY(hip_check(ctx, launch_argmax(C, crop, n, posx, posy, part, ctx->stream), "argmax"));
  }
  MTD_TRY(hip_check(ctx, launch_demod(C, crop, posx, posy, n, twC, work, iu, ctx->stream), "demodulation"));
  const int mode = field_mode(g);
  // DEROUL 1 / VOLKOV 1: the phase is unwrapped before the fit; the
  // reference only unwraps when the aberration correction runs (main.cpp:253-277)
  float* phu = nullptr;
  if (g->c_aber && g->deroul == MTD_DEROUL_VOLKOV) {
    FftTab tw2;
    const bool tabs2 = fft_tab(ctx, 2 * C, unwrap_len_supported(C), &tw2);
    phu = (float*)ws_get(ctx, WS_PHU, (size_t)n * C * C * sizeof(float));
    char* us = (char*)ws_get(ctx, WS_UNWRAP, unwrap_scratch_bytes(C, n));
    if (!tabs2 || !phu || !us)
      return set_err(ctx, MTD_ERR_HIP, "out of device memory (unwrapping)");
    const size_t D2 = (size_t)4 * C * C * n;  // (2C)^2 per angle
    float* gx = (float*)us;
    float* gy = gx + D2;
    float2* P = (float2*)(gy + D2);
    float2* Q = P + D2;
    MTD_TRY(hip_check(ctx, launch_unwrap(C, iu, n, tw2, gx, gy, P, Q, phu, ctx->stream),
                      "phase unwrapping"));
  } else if (g->c_aber && g->deroul == MTD_DEROUL_HERRAEZ) {
    const float* border = herraez_border_table(ctx, C, a0, n);
    phu = (float*)ws_get(ctx, WS_PHU, (size_t)n * C * C * sizeof(float));
    void* hs = ws_get(ctx, WS_UNWRAP, herraez_scratch_bytes(C, n));
    if (!border || !phu || !hs)
      return set_err(ctx, MTD_ERR_HIP, "out of device memory (unwrapping)");
    MTD_TRY(hip_check(ctx, launch_herraez(C, iu, n, border, hs, phu, ctx->stream),
                      "phase unwrapping"));
  }
  if (g->c_aber)
    MTD_TRY(hip_check(ctx, launch_lsq(C, iu, phu, n, *ctx->lsq, coef, ctx->stream), "aberration fit"));
  if (out_fields)
    MTD_TRY(hip_check(ctx, launch_form_field(C, iu, crop, posx, posy, n, coef, mode, phu, out_fields,
                                             ctx->stream), "field formation"));
  if (out_F) {
    const double* winC = tukey_table(ctx, C, g->alpha_field);
    if (!winC) return set_err(ctx, MTD_ERR_HIP, "out of device memory (window)");
    const float dx2 = (float)(g->pix_uborn * g->pix_uborn);
    // the spectra go to the gather: pre-weighted by the Ewald factor
    EwaldTables t;
    MTD_TRY(ewald_tables(ctx, g, &t));
    MTD_TRY(hip_check(ctx, launch_fwd_spectrum(C, iu, 1, crop, coef, mode, phu, posx, posy, n, winC,
                                               twC, dx2, 0, t.beta2, work, out_F, ctx->stream),
                      "forward spectrum"));
  }
  return MTD_OK;
}

// The fused field stage of the plain method: the hologram column pass
// writes Q, the crops' column inverse DFTs, instead of the crops, and the
// field stage runs from Q (launch_fused_field): the wrapped-phase fit
// (C_ABER 1, DEROUL 0), no crop window, compiled crop side whose column
// transform the hologram column kernel holds (holo_q_supported). Returns
// the C-point twiddle table for launch_holo_cols, nullptr where it does not
// apply.
static const float2* fused_qtw(mtd_ctx* ctx, const mtd_geom* g) {
  const int C = 2 * g->nxmax;
  if (!g->c_aber || g->deroul != MTD_DEROUL_OFF || g->alpha_crop > 0) return nullptr;
  if (!crop_len_supported(C) || !holo_q_supported(g->n, C)) return nullptr;
  return twiddle(ctx, C);
}

// argmax from the column pass's partials, then the fused field stage from
// Q to the Ewald-weighted spectra
static int front_fused(mtd_ctx* ctx, const mtd_geom* g, const float2* q, int n, int* posx,
                       int* posy, float2* out_F, const float2* part, int S) {
  const int C = 2 * g->nxmax;
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, true, &twC);
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)n * C * C * sizeof(float2));
  float2* ius = (float2*)ws_get(ctx, WS_IU, (size_t)n * C * C * sizeof(float2));
  double* coef = (double*)ws_get(ctx, WS_COEF, (size_t)n * 30 * sizeof(double));
  const double* winC = tukey_table(ctx, C, g->alpha_field);
  if (!tabs || !work || !ius || !coef || !winC)
    return set_err(ctx, MTD_ERR_HIP, "out of device memory (field stage)");
  MTD_TRY(hip_check(ctx, launch_argmax_fin(C, part, S, n, posx, posy, ctx->stream), "argmax"));
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  const float dx2 = (float)(g->pix_uborn * g->pix_uborn);
  return hip_check(ctx, launch_fused_field(C, q, posx, posy, n, *ctx->lsq, coef, ius,
                                           field_mode(g), winC, twC, dx2, t.beta2, work, out_F,
                                           ctx->stream),
                   "field stage");
}

// Hologram stage: up to 2 GiB of the N x C row-pass intermediate per chunk
// for stacks in device memory (fewer, larger launch pairs: measured over
// 600 holograms of 1024^2, crop 256, chunks of 8 to 600 holograms ran
// fastest at the largest, 600 in one launch pair 10 us per volume faster
// than two of 300; at 2048^2, crop 512, the field stage's chunks bound the
// hologram chunks; a chunk small enough for its intermediate to stay in
// cache between the passes is slower, the launch tails dominate), 128 MiB
// for host stacks (their chunks are the upload granularity, overlapped with
// the passes). Field stage: up to 4 GiB of per-angle buffers per launch, so
// that its small kernels get hundreds of angles at once (1000 angles at
// crop 512, 6 MiB each: two chunks instead of six at 1 GiB, the field
// stage's kernels 4.46 -> 3.97 ms and the hologram stage's 8.88 -> 8.80 ms
// per volume).
static const size_t kChunkBudget = 128ull << 20;
static const size_t kChunkBudgetDevice = 2ull << 30;
static const size_t kFieldBudget = 4ull << 30;

// holograms per hologram-stage chunk when no schedule sets it: the budget
// of row-pass intermediate, split into even chunks
static int default_holo_chunk(int N, int C, int n, bool host) {
  int hc = chunk_of((size_t)N * C * sizeof(float2), host ? kChunkBudget : kChunkBudgetDevice, n);
  if (hc < n) {
    const int nch = (n + hc - 1) / hc;
    hc = (n + nch - 1) / nch;
  }
  return hc;
}

// device bytes per angle of the field stage (crop, work, IU, and with
// unwrapping its scratch and the phase)
static size_t field_bytes(const mtd_geom* g) {
  const size_t C = 2 * (size_t)g->nxmax;
  size_t b = 3 * C * C * sizeof(float2);
  if (g->c_aber && g->deroul == MTD_DEROUL_VOLKOV)
    b += unwrap_scratch_bytes((int)C, 1) + C * C * sizeof(float);
  if (g->c_aber && g->deroul == MTD_DEROUL_HERRAEZ)
    b += herraez_scratch_bytes((int)C, 1) + C * C * sizeof(float);
  return b;
}

// holograms [f0, f0+n) -> crops (and argmax partials, S per angle), in
// hologram-stage chunks
// Host memory (pinned, or pageable: the runtime then stages each copy) as
// reported by the HIP runtime; device and managed memory are read in place.
static bool is_host_ptr(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // unregistered pointer: pageable host memory
    return true;
  }
  return at.type == hipMemoryTypeHost || at.type == hipMemoryTypeUnregistered;
}

// The context's copy stream and its slot events, created on first use
// (default priority, non-blocking with respect to the null stream).
static int copy_engine(mtd_ctx* ctx) {
  if (ctx->copy_stream) return MTD_OK;
  MTD_TRY(hip_check(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking),
                    "copy stream"));
  for (int k = 0; k < 2; ++k) {
    MTD_TRY(hip_check(ctx, hipEventCreateWithFlags(&ctx->ev_ready[k], hipEventDisableTiming),
                      "copy event"));
    MTD_TRY(hip_check(ctx, hipEventCreateWithFlags(&ctx->ev_free[k], hipEventDisableTiming),
                      "copy event"));
  }
  return MTD_OK;
}

// Hologram stage over angles f0 .. f0+n-1 in chunks of at most
// kChunkBudget of row-pass intermediate. Host holograms are streamed: chunk
// i+1 is copied into one of two device slots on the copy stream while chunk
// i is transformed; the copy into a slot waits for the row pass that last
// read it, the row pass waits for its copy.
static int holo_range(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int f0,
                      int n, float2* crop, float2* part, int S, const float2* qtw = nullptr) {
  const int N = g->n, C = 2 * g->nxmax;
  const bool host = is_host_ptr(holos);
  const int hc = ctx->holo_chunk_override > 0
                     ? std::min(ctx->holo_chunk_override, std::max(n, 1))
                     : default_holo_chunk(N, C, n, host);
  auto part_at = [&](int h0) { return part ? part + (size_t)h0 * S : nullptr; };
  if (!host) {
    for (int h0 = 0; h0 < n; h0 += hc)
      MTD_TRY(holo_chunk(ctx, g, holos, dtype, f0 + h0, std::min(hc, n - h0),
                         crop + (size_t)h0 * C * C, part_at(h0), qtw));
    return MTD_OK;
  }
  MTD_TRY(copy_engine(ctx));
  const size_t hb = (size_t)N * N * dtype_bytes(dtype);
  char* slots = (char*)ws_get(ctx, WS_HSTAGE, 2 * (size_t)hc * hb);
  if (!slots) return set_err(ctx, MTD_ERR_HIP, "out of device memory (hologram staging)");
  const int nchunks = (n + hc - 1) / hc;
  // A previous call may have used another chunk size, so its slots can
  // overlap either slot of this call: the first upload waits for the last
  // readers of both. Later uploads into slot k wait for the pass that last
  // read slot k.
  for (int k = 0; k < 2; ++k)
    if (ctx->slot_used[k])
      MTD_TRY(hip_check(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->ev_free[k], 0),
                        "copy wait"));
  bool used[2] = {false, false};
  auto prefetch = [&](int i) -> int {
    const int k = i % 2, h0 = i * hc, m = std::min(hc, n - h0);
    if (used[k])
      MTD_TRY(hip_check(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->ev_free[k], 0),
                        "copy wait"));
    MTD_TRY(hip_check(ctx, hipMemcpyAsync(slots + (size_t)k * hc * hb,
                                          (const char*)holos + (size_t)(f0 + h0) * hb,
                                          (size_t)m * hb, hipMemcpyHostToDevice,
                                          ctx->copy_stream),
                      "hologram upload"));
    MTD_TRY(hip_check(ctx, hipEventRecord(ctx->ev_ready[k], ctx->copy_stream), "copy event"));
    used[k] = true;
    return MTD_OK;
  };
  MTD_TRY(prefetch(0));
  for (int i = 0; i < nchunks; ++i) {
    if (i + 1 < nchunks) MTD_TRY(prefetch(i + 1));
    const int k = i % 2, h0 = i * hc, m = std::min(hc, n - h0);
    MTD_TRY(hip_check(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_ready[k], 0), "compute wait"));
    MTD_TRY(holo_chunk(ctx, g, slots + (size_t)k * hc * hb, dtype, 0, m,
                       crop + (size_t)h0 * C * C, part_at(h0), qtw));
    MTD_TRY(hip_check(ctx, hipEventRecord(ctx->ev_free[k], ctx->stream), "slot event"));
    ctx->slot_used[k] = true;
  }
  return MTD_OK;
}

}  // namespace mtd

using namespace mtd;

extern "C" {

const char* mtd_version(void) { return "mtd-mi355x 0.1 (gfx950)"; }

static int tune_cache_load(mtd_ctx* ctx);

// message of the last failed mtd_ctx_create on this thread
static thread_local std::string tl_create_err;

int mtd_ctx_create(int device, mtd_ctx** out) {
  if (!out) return MTD_ERR_ARG;
  *out = nullptr;
  const hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    tl_create_err = "hipSetDevice(" + std::to_string(device) + "): " + hipGetErrorName(e) +
                    " (" + hipGetErrorString(e) + ")";
    return MTD_ERR_HIP;
  }
  // the allocation-free entry points' scratch address on this device
  std::string nerr;
  if (noalloc_init(device, &nerr) != MTD_OK) {
    tl_create_err = nerr;
    return MTD_ERR_HIP;
  }
  mtd_ctx* c = new mtd_ctx();
  c->device = device;
  c->stream = nullptr;
  if (const char* cache = std::getenv("MTD_TUNE_CACHE")) {
    c->tune_cache = cache;
    (void)tune_cache_load(c);  // an unreadable file only means nothing is cached
  }
  *out = c;
  return MTD_OK;
}

int mtd_ctx_destroy(mtd_ctx* ctx) {
  if (!ctx) return MTD_OK;
  // teardown is best effort: errors here have no caller left to act on them
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  timing_collect(ctx);
  for (hipEvent_t e : ctx->pool) (void)hipEventDestroy(e);
  for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
  for (auto& kv : ctx->tile_orders) (void)hipFree(kv.second);
  for (auto& kv : ctx->gen_plans) delete kv.second;
  for (auto& kv : ctx->herraez_border) (void)hipFree(kv.second);
  for (auto& kv : ctx->bluestein) {
    (void)hipFree(kv.second.first);
    (void)hipFree(kv.second.second);
  }
  for (auto& kv : ctx->tukey) (void)hipFree(kv.second);
  for (auto& kv : ctx->tukey_signed) (void)hipFree(kv.second);
  for (auto& kv : ctx->htab) (void)hipFree(kv.second);
  for (auto& kv : ctx->betatab) (void)hipFree(kv.second);
  for (auto& kv : ctx->beta2tab) (void)hipFree(kv.second);
  for (auto& kv : ctx->qtab) (void)hipFree(kv.second);
  for (auto& kv : ctx->q2tab) (void)hipFree(kv.second);
  for (auto& kv : ctx->ws) (void)hipFree(kv.second.ptr);
  if (ctx->copy_stream) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    for (int k = 0; k < 2; ++k) {
      (void)hipEventDestroy(ctx->ev_ready[k]);
      (void)hipEventDestroy(ctx->ev_free[k]);
    }
    (void)hipStreamDestroy(ctx->copy_stream);
  }
  for (LsqPlan& p : ctx->lsq_plans) {
    (void)hipFree(p.d_samples);
    (void)hipFree(p.d_lu);
    (void)hipFree(p.d_perm);
  }
  if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
  if (ctx->batch_graph) (void)hipGraphDestroy(ctx->batch_graph);
  if (ctx->tv_exec) (void)hipGraphExecDestroy(ctx->tv_exec);
  if (ctx->tv_graph) (void)hipGraphDestroy(ctx->tv_graph);
  if (ctx->graph_stream) (void)hipStreamDestroy(ctx->graph_stream);
  delete ctx;
  return MTD_OK;
}

int mtd_ctx_set_stream(mtd_ctx* ctx, void* hip_stream) {
  if (!ctx) return MTD_ERR_ARG;
  ctx->stream = (hipStream_t)hip_stream;
  return MTD_OK;
}

int mtd_ctx_sync(mtd_ctx* ctx) {
  if (!ctx) return MTD_ERR_ARG;
  return hip_check(ctx, hipStreamSynchronize(ctx->stream), "stream synchronize");
}

const char* mtd_last_error(const mtd_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return tl_create_err.empty() ? "null context" : tl_create_err.c_str();
}

int mtd_ctx_set_timing(mtd_ctx* ctx, int on) {
  if (!ctx) return MTD_ERR_ARG;
  ctx->timing = on != 0;
  return MTD_OK;
}

int mtd_ctx_set_schedule(mtd_ctx* ctx, int holo_chunk, int gather_window) {
  if (!ctx) return MTD_ERR_ARG;
  if (holo_chunk < 0) return set_err(ctx, MTD_ERR_ARG, "negative hologram chunk");
  if (gather_window != 0 && !gather_window_supported(gather_window))
    return set_err(ctx, MTD_ERR_UNSUPPORTED, "gather window " + std::to_string(gather_window) +
                                                 " not built (8, 16, 24)");
  ctx->holo_chunk_override = holo_chunk;
  ctx->gather_zw_override = gather_window;
  return MTD_OK;
}

// the named schedule parameters: their context slot and accepted values
static int* sched_slot(mtd_ctx* ctx, const std::string& name) {
  if (name == "holo_chunk") return &ctx->holo_chunk_override;
  if (name == "gather_window") return &ctx->gather_zw_override;
  if (name == "gather_batch") return &ctx->gather_kb_override;
  if (name == "vol_rows_tile") return &ctx->vol_rows_tile;
  if (name == "vol_cols_tile") return &ctx->vol_cols_tile;
  return nullptr;
}

static bool sched_value_ok(const std::string& name, int v) {
  if (v == 0) return true;  // the default
  if (name == "holo_chunk") return v > 0;
  if (name == "gather_window") return gather_window_supported(v);
  if (name == "gather_batch") return v == 8 || v == 16;
  if (name == "vol_rows_tile") return v == 1;
  if (name == "vol_cols_tile") return v == 1 || v == 2;
  return false;
}

int mtd_ctx_set_schedule_param(mtd_ctx* ctx, const char* name, int value) {
  if (!ctx || !name) return MTD_ERR_ARG;
  int* slot = sched_slot(ctx, name);
  if (!slot) return set_err(ctx, MTD_ERR_ARG, std::string("unknown schedule parameter ") + name);
  if (!sched_value_ok(name, value))
    return set_err(ctx, MTD_ERR_UNSUPPORTED,
                   std::string("schedule parameter ") + name + " = " + std::to_string(value) +
                       " not built");
  *slot = value;
  return MTD_OK;
}

int mtd_ctx_get_schedule_param(mtd_ctx* ctx, const char* name, int* value) {
  if (!ctx || !name || !value) return MTD_ERR_ARG;
  int* slot = sched_slot(ctx, name);
  if (!slot) return set_err(ctx, MTD_ERR_ARG, std::string("unknown schedule parameter ") + name);
  *value = *slot;
  return MTD_OK;
}

int mtd_ctx_set_timing_only(mtd_ctx* ctx, const char* kernel) {
  if (!ctx) return MTD_ERR_ARG;
  ctx->timing_only = kernel ? kernel : "";
  return MTD_OK;
}

int mtd_ctx_timing_report(mtd_ctx* ctx, char* buf, size_t cap) {
  if (!ctx || !buf || cap == 0) return -1;
  timing_collect(ctx);
  std::string out;
  for (const auto& kv : ctx->ktimes) {
    char line[256];
    snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.n, kv.second.ms);
    out += line;
  }
  std::strncpy(buf, out.c_str(), cap - 1);
  buf[cap - 1] = 0;
  return (int)ctx->ktimes.size();
}

int mtd_ctx_timing_reset(mtd_ctx* ctx) {
  if (!ctx) return MTD_ERR_ARG;
  timing_collect(ctx);
  ctx->ktimes.clear();
  return MTD_OK;
}

int mtd_tukey1d(int n, double alpha, double* out) {
  if (n < 2 || !out) return MTD_ERR_ARG;
  tukey1d_host(n, alpha, out);
  return MTD_OK;
}

int mtd_geom_derive(double ccd_roix, double nxmax, double dim_final, double circle_cx,
                    double circle_cy, double na, double n0, double lambda, double tpcam,
                    double f_tube, double f_obj, double rf, double coef_na, int born,
                    int c_aber, int deroul, mtd_geom* out) {
  if (!out) return MTD_ERR_ARG;
  // every value passes through a C float (extract_val), then the double
  // derivations of manip.cpp:37-57,104 / Reconstruction/src/manip.cpp:52-94
  const double n0d = (float)n0, nad = (float)na, lamd = (float)lambda;
  const double G = (double)(float)f_tube / (double)(float)f_obj;
  const double Gt = G / (double)(float)rf;
  const double pix = (double)(float)tpcam / Gt;
  const size_t NX = (size_t)(float)nxmax;
  const int N = (int)(size_t)(float)ccd_roix;
  mtd_geom g{};
  g.n = N;
  g.nxmax = (int)NX;
  g.dim_final = (int)(size_t)(float)dim_final;
  g.circle_cx = (int)(size_t)(float)circle_cx;
  g.circle_cy = (int)(size_t)(float)circle_cy;
  g.born = born;
  g.c_aber = c_aber;
  g.deroul = deroul;
  g.n0 = n0d;
  g.na = nad;
  g.lambda = lamd;
  g.pix_holo = pix;
  g.rayon = std::round((double)NX * n0d / nad);
  g.pix_uborn = NX ? pix * N / (double)(2 * NX) : 0.0;
  g.delta_f = NX ? 1.0 / (g.pix_uborn * 2 * (double)NX) : 0.0;
  g.coef_na = (float)coef_na;
  g.alpha_holo = 0.05;
  g.alpha_field = 0.1;
  g.alpha_crop = 0.0;
  *out = g;
  return NX ? MTD_OK : MTD_ERR_ARG;
}

int mtd_holo_to_crop(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype,
                     int A, float* crops) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_HOLO));
  if (A == 0) return MTD_OK;
  if (!holos || !crops) return set_err(ctx, MTD_ERR_ARG, "null array");
  return holo_range(ctx, g, holos, dtype, 0, A, (float2*)crops, nullptr, 0);
}

int mtd_crop_to_field(mtd_ctx* ctx, const mtd_geom* g, const float* crops, int A,
                      const uint8_t* mask, int32_t* posspec, float* fields) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, 0));
  if (A == 0) return MTD_OK;
  if (!crops || !posspec || !fields) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  if (g->c_aber) MTD_TRY(lsq_prepare(ctx, C, mask));
  const int chunk = chunk_of(field_bytes(g), kFieldBudget, A);
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(front_chunk(ctx, g, (const float2*)crops + (size_t)a0 * C * C, n, a0, posspec + a0,
                        posspec + A + a0, (float2*)fields + (size_t)a0 * C * C, nullptr));
  }
  return MTD_OK;
}

int mtd_demodulate(mtd_ctx* ctx, const mtd_geom* g, const float* crops, int A,
                   const int32_t* posspec, float* u) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, 0));
  if (A == 0) return MTD_OK;
  if (!crops || !posspec || !u) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, crop_len_supported(C), &twC);
  const int chunk = chunk_of(2 * (size_t)C * C * sizeof(float2), kFieldBudget, A);
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)chunk * C * C * sizeof(float2));
  float2* iu = (float2*)ws_get(ctx, WS_IU, (size_t)chunk * C * C * sizeof(float2));
  if (!tabs || !work || !iu) return set_err(ctx, MTD_ERR_HIP, "out of device memory (demodulation)");
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(hip_check(ctx, launch_demod(C, (const float2*)crops + (size_t)a0 * C * C,
                                        posspec + a0, posspec + A + a0, n, twC, work, iu,
                                        ctx->stream),
                      "demodulation"));
    MTD_TRY(hip_check(ctx, launch_iu_to_u(C, iu, n, (float2*)u + (size_t)a0 * C * C,
                                          ctx->stream),
                      "demodulation frame"));
  }
  return MTD_OK;
}

int mtd_preprocess(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int A,
                   const uint8_t* mask, int32_t* posspec, float* fields) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_HOLO));
  if (A == 0) return MTD_OK;
  if (!holos || !posspec || !fields) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  if (g->c_aber) MTD_TRY(lsq_prepare(ctx, C, mask));
  const int chunk = chunk_of(field_bytes(g), kFieldBudget, A);
  const int S = fused_argmax_parts(g);
  float2* crop = (float2*)ws_get(ctx, WS_CROP, (size_t)chunk * C * C * sizeof(float2));
  float2* part =
      (float2*)ws_get(ctx, WS_HOLOPART, (size_t)chunk * std::max(S, 1) * sizeof(float2));
  if (!crop || !part) return set_err(ctx, MTD_ERR_HIP, "out of device memory (crop stack)");
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(holo_range(ctx, g, holos, dtype, a0, n, crop, S > 0 ? part : nullptr, S));
    MTD_TRY(front_chunk(ctx, g, crop, n, a0, posspec + a0, posspec + A + a0,
                        (float2*)fields + (size_t)a0 * C * C, nullptr, part, S));
  }
  return MTD_OK;
}

int mtd_field_to_spectrum(mtd_ctx* ctx, const mtd_geom* g, const float* fields,
                          const int32_t* posspec, int A, float* spectra) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, 0));
  if (A == 0) return MTD_OK;
  if (!fields || !posspec || !spectra) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, crop_len_supported(C), &twC);
  const double* winC = tukey_table(ctx, C, g->alpha_field);
  const int chunk = chunk_of((size_t)C * C * sizeof(float2), kFieldBudget, A);
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)chunk * C * C * sizeof(float2));
  if (!tabs || !winC || !work) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  const float dx2 = (float)(g->pix_uborn * g->pix_uborn);
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(hip_check(ctx, launch_fwd_spectrum(C, (const float2*)fields + (size_t)a0 * C * C, 0,
                                               nullptr, nullptr, field_mode(g), nullptr, posspec + a0,
                                               posspec + A + a0, n, winC, twC, dx2, 1, nullptr,
                                               work, (float2*)spectra + (size_t)a0 * C * C,
                                               ctx->stream),
                      "forward spectrum"));
  }
  return MTD_OK;
}

// Born forward-model weight of pupil radius r2 (calcHolo,
// Simul_holo/src/fonctions.cpp:115-122,136, PI = 3.14159 there): the product
// cteNormalisation * ctePot2UBorn * cteInd2Pot = -i kv (FP64), divided by
// sdz = sqrt(R^2 - r2) / R; q(r2) = (float)(kv / sdz), the bin value being
// -i q TF3D. The adjoint gather uses the same q: its built-in factor i gives
// conj(-i q) = i q.
static int born_weights(mtd_ctx* ctx, const mtd_geom* g, EwaldTables* t) {
  const double PI = 3.14159;
  const double kv = 2 * PI / g->lambda, k0 = kv * g->n0;
  // |(-1/(2 PI)) * (-i PI/k0) * (-2 kv^2 n0)|
  const double c = (1.0 / (2 * PI)) * (PI / k0) * (2 * kv * kv * g->n0);
  const double R = g->rayon;
  const auto key = std::make_tuple(bits_of(R), t->r2max, bits_of(g->lambda), bits_of(g->n0));
  const int nx = g->nxmax, C = 2 * nx;
  const auto key2 = std::make_tuple(bits_of(R), t->r2max, bits_of(g->lambda), bits_of(g->n0), nx);
  auto it = ctx->qtab.find(key);
  auto it2 = ctx->q2tab.find(key2);
  if (it != ctx->qtab.end() && it2 != ctx->q2tab.end()) {
    t->q = it->second;
    t->q2 = it2->second;
    return MTD_OK;
  }
  std::vector<float> q(t->r2max + 1);
  for (int r = 0; r <= t->r2max; ++r) {
    const double sdz = std::sqrt(t->R2 - (double)r) / R;
    q[r] = (float)(c / sdz);
  }
  if (it == ctx->qtab.end()) {
    float* d_q = upload(q.data(), q.size());
    if (!d_q) return set_err(ctx, MTD_ERR_HIP, "out of device memory (forward weights)");
    it = ctx->qtab.emplace(key, d_q).first;
  }
  if (it2 == ctx->q2tab.end()) {
    // the pupil-frame layout of the spectra: bin (c, r) is fd = (c - NXMAX, r - NXMAX)
    std::vector<float> q2((size_t)C * C, 0.f);
    for (int r = 0; r < C; ++r)
      for (int cc = 0; cc < C; ++cc) {
        const int fdx = cc - nx, fdy = r - nx, r2 = fdx * fdx + fdy * fdy;
        if (r2 <= t->r2max) q2[(size_t)r * C + cc] = q[r2];
      }
    float* d_q2 = upload(q2.data(), q2.size());
    if (!d_q2) return set_err(ctx, MTD_ERR_HIP, "out of device memory (forward weights)");
    it2 = ctx->q2tab.emplace(key2, d_q2).first;
  }
  t->q = it->second;
  t->q2 = it2->second;
  return MTD_OK;
}

// flags: 1 = write the count-normalised, centring-signed spectrum (no counts);
// 2 = spectra already multiplied by the Ewald factor of their pupil pixel;
// 4 = the centring-signed plain sums (no counts).
// band != nullptr: the hit band is computed into it and only the heights that
// meet it are written. weights: per-r2 factors in place of beta (the
// adjoint of the Born forward model), or nullptr.
// The geometry of one set of speculars, computed per call or taken from the
// context's plan (mtd_born_plan): posspec, angle table, hit band, tile lists.
struct GeoBufs {
  const int32_t* pos = nullptr;
  const AngleGeo* ag = nullptr;
  int2* band = nullptr;
  void* scratch = nullptr;  // the gather's tile lists, built
};

static int plan_bufs(mtd_ctx* ctx, const mtd_geom* g, int A, GeoBufs* b) {
  if (!ctx->plan_valid || A != ctx->plan_A || std::memcmp(g, &ctx->plan_g, sizeof(mtd_geom)))
    return set_err(ctx, MTD_ERR_ARG,
                   "posspec NULL needs a geometry plan (mtd_born_plan) of the same geometry "
                   "and angle count");
  b->pos = (const int32_t*)ctx->ws[WS_PLANPOS].ptr;
  b->ag = (const AngleGeo*)ctx->ws[WS_PLANAGEO].ptr;
  b->band = (int2*)ctx->ws[WS_PLANBAND].ptr;
  b->scratch = ctx->ws[WS_PLANGATHER].ptr;
  return MTD_OK;
}

static int scatter_impl(mtd_ctx* ctx, const mtd_geom* g, const float2* spec, int frame_pos,
                        const int32_t* posspec, int A, float* vol, int32_t* counts,
                        int flags = 0, int2* band = nullptr, const float* weights = nullptr,
                        const GeoBufs* plan = nullptr) {
  // per-voxel hit counts are accumulated in 16 bits (one hit per angle)
  if (A > 65535) return set_err(ctx, MTD_ERR_UNSUPPORTED, "more than 65535 angles per volume");
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  const AngleGeo* ag;
  void* scratch;
  if (plan) {  // geometry, band and tile lists already built
    ag = plan->ag;
    scratch = plan->scratch;
    posspec = plan->pos;
    if (band && band != plan->band) return set_err(ctx, MTD_ERR_ARG, "band of another plan");
  } else {
    AngleGeo* agw = (AngleGeo*)ws_get(ctx, WS_AGEO, (size_t)std::max(A, 1) * sizeof(AngleGeo));
    scratch = ws_get(ctx, WS_GATHER, gather_scratch_bytes(g->dim_final, A));
    if (!agw || !scratch) return set_err(ctx, MTD_ERR_HIP, "out of device memory (angle table)");
    if (A > 0)
      MTD_TRY(hip_check(ctx, launch_angle_geo(posspec, posspec + A, A, g->nxmax, t.R2, t.htab,
                                              t.r2tab, agw, ctx->stream), "angle geometry"));
    if (band)
      MTD_TRY(hip_check(ctx, launch_band(agw, A, g->dim_final, t.htab, t.r2max, band,
                                         ctx->stream), "hit band"));
    ag = agw;
  }
  // Host-side bound of the hit band: heights h(r2) - fz with h(r2) in
  // [h(r2max), R] and fz = h(|fm|^2) in [0, R] for every non-skipped angle.
  const int V = g->dim_final;
  const int zlo = V / 2 + (int)std::round(std::sqrt(t.R2 - (double)t.r2max) - g->rayon);
  const int zhi = V / 2 + (int)std::round(g->rayon);
  const int* order = tile_order(ctx, V);
  if (!order) return set_err(ctx, MTD_ERR_HIP, "out of device memory (tile order)");
  MTD_TRY(hip_check(ctx, launch_gather(spec, frame_pos, ag, A, 2 * g->nxmax, g->nxmax, V, t.htab,
                                       weights ? weights : t.beta, t.r2max, (float2*)vol, counts,
                                       scratch, order, flags, band, zlo, zhi, plan == nullptr,
                                       ctx->gather_zw_override > 0 ? ctx->gather_zw_override : 16,
                                       ctx->gather_kb_override > 0 ? ctx->gather_kb_override : 8,
                                       ctx->stream),
                    "Ewald gather"));
  return MTD_OK;
}

int mtd_ewald_scatter(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                      const int32_t* posspec, int A, float* vol, int32_t* counts) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (!vol || !counts || (A > 0 && (!spectra || !posspec)))
    return set_err(ctx, MTD_ERR_ARG, "null array");
  return scatter_impl(ctx, g, (const float2*)spectra, 1, posspec, A, vol, counts);
}

int mtd_ewald_accumulate(mtd_ctx* ctx, const mtd_geom* g, const float* fields,
                         const int32_t* posspec, int A, float* vol, int32_t* counts) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (!vol || !counts || (A > 0 && (!fields || !posspec)))
    return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  float2* F = (float2*)ws_get(ctx, WS_F, (size_t)std::max(A, 1) * C * C * sizeof(float2));
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, crop_len_supported(C), &twC);
  const double* winC = tukey_table(ctx, C, g->alpha_field);
  const int chunk = chunk_of((size_t)C * C * sizeof(float2), kFieldBudget, std::max(A, 1));
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)chunk * C * C * sizeof(float2));
  if (!F || !tabs || !winC || !work) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  const float dx2 = (float)(g->pix_uborn * g->pix_uborn);
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(hip_check(ctx, launch_fwd_spectrum(C, (const float2*)fields + (size_t)a0 * C * C, 0,
                                               nullptr, nullptr, field_mode(g), nullptr, posspec + a0,
                                               posspec + A + a0, n, winC, twC, dx2, 0, t.beta2,
                                               work, F + (size_t)a0 * C * C, ctx->stream),
                      "forward spectrum"));
  }
  // spectra pre-weighted by the Ewald factor (gather flag 2)
  return scatter_impl(ctx, g, F, 0, posspec, A, vol, counts, 2);
}

// scale > 0: output factor in place of delta_f^3 * (-1/(2 n0 kv^2)).
static int finalize_impl(mtd_ctx* ctx, const mtd_geom* g, const float* vol,
                         const int32_t* counts, int prenorm, const int2* band, float* index_re,
                         float* index_im, uint8_t* otf, double scale = 0) {
  const int V = g->dim_final;
  float2* work = (float2*)ws_get(ctx, WS_VWORK, (size_t)V * V * V * sizeof(float2));
  FftTab twV;
  const bool tabs = fft_tab(ctx, V, vol_len_supported(V), &twV);
  if (!work || !tabs) return set_err(ctx, MTD_ERR_HIP, "out of device memory (volume work)");
  // Reconstruction/main.cpp:346-347 (PI = M_PI there) and the delta_f^3 of
  // TF3Dcplx_INV (src/FFT_fonctions.cpp:399)
  const double kv = 2 * M_PI / g->lambda;
  const double cte = scale > 0 ? 1.0 : -1.0 / (2 * g->n0 * kv * kv);
  const float df3 = scale > 0 ? (float)scale : (float)std::pow(g->delta_f, 3);
  const float gsign = ((V / 2) & 1) ? -1.f : 1.f;
  return hip_check(ctx, launch_finalize(V, (const float2*)vol, counts, prenorm, band, work, df3,
                                        (float)cte, gsign, index_re, index_im, otf, twV,
                                        make_int2(ctx->vol_rows_tile, ctx->vol_cols_tile),
                                        ctx->stream),
                   "finalize");
}

int mtd_finalize(mtd_ctx* ctx, const mtd_geom* g, const float* vol, const int32_t* counts,
                 float* index_re, float* index_im, uint8_t* otf) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, 0, NEED_VOL));
  if (!vol || !counts || !index_re || !index_im) return set_err(ctx, MTD_ERR_ARG, "null array");
  return finalize_impl(ctx, g, vol, counts, 0, nullptr, index_re, index_im, otf);
}

// ---- Born forward model and its adjoint ----------------------------------
// voxel size of the reconstruction, 1/(V delta_f): Simul_holo's Tp_Tomo
static double tomo_pixel(const mtd_geom* g) { return 1.0 / (g->dim_final * g->delta_f); }

// Geometry of a forward / adjoint / weights call: the context's plan when
// posspec is NULL, else computed now into the per-call slots (band and angle
// table; the tile lists are built by the gather itself).
static int call_geometry(mtd_ctx* ctx, const mtd_geom* g, const int32_t* posspec, int A,
                         const EwaldTables& t, bool with_band, GeoBufs* b, bool* planned) {
  *planned = posspec == nullptr;
  if (*planned) return plan_bufs(ctx, g, A, b);
  AngleGeo* ag = (AngleGeo*)ws_get(ctx, WS_AGEO, (size_t)std::max(A, 1) * sizeof(AngleGeo));
  int2* band = (int2*)ws_get(ctx, WS_BAND, sizeof(int2));
  if (!ag || !band) return set_err(ctx, MTD_ERR_HIP, "out of device memory (angle table)");
  MTD_TRY(hip_check(ctx, launch_angle_geo(posspec, posspec + A, A, g->nxmax, t.R2, t.htab,
                                          t.r2tab, ag, ctx->stream), "angle geometry"));
  if (with_band)
    MTD_TRY(hip_check(ctx, launch_band(ag, A, g->dim_final, t.htab, t.r2max, band, ctx->stream),
                      "hit band"));
  b->pos = posspec;
  b->ag = ag;
  b->band = band;
  b->scratch = nullptr;
  return MTD_OK;
}

int mtd_born_plan(mtd_ctx* ctx, const mtd_geom* g, const int32_t* posspec, int A) {
  CtxScope scope(ctx);
  ctx->plan_valid = false;
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (A <= 0) return set_err(ctx, MTD_ERR_ARG, "a plan needs at least one angle");
  if (A > 65535) return set_err(ctx, MTD_ERR_UNSUPPORTED, "more than 65535 angles per volume");
  if (!posspec) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int V = g->dim_final;
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  MTD_TRY(born_weights(ctx, g, &t));
  int32_t* pos = (int32_t*)ws_get(ctx, WS_PLANPOS, (size_t)2 * A * sizeof(int32_t));
  AngleGeo* ag = (AngleGeo*)ws_get(ctx, WS_PLANAGEO, (size_t)A * sizeof(AngleGeo));
  int2* band = (int2*)ws_get(ctx, WS_PLANBAND, sizeof(int2));
  void* scratch = ws_get(ctx, WS_PLANGATHER, gather_scratch_bytes(V, A));
  if (!pos || !ag || !band || !scratch)
    return set_err(ctx, MTD_ERR_HIP, "out of device memory (geometry plan)");
  MTD_TRY(hip_check(ctx, hipMemcpyAsync(pos, posspec, (size_t)2 * A * sizeof(int32_t),
                                        hipMemcpyDeviceToDevice, ctx->stream),
                    "plan speculars"));
  MTD_TRY(hip_check(ctx, launch_angle_geo(pos, pos + A, A, g->nxmax, t.R2, t.htab, t.r2tab, ag,
                                          ctx->stream), "angle geometry"));
  MTD_TRY(hip_check(ctx, launch_band(ag, A, V, t.htab, t.r2max, band, ctx->stream), "hit band"));
  MTD_TRY(hip_check(ctx, launch_gather_prepass(ag, A, 2 * g->nxmax, g->nxmax, V, t.htab,
                                               t.r2max, scratch, ctx->stream),
                    "gather prepass"));
  ctx->plan_g = *g;
  ctx->plan_A = A;
  ctx->plan_valid = true;
  return MTD_OK;
}

int mtd_forward_spectra(mtd_ctx* ctx, const mtd_geom* g, const float* dn_re, const float* dn_im,
                        const int32_t* posspec, int A, float* spectra) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (A == 0) return MTD_OK;
  if (!dn_re || !spectra) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int V = g->dim_final, C = 2 * g->nxmax;
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  MTD_TRY(born_weights(ctx, g, &t));
  GeoBufs geo;
  bool planned;
  MTD_TRY(call_geometry(ctx, g, posspec, A, t, true, &geo, &planned));
  float2* work = (float2*)ws_get(ctx, WS_VWORK, (size_t)V * V * V * sizeof(float2));
  FftTab twV;
  const bool tabs = fft_tab(ctx, V, vol_len_supported(V), &twV);
  if (!work || !tabs) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  const double tp = tomo_pixel(g);
  const float gsign = ((V / 2) & 1) ? -1.f : 1.f;
  MTD_TRY(hip_check(ctx, launch_contrast_spectrum(V, dn_re, dn_im, work,
                                                  (float)(tp * tp * tp) * gsign, geo.band, twV,
                                                  ctx->stream),
                    "contrast spectrum"));
  return hip_check(ctx, launch_cap_sample(work, geo.ag, A, t.htab, t.q2, t.r2max, g->nxmax, C, V,
                                          (float2*)spectra, ctx->stream),
                   "cap sampling");
}

// pre: the spectra already carry the forward factor q of their bin (the
// solver's residual applies it), so the gather adds no per-hit factor
static int adjoint_impl(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                        const int32_t* posspec, int A, float* out_re, float* out_im, bool pre);

int mtd_adjoint_spectra(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                        const int32_t* posspec, int A, float* out_re, float* out_im) {
  CtxScope scope(ctx);
  return adjoint_impl(ctx, g, spectra, posspec, A, out_re, out_im, false);
}

static int adjoint_impl(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                        const int32_t* posspec, int A, float* out_re, float* out_im, bool pre) {
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (!out_re || (A > 0 && !spectra)) return set_err(ctx, MTD_ERR_ARG, "null array");
  if (A > 0 && !posspec && !ctx->plan_valid)
    return set_err(ctx, MTD_ERR_ARG, "null array (posspec, and no geometry plan)");
  const int V = g->dim_final;
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  MTD_TRY(born_weights(ctx, g, &t));
  float2* vol = (float2*)ws_get(ctx, WS_VOL, (size_t)V * V * V * sizeof(float2));
  if (!vol) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  // sum over angles of i q S on each cap voxel (conj of the forward weight),
  // centring-signed, then the unnormalised backward 3-D FFT times Tp^3
  int2* band;
  if (A > 0 && !posspec) {
    GeoBufs geo;
    MTD_TRY(plan_bufs(ctx, g, A, &geo));
    band = geo.band;
    MTD_TRY(scatter_impl(ctx, g, (const float2*)spectra, 1, nullptr, A, (float*)vol, nullptr,
                         pre ? 4 | 2 : 4, band, t.q, &geo));
  } else {
    band = (int2*)ws_get(ctx, WS_BAND, sizeof(int2));
    if (!band) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
    MTD_TRY(scatter_impl(ctx, g, (const float2*)spectra, 1, posspec, A, (float*)vol, nullptr,
                         pre ? 4 | 2 : 4, band, t.q));
  }
  const double tp = tomo_pixel(g);
  return finalize_impl(ctx, g, (const float*)vol, nullptr, 1, band, out_re, out_im, nullptr,
                       tp * tp * tp);
}

int mtd_cap_weights(mtd_ctx* ctx, const mtd_geom* g, const int32_t* posspec, int A,
                    float* weights) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (A == 0) return MTD_OK;
  if (!weights) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int V = g->dim_final, C = 2 * g->nxmax;
  EwaldTables t;
  MTD_TRY(ewald_tables(ctx, g, &t));
  MTD_TRY(born_weights(ctx, g, &t));
  GeoBufs geo;
  bool planned = posspec == nullptr;
  if (planned) MTD_TRY(plan_bufs(ctx, g, A, &geo));
  // hit counts from the accumulation itself (their values need no spectra:
  // the gather reads the stage buffer and the sums are discarded)
  float2* spec = (float2*)ws_get(ctx, WS_F, (size_t)A * C * C * sizeof(float2));
  float2* vol = (float2*)ws_get(ctx, WS_VOL, (size_t)V * V * V * sizeof(float2));
  int32_t* cnt = (int32_t*)ws_get(ctx, WS_CNT, (size_t)V * V * V * sizeof(int32_t));
  if (!spec || !vol || !cnt) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  MTD_TRY(scatter_impl(ctx, g, spec, 0, posspec, A, (float*)vol, cnt, 0, nullptr, nullptr,
                       planned ? &geo : nullptr));
  const AngleGeo* ag =
      planned ? geo.ag : (const AngleGeo*)ws_get(ctx, WS_AGEO, (size_t)A * sizeof(AngleGeo));
  return hip_check(ctx, launch_cap_weights(cnt, ag, A, t.htab, t.q, t.r2max, g->nxmax, C, V,
                                           weights, ctx->stream),
                   "cap weights");
}

int mtd_spectra_to_fields(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                          const int32_t* posspec, int A, float* fields) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, 0));
  if (A == 0) return MTD_OK;
  if (!spectra || !posspec || !fields) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax;
  FftTab twC;
  const bool tabs = fft_tab(ctx, C, crop_len_supported(C), &twC);
  const int chunk = chunk_of((size_t)C * C * sizeof(float2), kFieldBudget, A);
  float2* work = (float2*)ws_get(ctx, WS_WORK, (size_t)chunk * C * C * sizeof(float2));
  if (!tabs || !work) return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  // TF2Dcplx_INV's factor delta_f^2 (Simul_holo/src/FFT_fonctions.cpp:280-299)
  const float scale = (float)(g->delta_f * g->delta_f);
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(hip_check(ctx, launch_spectra_to_fields(C, (const float2*)spectra + (size_t)a0 * C * C,
                                                    posspec + a0, posspec + A + a0, n, twC, work,
                                                    scale, (float2*)fields + (size_t)a0 * C * C,
                                                    ctx->stream),
                      "field synthesis"));
  }
  return MTD_OK;
}

int mtd_tv_fista_step(mtd_ctx* ctx, int V, const float* z, const float* grad_data, float* x,
                      float* z_next, float data_scale, float lam, float eps, float step,
                      float beta, int positivity) {
  CtxScope scope(ctx);
  if (!ctx) return MTD_ERR_ARG;
  if (V <= 0 || V > 2048) return set_err(ctx, MTD_ERR_ARG, "volume side out of range");
  if (!z || !grad_data || !x || !z_next) return set_err(ctx, MTD_ERR_ARG, "null array");
  if (z == z_next) return set_err(ctx, MTD_ERR_ARG, "z and z_next must be distinct buffers");
  if (!(eps > 0.f) || !(step > 0.f)) return set_err(ctx, MTD_ERR_ARG, "eps and step must be > 0");
  return hip_check(ctx, launch_tv_fista(V, z, grad_data, x, z_next, data_scale, lam, eps, step,
                                        beta, positivity, ctx->stream),
                   "TV step");
}

// The solver loop of the adjoint-iterative form: n_iter iterations of
// forward, weighted residual, adjoint (real part) and the fused TV/FISTA
// step, on the geometry plan. The momentum coefficients are the host's
// t-sequence, so a captured loop bakes them in as kernel arguments.
static int tv_loop(mtd_ctx* ctx, const mtd_geom* g, const float* y, const float* w, int A,
                   int n_iter, float lam, float eps, float inv_kappa, int positivity, float* x,
                   float* z, float* zn, float* gd, float* S) {
  const int V = g->dim_final, C = 2 * g->nxmax;
  const size_t nv = (size_t)V * V * V;
  const float step = (float)(1.0 / (1.0 + 12.0 * (double)lam / (double)eps));
  MTD_TRY(hip_check(ctx, launch_tv_start(x, z, nv, positivity, ctx->stream), "solver start"));
  double t = 1.0;
  EwaldTables et;
  MTD_TRY(ewald_tables(ctx, g, &et));
  MTD_TRY(born_weights(ctx, g, &et));
  for (int k = 1; k <= n_iter; ++k) {
    MTD_TRY(mtd_forward_spectra(ctx, g, z, nullptr, nullptr, A, S));
    // W (S - y) times q: the adjoint's gather then takes the spectra as is
    MTD_TRY(hip_check(ctx, launch_residual((float2*)S, (const float2*)y, w, et.q2, A, C * C,
                                           ctx->stream),
                      "residual"));
    MTD_TRY(adjoint_impl(ctx, g, S, nullptr, A, gd, nullptr, true));
    const double tn = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
    MTD_TRY(mtd_tv_fista_step(ctx, V, z, gd, x, zn, inv_kappa, lam, eps, step,
                              (float)((t - 1.0) / tn), positivity));
    std::swap(z, zn);
    t = tn;
  }
  return MTD_OK;
}

int mtd_tv_solve(mtd_ctx* ctx, const mtd_geom* g, const float* spectra, const float* weights,
                 int A, int n_iter, float lam, float eps, float inv_kappa, int positivity,
                 float* x, int flags) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_VOL));
  if (!x || (A > 0 && !spectra)) return set_err(ctx, MTD_ERR_ARG, "null array");
  if (n_iter < 0) return set_err(ctx, MTD_ERR_ARG, "negative iteration count");
  if (!(eps > 0.f)) return set_err(ctx, MTD_ERR_ARG, "eps must be > 0");
  if (A > 0 && !(ctx->plan_valid && ctx->plan_A == A &&
                 std::memcmp(&ctx->plan_g, g, sizeof *g) == 0))
    return set_err(ctx, MTD_ERR_ARG, "no geometry plan for this geometry and angle count "
                                     "(mtd_born_plan first)");
  const int V = g->dim_final, C = 2 * g->nxmax;
  const size_t nv = (size_t)V * V * V;
  float* z = (float*)ws_get(ctx, WS_TVZ, nv * sizeof(float));
  float* zn = (float*)ws_get(ctx, WS_TVZN, nv * sizeof(float));
  float* gd = (float*)ws_get(ctx, WS_TVGD, nv * sizeof(float));
  float* S = (float*)ws_get(ctx, WS_TVS, (size_t)std::max(A, 1) * C * C * sizeof(float2));
  if (!z || !zn || !gd || !S) return set_err(ctx, MTD_ERR_HIP, "out of device memory (solver)");
  if (!(flags & MTD_TV_GRAPH) || n_iter == 0)
    return tv_loop(ctx, g, spectra, weights, A, n_iter, lam, eps, inv_kappa, positivity, x, z,
                   zn, gd, S);
  // one forward and one adjoint on scratch first: every table and workspace
  // the loop uses exists before the capture (nothing allocates inside it)
  MTD_TRY(mtd_forward_spectra(ctx, g, x, nullptr, nullptr, A, S));
  MTD_TRY(adjoint_impl(ctx, g, S, nullptr, A, gd, nullptr, true));
  MTD_TRY(hip_check(ctx, hipStreamSynchronize(ctx->stream), "solver sync"));
  if (ctx->tv_exec) (void)hipGraphExecDestroy(ctx->tv_exec);
  if (ctx->tv_graph) (void)hipGraphDestroy(ctx->tv_graph);
  ctx->tv_exec = nullptr;
  ctx->tv_graph = nullptr;
  if (!ctx->graph_stream)
    MTD_TRY(hip_check(ctx, hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking),
                      "graph stream"));
  MTD_TRY(hip_check(ctx, hipStreamBeginCapture(ctx->graph_stream, hipStreamCaptureModeThreadLocal),
                    "solver capture"));
  const hipStream_t saved = ctx->stream;
  const bool saved_timing = ctx->timing;
  ctx->stream = ctx->graph_stream;
  ctx->timing = false;
  const int rc = tv_loop(ctx, g, spectra, weights, A, n_iter, lam, eps, inv_kappa, positivity, x,
                         z, zn, gd, S);
  ctx->stream = saved;
  ctx->timing = saved_timing;
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->graph_stream, &gr);
  if (rc != MTD_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  MTD_TRY(hip_check(ctx, ec, "solver capture"));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(gr);
    return hip_check(ctx, ei, "solver graph instantiate");
  }
  ctx->tv_graph = gr;
  ctx->tv_exec = ex;
  return hip_check(ctx, hipGraphLaunch(ex, ctx->stream), "solver graph launch");
}

int mtd_reconstruct_plain(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int A,
                          const uint8_t* mask, int32_t* posspec, float* index_re,
                          float* index_im) {
  CtxScope scope(ctx);
  MTD_TRY(validate(ctx, g, A, NEED_HOLO | NEED_VOL));
  if (!index_re || !index_im || (A > 0 && !holos)) return set_err(ctx, MTD_ERR_ARG, "null array");
  const int C = 2 * g->nxmax, V = g->dim_final;
  if (g->c_aber) MTD_TRY(lsq_prepare(ctx, C, mask));
  int32_t* pos = posspec;
  if (!pos) {
    pos = (int32_t*)ws_get(ctx, WS_POS, (size_t)std::max(2 * A, 1) * sizeof(int32_t));
    if (!pos) return set_err(ctx, MTD_ERR_HIP, "out of device memory (posspec)");
  }
  float2* F = (float2*)ws_get(ctx, WS_F, (size_t)std::max(A, 1) * C * C * sizeof(float2));
  const int chunk = chunk_of(field_bytes(g), kFieldBudget, std::max(A, 1));
  const int S = fused_argmax_parts(g);
  float2* crop = (float2*)ws_get(ctx, WS_CROP, (size_t)chunk * C * C * sizeof(float2));
  float2* part =
      (float2*)ws_get(ctx, WS_HOLOPART, (size_t)chunk * std::max(S, 1) * sizeof(float2));
  float2* vol = (float2*)ws_get(ctx, WS_VOL, (size_t)V * V * V * sizeof(float2));
  int2* band = (int2*)ws_get(ctx, WS_BAND, sizeof(int2));
  if (!F || !crop || !part || !vol || !band)
    return set_err(ctx, MTD_ERR_HIP, "out of device memory");
  const float2* qp = S > 0 ? fused_qtw(ctx, g) : nullptr;
  for (int a0 = 0; a0 < A; a0 += chunk) {
    const int n = std::min(chunk, A - a0);
    MTD_TRY(holo_range(ctx, g, holos, dtype, a0, n, crop, S > 0 ? part : nullptr, S, qp));
    if (qp)
      MTD_TRY(front_fused(ctx, g, crop, n, pos + a0, pos + A + a0, F + (size_t)a0 * C * C, part,
                          S));
    else
      MTD_TRY(front_chunk(ctx, g, crop, n, a0, pos + a0, pos + A + a0, nullptr,
                          F + (size_t)a0 * C * C, part, S));
  }
  // The gather writes the count-normalised, centring-signed spectrum, and
  // only in the height band that holds hits; the inverse transform treats the
  // rest as the zeros it is and never reads it.
  MTD_TRY(scatter_impl(ctx, g, F, 0, pos, A, (float*)vol, nullptr, 1 | 2, band));
  return finalize_impl(ctx, g, (const float*)vol, nullptr, 1, band, index_re, index_im, nullptr);
}

// ---------------------------------------------------------------------------
// Batched form: independent acquisitions of their own shapes in one graph
// ---------------------------------------------------------------------------
static void key_put(std::vector<uint8_t>& k, const void* p, size_t n) {
  k.insert(k.end(), (const uint8_t*)p, (const uint8_t*)p + n);
}

static uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
  return h;
}

// the context's schedule parameters, read or applied as one record
static TunedSched sched_get(const mtd_ctx* ctx) {
  return TunedSched{ctx->holo_chunk_override, ctx->gather_zw_override, ctx->gather_kb_override,
                    ctx->vol_rows_tile, ctx->vol_cols_tile, 0.f, false};
}
static void sched_set(mtd_ctx* ctx, const TunedSched& t) {
  ctx->holo_chunk_override = t.chunk;
  ctx->gather_zw_override = t.zw;
  ctx->gather_kb_override = t.kb;
  ctx->vol_rows_tile = t.rows_tile;
  ctx->vol_cols_tile = t.cols_tile;
}

static TuneKey tune_key(const mtd_geom* g, int dtype) {
  return std::make_tuple(g->n, 2 * g->nxmax, g->dim_final, dtype);
}

// ---------------------------------------------------------------------------
// Schedule cache file. One text line per tuned shape:
//   mtd-tune <library version> | <device arch> | N C V dtype chunk zw kb rows cols ms
// Lines of another library version or device are skipped, the last line of
// a shape wins; new entries are appended one line per write.
// ---------------------------------------------------------------------------
static std::string device_arch(int device) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
  std::string a = p.gcnArchName;
  const size_t c = a.find(':');  // drop feature flags (":sramecc+:xnack-")
  return c == std::string::npos ? a : a.substr(0, c);
}

static std::string tune_prefix(int device) {
  return std::string("mtd-tune ") + mtd_version() + " | " + device_arch(device) + " | ";
}

static int tune_cache_load(mtd_ctx* ctx) {
  std::FILE* f = std::fopen(ctx->tune_cache.c_str(), "r");
  if (!f) return MTD_OK;  // no file yet: nothing tuned
  const std::string pre = tune_prefix(ctx->device);
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    if (std::strncmp(line, pre.c_str(), pre.size()) != 0) continue;
    int n, C, V, dt, ch, zw, kb, rt, ct;
    float ms;
    if (std::sscanf(line + pre.size(), "%d %d %d %d %d %d %d %d %d %f", &n, &C, &V, &dt, &ch,
                    &zw, &kb, &rt, &ct, &ms) != 10)
      continue;
    const TunedSched t{ch, zw, kb, rt, ct, ms, true};
    // an entry naming a variant this library does not build is ignored
    if (ch < 0 || (zw && !gather_window_supported(zw)) || (kb && kb != 8 && kb != 16) ||
        rt < 0 || rt > 1 || ct < 0 || ct > 2)
      continue;
    ctx->batch_tuned[std::make_tuple(n, C, V, dt)] = t;
  }
  std::fclose(f);
  return MTD_OK;
}

static void tune_cache_append(mtd_ctx* ctx, const TuneKey& k, const TunedSched& t) {
  if (ctx->tune_cache.empty()) return;
  std::FILE* f = std::fopen(ctx->tune_cache.c_str(), "a");
  if (!f) return;  // a read-only cache still serves what it holds
  char line[512];
  std::snprintf(line, sizeof line, "%s%d %d %d %d %d %d %d %d %d %.4f\n",
                tune_prefix(ctx->device).c_str(), std::get<0>(k), std::get<1>(k),
                std::get<2>(k), std::get<3>(k), t.chunk, t.zw, t.kb, t.rows_tile, t.cols_tile,
                t.ms);
  std::fputs(line, f);
  std::fclose(f);
}

// one acquisition with the schedule picked for its shape
static int batch_one(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int A,
                     const uint8_t* mask, int32_t* pos, float* re, float* im) {
  const auto it = ctx->batch_tuned.find(tune_key(g, dtype));
  const TunedSched user = sched_get(ctx);
  if (it != ctx->batch_tuned.end()) sched_set(ctx, it->second);
  const int rc = mtd_reconstruct_plain(ctx, g, holos, dtype, A, mask, pos, re, im);
  sched_set(ctx, user);
  return rc;
}

// Picks the plain reconstruction's schedule for one (N, C, V, dtype) by
// timing candidates on that acquisition, one axis at a time, each axis
// starting from the best so far: the hologram chunk (holograms per
// row/column launch pair: the launch count against the tail of each
// launch), the gather's height window (8, 16, 24: LDS per block, and so
// blocks per CU, against the entries each window rescans) and its load
// batch (8, 16 list entries in flight), then the volume passes' tile
// variants (row pass: 1 or 1/2 the lines per block; column passes: 1, 1/2
// or 2x the columns per block). Every candidate gives the same results. A
// candidate runs once untimed and twice timed (the faster counts). The pick
// is appended to the cache file when the context has one.
static int tune_schedule(mtd_ctx* ctx, const mtd_geom* g, const void* holos, int dtype, int A,
                         const uint8_t* mask, int32_t* pos, float* re, float* im) {
  const TuneKey key = tune_key(g, dtype);
  if (ctx->batch_tuned.count(key) || A <= 0) return MTD_OK;
  // chunk candidates: the default (0: the device-stack budget, even
  // chunks), smaller chunks, and the whole stack up to twice that budget
  const int cap = (int)std::max<size_t>(
      1, (size_t)2 * kChunkBudgetDevice / ((size_t)g->n * 2 * g->nxmax * sizeof(float2)));
  std::vector<int> chunks{0};
  for (int c : {32, 128})
    if (c < A) chunks.push_back(c);
  if (std::min(A, cap) > 128) chunks.push_back(std::min(A, cap));
  hipEvent_t e0, e1;
  MTD_TRY(hip_check(ctx, hipEventCreate(&e0), "tuner event"));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_err(ctx, MTD_ERR_HIP, "tuner event");
  }
  auto time_one = [&](float* ms) -> int {
    MTD_TRY(mtd_reconstruct_plain(ctx, g, holos, dtype, A, mask, pos, re, im));
    *ms = 0.f;
    for (int r = 0; r < 2; ++r) {
      MTD_TRY(hip_check(ctx, hipEventRecord(e0, ctx->stream), "tuner"));
      MTD_TRY(mtd_reconstruct_plain(ctx, g, holos, dtype, A, mask, pos, re, im));
      MTD_TRY(hip_check(ctx, hipEventRecord(e1, ctx->stream), "tuner"));
      MTD_TRY(hip_check(ctx, hipEventSynchronize(e1), "tuner"));
      float t = 0.f;
      MTD_TRY(hip_check(ctx, hipEventElapsedTime(&t, e0, e1), "tuner"));
      if (r == 0 || t < *ms) *ms = t;
    }
    return MTD_OK;
  };
  const TunedSched user = sched_get(ctx);
  TunedSched best{0, 0, 0, 0, 0, 0.f, false};
  int rc = MTD_OK;
  // one axis: every candidate value v of field f, keeping the fastest
  auto axis = [&](int TunedSched::*f, const std::vector<int>& vals) {
    for (int v : vals) {
      if (rc != MTD_OK) return;
      if (best.ms > 0.f && best.*f == v) continue;  // already timed
      TunedSched t = best;
      t.*f = v;
      sched_set(ctx, t);
      float ms = 0.f;
      if ((rc = time_one(&ms)) != MTD_OK) return;
      if (best.ms <= 0.f || ms < best.ms) {
        best = t;
        best.ms = ms;
      }
    }
  };
  axis(&TunedSched::chunk, chunks);
  axis(&TunedSched::zw, {0, 8, 24});  // 0: the default (16)
  axis(&TunedSched::kb, {0, 16});     // 0: the default (8)
  axis(&TunedSched::rows_tile, {0, 1});
  axis(&TunedSched::cols_tile, {0, 1, 2});
  sched_set(ctx, user);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == MTD_OK) {
    // record the defaults the zero entries stand for (same launches)
    if (best.chunk == 0)
      best.chunk = default_holo_chunk(g->n, 2 * g->nxmax, A, is_host_ptr(holos));
    if (best.zw == 0) best.zw = 16;
    if (best.kb == 0) best.kb = 8;
    ctx->batch_tuned[key] = best;
    tune_cache_append(ctx, key, best);
  }
  return rc;
}

// per-acquisition times: wall-clock stamps written by a one-thread kernel
// before the first and after every acquisition (WS_BATCHTS, n + 1 stamps)
static unsigned long long* batch_stamps(mtd_ctx* ctx, int n) {
  return (unsigned long long*)ws_get(ctx, WS_BATCHTS, (size_t)(n + 1) * 8);
}

static int batch_stamp(mtd_ctx* ctx, int i) {
  return hip_check(ctx, launch_stamp(batch_stamps(ctx, i) + i, ctx->stream), "batch stamp");
}

static int batch_times(mtd_ctx* ctx, int n, float* acq_ms) {
  std::vector<unsigned long long> t(n + 1);
  MTD_TRY(hip_check(ctx, hipStreamSynchronize(ctx->stream), "batch wait"));
  MTD_TRY(hip_check(ctx, hipMemcpy(t.data(), batch_stamps(ctx, n), t.size() * 8,
                                   hipMemcpyDeviceToHost), "batch stamps"));
  int khz = 0;
  MTD_TRY(hip_check(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                                               ctx->device), "wall clock rate"));
  for (int i = 0; i < n; ++i) acq_ms[i] = (float)((double)(t[i + 1] - t[i]) / khz);
  return MTD_OK;
}

int mtd_reconstruct_batch(mtd_ctx* ctx, int n, const mtd_geom* geoms, const void* const* holos,
                          const int* dtypes, const int* counts, const uint8_t* const* masks,
                          int32_t* const* posspec, float* const* index_re,
                          float* const* index_im, int flags, float* acq_ms) {
  CtxScope scope(ctx);
  if (!ctx) return MTD_ERR_ARG;
  if (n < 0 || (n > 0 && (!geoms || !holos || !dtypes || !counts || !index_re || !index_im)))
    return set_err(ctx, MTD_ERR_ARG, "null batch array");
  if (n == 0) return MTD_OK;
  const bool graph = flags & MTD_BATCH_GRAPH;
  // the arguments a captured graph holds: geometry bytes, every pointer,
  // counts, dtypes and mask contents (schedules are tuned before a capture,
  // and every schedule gives the same results)
  std::vector<uint8_t> key;
  key_put(key, &n, sizeof n);
  for (int i = 0; i < n; ++i) {
    const mtd_geom& g = geoms[i];
    MTD_TRY(validate(ctx, &g, counts[i], NEED_HOLO | NEED_VOL));
    if (counts[i] > 0 && !holos[i]) return set_err(ctx, MTD_ERR_ARG, "null hologram stack");
    if (!index_re[i] || !index_im[i]) return set_err(ctx, MTD_ERR_ARG, "null output volume");
    if (graph && counts[i] > 0 && is_host_ptr(holos[i]))
      return set_err(ctx, MTD_ERR_ARG,
                     "batched graph form: hologram stacks must be in device memory");
    key_put(key, &g, sizeof g);
    key_put(key, &holos[i], sizeof(void*));
    key_put(key, &dtypes[i], sizeof(int));
    key_put(key, &counts[i], sizeof(int));
    const uint8_t* m = masks ? masks[i] : nullptr;
    const uint64_t mh = m ? fnv1a(m, (size_t)4 * g.nxmax * g.nxmax) : 0;
    key_put(key, &mh, sizeof mh);
    int32_t* p = posspec ? posspec[i] : nullptr;
    key_put(key, &p, sizeof p);
    key_put(key, &index_re[i], sizeof(float*));
    key_put(key, &index_im[i], sizeof(float*));
  }
  // stamps sized before any capture (no allocation inside one)
  if (!batch_stamps(ctx, n)) return set_err(ctx, MTD_ERR_HIP, "out of device memory (stamps)");
  auto mask_of = [&](int i) { return masks ? masks[i] : nullptr; };
  auto pos_of = [&](int i) { return posspec ? posspec[i] : nullptr; };
  if (graph && ctx->batch_exec && key == ctx->batch_key && ctx->batch_gen == ctx->ws_gen) {
    MTD_TRY(hip_check(ctx, hipGraphLaunch(ctx->batch_exec, ctx->stream), "batch graph launch"));
    return acq_ms ? batch_times(ctx, n, acq_ms) : MTD_OK;
  }
  if (flags & MTD_BATCH_AUTOTUNE)
    for (int i = 0; i < n; ++i)
      MTD_TRY(tune_schedule(ctx, &geoms[i], holos[i], dtypes[i], counts[i], mask_of(i),
                            pos_of(i), index_re[i], index_im[i]));
  // eager pass: the results of this call, and every table, plan and
  // workspace the graph will use is in place before the capture
  MTD_TRY(batch_stamp(ctx, 0));
  for (int i = 0; i < n; ++i) {
    MTD_TRY(batch_one(ctx, &geoms[i], holos[i], dtypes[i], counts[i], mask_of(i), pos_of(i),
                      index_re[i], index_im[i]));
    MTD_TRY(batch_stamp(ctx, i + 1));
  }
  if (acq_ms) MTD_TRY(batch_times(ctx, n, acq_ms));
  if (!graph) return MTD_OK;
  // capture on a private stream (a graph cannot be captured from the legacy
  // default stream); launches then go to the context's stream. A replay of
  // the previous graph may still run on that stream: wait before freeing it.
  MTD_TRY(hip_check(ctx, hipStreamSynchronize(ctx->stream), "batch sync"));
  if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
  if (ctx->batch_graph) (void)hipGraphDestroy(ctx->batch_graph);
  ctx->batch_exec = nullptr;
  ctx->batch_graph = nullptr;
  ctx->batch_key.clear();
  if (!ctx->graph_stream)
    MTD_TRY(hip_check(ctx, hipStreamCreateWithFlags(&ctx->graph_stream, hipStreamNonBlocking),
                      "graph stream"));
  MTD_TRY(hip_check(ctx, hipStreamBeginCapture(ctx->graph_stream, hipStreamCaptureModeThreadLocal),
                    "batch capture"));
  const hipStream_t saved = ctx->stream;
  const bool saved_timing = ctx->timing;
  ctx->stream = ctx->graph_stream;
  ctx->timing = false;
  int rc = batch_stamp(ctx, 0);
  for (int i = 0; i < n && rc == MTD_OK; ++i) {
    rc = batch_one(ctx, &geoms[i], holos[i], dtypes[i], counts[i], mask_of(i), pos_of(i),
                   index_re[i], index_im[i]);
    if (rc == MTD_OK) rc = batch_stamp(ctx, i + 1);
  }
  ctx->stream = saved;
  ctx->timing = saved_timing;
  hipGraph_t gr = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->graph_stream, &gr);
  if (rc != MTD_OK) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  MTD_TRY(hip_check(ctx, ec, "batch capture"));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(gr);
    return hip_check(ctx, ei, "batch graph instantiate");
  }
  ctx->batch_graph = gr;
  ctx->batch_exec = ex;
  ctx->batch_key = std::move(key);
  ctx->batch_gen = ctx->ws_gen;
  return MTD_OK;
}

// the tuned schedule of the first volume side tuned for (n, C, dtype)
static const TunedSched* tuned_any_v(mtd_ctx* ctx, int n, int C, int dtype) {
  for (const auto& kv : ctx->batch_tuned)
    if (std::get<0>(kv.first) == n && std::get<1>(kv.first) == C &&
        std::get<3>(kv.first) == dtype)
      return &kv.second;
  return nullptr;
}

int mtd_batch_tuned_chunk(mtd_ctx* ctx, int n, int C, int dtype) {
  if (!ctx) return -1;
  const TunedSched* t = tuned_any_v(ctx, n, C, dtype);
  return t ? t->chunk : 0;
}

int mtd_batch_tuned_window(mtd_ctx* ctx, int n, int C, int dtype) {
  if (!ctx) return -1;
  const TunedSched* t = tuned_any_v(ctx, n, C, dtype);
  return t ? t->zw : 0;
}

int mtd_batch_tuned_schedule(mtd_ctx* ctx, int n, int C, int V, int dtype, int* sched) {
  if (!ctx || !sched) return MTD_ERR_ARG;
  const auto it = ctx->batch_tuned.find(std::make_tuple(n, C, V, dtype));
  if (it == ctx->batch_tuned.end()) {
    for (int i = 0; i < 6; ++i) sched[i] = 0;
    return 0;
  }
  const TunedSched& t = it->second;
  sched[0] = t.chunk;
  sched[1] = t.zw;
  sched[2] = t.kb;
  sched[3] = t.rows_tile;
  sched[4] = t.cols_tile;
  sched[5] = t.from_cache ? 1 : 0;
  return 1;
}

int mtd_ctx_set_tune_cache(mtd_ctx* ctx, const char* path) {
  if (!ctx) return MTD_ERR_ARG;
  ctx->tune_cache = path ? path : "";
  if (ctx->tune_cache.empty()) return MTD_OK;
  return tune_cache_load(ctx);
}

}  // extern "C"
