/* mtd.h -- C ABI of the MI355X-native off-axis preprocessing and Fourier
 * (Ewald-cap) reconstruction of madeba/MTD_transmission.
 *
 * The reference has no library layer: its hot path is two executables,
 * tomo_pretraitement (Tomo_prog/pretraitement_hors_axe_CPU/main.cpp) and
 * tomo_reconstruction (Tomo_prog/Reconstruction/main.cpp), whose internal
 * C++ functions take std::vector buffers. Each entry point below replaces one
 * of those internal functions or loops; the reference line it replaces is
 * cited next to it. The drop-in executables built from
 * mtd_transmission_amd/tools/ are thin mains over this ABI.
 *
 * Conventions
 *  - Every call returns 0 (MTD_OK) or an error code; mtd_last_error() gives
 *    the message of the last failure on that context (mtd_last_error(NULL):
 *    of the last failed mtd_ctx_create on the calling thread).
 *  - Array arguments are DEVICE pointers (hipMalloc / torch CUDA tensors)
 *    unless the comment says "host". The caller owns them; the context owns
 *    its scratch workspaces, twiddle tables and fit plans.
 *  - Hologram stacks (`holos` of mtd_holo_to_crop, mtd_preprocess,
 *    mtd_reconstruct_plain) may also be HOST memory: pinned (hipHostMalloc)
 *    or pageable. They are then copied chunk by chunk on a second stream of
 *    the context, each copy overlapped with the transform of the previous
 *    chunk (pageable memory is staged by the runtime and overlaps less).
 *  - Calls are asynchronous on the context's stream (mtd_ctx_set_stream);
 *    mtd_ctx_sync() waits. One context per host thread.
 *  - Complex arrays are interleaved float pairs (re, im) -- "c32".
 *  - posspec arrays hold 2*A int32: all x first, then all y (tab_posSpec.raw
 *    order, pretraitement_hors_axe_CPU/main.cpp:250-251), in crop pixels.
 *  - Images are row-major, x fastest (cpt = x + y*dim); volumes
 *    x + y*V + z*V*V.
 */
#ifndef MTD_H
#define MTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mtd_status {
  MTD_OK = 0,
  MTD_ERR_ARG = 1,          /* invalid argument or geometry            */
  MTD_ERR_UNSUPPORTED = 2,  /* size/mode beyond what is built         */
  MTD_ERR_HIP = 3,          /* HIP runtime error                       */
  MTD_ERR_IO = 4            /* file could not be read or written       */
};

enum mtd_dtype { MTD_U8 = 0, MTD_U16 = 1, MTD_F32 = 2 };

/* Phase unwrapping (pretraitement_hors_axe_CPU/main.cpp:260-277; only applied
 * with C_ABER 1): DEROUL 0 -> OFF; DEROUL 1 with VOLKOV 1 -> VOLKOV
 * (deroul_volkov4_total_sym_paire, src/deroulement_volkov4.cpp:14-92);
 * DEROUL 1 with VOLKOV 0 -> HERRAEZ (phaseUnwrapping_Mat,
 * deroulement_herraez.cpp:32-75: the reference's edge order and grouping
 * tree; the border reliabilities come from rand()'s stream, in which angle
 * a of a call's stack is the a-th image unwrapped, as in one run of the
 * program). */
enum mtd_deroul { MTD_DEROUL_OFF = 0, MTD_DEROUL_VOLKOV = 1, MTD_DEROUL_HERRAEZ = 2 };

typedef struct mtd_ctx mtd_ctx;

/* Acquisition geometry: the values both `manip` constructors derive
 * (pretraitement_hors_axe_CPU/manip.cpp:9-157,
 * Reconstruction/src/manip.cpp:10-163). Fill it with mtd_geom_derive(). */
typedef struct mtd_geom {
  int32_t n;           /* CCD_ROIX: hologram side N                        */
  int32_t nxmax;       /* NXMAX: crop side C = 2*NXMAX                     */
  int32_t dim_final;   /* DIM_FINAL: volume side V                         */
  int32_t circle_cx;   /* CIRCLE_CX: carrier column of the +1 order        */
  int32_t circle_cy;   /* CIRCLE_CY                                        */
  int32_t born;        /* BORN 1: Born field, 0: Rytov                     */
  int32_t c_aber;      /* C_ABER: polynomial aberration correction         */
  int32_t deroul;      /* enum mtd_deroul: DEROUL / VOLKOV of recon.txt     */
  double n0;           /* N0, float-rounded                                */
  double na;           /* NA, float-rounded                                */
  double lambda;       /* LAMBDA (m), float-rounded                        */
  double pix_holo;     /* tailleTheoPixelHolo = TPCAM/Gt                   */
  double rayon;        /* round(NXMAX*n0/NA)                               */
  double pix_uborn;    /* tailleTheoPixelUborn = pix_holo*N/(2*NXMAX)      */
  double delta_f;      /* Delta_fUborn = 1/(pix_uborn*2*NXMAX)             */
  double coef_na;      /* COEF_NA_OBJ_LIMIT (default 1)                    */
  double alpha_holo;   /* Tukey alpha of the hologram window (0.05)        */
  double alpha_field;  /* Tukey alpha of the field window (0.1)            */
  double alpha_crop;   /* ArrayFire variant (pretraitement_hors_axe_GPU/
                          main.cpp:74-79, src/GPU_fonctions.cpp:122): the crop
                          times (1+i) tukey2D(C, alpha_crop); 0 = off (the
                          CPU program, mtd_geom_derive's default)          */
} mtd_geom;

/* ---- context ---------------------------------------------------------- */
/* Replaces the FFTW_init / FFT_encaps plan objects
 * (pretraitement_hors_axe_CPU/src/FFTW_init.cpp, Reconstruction/src/FFT_encaps.cpp). */
int mtd_ctx_create(int device, mtd_ctx** out);
int mtd_ctx_destroy(mtd_ctx* ctx);
/* hip_stream: a hipStream_t (NULL = the legacy default stream). */
int mtd_ctx_set_stream(mtd_ctx* ctx, void* hip_stream);
int mtd_ctx_sync(mtd_ctx* ctx);
const char* mtd_last_error(const mtd_ctx* ctx);
const char* mtd_version(void);
/* Per-kernel timing with HIP events on the context's stream; replaces the
 * std::chrono prints at the reference's stage boundaries
 * (pretraitement_hors_axe_CPU/main.cpp:122-169,226-346;
 * Reconstruction/main.cpp:200-229,298-330). on=1 starts recording. */
int mtd_ctx_set_timing(mtd_ctx* ctx, int on);
/* Launch schedule of the plain path (no reference counterpart; results are
 * the same for every schedule): holo_chunk = holograms per hologram-stage
 * row/column launch pair (0: as many as fit 128 MiB of row intermediate),
 * gather_window = height of the Ewald gather's windows, 8, 16 or 24 voxels
 * (0: 16). The batched form's autotuner picks both per shape. */
int mtd_ctx_set_schedule(mtd_ctx* ctx, int holo_chunk, int gather_window);
/* Named schedule parameters of the plain path (no reference counterpart;
 * results are the same for every value): "holo_chunk" (as above),
 * "gather_window" (8, 16, 24), "gather_batch" (list entries whose spectrum
 * loads the gather issues together: 8, 16), "vol_rows_tile" (the volume
 * row pass with 1/2 the lines per block: 1), "vol_cols_tile" (the volume
 * column passes with 1/2 or 2x the columns per block: 1, 2). 0 restores the
 * default. An unknown name is MTD_ERR_ARG, a value not built
 * MTD_ERR_UNSUPPORTED. */
int mtd_ctx_set_schedule_param(mtd_ctx* ctx, const char* name, int value);
int mtd_ctx_get_schedule_param(mtd_ctx* ctx, const char* name, int* value);
/* Restricts the timing to the kernel of that report name (NULL or "": all
 * kernels). Each timed launch adds a pair of event markers to the stream,
 * which delays the next dispatch by a few microseconds; timing one kernel
 * keeps the step close to its untimed duration. No reference counterpart. */
int mtd_ctx_set_timing_only(mtd_ctx* ctx, const char* kernel);
/* Synchronises, then writes "name launches total_ms\n" lines (one per kernel,
 * totals since the last reset) into buf. Returns the number of lines. */
int mtd_ctx_timing_report(mtd_ctx* ctx, char* buf, size_t cap);
int mtd_ctx_timing_reset(mtd_ctx* ctx);

/* ---- host-only helpers (no GPU needed) -------------------------------- */
/* extract_val: last "KEY value" line, value through atof into a C float
 * (pretraitement_hors_axe_CPU/src/IO_fonctions.cpp:81-118; default form
 * Reconstruction/src/fonctions.cpp:137-180). has_default=0: missing -> 0.
 * Returns 1 if the key was found, 0 if not, -1 if the file is unreadable. */
int mtd_config_extract_val(const char* path, const char* token, int has_default,
                           double default_value, float* out);
/* extract_string (IO_fonctions.cpp:44-79). Copies at most cap-1 chars. */
int mtd_config_extract_string(const char* path, const char* token, char* out,
                              size_t cap);
/* Derived geometry from config values (manip.cpp:32-69,104;
 * Reconstruction/src/manip.cpp:52-94). Inputs are the raw config values;
 * they are rounded through float as extract_val does. */
int mtd_geom_derive(double ccd_roix, double nxmax, double dim_final,
                    double circle_cx, double circle_cy, double na, double n0,
                    double lambda, double tpcam, double f_tube, double f_obj,
                    double rf, double coef_na, int born, int c_aber, int deroul,
                    mtd_geom* out);
/* tukey2D 1-D factor (FFT_fonctions.cpp:61-87), host arrays of n doubles. */
int mtd_tukey1d(int n, double alpha, double* out);
/* Float32 multipage TIFF writer of indice.tif / absorption.tif / OTF3D.tif
 * (SAV3D_Tiff_Optimized, Reconstruction/src/fonctions.cpp:2286-2357: one
 * page per z, one strip per row, resolution 0.01/pixel_m per cm; orientation
 * 4 = ORIENTATION_BOTLEFT there, 1 for 2-D images). data: host float
 * [pages][h][w]. format 0: classic TIFF, or BigTIFF (64-bit offsets) when the
 * file would pass 4 GiB (a 1024^3 volume), where libtiff's classic mode
 * fails; 1: classic only (returns -2 and writes nothing past 4 GiB);
 * 2: BigTIFF. Returns 0 (classic written), 1 (BigTIFF written), < 0 error. */
int mtd_write_tiff_f32(const char* path, const float* data, int w, int h,
                       int pages, double pixel_m, int orientation, int format);

/* ---- preprocessing (tomo_pretraitement) -------------------------------- */
/* holo2TF_UBorn over a stack (fonctions.cpp:279-304, loop main.cpp:130-165):
 * holos[A][N][N] (dtype) -> crops[A][C][C] c32, the centred spectrum
 * /N^2 cropped at (CIRCLE_CX-NXMAX, CIRCLE_CY-NXMAX). */
int mtd_holo_to_crop(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                     int dtype, int A, float* crops);

/* Allocation-free variant of mtd_holo_to_crop. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same kernels and results; the twiddles, the window and the row
 * intermediate live in one static device array of 64 MiB per GPU (held until
 * the process ends), so one hologram's N x C row intermediate must fit 63 MiB
 * (any compiled hologram side does; sides without a compiled FFT are
 * refused) and holos must be device memory. It calls no hipMalloc/hipFree.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, threads or contexts) race. */
int mtd_holo_to_crop_noalloc(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                             int dtype, int A, float* crops);

/* Field formation from crops (main.cpp:230-342): coordSpec, calc_Uborn2,
 * atan2 phase, aberCorr2/ampliCorr2 polynomial fits (FP64), amplitude clamp,
 * Born/Rytov (or the C_ABER 0 branch). mask: host uint8 [C][C]
 * (Image_mask.pgm) or NULL (all 255, Correction_aberration.cpp:48-52).
 * Outputs posspec[2A] int32 and fields[A][C][C] c32 (UBornfinal order). */
int mtd_crop_to_field(mtd_ctx* ctx, const mtd_geom* g, const float* crops,
                      int A, const uint8_t* mask, int32_t* posspec,
                      float* fields);

/* calc_Uborn2 (pretraitement_hors_axe_CPU/fonctions.cpp:348-374): the
 * demodulated field U of each crop (circular shift by -posSpec, unnormalised
 * backward 2-D FFT, shift by (NXMAX, NXMAX)), u[A][C][C] complex64 in the
 * reference's frame; posspec as mtd_crop_to_field writes it (x then y). The
 * ArrayFire variant saves its phase atan2(U) (phasePI_atan2.raw,
 * pretraitement_hors_axe_GPU/main.cpp:212-214). Device arrays. */
int mtd_demodulate(mtd_ctx* ctx, const mtd_geom* g, const float* crops, int A,
                   const int32_t* posspec, float* u);

/* Both preprocessing loops (main.cpp:130-342) from holograms. */
int mtd_preprocess(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                   int dtype, int A, const uint8_t* mask, int32_t* posspec,
                   float* fields);

/* ---- reconstruction (tomo_reconstruction) ------------------------------ */
/* Per-angle spectrum (Reconstruction/main.cpp:207-217): U_B*tukey(C,0.1),
 * shift by NXMAX, forward FFT * pixUBorn^2, shift by +posSpec.
 * spectra[A][C][C] c32 in the frame retroPropag_Born reads. */
int mtd_field_to_spectrum(mtd_ctx* ctx, const mtd_geom* g, const float* fields,
                          const int32_t* posspec, int A, float* spectra);

/* retroPropag_Born over all angles in order (Reconstruction/src/fonctions.cpp:875-935,
 * loop main.cpp:203-225): deterministic Ewald-cap accumulation of spectra
 * (frame of mtd_field_to_spectrum) into vol[V^3] c32 and counts[V^3] int32.
 * Both outputs are fully overwritten. */
int mtd_ewald_scatter(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                      const int32_t* posspec, int A, float* vol,
                      int32_t* counts);

/* main.cpp:203-225 from fields: window + FFT + Ewald accumulation. */
int mtd_ewald_accumulate(mtd_ctx* ctx, const mtd_geom* g, const float* fields,
                         const int32_t* posspec, int A, float* vol,
                         int32_t* counts);

/* main.cpp:243-352: counts 0->1 and divide, fftshift3D, backward 3-D FFT
 * * delta_f^3, fftshift3D, * -1/(2 n0 kv^2). index_re/index_im: float V^3
 * (indice.tif / absorption.tif). otf: uint8 V^3 (Re != 0, OTF3D.tif) or NULL. */
int mtd_finalize(mtd_ctx* ctx, const mtd_geom* g, const float* vol,
                 const int32_t* counts, float* index_re, float* index_im,
                 uint8_t* otf);

/* The whole plain method, holograms to index volume, fused where stages
 * allow (both programs' compute: pretraitement main.cpp:130-342 and
 * Reconstruction/main.cpp:203-352). posspec may be NULL. */
int mtd_reconstruct_plain(mtd_ctx* ctx, const mtd_geom* g, const void* holos,
                          int dtype, int A, const uint8_t* mask,
                          int32_t* posspec, float* index_re, float* index_im);

/* ---- batched form ---------------------------------------------------------
 * No reference counterpart: the reference reconstructs one acquisition per
 * run of its two programs. n independent acquisitions, each with its own
 * geometry (hologram side, crop, volume), angle count, hologram stack
 * (device memory), host mask (or NULL), posspec (or NULL) and output
 * volumes, reconstructed by the plain method in order on the context's
 * stream; each result equals mtd_reconstruct_plain's on its own.
 * MTD_BATCH_GRAPH: the first call runs eagerly, then captures the whole batch
 * (every kernel of every acquisition) into one HIP graph; a later call with
 * the same arguments (geometries, pointers, counts, dtypes, mask contents)
 * replays it with a single launch, reading the stacks' current contents.
 * Other arguments recapture, and so does any call after which the context
 * moved a workspace or fit plan the graph uses (another entry point needed a
 * larger buffer). MTD_BATCH_AUTOTUNE: at the first sight of a
 * (hologram side, crop, dtype), the hologram-stage chunk (holograms per
 * row/column launch pair) and then the Ewald gather's height window (8, 16
 * or 24 voxels) are picked by timing candidates on that acquisition; graphs
 * then use the picked schedules (results are the same for every schedule). acq_ms (host, n floats, or
 * NULL): each acquisition's device time in this call, from wall-clock stamps
 * a one-thread kernel writes between acquisitions; the call then waits for
 * the batch to finish. */
enum mtd_batch_flags { MTD_BATCH_GRAPH = 1, MTD_BATCH_AUTOTUNE = 2 };
int mtd_reconstruct_batch(mtd_ctx* ctx, int n, const mtd_geom* geoms,
                          const void* const* holos, const int* dtypes,
                          const int* counts, const uint8_t* const* masks,
                          int32_t* const* posspec, float* const* index_re,
                          float* const* index_im, int flags, float* acq_ms);
/* The hologram chunk and the gather height window (8, 16 or 24 voxels) the
 * autotuner picked for (hologram side n, crop C, dtype), of the first volume
 * side tuned for them; 0 if that shape was not tuned. */
int mtd_batch_tuned_chunk(mtd_ctx* ctx, int n, int C, int dtype);
int mtd_batch_tuned_window(mtd_ctx* ctx, int n, int C, int dtype);
/* The whole schedule picked for (n, C, V, dtype), into sched[6]: hologram
 * chunk, gather window, gather load batch, volume row-pass tile variant
 * (0, 1), volume column-pass tile variant (0, 1, 2), and 1 if it was read
 * from the cache file (0 if tuned in this process). Returns 1 if the shape
 * is tuned, 0 if not (sched all 0). */
int mtd_batch_tuned_schedule(mtd_ctx* ctx, int n, int C, int V, int dtype, int* sched);
/* Schedule cache file of the autotuner (no reference counterpart): its
 * entries for this library version and device are loaded now, and every
 * shape tuned from now on is appended to it, keyed by (N, C, V, dtype),
 * library version and device. NULL or "": no file. A context starts with
 * the file named by the environment variable MTD_TUNE_CACHE, if set. A
 * missing file is not an error (nothing tuned yet). */
int mtd_ctx_set_tune_cache(mtd_ctx* ctx, const char* path);

/* ---- Born forward model (Simul_holo) and its adjoint -------------------- */
/* The simulator's hologram-spectrum synthesis (Simul_holo/main.cpp:106-109,
 * 162-171; calcHolo, Simul_holo/src/fonctions.cpp:109-140): TF3D =
 * fftshift3D(FFT3D(fftshift3D(dn))) * Tp^3 (TF3Dcplx,
 * Simul_holo/src/FFT_fonctions.cpp:301-318; Tp = 1/(V*delta_f)), then for
 * every angle and pupil pixel fd (|fd|^2 < COEF_NA_OBJ_LIMIT*NXMAX^2)
 * S[fd] = -i kv / (sqrt(R^2 - |fd|^2)/R) * TF3D[voxel], kv = 2*3.14159/LAMBDA,
 * at retroPropag_Born's voxel (the same integer map as calcHolo's
 * round(fd - fi)). dn_re/dn_im: index contrast n - n0, float V^3 (dn_im may
 * be NULL). spectra: [A][C][C] c32 in the frame of mtd_field_to_spectrum,
 * zero outside the pupil and for skipped angles. */
int mtd_forward_spectra(mtd_ctx* ctx, const mtd_geom* g, const float* dn_re,
                        const float* dn_im, const int32_t* posspec, int A,
                        float* spectra);
/* (posspec == NULL: the geometry plan of mtd_born_plan, below.) */

/* The simulator's per-angle field synthesis (Simul_holo/main.cpp:174-180:
 * decal2DCplxGen by -spec + dim/2, TF2Dcplx_INV x Delta_f_Uborn^2,
 * fftshift2D before SAVCplx): spectra[A][C][C] c32 in the frame of
 * mtd_forward_spectra (specular at posspec) -> fields[A][C][C] c32, the
 * UBornfinal layout. The inverse of mtd_field_to_spectrum without its
 * window. */
int mtd_spectra_to_fields(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                          const int32_t* posspec, int A, float* fields);

/* Geometry plan for iterative callers (no counterpart in the reference):
 * copies posspec[2A] and builds the per-angle cap geometry, the hit band and
 * the accumulation's per-tile angle lists once. mtd_forward_spectra,
 * mtd_adjoint_spectra and mtd_cap_weights called with posspec == NULL then
 * use it (same geometry and A, else MTD_ERR_ARG). One plan per context; the
 * next call replaces it. */
int mtd_born_plan(mtd_ctx* ctx, const mtd_geom* g, const int32_t* posspec, int A);

/* Adjoint of mtd_forward_spectra (no counterpart in the reference; the
 * gradient of the adjoint-iterative reconstruction): out = A^H spectra, the
 * sum over angles of i q S on each cap voxel (ascending angle order, as
 * retroPropag_Born) followed by the centred, unnormalised backward 3-D FFT
 * times Tp^3. out_re/out_im: float V^3; out_im may be NULL (the imaginary
 * part is then not stored, as the gradient of a real unknown needs only the
 * real part). */
int mtd_adjoint_spectra(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                        const int32_t* posspec, int A, float* out_re,
                        float* out_im);

/* Data weights of the adjoint-iterative form (no counterpart in the
 * reference): weights[A][C][C] float = 1 / (q^2 * hits of the sample's
 * voxel), q = kv / (sqrt(R^2 - |fd|^2)/R) the forward model's factor, on the
 * cap samples of mtd_forward_spectra (0 elsewhere). With them A^H W A is
 * Tp^6 times the projection of the centred 3-D spectrum on the hit voxels,
 * the normalisation the plain method applies by dividing by the counts. */
int mtd_cap_weights(mtd_ctx* ctx, const mtd_geom* g, const int32_t* posspec,
                    int A, float* weights);

/* One fused step of the adjoint-iterative solver (no counterpart in the
 * reference): float V^3 volumes, x' = max(z - step*(data_scale*grad_data +
 * lam*grad TV_eps(z)), 0) (no max when positivity == 0), z_next = x' +
 * beta*(x' - x), x <- x'. TV_eps is the isotropic total variation with
 * forward differences (zero across the far faces), smoothed by eps. z and
 * z_next must be distinct buffers. */
int mtd_tv_fista_step(mtd_ctx* ctx, int V, const float* z, const float* grad_data,
                      float* x, float* z_next, float data_scale, float lam,
                      float eps, float step, float beta, int positivity);

/* The whole adjoint-iterative solver loop in the library (no counterpart in
 * the reference): FISTA on 1/2 inv_kappa ||A x - spectra||_W^2 + lam
 * TV_eps(x), x >= 0 when positivity, n_iter iterations from x (clamped to
 * >= 0 first), on the geometry plan of mtd_born_plan (same geometry and A).
 * Each iteration: mtd_forward_spectra, the weighted residual W (S - y) in
 * place (weights: float [A][C][C] of mtd_cap_weights, or NULL for W = 1),
 * the real part of mtd_adjoint_spectra, and mtd_tv_fista_step with step
 * 1/(1 + 12 lam/eps) and the momentum t_{k+1} = (1 + sqrt(1 + 4 t_k^2))/2,
 * beta = (t_k - 1)/t_{k+1}, t_1 = 1. x: float V^3, in/out.
 * MTD_TV_GRAPH: the n_iter iterations are captured into one HIP graph (its
 * momentum coefficients baked in) and launched once. The iterates equal
 * the step-by-step calls bit for bit. */
enum mtd_tv_flags { MTD_TV_GRAPH = 1 };
int mtd_tv_solve(mtd_ctx* ctx, const mtd_geom* g, const float* spectra,
                 const float* weights, int A, int n_iter, float lam, float eps,
                 float inv_kappa, int positivity, float* x, int flags);

#ifdef __cplusplus
}
#endif
#endif /* MTD_H */
