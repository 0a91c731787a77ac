/*
 * gbm3d.h -- C ABI of libgbm3d.so, the MI355X (gfx950) implementation of the
 * GBM3D two-step denoiser (block matching -> global 3-D wavelet hard
 * threshold -> aggregation -> re-matching -> empirical Wiener -> aggregation).
 *
 * Conventions shared by every entry point:
 *   - Images are single precision, COLUMN-MAJOR (MATLAB layout): element
 *     (r, c) of an m x n image lives at r + c*m.
 *   - Block-match index sets S use the reference's encoding: a 1-based
 *     column-major index into the matchSize x matchSize search window,
 *     idx = a + matchSize*b + 1 with (a, b) the 0-based window offset; the
 *     block's image offset from its reference tile is (a - matchSize/2,
 *     b - matchSize/2).  S is stored K x Mt x Nt column-major (the cell array
 *     S{i,j} of uint32 K-vectors, flattened).
 *   - Every function returns 0 on success or a gbm3d_status code; the message
 *     of the last failure on the calling thread is gbm3d_last_error().  No
 *     function throws or exits across the ABI.
 *   - Host-buffer entry points copy in, compute on the context's stream and
 *     copy out (synchronous).  *_device entry points take device pointers and
 *     are asynchronous on the context's stream.
 */
#ifndef GBM3D_H
#define GBM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBM3D_MAX_SPINS 16
/* Field value meaning "absent" (MATLAB isfield(opts, f) == false). */
#define GBM3D_UNSET (-2147483647 - 1)

enum gbm3d_status {
  GBM3D_OK = 0,
  GBM3D_ERR_ARG = 1,           /* bad argument (size, NULL, range)             */
  GBM3D_ERR_UNSUPPORTED = 2,   /* option combination this build does not run   */
  GBM3D_ERR_HIP = 3,           /* HIP runtime failure                          */
  GBM3D_ERR_NOMEM = 4,         /* device allocation failed                     */
  GBM3D_ERR_MISSING_FIELD = 5  /* a required opts field is absent              */
};

enum gbm3d_filt { GBM3D_FILT_HT = 0, GBM3D_FILT_MEDIAN = 1 };

/*
 * Mirror of the reference's opts struct (utilities/setBM3Dopts.m:19-34,
 * utilities/checkBM3Dopts.m:10-24).  Integer fields set to GBM3D_UNSET and
 * spin lists with n* == GBM3D_UNSET are "absent".  profile[0] == 0 means no
 * profile field; a non-empty profile replaces the whole struct
 * (source/GBM3D.m:16-18).  wname/wnamez/order/orderz exist in the reference
 * but are never read on the GBM3D path (source/GBM3D.m:38-40), so they are
 * not carried.
 */
typedef struct gbm3d_opts {
  char profile[16];
  int32_t numMax, numMin;
  int32_t levels, cycleSpin;
  int32_t nMatchSpins;
  int32_t matchSpins[GBM3D_MAX_SPINS];
  int32_t nMatchSpinsWie;
  int32_t matchSpinsWie[GBM3D_MAX_SPINS];
  int32_t tauMode;
  int32_t blockSize, blockSizeWie, matchSize;
  int32_t Wiener;   /* 0 / 1 */
  int32_t filtType; /* enum gbm3d_filt */
} gbm3d_opts;

/* ---------------------------------------------------------------- options */

/* Every field absent (an empty MATLAB struct). */
int gbm3d_opts_clear(gbm3d_opts* opts);

/* opts = setBM3Dopts(profile)   -- utilities/setBM3Dopts.m:1-86.
 * Unknown profile names fall back to 'default' (:12-14). */
int gbm3d_set_opts(const char* profile, gbm3d_opts* opts);

/* opts = checkBM3Dopts(opts)    -- utilities/checkBM3Dopts.m:10-24.
 * Fills absent fields with the reference defaults in place. */
int gbm3d_check_opts(gbm3d_opts* opts);

/* The option handling at the top of GBM3D: opts == NULL behaves as
 * nargin<3 (profile 'default'); a profile replaces the struct; then
 * checkBM3Dopts.  source/GBM3D.m:15-19. */
int gbm3d_resolve_opts(const gbm3d_opts* in, gbm3d_opts* out);

/* tau = sigma*getTau(levels, mode): levels x 7, column-major
 * (tau[i + levels*j], i = level-1, j = subband-1). utilities/getTau.m:9-51. */
int gbm3d_get_tau(int levels, int mode, double sigma, double* tau);

/* Thread-local message of the last failure ("" when none). */
const char* gbm3d_last_error(void);

/* ---------------------------------------------------------------- context */

typedef struct gbm3d_ctx gbm3d_ctx;

/* One context per (device, stream).  Device workspace is owned by the
 * context, grows on demand and is released by gbm3d_destroy. */
int gbm3d_create(int device, gbm3d_ctx** ctx);
int gbm3d_destroy(gbm3d_ctx* ctx);
/* Use a caller-owned hipStream_t (NULL = the context's own stream). */
int gbm3d_set_stream(gbm3d_ctx* ctx, void* hip_stream);
int gbm3d_synchronize(gbm3d_ctx* ctx);

/* ------------------------------------------------------------- variants
 * Forms beside the reference path, set per context; a zeroed struct (or
 * NULL) is the reference behaviour.  No reference counterpart. */
typedef struct gbm3d_ext {
  int32_t match_bf16; /* 1: block distances on BF16 MFMA (FP32 accumulation),
                         then exact FP32 rescoring of every candidate the
                         error bounds cannot exclude: S equals the FP32
                         path's bit for bit (8 x 8 and 16 x 16 blocks; other
                         block sizes run the FP32 search)                    */
  int32_t rescore;    /* first rescored set per tile, numMax..64 (0: 2*numMax);
                         a launch-shape choice: results do not depend on it */
  int32_t learned;    /* 1: learned shrinkage replaces the step-1 wavelet hard
                         threshold (blockSize 8, numMax 16; weights from
                         gbm3d_shrink_set_weights)                            */
  int32_t learned_bound; /* fallback bound in 1/1000: per match spin, when
                         the learned gains remove more than bound x the noise
                         energy (coefficients x sigma^2) the reference filter
                         runs for that spin instead, decided on the device
                         (the call stays asynchronous; 0: 1000)               */
  int32_t match_stats; /* 1: count the BF16 matcher's paths for
                         gbm3d_bf16_match_stats (diagnostics: a few atomic
                         adds per reference tile)                             */
  int32_t reserved[3];
} gbm3d_ext;
int gbm3d_set_ext(gbm3d_ctx* ctx, const gbm3d_ext* ext);
int gbm3d_get_ext(gbm3d_ctx* ctx, gbm3d_ext* ext);
/* Path counters of the BF16 matcher since the last reset (synchronises the
 * context stream), out[8]: [0] tiles settled on the first rescored set, [1]
 * on a widened set, [2] by the exact search over every candidate, [3]
 * candidates rescored in total; why the exact search ran: [4] first set above
 * 128 candidates, [5] no finite K-th key in it, [6] widened set above 128,
 * [7] reference block of zero or non-finite norm.  reset != 0 clears them. */
int gbm3d_bf16_match_stats(gbm3d_ctx* ctx, unsigned long long* out, int reset);

/* --------------------------------------------------------------- hot path */

/*
 * The checks GBM3D's stages make on opts and the image size (the option rule
 * of source/GBM3D.m:15-19, the reference's own errors, and the block / match
 * / group sizes this library is built for), on the host only: no GPU is
 * touched.  workspace_bytes (nullable) receives the device workspace of one
 * m x n image with its match spins run one after another.  Every gbm3d_denoise*
 * call makes the same checks before it launches anything.  Returns
 * GBM3D_ERR_UNSUPPORTED for a combination the kernels are not built for.
 */
int gbm3d_validate(const gbm3d_opts* opts, int m, int n, size_t* workspace_bytes);

/*
 * [U, out] = GBM3D(I, sigma, opts)            -- source/GBM3D.m:1-84
 * I: m x n input (host), U: m x n output, U1 (nullable): m x n pilot
 * (out.U1, source/GBM3D.m:81).  opts may be NULL (profile 'default').
 */
int gbm3d_denoise(gbm3d_ctx* ctx, const float* I, int m, int n, float sigma,
                  const gbm3d_opts* opts, float* U, float* U1);

/* Same on device buffers, asynchronous on the context's stream. */
int gbm3d_denoise_device(gbm3d_ctx* ctx, const float* dI, int m, int n,
                         float sigma, const gbm3d_opts* opts, float* dU,
                         float* dU1);

/*
 * [U, out] = GBM3D_distributed(I, sigma, opts) -- source/GBM3D_distributed.m:1-122
 * Tiles of 256 + blockSize (utilities/getBM3D_distribution.m), each tile and
 * match spin denoised on its own, stitched with ramp windows
 * (utilities/makeInterpWindow.m).  U2 (nullable): out.U2, the first estimate
 * (out.U1 of GBM3D for images below 256 + blockSize, which run GBM3D).
 * The learned form is not supported here.
 */
int gbm3d_denoise_distributed(gbm3d_ctx* ctx, const float* I, int m, int n,
                              float sigma, const gbm3d_opts* opts, float* U,
                              float* U2);
int gbm3d_denoise_distributed_device(gbm3d_ctx* ctx, const float* dI, int m,
                                     int n, float sigma,
                                     const gbm3d_opts* opts, float* dU,
                                     float* dU2);

/*
 * Coloured noise: [U, out] = GBM3D(I, sigmaPSD, opts) with a matrix sigma --
 * source/GBM3D.m:29-36.  sigmaPSD: d1 x d2 (sigma^2 x the noise power
 * spectrum, column-major); thresholds from getColoredTau
 * (utilities/getColoredTau.m); sigmaWie: wm x wn per-coefficient Wiener
 * sigma of getColoredWie (utilities/getColoredWie.m), exactly blockSizeWie x
 * blockSizeWie when given (any other size is GBM3D_ERR_ARG); NULL computes it
 * with gbm3d_get_colored_wie, as source/GBM3D.m:35 does.  Host buffers.
 */
int gbm3d_denoise_colored(gbm3d_ctx* ctx, const float* I, int m, int n,
                          const double* sigmaPSD, int d1, int d2,
                          const double* sigmaWie, int wm, int wn,
                          const gbm3d_opts* opts, float* U, float* U1);
/* tau = getColoredTau(levels, sigmaPSD) (levels x 7, column-major). */
int gbm3d_get_colored_tau(int levels, const double* sigmaPSD, int d1, int d2,
                          double* tau);
/* sigmaWie = getColoredWie(sigmaPSD, d) -- utilities/getColoredWie.m:6-19
 * (d x d, column-major; d <= 64): the PSD smoothed by makeGausPSF's Gaussian,
 * resized to d x d as imresize(., [d d], 'bilinear'), then per DCT-II
 * coefficient (j, i) sqrt(sum(PSD_d .* |fft2(idct2(e_ji))|^2)) / d.
 * makeGausPSF is not part of the reference: the Gaussian is taken as the
 * circular one of standard deviation 3 px at the origin with unit sum. */
int gbm3d_get_colored_wie(const double* sigmaPSD, int d1, int d2, int d,
                          double* sigmaWie);

/*
 * gbm3d_denoise_device with no device allocation: no hipMalloc, hipFree or
 * allocating library call, for callers that cannot allocate (code under
 * stream capture, tight serving loops, harnesses that forbid it).  No
 * counterpart in the reference.  The workspace is a static device array of
 * GBM3D_NOALLOC_SCRATCH_BYTES per GPU, held from library load to process end
 * (gbm3d_validate reports what an image needs: 512 x 512 fits with every
 * profile); an image that needs more is GBM3D_ERR_ARG.  The array is shared
 * by every call on that GPU: calls that overlap on one GPU race, calls in
 * order on one stream do not.  Match spins run one after another on the
 * context stream; the learned form is GBM3D_ERR_UNSUPPORTED.  Results equal
 * gbm3d_denoise_device's bit for bit.
 */
#define GBM3D_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int gbm3d_denoise_device_noalloc(gbm3d_ctx* ctx, const float* dI, int m, int n,
                                 float sigma, const gbm3d_opts* opts,
                                 float* dU, float* dU1);

/* ------------------------------------------------------------- batches
 * [U_i, out_i] = GBM3D(I_i, sigma_i, opts) for `count` independent images
 * (device buffers, sizes m[i] x n[i]) sharing one opts.  Images run on
 * parallel streams (largest first onto the least-loaded stream), each stream
 * with its own workspace.  GBM3D_BATCH_GRAPH captures the whole batch into
 * one hipGraph on the first call and replays it while the batch (sizes,
 * sigmas, buffers, opts, launch shape) is unchanged.  GBM3D_BATCH_TUNE times
 * the launch-shape candidates (parallel streams 1/2/4/8 x reference tiles
 * per matching workgroup 1/2/4) once per batch and keeps the fastest.
 * Asynchronous on the context stream.  No reference counterpart: the nearest
 * reference form is the parfor over tiles of source/GBM3D_distributed.m:54. */
enum { GBM3D_BATCH_GRAPH = 1, GBM3D_BATCH_TUNE = 2 };
int gbm3d_denoise_batch_device(gbm3d_ctx* ctx, int count, const float* const* dI, const int* m, const int* n,
                               const float* sigma, const gbm3d_opts* opts, float* const* dU,
                               float* const* dU1, int flags);
/* Launch shape in use (after tuning) and the tuned batch time in ms. */
int gbm3d_batch_info(gbm3d_ctx* ctx, int* streams, int* match_waves, double* tuned_ms);

/* ------------------------------------------------- stage entry points
 * Each replaces one reference stage for stage-isolated use; m, n are the
 * (already block-multiple) image sizes the stage sees. */

/* [S, Volume] = matchBlocksCPU(I, sigma, opts)   -- source/matchBlocksCPU.m:1-116
 * (numMax == numMin branch).  d = blockSize, d2 = matchSize, K = numMax.
 * S: K x (m/d) x (n/d) uint32; Volume (nullable): m x n x K. */
int gbm3d_match(gbm3d_ctx* ctx, const float* I, int m, int n, int d, int d2,
                int K, uint32_t* S, float* Volume);

/* [S, Volume] = matchBlocksCPU(I, sigma, opts) with numMax ~= numMin: the
 * tolerance branch (source/matchBlocksCPU.m:21-28, 89-111), tol = 0.005
 * sigma^2 d^2.  Per tile, with `below` the window positions of distance <
 * tol: more than numMax -> the numMax nearest; exactly numMax -> those in
 * index order (find); then slot 1 := the window centre.  Fewer than numMax
 * (mink of numMin or find is shorter than numMax) is where the reference
 * fails (it assigns numMax blocks into numel(S) slices, :111): returns
 * GBM3D_ERR_ARG, S and Volume then holding the numMax nearest.  numMin ==
 * numMax is gbm3d_match.  Shapes as gbm3d_match. */
int gbm3d_match_tol(gbm3d_ctx* ctx, const float* I, int m, int n, int d, int d2,
                    int numMax, int numMin, double sigma, uint32_t* S,
                    float* Volume);

/* S = matchBlocksGPU(I, sigma, opts) -- source/matchBlocksGPU.m:18-91, the
 * whole-image search (not on the GBM3D path, which matches in windows): for
 * each reference tile (stride d, floor(m/d) x floor(n/d), column-major) the
 * positions of the m x n image whose d x d block, circularly wrapped, is
 * nearest by (|x|^2 + |y|^2 - 2<x,y>) / |x|^2.  S: numMax entries per tile,
 * 1-based column-major indices into m x n (the block's top-left pixel), the
 * tile's first counts[tile] of them its list, the rest 0.  numMax == numMin:
 * the numMax nearest (mink; counts all numMax).  numMin < numMax: the
 * positions with distance < 0.005 sigma^2 d^2 in index order, or the
 * numMax / numMin nearest when there are more / fewer (:71-81).  When
 * sigma > 30/255 the search runs on the image prefiltered by the mean over 5
 * cycle shifts of a 3-level 2-D sym4 wavelet hard threshold at 2 sigma
 * (:34-50; m and n multiples of 8).  counts, Ifilt (m x n, the image
 * searched): nullable.  d in {4, 8, 16, 32}; 1 <= numMin <= numMax <= 64.
 * Host buffers. */
int gbm3d_match_global(gbm3d_ctx* ctx, const float* I, int m, int n,
                       double sigma, int d, int numMax, int numMin,
                       uint32_t* S, uint32_t* counts, float* Ifilt);

/* Vavg = the 'ht' filter of denoiseAndAggregateGPU (cycle-spun 3-D
 * bior1.5 x Haar hard threshold, mean over spins, clip at 0)
 * -- source/denoiseAndAggregateGPU.m:19-47.  tau: levels x 7 (see
 * gbm3d_get_tau).  Volume, Vavg: m x n x K. */
int gbm3d_ht_filter(gbm3d_ctx* ctx, const float* Volume, int m, int n, int K,
                    int levels, int cycleSpin, const double* tau, float* Vavg);

/* Vavg = medfilt3(Volume, [3 3 3]): the filtType 'median' branch of
 * denoiseAndAggregateGPU -- source/denoiseAndAggregateGPU.m:49-52.  Volume,
 * Vavg: m x n x K (host); outside the volume the edge value repeats. */
int gbm3d_median_filter(gbm3d_ctx* ctx, const float* Volume, int m, int n,
                        int K, float* Vavg);

/* [U, W] = aggregation of denoiseAndAggregateGPU   -- :59-92.
 * U, W: m x n (crop of the padded canvas). */
int gbm3d_aggregate_ht(gbm3d_ctx* ctx, const float* Vavg, const uint32_t* S,
                       int m, int n, int d, int d2, int K, float* U, float* W);

/* [S, V1, V2] = matchBlocksWie(I, In, sigma, opts) -- source/matchBlocksWie.m:1-92
 * I = pilot, In = noisy.  V1, V2 (nullable): d x d x K x (m/d) x (n/d). */
int gbm3d_match_wie(gbm3d_ctx* ctx, const float* I, const float* In, int m,
                    int n, int d, int d2, int K, uint32_t* S, float* V1,
                    float* V2);

/* Filter part of denoiseAndAggregateWieFast: 2-D DCT x Haar-K empirical
 * Wiener, clip at 0 -- source/denoiseAndAggregateWieFast.m:19-66.
 * V1, V2, V2f: d x d x K x G.  gain2 (G): squared Frobenius norm of each
 * group's Wiener gains (the aggregation weight denominator, :81). */
int gbm3d_wiener(gbm3d_ctx* ctx, const float* V1, const float* V2, int d,
                 int K, int G, float sigma, float* V2f, float* gain2);

/* [U, W] = aggregation of denoiseAndAggregateWieFast -- :68-100. */
int gbm3d_aggregate_wie(gbm3d_ctx* ctx, const float* V2f, const float* gain2,
                        const uint32_t* S, int m, int n, int d, int d2, int K,
                        float* U, float* W);

/* [U, W] = denoiseAndAggregateWie(V1, V2, S, sigma, opts) -- clip01 = 1:
 * source/denoiseAndAggregateWie.m (dct()/idct() pages, filtered groups
 * clipped to [0, 1], so meant for images on [0, 1]); clip01 = 0: the same
 * call as denoiseAndAggregateWieFast (source/denoiseAndAggregateWieFast.m,
 * clip at 0 only).  Filter and aggregation in one call; layouts as
 * gbm3d_wiener / gbm3d_aggregate_wie, U and W m x n (unnormalised). */
int gbm3d_denoise_aggregate_wie(gbm3d_ctx* ctx, const float* V1, const float* V2,
                                const uint32_t* S, int m, int n, int d, int d2,
                                int K, float sigma, int clip01, float* U,
                                float* W);

/* ------------------------------------------------------- instrumentation
 * Per-stage device time of the last gbm3d_denoise* call, measured with HIP
 * events on the stream the kernels ran on (enabled by gbm3d_set_timing). */
enum gbm3d_stage {
  GBM3D_STAGE_PAD = 0,        /* mirror pad, padded search images, block-norm maps   */
  GBM3D_STAGE_MATCH1 = 1,     /* step-1 distances + top-K + volume (one kernel)       */
  GBM3D_STAGE_HT_LO = 2,      /* step-1 wavelet analysis low-pass, all levels         */
  GBM3D_STAGE_HT_SYN = 3,     /* step-1 threshold + synthesis, levels >= 2            */
  GBM3D_STAGE_HT_SYN0 = 4,    /* step-1 threshold + synthesis of the finest level     */
  GBM3D_STAGE_AGG1_PREP = 5,  /* step-1 cycle-spin mean, clip and group weights       */
  GBM3D_STAGE_AGG1 = 6,       /* step-1 gather aggregation (one kernel)               */
  GBM3D_STAGE_MATCH2 = 7,     /* step-2 distances + top-K (one kernel)                */
  GBM3D_STAGE_WIENER = 8,     /* step-2 Wiener filter (one kernel)                    */
  GBM3D_STAGE_AGG2 = 9,       /* step-2 gather aggregation (one kernel)               */
  GBM3D_STAGE_NORM = 10,      /* spin sums / normalisation / crop                     */
  GBM3D_STAGE_SHRINK = 11,    /* step-1 learned shrinkage (one kernel, variant)       */
  GBM3D_NUM_STAGES = 12
};
int gbm3d_set_timing(gbm3d_ctx* ctx, int enable);
/* Copy nbytes (a multiple of 16) between device buffers with a plain
 * streaming kernel on the context stream: a kernel of known HBM traffic
 * (nbytes read + nbytes written) that calibrates hardware byte counters. */
int gbm3d_device_copy(gbm3d_ctx* ctx, void* dst, const void* src, size_t nbytes);
/* ms[GBM3D_NUM_STAGES], launches[GBM3D_NUM_STAGES] (either may be NULL). */
int gbm3d_get_stage_times(gbm3d_ctx* ctx, double* ms, int* launches);

/* ---------------------------------------------------- learned shrinkage
 * A variant beside the reference path (no reference counterpart): a 4-layer
 * 1-D convolutional network (width 3, 64 channels, BF16 weights, FP32
 * accumulation) over the DCT-II x Haar-K coefficients of each 8 x 8 x 16
 * step-1 group, with sigma/50 as an input channel, predicts per-coefficient
 * gains that replace the step-1 hard threshold.  Targets are the Wiener gains
 * |c_clean|^2/(|c_clean|^2 + sigma^2) of the clean group (the gain formula of
 * source/denoiseAndAggregateWieFast.m:49).
 * Parameter vector (FP32): W0[3][64][64] | wsig[3][64] | b0[64] |
 * (W_l[3][64][64] | b_l[64]) for l = 1..3; W[tap][cin][cout], tap 0 reads
 * position k-1. */
#define GBM3D_SHRINK_PARAMS 49600
int gbm3d_shrink_set_weights(gbm3d_ctx* ctx, const float* params, int count);
/* Image `image` of the last denoise call with the learned form (waits for
 * the context stream): energy the learned gains removed from the noisy
 * coefficients and energy a 2.7 sigma hard threshold would remove, both
 * relative to the noise energy (coefficients x sigma^2) and averaged over the
 * match spins, and the number of spins that fell back. */
int gbm3d_shrink_stats(gbm3d_ctx* ctx, int image, double* est_learned,
                       double* est_ht, int* fell_back);
/* Learned shrinkage of G groups (host buffers, 8 x 8 x 16 x G): filtered
 * groups (coefficients times the gains, clipped at 0) and, when gains !=
 * NULL, the network gains in coefficient order (t = u + 8v, k). */
int gbm3d_shrink_groups(gbm3d_ctx* ctx, const float* V, int G, float sigma,
                        float* Vout, float* gains);
/* Training on device buffers (context stream).  Rows r = 16 i + k of group
 * i, 64 columns (coefficient t).
 * features: step-1 matching (spin 0, 8x8 blocks, matchSize 32, K 16) on the
 * noisy image, then for the nsel groups listed in dSel: X = c_noisy/sigma and
 * per target
 *   GBM3D_SHRINK_TARGET_WIENER: T = Wiener gains of the clean group,
 *     Wt = min(1 + c_clean^2/sigma^2, 100);
 *   GBM3D_SHRINK_TARGET_ENERGY: T = min(|c_clean| / |c_noisy|, 1) (the gain
 *     whose pilot coefficient has the clean magnitude, so that step 2's gain
 *     computed from the pilot is the oracle Wiener gain),
 *     Wt = min(c_noisy^2/sigma^2, 100);
 * Wt is 0 on the fixed DC approximation.  m, n: multiples of 8, >= 32. */
enum { GBM3D_SHRINK_TARGET_WIENER = 0, GBM3D_SHRINK_TARGET_ENERGY = 1 };
int gbm3d_shrink_features(gbm3d_ctx* ctx, const float* dClean,
                          const float* dNoisy, int m, int n, float sigma,
                          const uint32_t* dSel, int nsel, int target,
                          float* dX, float* dT, float* dWt);
/* Gains of ngroups groups (sigma per group in dSig). */
int gbm3d_shrink_forward(gbm3d_ctx* ctx, const float* dParams, const float* dX,
                         const float* dSig, int ngroups, float* dGains);
/* Gradient of sum(Wt (g - T)^2)/ngroups; *loss (nullable) synchronises. */
int gbm3d_shrink_grad(gbm3d_ctx* ctx, const float* dParams, const float* dX,
                      const float* dT, const float* dWt, const float* dSig,
                      int ngroups, float* dGrad, double* loss);
/* One Adam step (bias-corrected, step >= 1) on the parameter vector. */
int gbm3d_shrink_adam(gbm3d_ctx* ctx, float* dParams, float* dM, float* dV,
                      const float* dGrad, int step, float lr, float beta1,
                      float beta2, float eps);

/* ------------------------------------------------------ synthetic inputs
 * Seeded clean / noisy image pair on device buffers (m x n, column-major):
 * background 0, 12 rectangles then 8 disks of intensity U[0,255], a
 * sinusoid of amplitude 20 and period U[6,16] px on a rectangle, clip to
 * [0,255]; noisy = clean + sigma N(0,1), not clipped.  Either output may be
 * NULL.  No reference counterpart (test and training data). */
int gbm3d_synth_pair(gbm3d_ctx* ctx, uint64_t seed, int m, int n, float sigma,
                     float* dClean, float* dNoisy);

#ifdef __cplusplus
}
#endif
#endif /* GBM3D_H */
