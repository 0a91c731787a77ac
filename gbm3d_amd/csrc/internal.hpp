// internal.hpp -- shared host/device declarations of libgbm3d (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gbm3d.h"

namespace gbm3d {

// ------------------------------------------------------------ error plumbing
void set_error(const std::string& msg);
void clear_error();

struct Status {
  int code = GBM3D_OK;
  Status() = default;
  Status(int c) : code(c) {}
  bool ok() const { return code == GBM3D_OK; }
};

int fail(int code, const std::string& msg);

#define GBM3D_HIP(expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::gbm3d::fail(GBM3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GBM3D_TRY(expr)               \
  do {                                \
    int s_ = (expr);                  \
    if (s_ != GBM3D_OK) return s_;    \
  } while (0)

// Launch-error check right after a kernel launch.
int check_launch(const char* what);

// ------------------------------------------------------------------ options
int resolve(const gbm3d_opts* in, gbm3d_opts* out);
void tau_table(int levels, int mode, double sigma, double* tau);
// Coloured-noise thresholds (levels x 7, column-major) from sigma^2 * PSD (d1 x d2).
void colored_tau(int levels, const double* psd, int d1, int d2, double* tau);
void colored_wie(const double* psd, int d1, int d2, int d, double* sigmaWie);

// ------------------------------------------------------------ the filters
// bior1.5 (x, y) x Haar (z) of waveTransforms3D/getWaveFilters3Dbior.m:16-27,49-50.
constexpr double kCa = 0.016572815184060;
constexpr double kCb = 0.121533978016438;
constexpr double kCc = 0.70710678118654752440;  // sqrt(2)/2

// ------------------------------------------------------------ kernel launchers
// All launchers validate their own grid assumptions and return a status.
struct PadGeom {
  int M, N;   // image size (block multiple)
  int d2;     // match size; padded image is (M+d2) x (N+d2)
  int P() const { return M + d2; }
  int Q() const { return N + d2; }
};

// Block / match / group sizes the kernels are built for (checked by the
// pipeline's option validation before anything is launched).
bool match_supported(int d, int d2);       // FP32 matcher: d in {4, 8, 16, 32}, d2 in {32, 48}
bool match_bf16_supported(int d, int d2);  // BF16-MFMA matcher: d in {8, 16}, d2 in {32, 48}
bool wiener_supported(int d, int K);       // step-2 Wiener: d in {4, 8, 16}, K in {8, 16, 32}

int launch_mirror_pad(const float* src, int m, int n, int M, int N, float* dst, hipStream_t st);
int launch_search_pad(const float* img, int M, int N, int spin, int d2, float* pad, hipStream_t st);
int launch_norm_map(const float* pad, int P, int Q, int d, float* bn, hipStream_t st);
// launch_search_pad and launch_norm_map (block size d) of its result in one pass
int launch_search_pad_norm(const float* img, int M, int N, int spin, int d2, int d, float* pad, float* bn,
                           hipStream_t st);
// Block matching: S (K x Mt x Nt, 1-based window indices) and optionally the
// mosaic volume (M x N x K).
int launch_match(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S,
                 float* vol, hipStream_t st);
// d x d box sums of the padded image (the centring term of the BF16 matcher).
int launch_sum_map(const float* pad, int P, int Q, int d, float* sm, hipStream_t st);
// BF16-MFMA matching: approximate scores with rigorous error bounds; every
// candidate that can reach the top K is rescored with the FP32 chain, so S
// and the volume equal launch_match's bit for bit.  R (K..64): size of the
// first rescored set.  stats (nullable, kBf16Stats uint32 counters): tiles
// finished on the first set [0], on a widened set [1], by the exact search
// over all candidates [2]; rescored candidates [3]; why the exact search ran:
// first set above 128 candidates [4], no finite K-th key in it [5], widened
// set above 128 [6], reference block of zero or non-finite norm [7].
constexpr int kBf16Stats = 8;
// The counters live in kBf16StatSlots copies kBf16StatStride words apart
// (one cache line each), summed on the host.
constexpr int kBf16StatSlots = 64, kBf16StatStride = 32;
// matchBlocksGPU (k_gmatch.hip).  Prefilter: Iavg (m x n) from I, work of
// (m n / 4)(4 (1 + 1/4 + 1/16) + 1 + 1/4) floats.  Search: S (K x G), counts
// (G, nullable), work of (m + d - 1)(n + d - 1) + m n floats.
int launch_wav2_prefilter(const float* I, int m, int n, double tau, float* Iavg, float* work, hipStream_t st);
int launch_match_global(const float* I, int m, int n, int d, int K, int Kmin, double tol, uint32_t* S,
                        uint32_t* counts, float* work, hipStream_t st);
bool match_global_supported(int d);

int launch_match_bf16(const float* pad, const float* bn, const float* smap, int M, int N, int d, int d2, int K,
                      int R, uint32_t* S, float* vol, hipStream_t st, uint32_t* stats = nullptr);
// Same with nw (1, 2, 4) reference tiles per workgroup.
// matchBlocksCPU's tolerance branch (numMax ~= numMin, :89-99) for the FP32
// matcher: Kmin = numMin (0: off), tolu = the smallest float >= tol; a tile
// whose list would be shorter than numMax (where the reference fails) sets
// *flag = 1.
struct TolArgs {
  int Kmin;
  float tolu;
  int* flag;
};
int launch_set_i32(int* p, int v, hipStream_t st);
int launch_match_w(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                   int nw, hipStream_t st, const TolArgs* ta = nullptr);
// Gather V (d x d x K x G) from a padded image through S.
int launch_mosaic_gather(const float* pad, const uint32_t* S, int M, int N, int d, int d2, int K, float* V,
                         const int* run, hipStream_t st);
int launch_gather_groups(const float* pad, const uint32_t* S, int M, int N, int d, int d2, int K,
                         float* V, hipStream_t st);

// Wavelet hard threshold (all levels, all cycle spins) on the mosaic volume.
struct HtWork {
  float* vrec;                 // M*N*K cycle-spin accumulator (nullptr when cycleSpin == 1)
  std::vector<float*> lo[4];   // lo[l][spin]: level-(l+1) approximation
  std::vector<float*> rec[4];  // rec[l][spin]: reconstructed level-(l+1) approximation
};
size_t ht_work_floats(int M, int N, int K, int levels, int cycleSpin);
void ht_work_bind(float* base, int M, int N, int K, int levels, int cycleSpin, HtWork* w);
// hook(user, stage, 1) before and hook(user, stage, 0) after each kernel group
// (stages GBM3D_STAGE_HT_LO / _HT_SYN / _HT_SYN0); may be null.
typedef void (*StageHook)(void* user, int stage, int begin);
// Writes Vavg = max(mean over cycle spins of the rebuilt volume, 0), in the
// mosaic layout (gd == 0) or group-contiguous with gd x gd blocks
// (Vavg[u + gd*(v + gd*(k + K*g))], the aggregation's layout).
// run (nullable, device): when it holds 0 every kernel returns at once.
// wgt (nullable, device, (M/gd)(N/gd) floats): where the fused finest level
// can (gd == 8), it also writes the step-1 weights of the gd x gd groups
// (launch_agg1_prep's values, bit for bit) and sets *wgt_done; otherwise
// *wgt_done is false and the caller runs launch_agg1_prep.
int run_ht(const float* vol, int M, int N, int K, int levels, int cycleSpin, const double* tau,
           const HtWork& w, float* vavg, int gd, hipStream_t st, StageHook hook = nullptr, void* user = nullptr,
           const int* run = nullptr, float* wgt = nullptr, bool* wgt_done = nullptr);

// The 'median' step-1 filter: Vavg = medfilt3(vol, [3 3 3]) (edge values
// repeated), mosaic (gd == 0) or group-contiguous output.
int launch_median3(const float* vol, int M, int N, int K, int gd, float* out, hipStream_t st);

// Step-1 group weights 1/(TV(Vavg)+eps) from the group-contiguous Vavg.
int launch_agg1_prep(const float* vavg_groups, int M, int N, int K, int d, float* wgt, hipStream_t st);
// Mosaic (M x N x K) -> group-contiguous (d x d x K x G) copy.
int launch_mosaic_to_groups(const float* src, int M, int N, int K, int d, float* dst, hipStream_t st);

// Gather aggregation. step 1: blocks from the group-contiguous Vavg, weight wgt[g] * sine mask (.1,.9);
// step 2: blocks from V2f (d x d x K x G), weight mask(.2,.8)/gain2[g].
// masks: device array [2][d*d] (step-1 mask, then step-2 mask).
// Writes U, W for one match spin in the unshifted frame (shift 'spin').
int launch_aggregate_m(int step, const float* blocks, const float* wgt, const uint32_t* S, const float* masks,
                       int M, int N, int d, int d2, int K, int spin, float* U, float* W, hipStream_t st);
// Step-2 Wiener filter on groups gathered from the padded pilot / noisy images
// (pads != nullptr) or read from V1/V2 arrays.
// sig2tab (host, d x d, nullable): per-coefficient noise variance (coloured noise).
// clip01: clip the filtered groups to [0, 1] (denoiseAndAggregateWie.m:55)
// instead of at 0 only (denoiseAndAggregateWieFast.m:64).
int launch_wiener(const float* pad_pilot, const float* pad_noisy, const uint32_t* S, const float* V1,
                  const float* V2, int M, int N, int d, int d2, int K, long G, float sigma, float* V2f,
                  float* gain2, hipStream_t st, const float* sig2tab = nullptr, bool clip01 = false);
// out = sum_s U_s / (sum_s W_s + sum_t Wextra_t), cropped to m x n, double math;
// out2 (nullable) receives the same values cropped to m2 x n2 (m2 <= m, n2 <= n).
int launch_normalize(const float* U, int nU, const float* W, int nW, const float* Wextra, int nWextra,
                     int M, int N, int m, int n, float* out, hipStream_t st, float* out2 = nullptr, int m2 = 0,
                     int n2 = 0);

// ------------------------------------------------------- learned shrinkage
// Network over the DCT-Haar coefficients of 8 x 8 x 16 groups (4 conv layers
// of width 3 along the 16 positions, 64 channels, sigma input channel).
namespace shrink {
constexpr int CH = 64, POS = 16, TAPS = 3, KW = TAPS * CH, NL = 4;
constexpr int OFF_WSIG = KW * CH, OFF_B0 = OFF_WSIG + TAPS * CH;
__host__ __device__ constexpr int off_w(int l) { return l == 0 ? 0 : OFF_B0 + CH + (l - 1) * (KW * CH + CH); }
__host__ __device__ constexpr int off_b(int l) { return l == 0 ? OFF_B0 : off_w(l) + KW * CH; }
constexpr int NPARAM = off_b(NL - 1) + CH;
static_assert(NPARAM == GBM3D_SHRINK_PARAMS, "parameter count");
constexpr int FB = NL * CH + TAPS * CH;  // biases then sigma-channel weights (FP32)
}  // namespace shrink
constexpr int kShrinkGrid = 256;    // persistent blocks of the fused kernel
constexpr int kShrinkChunk = 1024;  // data rows per weight-gradient block
// Blocks of the fused shrinkage launch for G groups (its stats entries).
int shrink_grid(long G);
// params -> BF16 weights wb [4][64 cout][192 (tap, cin)] and FP32 fb (host).
void shrink_pack_weights(const float* params, uint16_t* wb, float* fb);
// Fused shrinkage of G groups: gathered from the padded image through S
// (S != nullptr) or read from src as d x d x K x G; coefficients scaled by
// the gains; writes filtered groups (clipped at 0), optionally the gains, and
// per-block sums of the energy removed from the noisy coefficients,
// stats[2 * b] (learned gains) / stats[2 * b + 1] (hard threshold at 2.7 sigma).
int launch_shrink(const float* src, const uint32_t* S, int M, int d2, long G, float sigma, const uint16_t* wb,
                  const float* fb, float* vout, float* gout, float* stats, int* grid_used, hipStream_t st);
// Training rows of the selected groups (noisy features, oracle Wiener gains, loss weights).
int launch_learned_judge(const float* stats, int grid, double noise, double bound, int* run, double* res,
                         hipStream_t st);
int launch_shrink_features(const float* padC, const float* padN, const uint32_t* S, int M, int d2,
                           const uint32_t* sel, int nsel, float sigma, int target, float* X, float* Tg, float* Wt,
                           hipStream_t st);
size_t shrink_train_floats(int ngroups);
int shrink_forward(const float* par, const float* X, const float* sig, int ngroups, float* ws, float* gains,
                   hipStream_t st);
int shrink_grad(const float* par, const float* X, const float* Tg, const float* Wt, const float* sig, int ngroups,
                float* ws, float* grad, float* loss_part, int* loss_blocks, hipStream_t st);
int launch_shrink_adam(float* p, float* m, float* v, const float* g, int n, float lr_t, float b1, float b2,
                       float eps, hipStream_t st);

// ------------------------------------------------------ synthetic test images
// Shape parameters of one clean image (the construction of gbm3d_synth_pair).
struct SynthParams {
  int rect[12][4];     // r0, r1, c0, c1 (half-open)
  double rect_val[12];
  double disk[8][4];   // centre row, centre col, radius^2, value
  int tex[4];          // r0, r1, c0, c1
  double tex_theta, tex_period, tex_phase;
};
void synth_params(uint64_t seed, int m, int n, SynthParams* p);
int launch_synth_pair(const SynthParams& p, uint64_t seed, int m, int n, float sigma, float* clean, float* noisy,
                      hipStream_t st);

}  // namespace gbm3d
