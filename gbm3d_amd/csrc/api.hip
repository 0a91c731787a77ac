// api.hip -- the C ABI of libgbm3d: contexts, the two-step pipeline of
// source/GBM3D.m:21-83, and the stage entry points.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "wave.hpp"

using namespace gbm3d;

struct gbm3d_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  char* io = nullptr;  // staging for the host-buffer entry points
  size_t io_bytes = 0;
  float* masks = nullptr;  // [4 block sizes][2][32*32]
  gbm3d_ext ext{};
  bool timing = false;
  bool in_batch = false;  // stage timers off while a batch runs on several streams
  // batched launches
  std::vector<hipStream_t> bstreams;
  std::vector<hipEvent_t> bjoin;
  hipEvent_t bfork = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<int64_t> gkey;
  hipGraphExec_t dexec = nullptr;  // distributed driver
  std::vector<int64_t> dkey;
  int tuned_streams = 0, tuned_match_waves = 0;
  double tuned_ms = 0.0;
  std::vector<int64_t> tkey;
  int match_waves = 4;  // reference tiles per matching workgroup (single-image path)
  // coloured noise of the current call (gbm3d_denoise_colored)
  bool colored = false;
  std::vector<double> ctau;    // levels x 7
  std::vector<float> wie_s2;   // blockSizeWie^2: sigmaWie.^2
  // learned shrinkage
  uint16_t* shrink_wb = nullptr;  // BF16 weights [4][64][192]
  float* shrink_fb = nullptr;     // FP32 biases and sigma-channel weights
  float* sstats = nullptr;        // per image and spin: [kShrinkGrid][2] residual partials
  size_t sstats_floats = 0;
  std::vector<double> s_learned, s_ht;  // last call, per image
  std::vector<int> s_fell;
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
};

namespace {

constexpr int kMaskStride = 2 * 32 * 32;
int mask_slot(int d) { return d == 4 ? 0 : d == 8 ? 1 : d == 16 ? 2 : d == 32 ? 3 : -1; }

// MATLAB linspace(a,b,n) with the last element pinned to b.
std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> y(n);
  const double step = n > 1 ? (b - a) / (n - 1) : 0.0;
  for (int k = 0; k < n; ++k) y[k] = a + k * step;
  y[n - 1] = b;
  return y;
}

// Bump allocator over the context arena; a first pass with base == nullptr
// measures the layout.
struct Carve {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t n) {
    size_t a = (off + 255) & ~(size_t)255;
    off = a + n * sizeof(T);
    return base ? reinterpret_cast<T*>(base + a) : nullptr;
  }
};

int ensure(char** buf, size_t* have, size_t need) {
  if (need <= *have) return GBM3D_OK;
  if (*buf) GBM3D_HIP(hipFree(*buf));
  *buf = nullptr;
  *have = 0;
  size_t want = need + need / 8;
  if (hipMalloc(reinterpret_cast<void**>(buf), want) != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    return fail(GBM3D_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  *have = want;
  return GBM3D_OK;
}

void mark_begin(gbm3d_ctx* c, int stage) {
  if (!c->timing || c->in_batch) return;
  if (c->pool_used + 2 > c->pool.size()) {
    hipEvent_t e1, e2;
    if (hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess) {
      c->timing = false;
      return;
    }
    c->pool.push_back(e1);
    c->pool.push_back(e2);
  }
  hipEvent_t a = c->pool[c->pool_used++];
  hipEvent_t b = c->pool[c->pool_used++];
  (void)hipEventRecord(a, c->stream);
  c->marks.push_back({stage, a, b});
}

void mark_end(gbm3d_ctx* c, int stage) {
  if (!c->timing || c->in_batch || c->marks.empty() || c->marks.back().stage != stage) return;
  (void)hipEventRecord(c->marks.back().b, c->stream);
}

void ht_hook(void* user, int stage, int begin) {
  gbm3d_ctx* c = static_cast<gbm3d_ctx*>(user);
  if (begin)
    mark_begin(c, stage);
  else
    mark_end(c, stage);
}

// HIP events around a group of launches on the context stream (when enabled).
struct StageTimer {
  gbm3d_ctx* c;
  int stage;
  StageTimer(gbm3d_ctx* ctx, int s) : c(ctx), stage(s) { mark_begin(c, s); }
  ~StageTimer() { mark_end(c, stage); }
};

int validate(const gbm3d_opts& o, int m, int n, int* M, int* N) {
  if (m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "image must be non-empty");
  if (o.filtType != GBM3D_FILT_HT && o.filtType != GBM3D_FILT_MEDIAN)
    return fail(GBM3D_ERR_ARG, "unrecognized filtering");  // denoiseAndAggregateGPU.m:54
  if (o.numMax != o.numMin) return fail(GBM3D_ERR_UNSUPPORTED, "numMax ~= numMin (tolerance matching) is not built");
  if (o.nMatchSpins < 1 || o.nMatchSpins > GBM3D_MAX_SPINS) return fail(GBM3D_ERR_ARG, "matchSpins: 1..16 entries");
  if (o.Wiener) {
    if (o.nMatchSpinsWie == GBM3D_UNSET) return fail(GBM3D_ERR_MISSING_FIELD, "Reference to non-existent field 'matchSpinsWie'");
    if (o.nMatchSpinsWie < 1 || o.nMatchSpinsWie > GBM3D_MAX_SPINS) return fail(GBM3D_ERR_ARG, "matchSpinsWie: 1..16 entries");
  }
  const int bs = o.blockSize, d2 = o.matchSize, K = o.numMax;
  if (mask_slot(bs) < 0) return fail(GBM3D_ERR_UNSUPPORTED, "blockSize must be 4, 8, 16 or 32");
  if (o.Wiener && (o.blockSizeWie != 4 && o.blockSizeWie != 8 && o.blockSizeWie != 16))
    return fail(GBM3D_ERR_UNSUPPORTED, "blockSizeWie must be 4, 8 or 16");
  if (d2 != 32 && d2 != 48) return fail(GBM3D_ERR_UNSUPPORTED, "matchSize must be 32 or 48");
  if (o.levels < 1 || o.levels > 4) return fail(GBM3D_ERR_UNSUPPORTED, "levels must be 1..4");
  if (o.cycleSpin < 1 || o.cycleSpin > 8) return fail(GBM3D_ERR_ARG, "cycleSpin must be 1..8");
  if (K < 1 || K > 64 || K % (1 << o.levels)) return fail(GBM3D_ERR_UNSUPPORTED, "numMax must be <= 64 and a multiple of 2^levels");
  *M = (m + bs - 1) / bs * bs;
  *N = (n + bs - 1) / bs * bs;
  if (*M > 2 * m || *N > 2 * n) return fail(GBM3D_ERR_ARG, "image smaller than half a block");
  if (*M < d2 || *N < d2) return fail(GBM3D_ERR_ARG, "padded image smaller than matchSize");
  if (*M % (1 << o.levels) || *N % (1 << o.levels)) return fail(GBM3D_ERR_UNSUPPORTED, "image size not a multiple of 2^levels");
  if (o.Wiener && (*M % o.blockSizeWie || *N % o.blockSizeWie))
    return fail(GBM3D_ERR_UNSUPPORTED, "padded size must be a multiple of blockSizeWie");
  return GBM3D_OK;
}

struct Plan {
  int m, n, M, N, bs, bw, d2, K, n1, n2;
  size_t MN() const { return (size_t)M * N; }
};

constexpr int kMaxLanes = 8;  // match spins in flight at once (one stream each)

// Buffers of one match spin in flight.
struct SpinBufs {
  float *pad1, *pad2, *bn, *smap, *vol, *vavg, *w1, *V2f, *g2, *htw;
  uint32_t *S1, *S2;
};

// All device buffers of one denoise call: shared images and spin sums, and
// one SpinBufs per lane (spins run lane = spin % lanes).
struct Bufs {
  float *Ip, *U1s, *W1s, *pilot, *U2s, *W2s;
  SpinBufs sb[kMaxLanes];
};

void carve_spin(Carve& cv, const Plan& p, const gbm3d_opts& o, SpinBufs* b);

void carve(Carve& cv, const Plan& p, const gbm3d_opts& o, int lanes, Bufs* b) {
  const size_t MN = p.MN();
  b->Ip = cv.take<float>(MN);
  b->U1s = cv.take<float>(MN * p.n1);
  b->W1s = cv.take<float>(MN * p.n1);
  b->pilot = cv.take<float>(MN);
  b->U2s = cv.take<float>(MN * p.n2);
  b->W2s = cv.take<float>(MN * p.n2);
  for (int l = 0; l < lanes; ++l) carve_spin(cv, p, o, &b->sb[l]);
}

void carve_spin(Carve& cv, const Plan& p, const gbm3d_opts& o, SpinBufs* b) {
  const size_t MN = p.MN(), P = (size_t)p.M + p.d2, Q = (size_t)p.N + p.d2;
  const int dmin = std::min(p.bs, o.Wiener ? p.bw : p.bs);
  const size_t G1 = MN / ((size_t)p.bs * p.bs), G2 = o.Wiener ? MN / ((size_t)p.bw * p.bw) : 0;
  {
    SpinBufs& sb = *b;
    sb.pad1 = cv.take<float>(P * Q);
    sb.pad2 = cv.take<float>(P * Q);
    sb.bn = cv.take<float>((P - dmin + 1) * (Q - dmin + 1));
    sb.smap = cv.take<float>((P - dmin + 1) * (Q - dmin + 1));
    sb.S1 = cv.take<uint32_t>((size_t)p.K * G1);
    sb.vol = cv.take<float>(MN * p.K);
    sb.htw = cv.take<float>(ht_work_floats(p.M, p.N, p.K, o.levels, o.cycleSpin));
    sb.vavg = cv.take<float>(MN * p.K);
    sb.w1 = cv.take<float>(G1);
    sb.S2 = cv.take<uint32_t>((size_t)p.K * G2);
    sb.V2f = cv.take<float>(G2 ? MN * p.K : 0);
    sb.g2 = cv.take<float>(G2);
  }
}

// Lanes of a single-image call: its match spins are independent until the
// spin sums are normalised (source/GBM3D.m:50-57, 71-82), so they run on
// parallel streams; serial under stage timing (the timers record on the
// context stream).
int spin_lanes(const gbm3d_ctx* c, const Plan& p) {
  if (c->timing) return 1;
  return std::max(1, std::min(kMaxLanes, std::max(p.n1, p.n2)));
}

// Block matching on a padded image with the context's distance variant.
int do_match(gbm3d_ctx* c, const float* pad, float* bn, float* smap, int M, int N, int d, int d2, int K, uint32_t* S,
             float* vol, hipStream_t st, int stage) {
  const int P = M + d2, Q = N + d2;
  {
    StageTimer t(c, GBM3D_STAGE_PAD);
    GBM3D_TRY(launch_norm_map(pad, P, Q, d, bn, st));
    if (c->ext.match_bf16) GBM3D_TRY(launch_sum_map(pad, P, Q, d, smap, st));
  }
  StageTimer t(c, stage);
  if (c->ext.match_bf16) {
    const int R = c->ext.rescore > 0 ? std::max(c->ext.rescore, 2 * K) : 2 * K;
    return launch_match_bf16(pad, bn, smap, M, N, d, d2, K, std::min(R, 64), S, vol, st);
  }
  return launch_match_w(pad, bn, M, N, d, d2, K, S, vol, c->match_waves, st);
}

int make_plan(const gbm3d_opts& o, int m, int n, Plan* p) {
  p->m = m;
  p->n = n;
  GBM3D_TRY(validate(o, m, n, &p->M, &p->N));
  p->bs = o.blockSize;
  p->bw = o.blockSizeWie;
  p->d2 = o.matchSize;
  p->K = o.numMax;
  p->n1 = o.nMatchSpins;
  p->n2 = o.Wiener ? o.nMatchSpinsWie : 0;
  return GBM3D_OK;
}

// Device workspace bytes of one image's pipeline.
size_t pipeline_bytes(const Plan& p, const gbm3d_opts& o, int lanes = 1) {
  Carve measure{nullptr};
  Bufs b;
  carve(measure, p, o, lanes, &b);
  return measure.off + 256;
}

// One image through both steps on stream st, workspace at ws (pipeline_bytes).
// Fork `lanes` streams from st (lanes > 1), or run everything on st.
int lanes_fork(gbm3d_ctx* c, int lanes, hipStream_t st) {
  if (lanes <= 1) return GBM3D_OK;
  GBM3D_HIP(hipEventRecord(c->bfork, st));
  for (int l = 0; l < lanes; ++l) GBM3D_HIP(hipStreamWaitEvent(c->bstreams[l], c->bfork, 0));
  return GBM3D_OK;
}

int lanes_join(gbm3d_ctx* c, int lanes, hipStream_t st) {
  if (lanes <= 1) return GBM3D_OK;
  for (int l = 0; l < lanes; ++l) {
    GBM3D_HIP(hipEventRecord(c->bjoin[l], c->bstreams[l]));
    GBM3D_HIP(hipStreamWaitEvent(st, c->bjoin[l], 0));
  }
  return GBM3D_OK;
}

// One image through both steps on stream st, workspace at ws
// (pipeline_bytes(p, o, lanes)); with lanes > 1 the match spins of each step
// run on the context's batch streams (not while a batch is being issued).
int run_pipeline_on(gbm3d_ctx* c, const float* dI, const Plan& p, float sigma, const gbm3d_opts& o, float* dU,
                    float* dU1, hipStream_t st, char* ws, float* sstats, int lanes = 1) {
  const int m = p.m, n = p.n;
  Carve cv{ws};
  Bufs b;
  carve(cv, p, o, lanes, &b);
  const size_t MN = p.MN();
  std::vector<double> tau((size_t)o.levels * 7);
  if (c->colored)
    tau = c->ctau;  // getColoredTau (source/GBM3D.m:33-35)
  else
    tau_table(o.levels, o.tauMode, (double)sigma, tau.data());
  const float* masks1 = c->masks + kMaskStride * mask_slot(p.bs);
  // spin ii of a step with `nl` lanes runs on lane ii % nl
  auto lane_stream = [&](int ii, int nl) { return nl > 1 ? c->bstreams[ii % nl] : st; };

  {
    StageTimer t(c, GBM3D_STAGE_PAD);
    GBM3D_TRY(launch_mirror_pad(dI, m, n, p.M, p.N, b.Ip, st));  // source/GBM3D.m:22-25
  }
  // ---- step 1: source/GBM3D.m:50-57
  // The noisy image's search image of spin s is built into a lane's pad2 and
  // left there: a step-2 spin with the same s on that lane reuses it.
  int noisy_spin[kMaxLanes];
  for (int l = 0; l < kMaxLanes; ++l) noisy_spin[l] = INT_MIN;
  const int l1 = std::max(1, std::min(lanes, p.n1));
  GBM3D_TRY(lanes_fork(c, l1, st));
  for (int ii = 0; ii < p.n1; ++ii) {
    const int sp = o.matchSpins[ii];
    const hipStream_t ls = lane_stream(ii, l1);
    SpinBufs& sb = b.sb[ii % l1];
    {
      StageTimer t(c, GBM3D_STAGE_PAD);
      GBM3D_TRY(launch_search_pad(b.Ip, p.M, p.N, sp, p.d2, sb.pad2, ls));
      noisy_spin[ii % l1] = sp;
    }
    if (c->ext.learned) {
      // variant: learned gains on the DCT-Haar coefficients of each group,
      // gathered straight from the search image (no mosaic volume)
      GBM3D_TRY(do_match(c, sb.pad2, sb.bn, sb.smap, p.M, p.N, p.bs, p.d2, p.K, sb.S1, nullptr, ls,
                         GBM3D_STAGE_MATCH1));
      StageTimer t(c, GBM3D_STAGE_SHRINK);
      const long G1 = (long)(p.M / p.bs) * (p.N / p.bs);
      GBM3D_TRY(launch_shrink(sb.pad2, sb.S1, p.M, p.d2, G1, sigma, c->shrink_wb, c->shrink_fb, sb.vavg, nullptr,
                              sstats + (size_t)2 * kShrinkGrid * ii, nullptr, ls));
    } else {
      HtWork hw;
      ht_work_bind(sb.htw, p.M, p.N, p.K, o.levels, o.cycleSpin, &hw);
      GBM3D_TRY(do_match(c, sb.pad2, sb.bn, sb.smap, p.M, p.N, p.bs, p.d2, p.K, sb.S1, sb.vol, ls,
                         GBM3D_STAGE_MATCH1));
      if (o.filtType == GBM3D_FILT_MEDIAN) {
        StageTimer t(c, GBM3D_STAGE_HT_SYN0);
        GBM3D_TRY(launch_median3(sb.vol, p.M, p.N, p.K, p.bs, sb.vavg, ls));  // denoiseAndAggregateGPU.m:49-52
      } else {
        GBM3D_TRY(run_ht(sb.vol, p.M, p.N, p.K, o.levels, o.cycleSpin, tau.data(), hw, sb.vavg, p.bs, ls, ht_hook,
                         c));
      }
    }
    {
      StageTimer t(c, GBM3D_STAGE_AGG1_PREP);
      GBM3D_TRY(launch_agg1_prep(sb.vavg, p.M, p.N, p.K, p.bs, sb.w1, ls));
    }
    {
      StageTimer t(c, GBM3D_STAGE_AGG1);
      GBM3D_TRY(launch_aggregate_m(1, sb.vavg, sb.w1, sb.S1, masks1, p.M, p.N, p.bs, p.d2, p.K, sp,
                                   b.U1s + MN * ii, b.W1s + MN * ii, ls));
    }
  }
  GBM3D_TRY(lanes_join(c, l1, st));
  {
    StageTimer t(c, GBM3D_STAGE_NORM);
    // the pilot (padded size) and, from the same pass, U1 (cropped)
    GBM3D_TRY(launch_normalize(b.U1s, p.n1, b.W1s, p.n1, nullptr, 0, p.M, p.N, p.M, p.N, b.pilot, st, dU1, m, n));
    if (!o.Wiener) {  // :59-63
      GBM3D_TRY(launch_normalize(b.U1s, p.n1, b.W1s, p.n1, nullptr, 0, p.M, p.N, m, n, dU, st));
      return GBM3D_OK;
    }
  }
  // ---- step 2: source/GBM3D.m:70-82
  const float* masks2 = c->masks + kMaskStride * mask_slot(p.bw);
  const long G2 = (long)(p.M / p.bw) * (p.N / p.bw);
  const int l2 = std::max(1, std::min(lanes, p.n2));
  GBM3D_TRY(lanes_fork(c, l2, st));
  for (int ii = 0; ii < p.n2; ++ii) {
    const int sp = o.matchSpinsWie[ii];
    const hipStream_t ls = lane_stream(ii, l2);
    SpinBufs& sb = b.sb[ii % l2];
    {
      StageTimer t(c, GBM3D_STAGE_PAD);
      GBM3D_TRY(launch_search_pad(b.pilot, p.M, p.N, sp, p.d2, sb.pad1, ls));
      if (noisy_spin[ii % l2] != sp) {  // else step 1 left this search image in pad2 (the steps are joined)
        GBM3D_TRY(launch_search_pad(b.Ip, p.M, p.N, sp, p.d2, sb.pad2, ls));
        noisy_spin[ii % l2] = sp;
      }
    }
    GBM3D_TRY(do_match(c, sb.pad1, sb.bn, sb.smap, p.M, p.N, p.bw, p.d2, p.K, sb.S2, nullptr, ls,
                       GBM3D_STAGE_MATCH2));
    {
      StageTimer t(c, GBM3D_STAGE_WIENER);
      GBM3D_TRY(launch_wiener(sb.pad1, sb.pad2, sb.S2, nullptr, nullptr, p.M, p.N, p.bw, p.d2, p.K, G2, sigma,
                              sb.V2f, sb.g2, ls, c->colored ? c->wie_s2.data() : nullptr));
    }
    {
      StageTimer t(c, GBM3D_STAGE_AGG2);
      GBM3D_TRY(launch_aggregate_m(2, sb.V2f, sb.g2, sb.S2, masks2, p.M, p.N, p.bw, p.d2, p.K, sp,
                                   b.U2s + MN * ii, b.W2s + MN * ii, ls));
    }
  }
  GBM3D_TRY(lanes_join(c, l2, st));
  {
    // U = sum(U,3)./sum(W,3): slices past numel(matchSpinsWie) still hold the
    // step-1 weights (U, W are sized by numel(matchSpins), :26-27).
    StageTimer t(c, GBM3D_STAGE_NORM);
    const int extra = std::max(0, p.n1 - p.n2);
    GBM3D_TRY(launch_normalize(b.U2s, p.n2, b.W2s, p.n2, b.W1s + MN * p.n2, extra, p.M, p.N, m, n, dU, st));
  }
  return GBM3D_OK;
}

// The learned form needs loaded weights, 8 x 8 x 16 step-1 groups and a
// residual buffer of [images][spins][kShrinkGrid][2].
int learned_setup(gbm3d_ctx* c, const gbm3d_opts& o, int images) {
  if (!c->ext.learned) return GBM3D_OK;
  if (!c->shrink_wb) return fail(GBM3D_ERR_ARG, "learned shrinkage: no weights loaded (gbm3d_shrink_set_weights)");
  if (o.blockSize != 8 || o.numMax != 16)
    return fail(GBM3D_ERR_UNSUPPORTED, "learned shrinkage needs blockSize 8 and numMax 16");
  const size_t need = (size_t)images * GBM3D_MAX_SPINS * 2 * kShrinkGrid;
  if (need > c->sstats_floats) {
    if (c->sstats) GBM3D_HIP(hipFree(c->sstats));
    c->sstats = nullptr;
    c->sstats_floats = 0;
    GBM3D_HIP(hipMalloc(reinterpret_cast<void**>(&c->sstats), need * sizeof(float)));
    c->sstats_floats = need;
  }
  return GBM3D_OK;
}

float* image_stats(gbm3d_ctx* c, int i) {
  return c->ext.learned ? c->sstats + (size_t)i * GBM3D_MAX_SPINS * 2 * kShrinkGrid : nullptr;
}

int shrink_grid_of(const Plan& p) { return shrink_grid((long)(p.M / p.bs) * (p.N / p.bs)); }

// After a learned-form call has finished: per image, the energy the learned
// gains removed from the noisy coefficients relative to the noise energy
// (coefficients x sigma^2, all spins).  Removing more than `bound` times the
// noise energy means the shrinkage also removed signal (the discrepancy
// principle); such images are recomputed on the reference path (eagerly, on
// the context stream).
int learned_judge(gbm3d_ctx* c, const gbm3d_opts& o, int count, const std::vector<Plan>& plans,
                  const float* const* dI, const float* sigma, float* const* dU, float* const* dU1) {
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  const size_t per = (size_t)GBM3D_MAX_SPINS * 2 * kShrinkGrid;
  std::vector<float> h((size_t)count * per);
  GBM3D_HIP(hipMemcpy(h.data(), c->sstats, h.size() * sizeof(float), hipMemcpyDeviceToHost));
  const double bound = (c->ext.learned_bound > 0 ? c->ext.learned_bound : 1000) / 1000.0;
  c->s_learned.assign(count, 0.0);
  c->s_ht.assign(count, 0.0);
  c->s_fell.assign(count, 0);
  std::vector<int> redo;
  for (int i = 0; i < count; ++i) {
    const int grid = shrink_grid_of(plans[i]);
    for (int sp = 0; sp < plans[i].n1; ++sp)
      for (int b = 0; b < grid; ++b) {
        c->s_learned[i] += h[i * per + (size_t)2 * kShrinkGrid * sp + 2 * b];
        c->s_ht[i] += h[i * per + (size_t)2 * kShrinkGrid * sp + 2 * b + 1];
      }
    const double noise = (double)plans[i].n1 * plans[i].M * plans[i].N * plans[i].K * (double)sigma[i] * sigma[i];
    c->s_learned[i] /= noise;
    c->s_ht[i] /= noise;
    if (!(c->s_learned[i] <= bound)) redo.push_back(i);
  }
  if (redo.empty()) return GBM3D_OK;
  const gbm3d_ext keep = c->ext;
  c->ext.learned = 0;
  int st = GBM3D_OK;
  for (int i : redo) {
    c->s_fell[i] = 1;
    st = ensure(&c->arena, &c->arena_bytes, pipeline_bytes(plans[i], o));
    if (st == GBM3D_OK)
      st = run_pipeline_on(c, dI[i], plans[i], sigma[i], o, dU[i], dU1 ? dU1[i] : nullptr, c->stream, c->arena,
                           nullptr);
    if (st != GBM3D_OK) break;
  }
  c->ext = keep;
  return st;
}

int run_pipeline(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts& o, float* dU,
                 float* dU1) {
  Plan p;
  GBM3D_TRY(make_plan(o, m, n, &p));
  GBM3D_TRY(learned_setup(c, o, 1));
  const int lanes = spin_lanes(c, p);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, pipeline_bytes(p, o, lanes)));
  GBM3D_TRY(run_pipeline_on(c, dI, p, sigma, o, dU, dU1, c->stream, c->arena, image_stats(c, 0), lanes));
  if (!c->ext.learned) return GBM3D_OK;
  std::vector<Plan> plans{p};
  float* u1[1] = {dU1};
  return learned_judge(c, o, 1, plans, &dI, &sigma, &dU, dU1 ? u1 : nullptr);
}

int set_dev(gbm3d_ctx* c) {
  if (!c) return fail(GBM3D_ERR_ARG, "NULL context");
  GBM3D_HIP(hipSetDevice(c->device));
  clear_error();
  return GBM3D_OK;
}

template <class T>
int upload(gbm3d_ctx* c, T* dst, const T* src, size_t n) {
  GBM3D_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return GBM3D_OK;
}
template <class T>
int download(gbm3d_ctx* c, T* dst, const T* src, size_t n) {
  if (!dst) return GBM3D_OK;
  GBM3D_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return GBM3D_OK;
}

int resolve_checked(const gbm3d_opts* in, gbm3d_opts* out) {
  GBM3D_TRY(resolve(in, out));
  return GBM3D_OK;
}

bool supported_block(int d) { return d == 4 || d == 8 || d == 16 || d == 32; }

int stage_sizes(int m, int n, int d, int d2, int K) {
  if (m < 1 || n < 1 || !supported_block(d)) return fail(GBM3D_ERR_ARG, "stage: bad size or block size");
  if (m % d || n % d) return fail(GBM3D_ERR_ARG, "stage: image size must be a multiple of the block size");
  if (m < d2 || n < d2 || (d2 != 32 && d2 != 48)) return fail(GBM3D_ERR_ARG, "stage: matchSize must be 32 or 48 and <= image size");
  if (K < 1 || K > 64) return fail(GBM3D_ERR_ARG, "stage: numMax must be 1..64");
  return GBM3D_OK;
}

}  // namespace

extern "C" {

int gbm3d_create(int device, gbm3d_ctx** out) {
  clear_error();
  if (!out) return fail(GBM3D_ERR_ARG, "NULL output");
  *out = nullptr;
  int count = 0;
  GBM3D_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(GBM3D_ERR_ARG, "no such device");
  GBM3D_HIP(hipSetDevice(device));
  gbm3d_ctx* c = new gbm3d_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(GBM3D_ERR_HIP, "hipStreamCreate failed");
  }
  c->stream = c->own;
  c->bstreams.assign(8, nullptr);
  c->bjoin.assign(8, nullptr);
  bool ok = hipEventCreateWithFlags(&c->bfork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < 8 && ok; ++i)
    ok = hipStreamCreateWithFlags(&c->bstreams[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&c->bjoin[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    delete c;
    return fail(GBM3D_ERR_HIP, "batch stream/event creation failed");
  }
  // Aggregation masks for every supported block size (computed in double,
  // stored as float): step 1 sin(pi*linspace(.1,.9,d)) outer product,
  // step 2 sin(pi*linspace(.2,.8,d)) outer product.
  std::vector<float> h(4 * kMaskStride, 0.f);
  const int ds[4] = {4, 8, 16, 32};
  for (int i = 0; i < 4; ++i) {
    const int d = ds[i];
    for (int which = 0; which < 2; ++which) {
      std::vector<double> s = which == 0 ? linspace(0.1, 0.9, d) : linspace(0.2, 0.8, d);
      for (auto& x : s) x = std::sin(M_PI * x);
      for (int v = 0; v < d; ++v)
        for (int u = 0; u < d; ++u) h[i * kMaskStride + which * d * d + u + d * v] = (float)(s[u] * s[v]);
    }
  }
  if (hipMalloc(reinterpret_cast<void**>(&c->masks), h.size() * sizeof(float)) != hipSuccess ||
      hipMemcpy(c->masks, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return fail(GBM3D_ERR_HIP, "mask table upload failed");
  }
  *out = c;
  return GBM3D_OK;
}

int gbm3d_destroy(gbm3d_ctx* c) {
  if (!c) return GBM3D_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->arena) (void)hipFree(c->arena);
  if (c->io) (void)hipFree(c->io);
  if (c->masks) (void)hipFree(c->masks);
  if (c->shrink_wb) (void)hipFree(c->shrink_wb);
  if (c->shrink_fb) (void)hipFree(c->shrink_fb);
  if (c->sstats) (void)hipFree(c->sstats);
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->dexec) (void)hipGraphExecDestroy(c->dexec);
  for (auto sb : c->bstreams)
    if (sb) (void)hipStreamDestroy(sb);
  for (auto e : c->bjoin)
    if (e) (void)hipEventDestroy(e);
  if (c->bfork) (void)hipEventDestroy(c->bfork);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return GBM3D_OK;
}

int gbm3d_set_ext(gbm3d_ctx* c, const gbm3d_ext* e) {
  GBM3D_TRY(set_dev(c));
  if (e && (e->rescore < 0 || e->rescore > 64)) return fail(GBM3D_ERR_ARG, "ext: rescore must be 0..64");
  c->ext = e ? *e : gbm3d_ext{};
  return GBM3D_OK;
}

int gbm3d_get_ext(gbm3d_ctx* c, gbm3d_ext* e) {
  GBM3D_TRY(set_dev(c));
  if (!e) return fail(GBM3D_ERR_ARG, "NULL ext");
  *e = c->ext;
  return GBM3D_OK;
}

int gbm3d_set_stream(gbm3d_ctx* c, void* s) {
  GBM3D_TRY(set_dev(c));
  c->stream = s ? reinterpret_cast<hipStream_t>(s) : c->own;
  return GBM3D_OK;
}

int gbm3d_synchronize(gbm3d_ctx* c) {
  GBM3D_TRY(set_dev(c));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_set_timing(gbm3d_ctx* c, int enable) {
  GBM3D_TRY(set_dev(c));
  c->timing = enable != 0;
  return GBM3D_OK;
}

__global__ void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int gbm3d_device_copy(gbm3d_ctx* c, void* dst, const void* src, size_t nbytes) {
  GBM3D_TRY(set_dev(c));
  if (!dst || !src || nbytes % 16) return fail(GBM3D_ERR_ARG, "device_copy: NULL or size not a multiple of 16");
  const size_t n = nbytes / 16;
  if (n == 0) return GBM3D_OK;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, c->stream, static_cast<const float4*>(src),
                     static_cast<float4*>(dst), n);
  return check_launch("device_copy");
}

int gbm3d_get_stage_times(gbm3d_ctx* c, double* ms, int* launches) {
  GBM3D_TRY(set_dev(c));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> acc(GBM3D_NUM_STAGES, 0.0);
  std::vector<int> cnt(GBM3D_NUM_STAGES, 0);
  for (auto& mk : c->marks) {
    float t = 0.f;
    GBM3D_HIP(hipEventElapsedTime(&t, mk.a, mk.b));
    acc[mk.stage] += t;
    cnt[mk.stage] += 1;
  }
  for (int s = 0; s < GBM3D_NUM_STAGES; ++s) {
    if (ms) ms[s] = acc[s];
    if (launches) launches[s] = cnt[s];
  }
  return GBM3D_OK;
}

int gbm3d_denoise_device(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts* opts, float* dU,
                         float* dU1) {
  GBM3D_TRY(set_dev(c));
  if (!dI || !dU) return fail(GBM3D_ERR_ARG, "NULL image");
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  if (c->timing) {
    c->marks.clear();
    c->pool_used = 0;
  }
  return run_pipeline(c, dI, m, n, sigma, o, dU, dU1);
}

int gbm3d_denoise(gbm3d_ctx* c, const float* I, int m, int n, float sigma, const gbm3d_opts* opts, float* U,
                  float* U1) {
  GBM3D_TRY(set_dev(c));
  if (!I || !U || m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "NULL or empty image");
  const size_t mn = (size_t)m * n;
  GBM3D_TRY(ensure(&c->io, &c->io_bytes, 3 * mn * sizeof(float) + 1024));
  float* dI = reinterpret_cast<float*>(c->io);
  float* dU = dI + mn;
  float* dU1 = dU + mn;
  GBM3D_TRY(upload(c, dI, I, mn));
  GBM3D_TRY(gbm3d_denoise_device(c, dI, m, n, sigma, opts, dU, U1 ? dU1 : nullptr));
  GBM3D_TRY(download(c, U, dU, mn));
  if (U1) GBM3D_TRY(download(c, U1, dU1, mn));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

// ---------------------------------------------------------- stage entry points

int gbm3d_match(gbm3d_ctx* c, const float* I, int m, int n, int d, int d2, int K, uint32_t* S, float* Volume) {
  GBM3D_TRY(set_dev(c));
  if (!I || !S) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, K));
  const size_t mn = (size_t)m * n, P = (size_t)m + d2, Q = (size_t)n + d2, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** img, float** pad, float** bn, float** sm, uint32_t** dS, float** vol) {
    *img = k.take<float>(mn);
    *pad = k.take<float>(P * Q);
    *bn = k.take<float>((P - d + 1) * (Q - d + 1));
    *sm = k.take<float>((P - d + 1) * (Q - d + 1));
    *dS = k.take<uint32_t>(G * K);
    *vol = k.take<float>(Volume ? mn * K : 0);
  };
  float *img, *pad, *bn, *sm, *vol;
  uint32_t* dS;
  lay(cv, &img, &pad, &bn, &sm, &dS, &vol);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &img, &pad, &bn, &sm, &dS, &vol);
  GBM3D_TRY(upload(c, img, I, mn));
  GBM3D_TRY(launch_search_pad(img, m, n, 0, d2, pad, c->stream));
  GBM3D_TRY(do_match(c, pad, bn, sm, m, n, d, d2, K, dS, Volume ? vol : nullptr, c->stream, GBM3D_STAGE_MATCH1));
  GBM3D_TRY(download(c, S, dS, G * K));
  if (Volume) GBM3D_TRY(download(c, Volume, vol, mn * K));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_ht_filter(gbm3d_ctx* c, const float* Volume, int m, int n, int K, int levels, int cycleSpin,
                    const double* tau, float* Vavg) {
  GBM3D_TRY(set_dev(c));
  if (!Volume || !tau || !Vavg) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (m < 2 || n < 2 || K < 2) return fail(GBM3D_ERR_ARG, "ht_filter: size");
  if (levels < 1 || levels > 4 || cycleSpin < 1 || cycleSpin > 8) return fail(GBM3D_ERR_ARG, "ht_filter: levels/cycleSpin");
  const size_t vol = (size_t)m * n * K;
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** in, float** out, float** htw) {
    *in = k.take<float>(vol);
    *out = k.take<float>(vol);
    *htw = k.take<float>(ht_work_floats(m, n, K, levels, cycleSpin));
  };
  float *in, *out, *htw;
  lay(cv, &in, &out, &htw);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &in, &out, &htw);
  HtWork hw;
  ht_work_bind(htw, m, n, K, levels, cycleSpin, &hw);
  GBM3D_TRY(upload(c, in, Volume, vol));
  GBM3D_TRY(run_ht(in, m, n, K, levels, cycleSpin, tau, hw, out, 0, c->stream));
  GBM3D_TRY(download(c, Vavg, out, vol));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_median_filter(gbm3d_ctx* c, const float* Volume, int m, int n, int K, float* Vavg) {
  GBM3D_TRY(set_dev(c));
  if (!Volume || !Vavg) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (m < 1 || n < 1 || K < 1) return fail(GBM3D_ERR_ARG, "median_filter: size");
  const size_t vol = (size_t)m * n * K;
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** in, float** out) {
    *in = k.take<float>(vol);
    *out = k.take<float>(vol);
  };
  float *in, *out;
  lay(cv, &in, &out);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &in, &out);
  GBM3D_TRY(upload(c, in, Volume, vol));
  GBM3D_TRY(launch_median3(in, m, n, K, 0, out, c->stream));
  GBM3D_TRY(download(c, Vavg, out, vol));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_aggregate_ht(gbm3d_ctx* c, const float* Vavg, const uint32_t* S, int m, int n, int d, int d2, int K,
                       float* U, float* W) {
  GBM3D_TRY(set_dev(c));
  if (!Vavg || !S || !U || !W) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, K));
  const size_t mn = (size_t)m * n, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** va, float** vg, uint32_t** dS, float** w, float** dU, float** dW) {
    *va = k.take<float>(mn * K);
    *vg = k.take<float>(mn * K);
    *dS = k.take<uint32_t>(G * K);
    *w = k.take<float>(G);
    *dU = k.take<float>(mn);
    *dW = k.take<float>(mn);
  };
  float *va, *vg, *w, *dU, *dW;
  uint32_t* dS;
  lay(cv, &va, &vg, &dS, &w, &dU, &dW);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &va, &vg, &dS, &w, &dU, &dW);
  GBM3D_TRY(upload(c, va, Vavg, mn * K));
  GBM3D_TRY(upload(c, dS, S, G * K));
  GBM3D_TRY(launch_mosaic_to_groups(va, m, n, K, d, vg, c->stream));
  GBM3D_TRY(launch_agg1_prep(vg, m, n, K, d, w, c->stream));
  GBM3D_TRY(launch_aggregate_m(1, vg, w, dS, c->masks + kMaskStride * mask_slot(d), m, n, d, d2, K, 0, dU, dW,
                               c->stream));
  GBM3D_TRY(download(c, U, dU, mn));
  GBM3D_TRY(download(c, W, dW, mn));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_match_wie(gbm3d_ctx* c, const float* I, const float* In, int m, int n, int d, int d2, int K, uint32_t* S,
                    float* V1, float* V2) {
  GBM3D_TRY(set_dev(c));
  if (!I || !In || !S) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, K));
  const size_t mn = (size_t)m * n, P = (size_t)m + d2, Q = (size_t)n + d2, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** i1, float** i2, float** p1, float** p2, float** bn, float** sm, uint32_t** dS,
                 float** v) {
    *i1 = k.take<float>(mn);
    *i2 = k.take<float>(mn);
    *p1 = k.take<float>(P * Q);
    *p2 = k.take<float>(P * Q);
    *bn = k.take<float>((P - d + 1) * (Q - d + 1));
    *sm = k.take<float>((P - d + 1) * (Q - d + 1));
    *dS = k.take<uint32_t>(G * K);
    *v = k.take<float>(mn * K);
  };
  float *i1, *i2, *p1, *p2, *bn, *sm, *v;
  uint32_t* dS;
  lay(cv, &i1, &i2, &p1, &p2, &bn, &sm, &dS, &v);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &i1, &i2, &p1, &p2, &bn, &sm, &dS, &v);
  GBM3D_TRY(upload(c, i1, I, mn));
  GBM3D_TRY(upload(c, i2, In, mn));
  GBM3D_TRY(launch_search_pad(i1, m, n, 0, d2, p1, c->stream));
  GBM3D_TRY(launch_search_pad(i2, m, n, 0, d2, p2, c->stream));
  GBM3D_TRY(do_match(c, p1, bn, sm, m, n, d, d2, K, dS, nullptr, c->stream, GBM3D_STAGE_MATCH2));
  GBM3D_TRY(download(c, S, dS, G * K));
  if (V1) {
    GBM3D_TRY(launch_gather_groups(p1, dS, m, n, d, d2, K, v, c->stream));
    GBM3D_TRY(download(c, V1, v, mn * K));
    GBM3D_HIP(hipStreamSynchronize(c->stream));
  }
  if (V2) {
    GBM3D_TRY(launch_gather_groups(p2, dS, m, n, d, d2, K, v, c->stream));
    GBM3D_TRY(download(c, V2, v, mn * K));
  }
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_wiener(gbm3d_ctx* c, const float* V1, const float* V2, int d, int K, int G, float sigma, float* V2f,
                 float* gain2) {
  GBM3D_TRY(set_dev(c));
  if (!V1 || !V2 || !V2f || !gain2 || G < 0) return fail(GBM3D_ERR_ARG, "NULL argument");
  const size_t n = (size_t)d * d * K * G;
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** a, float** b, float** o, float** g) {
    *a = k.take<float>(n);
    *b = k.take<float>(n);
    *o = k.take<float>(n);
    *g = k.take<float>(G);
  };
  float *a, *bb, *o, *g;
  lay(cv, &a, &bb, &o, &g);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &a, &bb, &o, &g);
  GBM3D_TRY(upload(c, a, V1, n));
  GBM3D_TRY(upload(c, bb, V2, n));
  GBM3D_TRY(launch_wiener(nullptr, nullptr, nullptr, a, bb, 0, 0, d, 0, K, G, sigma, o, g, c->stream));
  GBM3D_TRY(download(c, V2f, o, n));
  GBM3D_TRY(download(c, gain2, g, (size_t)G));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_aggregate_wie(gbm3d_ctx* c, const float* V2f, const float* gain2, const uint32_t* S, int m, int n, int d,
                        int d2, int K, float* U, float* W) {
  GBM3D_TRY(set_dev(c));
  if (!V2f || !gain2 || !S || !U || !W) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, K));
  const size_t mn = (size_t)m * n, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** v, float** g, uint32_t** dS, float** dU, float** dW) {
    *v = k.take<float>(mn * K);
    *g = k.take<float>(G);
    *dS = k.take<uint32_t>(G * K);
    *dU = k.take<float>(mn);
    *dW = k.take<float>(mn);
  };
  float *v, *g, *dU, *dW;
  uint32_t* dS;
  lay(cv, &v, &g, &dS, &dU, &dW);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &v, &g, &dS, &dU, &dW);
  GBM3D_TRY(upload(c, v, V2f, mn * K));
  GBM3D_TRY(upload(c, g, gain2, G));
  GBM3D_TRY(upload(c, dS, S, G * K));
  GBM3D_TRY(launch_aggregate_m(2, v, g, dS, c->masks + kMaskStride * mask_slot(d), m, n, d, d2, K, 0, dU, dW,
                               c->stream));
  GBM3D_TRY(download(c, U, dU, mn));
  GBM3D_TRY(download(c, W, dW, mn));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
namespace {

struct BatchJob {
  int count;
  const float* const* dI;
  const float* const* dUo;
  const int *m, *n;
  const float* sigma;
  float* const* dU;
  float* const* dU1;
  std::vector<Plan> plans;
};

// Largest images first onto the least-loaded stream; per-stream workspace.
void assign_streams(const BatchJob& j, const gbm3d_opts& o, int S, std::vector<int>* assign, std::vector<size_t>* wsz) {
  std::vector<int> order(j.count);
  for (int i = 0; i < j.count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (int64_t)j.plans[a].M * j.plans[a].N > (int64_t)j.plans[b].M * j.plans[b].N;
  });
  std::vector<int64_t> load(S, 0);
  assign->assign(j.count, 0);
  wsz->assign(S, 0);
  for (int i : order) {
    int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    (*assign)[i] = s;
    load[s] += (int64_t)j.plans[i].M * j.plans[i].N;
    (*wsz)[s] = std::max((*wsz)[s], pipeline_bytes(j.plans[i], o));
  }
}

int issue_batch(gbm3d_ctx* c, const BatchJob& j, const gbm3d_opts& o, int S, int mw) {
  std::vector<int> assign;
  std::vector<size_t> wsz;
  assign_streams(j, o, S, &assign, &wsz);
  std::vector<size_t> off(S, 0);
  size_t tot = 0;
  for (int s = 0; s < S; ++s) {
    off[s] = tot;
    tot += (wsz[s] + 4095) & ~(size_t)4095;
  }
  if (tot > c->arena_bytes) return fail(GBM3D_ERR_NOMEM, "batch workspace not reserved");
  c->in_batch = true;
  c->match_waves = mw;
  int st = GBM3D_OK;
  if (hipEventRecord(c->bfork, c->stream) != hipSuccess) st = fail(GBM3D_ERR_HIP, "batch fork");
  for (int s = 0; s < S && st == GBM3D_OK; ++s)
    if (hipStreamWaitEvent(c->bstreams[s], c->bfork, 0) != hipSuccess) st = fail(GBM3D_ERR_HIP, "batch fork wait");
  for (int i = 0; i < j.count && st == GBM3D_OK; ++i) {
    const int s = assign[i];
    st = run_pipeline_on(c, j.dI[i], j.plans[i], j.sigma[i], o, j.dU[i], j.dU1 ? j.dU1[i] : nullptr, c->bstreams[s],
                         c->arena + off[s], image_stats(c, i));
  }
  for (int s = 0; s < S; ++s) {
    // join every branch even after an error so the capture stays well formed
    (void)hipEventRecord(c->bjoin[s], c->bstreams[s]);
    (void)hipStreamWaitEvent(c->stream, c->bjoin[s], 0);
  }
  c->in_batch = false;
  c->match_waves = 4;
  return st;
}

size_t batch_bytes(const BatchJob& j, const gbm3d_opts& o, int S) {
  std::vector<int> assign;
  std::vector<size_t> wsz;
  assign_streams(j, o, S, &assign, &wsz);
  size_t tot = 0;
  for (size_t w : wsz) tot += (w + 4095) & ~(size_t)4095;
  return tot;
}

std::vector<int64_t> batch_key(const BatchJob& j, const gbm3d_opts& o, const gbm3d_ext& e, int S, int mw) {
  std::vector<int64_t> k = {j.count, S, mw, e.match_bf16, e.rescore, e.learned, e.learned_bound};
  for (int i = 0; i < j.count; ++i) {
    uint32_t sb;
    std::memcpy(&sb, &j.sigma[i], 4);
    k.push_back(j.m[i]);
    k.push_back(j.n[i]);
    k.push_back(sb);
    k.push_back((int64_t)(intptr_t)j.dI[i]);
    k.push_back((int64_t)(intptr_t)j.dU[i]);
    k.push_back(j.dU1 ? (int64_t)(intptr_t)j.dU1[i] : 0);
  }
  const int32_t* ow = reinterpret_cast<const int32_t*>(&o);
  for (size_t i = 0; i < sizeof(o) / 4; ++i) k.push_back(ow[i]);
  return k;
}

int batch_run(gbm3d_ctx* c, BatchJob& j, const gbm3d_opts& o, int flags);

}  // namespace

extern "C" {

int gbm3d_denoise_batch_device(gbm3d_ctx* c, int count, const float* const* dI, const int* m, const int* n,
                               const float* sigma, const gbm3d_opts* opts, float* const* dU, float* const* dU1,
                               int flags) {
  GBM3D_TRY(set_dev(c));
  if (count < 1 || !dI || !m || !n || !sigma || !dU) return fail(GBM3D_ERR_ARG, "batch: NULL or empty");
  gbm3d_opts o;
  GBM3D_TRY(resolve(opts, &o));
  BatchJob j{count, dI, nullptr, m, n, sigma, dU, dU1, {}};
  j.plans.resize(count);
  for (int i = 0; i < count; ++i) {
    if (!dI[i] || !dU[i]) return fail(GBM3D_ERR_ARG, "batch: NULL image " + std::to_string(i));
    GBM3D_TRY(make_plan(o, m[i], n[i], &j.plans[i]));
  }
  GBM3D_TRY(learned_setup(c, o, count));
  if (c->ext.learned) {
    // learned form: run, then judge every image against the bound
    GBM3D_TRY(batch_run(c, j, o, flags));
    return learned_judge(c, o, count, j.plans, dI, sigma, dU, dU1);
  }
  return batch_run(c, j, o, flags);
}

int gbm3d_batch_info(gbm3d_ctx* c, int* streams, int* match_waves, double* tuned_ms) {
  GBM3D_TRY(set_dev(c));
  if (streams) *streams = c->tuned_streams;
  if (match_waves) *match_waves = c->tuned_match_waves;
  if (tuned_ms) *tuned_ms = c->tuned_ms;
  return GBM3D_OK;
}

}  // extern "C"

namespace {

int batch_run(gbm3d_ctx* c, BatchJob& j, const gbm3d_opts& o, int flags) {
  const int count = j.count;
  int S = std::min(count, 4), mw = 4;
  if (flags & GBM3D_BATCH_TUNE) {
    std::vector<int64_t> tk = batch_key(j, o, c->ext, 0, 0);
    if (tk != c->tkey) {
      hipEvent_t a, b;
      GBM3D_HIP(hipEventCreate(&a));
      GBM3D_HIP(hipEventCreate(&b));
      double best = 1e300;
      int bs = S, bm = mw;
      for (int cs : {1, 2, 4, 8}) {
        if (cs > count) continue;
        GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, batch_bytes(j, o, cs)));
        for (int cm : {1, 2, 4}) {
          GBM3D_TRY(issue_batch(c, j, o, cs, cm));  // warm
          GBM3D_HIP(hipEventRecord(a, c->stream));
          GBM3D_TRY(issue_batch(c, j, o, cs, cm));
          GBM3D_HIP(hipEventRecord(b, c->stream));
          GBM3D_HIP(hipEventSynchronize(b));
          float ms = 0.f;
          GBM3D_HIP(hipEventElapsedTime(&ms, a, b));
          if (ms < best) {
            best = ms;
            bs = cs;
            bm = cm;
          }
        }
      }
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      c->tkey = tk;
      c->tuned_streams = bs;
      c->tuned_match_waves = bm;
      c->tuned_ms = best;
    }
  }
  if (c->tuned_streams > 0 && c->tkey == batch_key(j, o, c->ext, 0, 0)) {
    S = c->tuned_streams;
    mw = c->tuned_match_waves;
  }
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, batch_bytes(j, o, S)));
  if (!(flags & GBM3D_BATCH_GRAPH)) return issue_batch(c, j, o, S, mw);
  std::vector<int64_t> key = batch_key(j, o, c->ext, S, mw);
  key.push_back((int64_t)(intptr_t)c->arena);  // a grown arena invalidates the captured addresses
  key.push_back((int64_t)(intptr_t)c->sstats);
  if (!c->gexec || key != c->gkey) {
    if (c->gexec) {
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    GBM3D_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
    const int st = issue_batch(c, j, o, S, mw);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (st != GBM3D_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (ec != hipSuccess) return fail(GBM3D_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      c->gexec = nullptr;
      return fail(GBM3D_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
    }
    c->gkey = key;
  }
  GBM3D_HIP(hipGraphLaunch(c->gexec, c->stream));
  return GBM3D_OK;
}

}  // namespace

// ---------------------------------------------------------- learned shrinkage
extern "C" {

int gbm3d_shrink_set_weights(gbm3d_ctx* c, const float* params, int count) {
  GBM3D_TRY(set_dev(c));
  if (!params || count != GBM3D_SHRINK_PARAMS)
    return fail(GBM3D_ERR_ARG, "shrink weights: expected " + std::to_string(GBM3D_SHRINK_PARAMS) + " values");
  for (int i = 0; i < count; ++i)
    if (!std::isfinite(params[i])) return fail(GBM3D_ERR_ARG, "shrink weights: non-finite value");
  std::vector<uint16_t> wb((size_t)shrink::NL * shrink::CH * shrink::KW);
  std::vector<float> fb(shrink::FB);
  shrink_pack_weights(params, wb.data(), fb.data());
  if (!c->shrink_wb) GBM3D_HIP(hipMalloc(reinterpret_cast<void**>(&c->shrink_wb), wb.size() * sizeof(uint16_t)));
  if (!c->shrink_fb) GBM3D_HIP(hipMalloc(reinterpret_cast<void**>(&c->shrink_fb), fb.size() * sizeof(float)));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  GBM3D_HIP(hipMemcpy(c->shrink_wb, wb.data(), wb.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  GBM3D_HIP(hipMemcpy(c->shrink_fb, fb.data(), fb.size() * sizeof(float), hipMemcpyHostToDevice));
  return GBM3D_OK;
}

int gbm3d_shrink_stats(gbm3d_ctx* c, int image, double* est_learned, double* est_ht, int* fell_back) {
  GBM3D_TRY(set_dev(c));
  if (image < 0 || image >= (int)c->s_learned.size()) return fail(GBM3D_ERR_ARG, "shrink stats: no such image");
  if (est_learned) *est_learned = c->s_learned[image];
  if (est_ht) *est_ht = c->s_ht[image];
  if (fell_back) *fell_back = c->s_fell[image];
  return GBM3D_OK;
}

int gbm3d_shrink_groups(gbm3d_ctx* c, const float* V, int G, float sigma, float* Vout, float* gains) {
  GBM3D_TRY(set_dev(c));
  if (!V || !Vout || G < 0) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (!c->shrink_wb) return fail(GBM3D_ERR_ARG, "learned shrinkage: no weights loaded (gbm3d_shrink_set_weights)");
  if (G == 0) return GBM3D_OK;
  const size_t n = (size_t)64 * 16 * G;
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** a, float** o, float** g) {
    *a = k.take<float>(n);
    *o = k.take<float>(n);
    *g = k.take<float>(gains ? n : 0);
  };
  float *a, *o, *g;
  lay(cv, &a, &o, &g);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &a, &o, &g);
  GBM3D_TRY(upload(c, a, V, n));
  GBM3D_TRY(launch_shrink(a, nullptr, 0, 0, G, sigma, c->shrink_wb, c->shrink_fb, o, gains ? g : nullptr, nullptr,
                          nullptr, c->stream));
  GBM3D_TRY(download(c, Vout, o, n));
  if (gains) GBM3D_TRY(download(c, gains, g, n));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_shrink_features(gbm3d_ctx* c, const float* dClean, const float* dNoisy, int m, int n, float sigma,
                          const uint32_t* dSel, int nsel, float* dX, float* dT, float* dWt) {
  GBM3D_TRY(set_dev(c));
  if (!dClean || !dNoisy || !dSel || !dX || !dT || !dWt || nsel < 0) return fail(GBM3D_ERR_ARG, "NULL argument");
  constexpr int d = 8, d2 = 32, K = 16;
  if (m % d || n % d || m < d2 || n < d2) return fail(GBM3D_ERR_ARG, "shrink features: m, n multiples of 8, >= 32");
  if (!(sigma > 0.f)) return fail(GBM3D_ERR_ARG, "shrink features: sigma must be > 0");
  const size_t P = (size_t)m + d2, Q = (size_t)n + d2, G = (size_t)m * n / (d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** pc, float** pn, float** bn, uint32_t** S) {
    *pc = k.take<float>(P * Q);
    *pn = k.take<float>(P * Q);
    *bn = k.take<float>((P - d + 1) * (Q - d + 1));
    *S = k.take<uint32_t>(G * K);
  };
  float *pc, *pn, *bn;
  uint32_t* S;
  lay(cv, &pc, &pn, &bn, &S);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &pc, &pn, &bn, &S);
  GBM3D_TRY(launch_search_pad(dClean, m, n, 0, d2, pc, c->stream));
  GBM3D_TRY(launch_search_pad(dNoisy, m, n, 0, d2, pn, c->stream));
  GBM3D_TRY(launch_norm_map(pn, (int)P, (int)Q, d, bn, c->stream));
  GBM3D_TRY(launch_match_w(pn, bn, m, n, d, d2, K, S, nullptr, c->match_waves, c->stream));
  return launch_shrink_features(pc, pn, S, m, d2, dSel, nsel, sigma, dX, dT, dWt, c->stream);
}

int gbm3d_shrink_forward(gbm3d_ctx* c, const float* dParams, const float* dX, const float* dSig, int ngroups,
                         float* dGains) {
  GBM3D_TRY(set_dev(c));
  if (!dParams || !dX || !dSig || !dGains || ngroups < 0) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, shrink_train_floats(ngroups) * sizeof(float)));
  return shrink_forward(dParams, dX, dSig, ngroups, reinterpret_cast<float*>(c->arena), dGains, c->stream);
}

int gbm3d_shrink_grad(gbm3d_ctx* c, const float* dParams, const float* dX, const float* dT, const float* dWt,
                      const float* dSig, int ngroups, float* dGrad, double* loss) {
  GBM3D_TRY(set_dev(c));
  if (!dParams || !dX || !dT || !dWt || !dSig || !dGrad) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (ngroups < 1) return fail(GBM3D_ERR_ARG, "shrink grad: empty batch");
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, shrink_train_floats(ngroups) * sizeof(float)));
  float* ws = reinterpret_cast<float*>(c->arena);
  int lb = 0;
  GBM3D_TRY(shrink_grad(dParams, dX, dT, dWt, dSig, ngroups, ws, dGrad, nullptr, &lb, c->stream));
  if (loss) {
    // the loss partials follow the gradient workspace (see shrink_train_floats)
    const size_t off = shrink_train_floats(ngroups) - 1024;
    std::vector<float> h(lb);
    GBM3D_HIP(hipMemcpyAsync(h.data(), ws + off, lb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    GBM3D_HIP(hipStreamSynchronize(c->stream));
    double acc = 0.0;
    for (float v : h) acc += v;
    *loss = acc / ngroups;
  }
  return GBM3D_OK;
}

int gbm3d_shrink_adam(gbm3d_ctx* c, float* dParams, float* dM, float* dV, const float* dGrad, int step, float lr,
                      float beta1, float beta2, float eps) {
  GBM3D_TRY(set_dev(c));
  if (!dParams || !dM || !dV || !dGrad) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (step < 1 || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return fail(GBM3D_ERR_ARG, "adam: step >= 1, 0 <= beta < 1, eps > 0");
  const double lr_t = lr * std::sqrt(1.0 - std::pow((double)beta2, step)) / (1.0 - std::pow((double)beta1, step));
  return launch_shrink_adam(dParams, dM, dV, dGrad, GBM3D_SHRINK_PARAMS, (float)lr_t, beta1, beta2, eps, c->stream);
}

int gbm3d_synth_pair(gbm3d_ctx* c, uint64_t seed, int m, int n, float sigma, float* dClean, float* dNoisy) {
  GBM3D_TRY(set_dev(c));
  if (m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "synth: empty image");
  SynthParams prm;
  synth_params(seed, m, n, &prm);
  return launch_synth_pair(prm, seed, m, n, sigma, dClean, dNoisy, c->stream);
}

}  // extern "C"

// ---------------------------------------------------------- distributed
// [U, out] = GBM3D_distributed(I, sigma, opts) (source/GBM3D_distributed.m):
// the image is cut into overlapping tiles of 256 + blockSize, each tile and
// match spin ("case") is denoised on its own, and the unnormalised U, W of
// the cases are stitched with ramp windows.  The cases are independent and
// run on the context's lane streams; one gather kernel per step stitches them
// in case order (deterministic).
namespace {

struct TileCase {
  int y1, y2, x1, x2, spin;  // 0-based, half-open
};

// getBM3D_distribution (utilities/getBM3D_distribution.m:10-48).
std::vector<TileCase> tile_cases(int m, int n, const int32_t* spins, int ns, int k, int* ny, int* nx) {
  const int d0 = 256 + k;
  const int Xcnt = (n - k + (d0 - k) - 1) / (d0 - k), Ycnt = (m - k + (d0 - k) - 1) / (d0 - k);
  *ny = Ycnt;
  *nx = Xcnt;
  std::vector<TileCase> R;
  for (int i = 0; i < Ycnt; ++i)
    for (int j = 0; j < Xcnt; ++j)
      for (int s = 0; s < ns; ++s) {
        int y1 = i * (d0 - k) + 1, y2 = i * (d0 - k) + d0, x1 = j * (d0 - k) + 1, x2 = j * (d0 - k) + d0;
        if (x2 >= n) x1 = std::min(n - (n - x1 + 1 + k - 1) / k * k + 1, n - 2 * k + 1);
        if (y2 >= m) y1 = std::min(m - (m - y1 + 1 + k - 1) / k * k + 1, m - 2 * k + 1);
        R.push_back({y1 - 1, std::min(y2, m), x1 - 1, std::min(x2, n), spins[s]});
      }
  return R;
}

constexpr int kMaxBands = 64;

struct StitchArgs {
  int m, n, k, ny, nx, ns;
  int ry1[kMaxBands], ry2[kMaxBands], cx1[kMaxBands], cx2[kMaxBands];
  int spins[GBM3D_MAX_SPINS];
  double ramp[32];  // linspace(.01, 1, k)
  size_t stride;    // floats per case in Uc / Wc
};

__global__ void tile_extract_kernel(const float* __restrict__ src, int m, int y1, int x1, int th, int tw, int spin,
                                    float* __restrict__ dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= th * tw) return;
  const int r = e % th, c = e / th;
  // circshift(I(tile), [spin, spin])
  dst[e] = src[(size_t)(y1 + pmod_i(r - spin, th)) + (size_t)m * (x1 + pmod_i(c - spin, tw))];
}

// U(y, x) = sum_cases w U_case / sum_cases w W_case over the cases covering
// (y, x), in case order; w = makeInterpWindow (utilities/makeInterpWindow.m),
// its factors applied in the reference's order; U_case read shifted back by
// the case's spin.
__global__ void stitch_kernel(StitchArgs a, const float* __restrict__ Uc, const float* __restrict__ Wc,
                              float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)a.m * a.n) return;
  const int y = (int)(e % a.m), x = (int)(e / a.m);
  double num = 0.0, den = 0.0;
  for (int i = 0; i < a.ny; ++i) {
    if (y < a.ry1[i] || y >= a.ry2[i]) continue;
    const int r = y - a.ry1[i], th = a.ry2[i] - a.ry1[i];
    for (int j = 0; j < a.nx; ++j) {
      if (x < a.cx1[j] || x >= a.cx2[j]) continue;
      const int c = x - a.cx1[j], tw = a.cx2[j] - a.cx1[j];
      double w = 1.0;
      if (r < a.k) w *= a.ramp[r];
      if (c < a.k) w *= a.ramp[c];
      if (r >= th - a.k) w *= a.ramp[th - 1 - r];
      if (c >= tw - a.k) w *= a.ramp[tw - 1 - c];
      for (int s = 0; s < a.ns; ++s) {
        const size_t cs = (size_t)((i * a.nx + j) * a.ns + s) * a.stride;
        const int sp = a.spins[s];
        const size_t src = (size_t)pmod_i(r + sp, th) + (size_t)th * pmod_i(c + sp, tw);
        num += (double)Uc[cs + src] * w;
        den += w * (double)Wc[cs + src];
      }
    }
  }
  out[e] = (float)(num / den);
}

int stitch(const std::vector<TileCase>& cases, int m, int n, int k, int ny, int nx, int ns, const int32_t* spins,
           size_t stride, const float* Uc, const float* Wc, float* out, hipStream_t st) {
  if (ny > kMaxBands || nx > kMaxBands) return fail(GBM3D_ERR_UNSUPPORTED, "distributed: more than 64 tile rows");
  if (k > 32) return fail(GBM3D_ERR_UNSUPPORTED, "distributed: blockSize > 32");
  StitchArgs a{};
  a.m = m;
  a.n = n;
  a.k = k;
  a.ny = ny;
  a.nx = nx;
  a.ns = ns;
  for (int i = 0; i < ny; ++i) {
    a.ry1[i] = cases[(size_t)i * nx * ns].y1;
    a.ry2[i] = cases[(size_t)i * nx * ns].y2;
  }
  for (int j = 0; j < nx; ++j) {
    a.cx1[j] = cases[(size_t)j * ns].x1;
    a.cx2[j] = cases[(size_t)j * ns].x2;
  }
  for (int s = 0; s < ns; ++s) a.spins[s] = spins[s];
  const std::vector<double> r = linspace(0.01, 1.0, k);
  for (int i = 0; i < k; ++i) a.ramp[i] = r[i];
  a.stride = stride;
  const size_t total = (size_t)m * n;
  hipLaunchKernelGGL(stitch_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a, Uc, Wc, out);
  return check_launch("stitch");
}

// need != nullptr: only report the workspace bytes (nothing is issued).
int run_distributed(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts& o, float* dU,
                    float* dU2, size_t* need = nullptr) {
  const int k = o.blockSize;
  if (m < 256 + k || n < 256 + k) {  // :22-25
    if (need) {
      *need = 0;
      return GBM3D_OK;
    }
    return run_pipeline(c, dI, m, n, sigma, o, dU, dU2);
  }
  if (c->ext.learned) return fail(GBM3D_ERR_UNSUPPORTED, "distributed: the learned form is not supported");
  int dm = 0, dn = 0;
  GBM3D_TRY(validate(o, 256 + k, 256 + k, &dm, &dn));
  int ny1, nx1, ny2 = 0, nx2 = 0;
  const std::vector<TileCase> c1 = tile_cases(m, n, o.matchSpins, o.nMatchSpins, k, &ny1, &nx1);
  std::vector<TileCase> c2;
  if (o.Wiener) c2 = tile_cases(m, n, o.matchSpinsWie, o.nMatchSpinsWie, k, &ny2, &nx2);
  // every tile must be a valid image for the per-tile stages
  for (const std::vector<TileCase>* cs : {&c1, static_cast<const std::vector<TileCase>*>(&c2)})
    for (const TileCase& t : *cs) {
      int tM, tN;
      GBM3D_TRY(validate(o, t.y2 - t.y1, t.x2 - t.x1, &tM, &tN));
      if (tM != t.y2 - t.y1 || tN != t.x2 - t.x1) return fail(GBM3D_ERR_UNSUPPORTED, "distributed: tile size");
    }
  const int d0 = 256 + k;
  Plan tp;  // the largest tile
  GBM3D_TRY(make_plan(o, d0, d0, &tp));
  const int lanes = c->timing ? 1 : kMaxLanes;
  const size_t stride = (size_t)d0 * d0, ncase = std::max(c1.size(), c2.size());
  Carve cv{nullptr};
  auto lay = [&](Carve& kv, float** Uc, float** Wc, float** U2, float** tI, float** tP, SpinBufs* sb) {
    *Uc = kv.take<float>(stride * ncase);
    *Wc = kv.take<float>(stride * ncase);
    *U2 = kv.take<float>((size_t)m * n);
    for (int l = 0; l < lanes; ++l) {
      tI[l] = kv.take<float>(stride);
      tP[l] = kv.take<float>(stride);
      carve_spin(kv, tp, o, &sb[l]);
    }
  };
  float *Uc, *Wc, *U2, *tI[kMaxLanes], *tP[kMaxLanes];
  SpinBufs sb[kMaxLanes];
  lay(cv, &Uc, &Wc, &U2, tI, tP, sb);
  if (need) {
    *need = cv.off + 256;
    return GBM3D_OK;
  }
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off + 256));
  Carve k2{c->arena};
  lay(k2, &Uc, &Wc, &U2, tI, tP, sb);
  const hipStream_t st = c->stream;
  std::vector<double> tau((size_t)o.levels * 7);
  tau_table(o.levels, o.tauMode, (double)sigma, tau.data());
  const float* masks1 = c->masks + kMaskStride * mask_slot(o.blockSize);
  auto extract = [&](const float* src, const TileCase& t, float* dst, hipStream_t ls) {
    const int th = t.y2 - t.y1, tw = t.x2 - t.x1;
    hipLaunchKernelGGL(tile_extract_kernel, dim3((th * tw + 255) / 256), dim3(256), 0, ls, src, m, t.y1, t.x1, th,
                       tw, t.spin, dst);
    return check_launch("tile_extract");
  };
  // ---- step 1 (:53-61): matchBlocksCPU + denoiseAndAggregateGPU per case
  const int l1 = std::max(1, std::min(lanes, (int)c1.size()));
  GBM3D_TRY(lanes_fork(c, l1, st));
  for (size_t i = 0; i < c1.size(); ++i) {
    const TileCase& t = c1[i];
    const int th = t.y2 - t.y1, tw = t.x2 - t.x1, l = (int)(i % l1);
    const hipStream_t ls = l1 > 1 ? c->bstreams[l] : st;
    SpinBufs& b = sb[l];
    GBM3D_TRY(extract(dI, t, tI[l], ls));
    {
      StageTimer tm(c, GBM3D_STAGE_PAD);
      GBM3D_TRY(launch_search_pad(tI[l], th, tw, 0, o.matchSize, b.pad1, ls));
    }
    GBM3D_TRY(do_match(c, b.pad1, b.bn, b.smap, th, tw, k, o.matchSize, o.numMax, b.S1, b.vol, ls,
                       GBM3D_STAGE_MATCH1));
    if (o.filtType == GBM3D_FILT_MEDIAN) {
      StageTimer tm(c, GBM3D_STAGE_HT_SYN0);
      GBM3D_TRY(launch_median3(b.vol, th, tw, o.numMax, k, b.vavg, ls));
    } else {
      HtWork hw;
      ht_work_bind(b.htw, th, tw, o.numMax, o.levels, o.cycleSpin, &hw);
      GBM3D_TRY(run_ht(b.vol, th, tw, o.numMax, o.levels, o.cycleSpin, tau.data(), hw, b.vavg, k, ls, ht_hook, c));
    }
    {
      StageTimer tm(c, GBM3D_STAGE_AGG1_PREP);
      GBM3D_TRY(launch_agg1_prep(b.vavg, th, tw, o.numMax, k, b.w1, ls));
    }
    StageTimer tm(c, GBM3D_STAGE_AGG1);
    GBM3D_TRY(launch_aggregate_m(1, b.vavg, b.w1, b.S1, masks1, th, tw, k, o.matchSize, o.numMax, 0,
                                 Uc + stride * i, Wc + stride * i, ls));
  }
  GBM3D_TRY(lanes_join(c, l1, st));
  {
    StageTimer tm(c, GBM3D_STAGE_NORM);
    GBM3D_TRY(stitch(c1, m, n, k, ny1, nx1, o.nMatchSpins, o.matchSpins, stride, Uc, Wc, U2, st));  // :65-79
    if (dU2) GBM3D_HIP(hipMemcpyAsync(dU2, U2, (size_t)m * n * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (!o.Wiener) {  // :81-84
      GBM3D_HIP(hipMemcpyAsync(dU, U2, (size_t)m * n * sizeof(float), hipMemcpyDeviceToDevice, st));
      return GBM3D_OK;
    }
  }
  // ---- step 2 (:94-104): matchBlocksWie on the U2 tile + denoiseAndAggregateWieFast
  const int bw = o.blockSizeWie;
  const float* masks2 = c->masks + kMaskStride * mask_slot(bw);
  const int l2 = std::max(1, std::min(lanes, (int)c2.size()));
  GBM3D_TRY(lanes_fork(c, l2, st));
  for (size_t i = 0; i < c2.size(); ++i) {
    const TileCase& t = c2[i];
    const int th = t.y2 - t.y1, tw = t.x2 - t.x1, l = (int)(i % l2);
    const hipStream_t ls = l2 > 1 ? c->bstreams[l] : st;
    SpinBufs& b = sb[l];
    GBM3D_TRY(extract(dI, t, tI[l], ls));
    GBM3D_TRY(extract(U2, t, tP[l], ls));
    {
      StageTimer tm(c, GBM3D_STAGE_PAD);
      GBM3D_TRY(launch_search_pad(tP[l], th, tw, 0, o.matchSize, b.pad1, ls));
      GBM3D_TRY(launch_search_pad(tI[l], th, tw, 0, o.matchSize, b.pad2, ls));
    }
    GBM3D_TRY(do_match(c, b.pad1, b.bn, b.smap, th, tw, bw, o.matchSize, o.numMax, b.S2, nullptr, ls,
                       GBM3D_STAGE_MATCH2));
    const long G2 = (long)(th / bw) * (tw / bw);
    {
      StageTimer tm(c, GBM3D_STAGE_WIENER);
      GBM3D_TRY(launch_wiener(b.pad1, b.pad2, b.S2, nullptr, nullptr, th, tw, bw, o.matchSize, o.numMax, G2, sigma,
                              b.V2f, b.g2, ls));
    }
    StageTimer tm(c, GBM3D_STAGE_AGG2);
    GBM3D_TRY(launch_aggregate_m(2, b.V2f, b.g2, b.S2, masks2, th, tw, bw, o.matchSize, o.numMax, 0,
                                 Uc + stride * i, Wc + stride * i, ls));
  }
  GBM3D_TRY(lanes_join(c, l2, st));
  StageTimer tm(c, GBM3D_STAGE_NORM);
  return stitch(c2, m, n, k, ny2, nx2, o.nMatchSpinsWie, o.matchSpinsWie, stride, Uc, Wc, dU, st);  // :107-121
}

}  // namespace

extern "C" {

int gbm3d_denoise_distributed_device(gbm3d_ctx* c, const float* dI, int m, int n, float sigma,
                                     const gbm3d_opts* opts, float* dU, float* dU2) {
  GBM3D_TRY(set_dev(c));
  if (!dI || !dU || m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "NULL or empty image");
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  if (c->timing) {
    c->marks.clear();
    c->pool_used = 0;
    return run_distributed(c, dI, m, n, sigma, o, dU, dU2);
  }
  size_t need = 0;
  GBM3D_TRY(run_distributed(c, dI, m, n, sigma, o, dU, dU2, &need));
  if (need == 0) return run_distributed(c, dI, m, n, sigma, o, dU, dU2);  // small image: GBM3D
  // the per-tile cases are hundreds of small launches: capture them once into
  // a graph (workspace reserved beforehand) and replay it
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, need));
  uint32_t sb;
  std::memcpy(&sb, &sigma, 4);
  std::vector<int64_t> key = {m, n, sb, (int64_t)(intptr_t)dI, (int64_t)(intptr_t)dU, (int64_t)(intptr_t)dU2,
                              (int64_t)(intptr_t)c->arena, c->ext.match_bf16, c->ext.rescore};
  const int32_t* ow = reinterpret_cast<const int32_t*>(&o);
  for (size_t i = 0; i < sizeof(o) / 4; ++i) key.push_back(ow[i]);
  if (!c->dexec || key != c->dkey) {
    if (c->dexec) {
      (void)hipGraphExecDestroy(c->dexec);
      c->dexec = nullptr;
    }
    GBM3D_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
    const int st = run_distributed(c, dI, m, n, sigma, o, dU, dU2);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (st != GBM3D_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (ec != hipSuccess) return fail(GBM3D_ERR_HIP, std::string("distributed capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&c->dexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      c->dexec = nullptr;
      return fail(GBM3D_ERR_HIP, std::string("distributed graph instantiate: ") + hipGetErrorString(ei));
    }
    c->dkey = key;
  }
  GBM3D_HIP(hipGraphLaunch(c->dexec, c->stream));
  return GBM3D_OK;
}

int gbm3d_denoise_distributed(gbm3d_ctx* c, const float* I, int m, int n, float sigma, const gbm3d_opts* opts,
                              float* U, float* U2) {
  GBM3D_TRY(set_dev(c));
  if (!I || !U || m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "NULL or empty image");
  const size_t mn = (size_t)m * n;
  GBM3D_TRY(ensure(&c->io, &c->io_bytes, 3 * mn * sizeof(float) + 1024));
  float* dI = reinterpret_cast<float*>(c->io);
  float* dU = dI + mn;
  float* dU2 = dU + mn;
  GBM3D_TRY(upload(c, dI, I, mn));
  GBM3D_TRY(gbm3d_denoise_distributed_device(c, dI, m, n, sigma, opts, dU, U2 ? dU2 : nullptr));
  GBM3D_TRY(download(c, U, dU, mn));
  if (U2) GBM3D_TRY(download(c, U2, dU2, mn));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------- coloured noise
extern "C" {

int gbm3d_denoise_colored(gbm3d_ctx* c, const float* I, int m, int n, const double* sigmaPSD, int d1, int d2,
                          const double* sigmaWie, const gbm3d_opts* opts, float* U, float* U1) {
  GBM3D_TRY(set_dev(c));
  if (!I || !U || !sigmaPSD || m < 1 || n < 1 || d1 < 1 || d2 < 1) return fail(GBM3D_ERR_ARG, "NULL or empty input");
  if (c->ext.learned) return fail(GBM3D_ERR_UNSUPPORTED, "coloured noise: the learned form is not supported");
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  if (o.Wiener && !sigmaWie)
    return fail(GBM3D_ERR_UNSUPPORTED, "coloured noise: sigmaWie (getColoredWie) must be given");
  std::vector<double> tau((size_t)o.levels * 7);
  colored_tau(o.levels, sigmaPSD, d1, d2, tau.data());
  std::vector<float> s2;
  if (o.Wiener) {
    const int bw = o.blockSizeWie;
    s2.resize((size_t)bw * bw);
    for (int e = 0; e < bw * bw; ++e) s2[e] = (float)(sigmaWie[e] * sigmaWie[e]);
  }
  c->colored = true;
  c->ctau = tau;
  c->wie_s2 = s2;
  // the scalar sigma is not read on this path (matching uses it only for the
  // inactive tolerance branch, matchBlocksCPU.m:20-22)
  const int st = gbm3d_denoise(c, I, m, n, 1.0f, &o, U, U1);
  c->colored = false;
  return st;
}

}  // extern "C"
