// pipeline.hip -- the two-step pipeline of source/GBM3D.m:21-83 on one
// image: option validation against the built kernels, workspace layout, the
// match spins of both steps on lane streams, and the gbm3d_denoise* /
// gbm3d_denoise_colored entry points.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"

namespace gbm3d {

std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> y(n);
  const double step = n > 1 ? (b - a) / (n - 1) : 0.0;
  for (int k = 0; k < n; ++k) y[k] = a + k * step;
  y[n - 1] = b;
  return y;
}

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

int validate(const gbm3d_opts& o, int m, int n, int* M, int* N) {
  if (m < 1 || n < 1) return fail(GBM3D_ERR_ARG, "image must be non-empty");
  if (o.filtType != GBM3D_FILT_HT && o.filtType != GBM3D_FILT_MEDIAN)
    return fail(GBM3D_ERR_ARG, "unrecognized filtering");  // denoiseAndAggregateGPU.m:54
  if (o.numMin < 1 || o.numMin > o.numMax) return fail(GBM3D_ERR_ARG, "numMin must be in [1, numMax]");
  if (o.nMatchSpins < 1 || o.nMatchSpins > GBM3D_MAX_SPINS) return fail(GBM3D_ERR_ARG, "matchSpins: 1..16 entries");
  if (o.Wiener) {
    if (o.nMatchSpinsWie == GBM3D_UNSET) return fail(GBM3D_ERR_MISSING_FIELD, "Reference to non-existent field 'matchSpinsWie'");
    if (o.nMatchSpinsWie < 1 || o.nMatchSpinsWie > GBM3D_MAX_SPINS) return fail(GBM3D_ERR_ARG, "matchSpinsWie: 1..16 entries");
  }
  const int bs = o.blockSize, d2 = o.matchSize, K = o.numMax;
  if (mask_slot(bs) < 0) return fail(GBM3D_ERR_UNSUPPORTED, "blockSize must be 4, 8, 16 or 32");
  if (d2 != 32 && d2 != 48) return fail(GBM3D_ERR_UNSUPPORTED, "matchSize must be 32 or 48");
  if (!match_supported(bs, d2)) return fail(GBM3D_ERR_UNSUPPORTED, "matching: (blockSize, matchSize) not built");
  if (o.levels < 1 || o.levels > 4) return fail(GBM3D_ERR_UNSUPPORTED, "levels must be 1..4");
  if (o.cycleSpin < 1 || o.cycleSpin > 8) return fail(GBM3D_ERR_ARG, "cycleSpin must be 1..8");
  if (K < 1 || K > 64 || K % (1 << o.levels)) return fail(GBM3D_ERR_UNSUPPORTED, "numMax must be <= 64 and a multiple of 2^levels");
  if (o.Wiener) {
    if (!match_supported(o.blockSizeWie, d2))
      return fail(GBM3D_ERR_UNSUPPORTED, "blockSizeWie must be 4, 8 or 16 (with matchSize 32 or 48)");
    if (!wiener_supported(o.blockSizeWie, K))
      return fail(GBM3D_ERR_UNSUPPORTED, "Wiener step: (blockSizeWie, numMax) must be (4|8|16, 8|16|32)");
  }
  *M = (m + bs - 1) / bs * bs;
  *N = (n + bs - 1) / bs * bs;
  if (*M > 2 * m || *N > 2 * n) return fail(GBM3D_ERR_ARG, "image smaller than half a block");
  if (*M < d2 || *N < d2) return fail(GBM3D_ERR_ARG, "padded image smaller than matchSize");
  if (*M % (1 << o.levels) || *N % (1 << o.levels)) return fail(GBM3D_ERR_UNSUPPORTED, "image size not a multiple of 2^levels");
  if (o.Wiener && (*M % o.blockSizeWie || *N % o.blockSizeWie))
    return fail(GBM3D_ERR_UNSUPPORTED, "padded size must be a multiple of blockSizeWie");
  return GBM3D_OK;
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

size_t pipeline_bytes(const Plan& p, const gbm3d_opts& o, int lanes) {
  Carve measure{nullptr};
  Bufs b;
  carve(measure, p, o, lanes, &b);
  return measure.off + 256;
}

// Lanes of a single-image call: its match spins are independent until the
// spin sums are normalised (source/GBM3D.m:50-57, 71-82), so they run on
// parallel streams; serial under stage timing (the timers record on the
// context stream).
int spin_lanes(const gbm3d_ctx* c, const Plan& p) {
  if (c->timing) return 1;
  return std::max(1, std::min(kMaxLanes, std::max(p.n1, p.n2)));
}

int do_match(gbm3d_ctx* c, const float* pad, float* bn, float* smap, int M, int N, int d, int d2, int K, uint32_t* S,
             float* vol, hipStream_t st, int stage, const TolArgs* ta, bool bn_ready) {
  const int P = M + d2, Q = N + d2;
  // the BF16 form is built for 8 x 8 and 16 x 16 blocks and the windowed
  // mink; other block sizes and the tolerance branch take the FP32 search,
  // whose index lists the BF16 form reproduces anyway
  const bool bf16 = c->ext.match_bf16 && match_bf16_supported(d, d2) && !(ta && ta->Kmin);
  {
    StageTimer t(c, GBM3D_STAGE_PAD);
    if (!bn_ready) GBM3D_TRY(launch_norm_map(pad, P, Q, d, bn, st));
    if (bf16) GBM3D_TRY(launch_sum_map(pad, P, Q, d, smap, st));
  }
  StageTimer t(c, stage);
  if (bf16) {
    const int R = c->ext.rescore > 0 ? std::max(c->ext.rescore, K) : 2 * K;
    return launch_match_bf16(pad, bn, smap, M, N, d, d2, K, std::min(R, 64), S, vol, st,
                             c->ext.match_stats ? c->bf16_stats : nullptr);
  }
  return launch_match_w(pad, bn, M, N, d, d2, K, S, vol, c->match_waves, st, ta);
}

int lanes_fork(gbm3d_ctx* c, int lanes, hipStream_t st) {
  if (lanes <= 1) return GBM3D_OK;
  GBM3D_HIP(hipEventRecord(c->bfork, st));
  for (int l = 0; l < lanes; ++l) GBM3D_HIP(hipStreamWaitEvent(c->bstreams[l], c->bfork, 0));
  return GBM3D_OK;
}

int lanes_join(gbm3d_ctx* c, int lanes, hipStream_t st) {
  if (lanes <= 1) return GBM3D_OK;
  int status = GBM3D_OK;
  for (int l = 0; l < lanes; ++l) {
    // every lane is joined even when one of the calls fails
    if (hipEventRecord(c->bjoin[l], c->bstreams[l]) != hipSuccess ||
        hipStreamWaitEvent(st, c->bjoin[l], 0) != hipSuccess)
      status = fail(GBM3D_ERR_HIP, "lane join");
  }
  return status;
}

int run_pipeline_on(gbm3d_ctx* c, const float* dI, const Plan& p, float sigma, const gbm3d_opts& o, float* dU,
                    float* dU1, hipStream_t st, char* ws, const LearnedSlot* lslot, int lanes) {
  const int m = p.m, n = p.n;
  Carve cv{ws};
  Bufs b;
  carve(cv, p, o, lanes, &b);
  const size_t MN = p.MN();
  double tau[4 * 7];
  if (c->colored)
    std::copy(c->ctau.begin(), c->ctau.end(), tau);  // getColoredTau (source/GBM3D.m:33-35)
  else
    tau_table(o.levels, o.tauMode, (double)sigma, tau);
  const float* masks1 = c->masks + kMaskStride * mask_slot(p.bs);
  // spin ii of a step with `nl` lanes runs on lane ii % nl
  auto lane_stream = [&](int ii, int nl) { return nl > 1 ? c->bstreams[ii % nl] : st; };

  if (p.M == m && p.N == n && dI != dU1 && dI != dU) {
    // already a block multiple: the padded image is the input itself (read
    // only; not when an output overwrites it before step 2 reads it)
    b.Ip = const_cast<float*>(dI);
  } else {
    StageTimer t(c, GBM3D_STAGE_PAD);
    GBM3D_TRY(launch_mirror_pad(dI, m, n, p.M, p.N, b.Ip, st));  // source/GBM3D.m:22-25
  }
  // ---- step 1: source/GBM3D.m:50-57
  // The noisy image's search image of spin s is built into a lane's pad2 and
  // left there: a step-2 spin with the same s on that lane reuses it.
  int noisy_spin[kMaxLanes];
  for (int l = 0; l < kMaxLanes; ++l) noisy_spin[l] = INT_MIN;
  // numMax ~= numMin: matchBlocksCPU's tolerance branch in step 1 (step 2,
  // matchBlocksWie, always keeps the numMax nearest), tol = 0.005 sigma^2 d^2
  // with sigma = mean(sqrt(PSD)) for coloured noise (matchBlocksCPU.m:21-28)
  TolArgs tol{0, 0.f, nullptr};
  if (o.numMin != o.numMax) {
    const double sg = c->colored ? c->tol_sigma : (double)sigma;
    const double t = 0.005 * sg * sg * p.bs * p.bs;
    float tu = (float)t;
    if ((double)tu < t) tu = nextafterf(tu, __builtin_inff());
    tol = TolArgs{o.numMin, tu, c->tolflag};
    GBM3D_TRY(launch_set_i32(c->tolflag, 0, st));
  }
  const int l1 = std::max(1, std::min(lanes, p.n1));
  GBM3D_TRY(on_lanes(c, p.n1, l1, st, [&](int ii) {
    const int sp = o.matchSpins[ii];
    const hipStream_t ls = lane_stream(ii, l1);
    SpinBufs& sb = b.sb[ii % l1];
    {
      StageTimer t(c, GBM3D_STAGE_PAD);
      // the search image and the block norms its matching takes
      GBM3D_TRY(launch_search_pad_norm(b.Ip, p.M, p.N, sp, p.d2, p.bs, sb.pad2, sb.bn, ls));
      noisy_spin[ii % l1] = sp;
    }
    bool wdone = false;  // the step-1 weights written with Vavg
    if (c->ext.learned) {
      // variant: learned gains on the DCT-Haar coefficients of each group,
      // gathered straight from the search image (no mosaic volume); then the
      // judge on the device, and the reference filter as gated launches that
      // run only when the judge falls back (they rebuild the volume and
      // overwrite vavg exactly as the reference path computes them)
      if (!lslot) return fail(GBM3D_ERR_ARG, "learned shrinkage: no device state for this image");
      GBM3D_TRY(do_match(c, sb.pad2, sb.bn, sb.smap, p.M, p.N, p.bs, p.d2, p.K, sb.S1, nullptr, ls,
                         GBM3D_STAGE_MATCH1, nullptr, true));
      {
        StageTimer t(c, GBM3D_STAGE_SHRINK);
        const long G1 = (long)(p.M / p.bs) * (p.N / p.bs);
        float* stats = lslot->stats + (size_t)2 * kShrinkGrid * ii;
        int grid = 0;
        GBM3D_TRY(launch_shrink(sb.pad2, sb.S1, p.M, p.d2, G1, sigma, c->shrink_wb, c->shrink_fb, sb.vavg, nullptr,
                                stats, &grid, ls));
        const double noise = (double)p.M * p.N * p.K * (double)sigma * sigma;
        const double bound = (c->ext.learned_bound > 0 ? c->ext.learned_bound : 1000) / 1000.0;
        GBM3D_TRY(launch_learned_judge(stats, grid, noise, bound, lslot->run + ii, lslot->res + 3 * ii, ls));
      }
      HtWork hw;
      ht_work_bind(sb.htw, p.M, p.N, p.K, o.levels, o.cycleSpin, &hw);
      GBM3D_TRY(launch_mosaic_gather(sb.pad2, sb.S1, p.M, p.N, p.bs, p.d2, p.K, sb.vol, lslot->run + ii, ls));
      GBM3D_TRY(run_ht(sb.vol, p.M, p.N, p.K, o.levels, o.cycleSpin, tau, hw, sb.vavg, p.bs, ls, ht_hook, c,
                       lslot->run + ii));
    } else {
      HtWork hw;
      ht_work_bind(sb.htw, p.M, p.N, p.K, o.levels, o.cycleSpin, &hw);
      GBM3D_TRY(do_match(c, sb.pad2, sb.bn, sb.smap, p.M, p.N, p.bs, p.d2, p.K, sb.S1, sb.vol, ls,
                         GBM3D_STAGE_MATCH1, tol.Kmin ? &tol : nullptr, true));
      if (o.filtType == GBM3D_FILT_MEDIAN) {
        StageTimer t(c, GBM3D_STAGE_HT_SYN0);
        GBM3D_TRY(launch_median3(sb.vol, p.M, p.N, p.K, p.bs, sb.vavg, ls));  // denoiseAndAggregateGPU.m:49-52
      } else {
        // the fused finest level writes the step-1 weights too where it can
        GBM3D_TRY(run_ht(sb.vol, p.M, p.N, p.K, o.levels, o.cycleSpin, tau, hw, sb.vavg, p.bs, ls, ht_hook, c,
                         nullptr, sb.w1, &wdone));
      }
    }
    if (!wdone) {
      StageTimer t(c, GBM3D_STAGE_AGG1_PREP);
      GBM3D_TRY(launch_agg1_prep(sb.vavg, p.M, p.N, p.K, p.bs, sb.w1, ls));
    }
    StageTimer t(c, GBM3D_STAGE_AGG1);
    return launch_aggregate_m(1, sb.vavg, sb.w1, sb.S1, masks1, p.M, p.N, p.bs, p.d2, p.K, sp, b.U1s + MN * ii,
                              b.W1s + MN * ii, ls);
  }));
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
  GBM3D_TRY(on_lanes(c, p.n2, l2, st, [&](int ii) {
    const int sp = o.matchSpinsWie[ii];
    const hipStream_t ls = lane_stream(ii, l2);
    SpinBufs& sb = b.sb[ii % l2];
    {
      StageTimer t(c, GBM3D_STAGE_PAD);
      GBM3D_TRY(launch_search_pad_norm(b.pilot, p.M, p.N, sp, p.d2, p.bw, sb.pad1, sb.bn, ls));
      if (noisy_spin[ii % l2] != sp) {  // else step 1 left this search image in pad2 (the steps are joined)
        GBM3D_TRY(launch_search_pad(b.Ip, p.M, p.N, sp, p.d2, sb.pad2, ls));
        noisy_spin[ii % l2] = sp;
      }
    }
    GBM3D_TRY(do_match(c, sb.pad1, sb.bn, sb.smap, p.M, p.N, p.bw, p.d2, p.K, sb.S2, nullptr, ls,
                       GBM3D_STAGE_MATCH2, nullptr, true));
    {
      StageTimer t(c, GBM3D_STAGE_WIENER);
      GBM3D_TRY(launch_wiener(sb.pad1, sb.pad2, sb.S2, nullptr, nullptr, p.M, p.N, p.bw, p.d2, p.K, G2, sigma,
                              sb.V2f, sb.g2, ls, c->colored ? c->wie_s2.data() : nullptr));
    }
    StageTimer t(c, GBM3D_STAGE_AGG2);
    return launch_aggregate_m(2, sb.V2f, sb.g2, sb.S2, masks2, p.M, p.N, p.bw, p.d2, p.K, sp, b.U2s + MN * ii,
                              b.W2s + MN * ii, ls);
  }));
  // U = sum(U,3)./sum(W,3): slices past numel(matchSpinsWie) still hold the
  // step-1 weights (U, W are sized by numel(matchSpins), :26-27).
  StageTimer t(c, GBM3D_STAGE_NORM);
  const int extra = std::max(0, p.n1 - p.n2);
  return launch_normalize(b.U2s, p.n2, b.W2s, p.n2, b.W1s + MN * p.n2, extra, p.M, p.N, m, n, dU, st);
}

int run_pipeline(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts& o, float* dU,
                 float* dU1) {
  Plan p;
  GBM3D_TRY(make_plan(o, m, n, &p));
  GBM3D_TRY(learned_setup(c, o, std::vector<Plan>{p}));
  const int lanes = spin_lanes(c, p);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, pipeline_bytes(p, o, lanes)));
  const LearnedSlot slot = c->ext.learned ? learned_slot(c, 0) : LearnedSlot{};
  GBM3D_TRY(run_pipeline_on(c, dI, p, sigma, o, dU, dU1, c->stream, c->arena, c->ext.learned ? &slot : nullptr, lanes));
  if (o.numMin == o.numMax) return GBM3D_OK;
  // the tolerance branch: whether a tile came out short is known only after
  // matching, so this call waits for the stream
  int short_tile = 0;
  GBM3D_HIP(hipMemcpyAsync(&short_tile, c->tolflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  if (short_tile)
    return fail(GBM3D_ERR_ARG,
                "tolerance matching: a reference block has fewer than numMax matches below tol, where "
                "matchBlocksCPU fails (Volume(.., 1:numel(S)) = blocks, matchBlocksCPU.m:111)");
  return GBM3D_OK;
}

int set_dev(gbm3d_ctx* c) {
  if (!c) return fail(GBM3D_ERR_ARG, "NULL context");
  GBM3D_HIP(hipSetDevice(c->device));
  clear_error();
  return GBM3D_OK;
}

int resolve_checked(const gbm3d_opts* in, gbm3d_opts* out) {
  GBM3D_TRY(resolve(in, out));
  return GBM3D_OK;
}

}  // namespace gbm3d

using namespace gbm3d;

extern "C" {

int gbm3d_validate(const gbm3d_opts* opts, int m, int n, size_t* workspace_bytes) {
  clear_error();
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  Plan p;
  GBM3D_TRY(make_plan(o, m, n, &p));
  if (workspace_bytes) *workspace_bytes = pipeline_bytes(p, o, 1);
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

// [U, out] = GBM3D(I, sigmaPSD, opts) with a matrix sigma (source/GBM3D.m:29-36).
int gbm3d_denoise_colored(gbm3d_ctx* c, const float* I, int m, int n, const double* sigmaPSD, int d1, int d2,
                          const double* sigmaWie, int wm, int wn, const gbm3d_opts* opts, float* U, float* U1) {
  GBM3D_TRY(set_dev(c));
  if (!I || !U || !sigmaPSD || m < 1 || n < 1 || d1 < 1 || d2 < 1) return fail(GBM3D_ERR_ARG, "NULL or empty input");
  if (c->ext.learned) return fail(GBM3D_ERR_UNSUPPORTED, "coloured noise: the learned form is not supported");
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  // sigmaWie = getColoredWie(sigma, blockSizeWie) when the caller gives none
  // (source/GBM3D.m:35)
  std::vector<double> swie;
  if (o.Wiener && !sigmaWie) {
    swie.resize((size_t)o.blockSizeWie * o.blockSizeWie);
    colored_wie(sigmaPSD, d1, d2, o.blockSizeWie, swie.data());
    sigmaWie = swie.data();
    wm = wn = o.blockSizeWie;
  }
  // sigmaWie multiplies the blockSizeWie x blockSizeWie DCT coefficients of
  // every slice (denoiseAndAggregateWieFast.m:49): its size must match
  if (o.Wiener && (wm != o.blockSizeWie || wn != o.blockSizeWie))
    return fail(GBM3D_ERR_ARG, "coloured noise: sigmaWie must be blockSizeWie x blockSizeWie (" +
                                   std::to_string(o.blockSizeWie) + " x " + std::to_string(o.blockSizeWie) +
                                   "), got " + std::to_string(wm) + " x " + std::to_string(wn));
  if (o.levels < 1 || o.levels > 4) return fail(GBM3D_ERR_UNSUPPORTED, "levels must be 1..4");
  std::vector<double> tau((size_t)o.levels * 7);
  colored_tau(o.levels, sigmaPSD, d1, d2, tau.data());
  std::vector<float> s2;
  if (o.Wiener) {
    const int bw = o.blockSizeWie;
    s2.resize((size_t)bw * bw);
    for (int e = 0; e < bw * bw; ++e) s2[e] = (float)(sigmaWie[e] * sigmaWie[e]);
  }
  double ssum = 0.0;
  for (int e = 0; e < d1 * d2; ++e) ssum += std::sqrt(sigmaPSD[e]);
  c->colored = true;
  c->ctau = tau;
  c->wie_s2 = s2;
  c->tol_sigma = ssum / (d1 * d2);  // mean(sqrt(sigma(:))), matchBlocksCPU.m:21
  // the scalar sigma is not read on this path (thresholds and Wiener weights
  // come from the PSD and sigmaWie)
  const int st = gbm3d_denoise(c, I, m, n, 1.0f, &o, U, U1);
  c->colored = false;
  return st;
}

}  // extern "C"
