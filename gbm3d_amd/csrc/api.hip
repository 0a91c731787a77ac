// api.hip -- the C ABI of libgbm3d: contexts and the stage entry points
// (matchBlocksCPU, denoiseAndAggregateGPU, matchBlocksWie,
// denoiseAndAggregateWieFast as separate calls on host buffers).
#include <algorithm>
#include <cmath>

#include "ctx.hpp"

using namespace gbm3d;

namespace {

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
  // (64 more words after the table: the tolerance branch's flag)
  if (hipMalloc(reinterpret_cast<void**>(&c->masks), (h.size() + 64) * sizeof(float)) != hipSuccess ||
      hipMemcpy(c->masks, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return fail(GBM3D_ERR_HIP, "mask table upload failed");
  }
  c->tolflag = reinterpret_cast<int*>(c->masks + h.size());
  if (noalloc_scratch_init(device, &c->noalloc_ws) != GBM3D_OK) {
    const std::string why = gbm3d_last_error();
    (void)hipFree(c->masks);
    (void)hipStreamDestroy(c->own);
    delete c;
    return fail(GBM3D_ERR_HIP, why);
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
  if (c->bf16_stats) (void)hipFree(c->bf16_stats);
  if (c->shrink_fb) (void)hipFree(c->shrink_fb);
  if (c->lslots) (void)hipFree(c->lslots);
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
  if (e && e->match_stats && !c->bf16_stats) {
    // the BF16 matcher's path counters, allocated and cleared here at set-up
    const std::vector<uint32_t> z((size_t)kBf16StatSlots * kBf16StatStride, 0u);
    GBM3D_HIP(hipMalloc(reinterpret_cast<void**>(&c->bf16_stats), z.size() * sizeof(uint32_t)));
    GBM3D_HIP(hipMemcpy(c->bf16_stats, z.data(), z.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  c->ext = e ? *e : gbm3d_ext{};
  return GBM3D_OK;
}

int gbm3d_bf16_match_stats(gbm3d_ctx* c, unsigned long long* out, int reset) {
  GBM3D_TRY(set_dev(c));
  if (!out) return fail(GBM3D_ERR_ARG, "NULL output");
  std::vector<uint32_t> h((size_t)kBf16StatSlots * kBf16StatStride, 0u);
  if (c->bf16_stats) {
    GBM3D_HIP(hipStreamSynchronize(c->stream));
    GBM3D_HIP(hipMemcpy(h.data(), c->bf16_stats, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (reset) {
      const std::vector<uint32_t> z(h.size(), 0u);
      GBM3D_HIP(hipMemcpy(c->bf16_stats, z.data(), z.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
  }
  for (int i = 0; i < kBf16Stats; ++i) {
    out[i] = 0;
    for (int sl = 0; sl < kBf16StatSlots; ++sl) out[i] += h[(size_t)sl * kBf16StatStride + i];
  }
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
  GBM3D_TRY(launch_search_pad_norm(img, m, n, 0, d2, d, pad, bn, c->stream));  // the pipeline's search image
  GBM3D_TRY(do_match(c, pad, bn, sm, m, n, d, d2, K, dS, Volume ? vol : nullptr, c->stream, GBM3D_STAGE_MATCH1,
                     nullptr, true));
  GBM3D_TRY(download(c, S, dS, G * K));
  if (Volume) GBM3D_TRY(download(c, Volume, vol, mn * K));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

int gbm3d_match_tol(gbm3d_ctx* c, const float* I, int m, int n, int d, int d2, int numMax, int numMin, double sigma,
                    uint32_t* S, float* Volume) {
  GBM3D_TRY(set_dev(c));
  if (!I || !S) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, numMax));
  if (numMin < 1 || numMin > numMax) return fail(GBM3D_ERR_ARG, "stage: numMin must be in [1, numMax]");
  if (!(sigma >= 0.0) || !std::isfinite(sigma)) return fail(GBM3D_ERR_ARG, "stage: sigma must be finite and >= 0");
  const size_t mn = (size_t)m * n, P = (size_t)m + d2, Q = (size_t)n + d2, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** img, float** pad, float** bn, uint32_t** dS, float** vol) {
    *img = k.take<float>(mn);
    *pad = k.take<float>(P * Q);
    *bn = k.take<float>((P - d + 1) * (Q - d + 1));
    *dS = k.take<uint32_t>(G * numMax);
    *vol = k.take<float>(Volume ? mn * numMax : 0);
  };
  float *img, *pad, *bn, *vol;
  uint32_t* dS;
  lay(cv, &img, &pad, &bn, &dS, &vol);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &img, &pad, &bn, &dS, &vol);
  const double t = 0.005 * sigma * sigma * d * d;  // matchBlocksCPU.m:22-28
  float tu = (float)t;
  if ((double)tu < t) tu = nextafterf(tu, __builtin_inff());
  const TolArgs ta{numMin == numMax ? 0 : numMin, tu, c->tolflag};
  GBM3D_TRY(upload(c, img, I, mn));
  GBM3D_TRY(launch_set_i32(c->tolflag, 0, c->stream));
  GBM3D_TRY(launch_search_pad_norm(img, m, n, 0, d2, d, pad, bn, c->stream));
  GBM3D_TRY(launch_match_w(pad, bn, m, n, d, d2, numMax, dS, Volume ? vol : nullptr, 4, c->stream,
                           ta.Kmin ? &ta : nullptr));
  int short_tile = 0;
  GBM3D_HIP(hipMemcpyAsync(&short_tile, c->tolflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GBM3D_TRY(download(c, S, dS, G * numMax));
  if (Volume) GBM3D_TRY(download(c, Volume, vol, mn * numMax));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  if (short_tile)
    return fail(GBM3D_ERR_ARG, "tolerance matching: a reference block has fewer than numMax matches below tol "
                               "(matchBlocksCPU fails here, matchBlocksCPU.m:111); S holds the numMax nearest");
  return GBM3D_OK;
}

int gbm3d_match_global(gbm3d_ctx* c, const float* I, int m, int n, double sigma, int d, int numMax, int numMin,
                       uint32_t* S, uint32_t* counts, float* Ifilt) {
  GBM3D_TRY(set_dev(c));
  if (!I || !S) return fail(GBM3D_ERR_ARG, "NULL argument");
  if (!match_global_supported(d)) return fail(GBM3D_ERR_UNSUPPORTED, "matchBlocksGPU: block size must be 4, 8, 16 or 32");
  if (m < d || n < d) return fail(GBM3D_ERR_ARG, "matchBlocksGPU: image smaller than one block");
  if ((size_t)m * n >= ((size_t)1 << 31)) return fail(GBM3D_ERR_ARG, "matchBlocksGPU: image of 2^31 pixels or more");
  if (numMax < 1 || numMax > 64 || numMin < 1 || numMin > numMax || (size_t)numMax > (size_t)m * n)
    return fail(GBM3D_ERR_ARG, "matchBlocksGPU: need 1 <= numMin <= numMax <= min(64, m n)");
  if (!(sigma >= 0.0) || !std::isfinite(sigma)) return fail(GBM3D_ERR_ARG, "matchBlocksGPU: sigma must be finite and >= 0");
  const bool pre = sigma > 30.0 / 255.0;  // matchBlocksGPU.m:34
  if (pre && (m % 8 || n % 8))
    return fail(GBM3D_ERR_ARG, "matchBlocksGPU: the prefilter's 3-level transform needs m and n multiples of 8");
  const size_t mn = (size_t)m * n, G = (size_t)(m / d) * (n / d);
  const size_t wpre = pre ? (mn / 4) * 4 + (mn / 16) * 4 + (mn / 64) * 4 + mn / 4 + mn / 16 : 0;
  const size_t wmatch = (size_t)(m + d - 1) * (n + d - 1) + mn;
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** img, float** filt, float** work, uint32_t** dS, uint32_t** dC) {
    *img = k.take<float>(mn);
    *filt = k.take<float>(pre ? mn : 0);
    *work = k.take<float>(std::max(wpre, wmatch));
    *dS = k.take<uint32_t>(G * numMax);
    *dC = k.take<uint32_t>(G);
  };
  float *img, *filt, *work;
  uint32_t *dS, *dC;
  lay(cv, &img, &filt, &work, &dS, &dC);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &img, &filt, &work, &dS, &dC);
  GBM3D_TRY(upload(c, img, I, mn));
  const float* src = img;
  if (pre) {
    GBM3D_TRY(launch_wav2_prefilter(img, m, n, 2.0 * sigma, filt, work, c->stream));  // tau = 2 sigma (:35)
    src = filt;
  }
  const double tol = 0.005 * sigma * sigma * d * d;  // :19-23
  GBM3D_TRY(launch_match_global(src, m, n, d, numMax, numMin, tol, dS, dC, work, c->stream));
  GBM3D_TRY(download(c, S, dS, G * numMax));
  if (counts) GBM3D_TRY(download(c, counts, dC, G));
  if (Ifilt) GBM3D_TRY(download(c, Ifilt, src, mn));
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

// [U, W] = denoiseAndAggregateWie(V1, V2, S, sigma, opts) (clip01 = 1,
// source/denoiseAndAggregateWie.m) or denoiseAndAggregateWieFast (clip01 = 0,
// source/denoiseAndAggregateWieFast.m): Wiener filter, then aggregation.
int gbm3d_denoise_aggregate_wie(gbm3d_ctx* c, const float* V1, const float* V2, const uint32_t* S, int m, int n,
                                int d, int d2, int K, float sigma, int clip01, float* U, float* W) {
  GBM3D_TRY(set_dev(c));
  if (!V1 || !V2 || !S || !U || !W) return fail(GBM3D_ERR_ARG, "NULL argument");
  GBM3D_TRY(stage_sizes(m, n, d, d2, K));
  if (!wiener_supported(d, K)) return fail(GBM3D_ERR_UNSUPPORTED, "Wiener step: (blockSizeWie, numMax) not built");
  const size_t mn = (size_t)m * n, G = mn / ((size_t)d * d);
  Carve cv{nullptr};
  auto lay = [&](Carve& k, float** a, float** b, float** v, float** g, uint32_t** dS, float** dU, float** dW) {
    *a = k.take<float>(mn * K);
    *b = k.take<float>(mn * K);
    *v = k.take<float>(mn * K);
    *g = k.take<float>(G);
    *dS = k.take<uint32_t>(G * K);
    *dU = k.take<float>(mn);
    *dW = k.take<float>(mn);
  };
  float *a, *b, *v, *g, *dU, *dW;
  uint32_t* dS;
  lay(cv, &a, &b, &v, &g, &dS, &dU, &dW);
  GBM3D_TRY(ensure(&c->arena, &c->arena_bytes, cv.off));
  Carve k2{c->arena};
  lay(k2, &a, &b, &v, &g, &dS, &dU, &dW);
  GBM3D_TRY(upload(c, a, V1, mn * K));
  GBM3D_TRY(upload(c, b, V2, mn * K));
  GBM3D_TRY(upload(c, dS, S, G * K));
  GBM3D_TRY(launch_wiener(nullptr, nullptr, nullptr, a, b, 0, 0, d, 0, K, (long)G, sigma, v, g, c->stream, nullptr,
                          clip01 != 0));
  GBM3D_TRY(launch_aggregate_m(2, v, g, dS, c->masks + kMaskStride * mask_slot(d), m, n, d, d2, K, 0, dU, dW,
                               c->stream));
  GBM3D_TRY(download(c, U, dU, mn));
  GBM3D_TRY(download(c, W, dW, mn));
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  return GBM3D_OK;
}

}  // extern "C"
