// learned.hip -- the learned-shrinkage form (a small BF16 conv network on
// the DCT-Haar coefficients of each step-1 group replaces the wavelet hard
// threshold), its per-image device state (residual partials, the per-spin
// fallback gate the judge kernel sets), and the entry points of its
// training (features, forward, gradient, Adam) and of the synthetic image
// generator.  No reference counterpart; the gain targets follow
// source/denoiseAndAggregateWieFast.m:49.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"

using namespace gbm3d;

namespace gbm3d {

namespace {

constexpr size_t kSlotStats = (size_t)GBM3D_MAX_SPINS * 2 * kShrinkGrid * sizeof(float);
constexpr size_t kSlotRun = ((GBM3D_MAX_SPINS * sizeof(int) + 255) / 256) * 256;
constexpr size_t kSlotRes = (size_t)GBM3D_MAX_SPINS * 3 * sizeof(double);
constexpr size_t kSlotBytes = kSlotStats + kSlotRun + kSlotRes;

}  // namespace

// The learned form needs loaded weights, 8 x 8 x 16 step-1 groups, the hard
// threshold filter (its fallback) and one LearnedSlot per image.
int learned_setup(gbm3d_ctx* c, const gbm3d_opts& o, const std::vector<Plan>& plans) {
  if (!c->ext.learned) return GBM3D_OK;
  if (!c->shrink_wb) return fail(GBM3D_ERR_ARG, "learned shrinkage: no weights loaded (gbm3d_shrink_set_weights)");
  if (o.blockSize != 8 || o.numMax != 16 || o.numMin != 16)
    return fail(GBM3D_ERR_UNSUPPORTED, "learned shrinkage needs blockSize 8 and numMax = numMin = 16");
  if (o.filtType != GBM3D_FILT_HT)
    return fail(GBM3D_ERR_UNSUPPORTED, "learned shrinkage falls back to the 'ht' filter: filtType must be 'ht'");
  const size_t images = plans.size();
  if (images > c->lslots_images) {
    if (c->lslots) GBM3D_HIP(hipFree(c->lslots));
    c->lslots = nullptr;
    c->lslots_images = 0;
    GBM3D_HIP(hipMalloc(reinterpret_cast<void**>(&c->lslots), images * kSlotBytes));
    c->lslots_images = images;
  }
  c->l_spins.resize(images);
  for (size_t i = 0; i < images; ++i) c->l_spins[i] = plans[i].n1;
  return GBM3D_OK;
}

LearnedSlot learned_slot(gbm3d_ctx* c, int image) {
  char* base = c->lslots + (size_t)image * kSlotBytes;
  return {reinterpret_cast<float*>(base), reinterpret_cast<int*>(base + kSlotStats),
          reinterpret_cast<double*>(base + kSlotStats + kSlotRun)};
}

}  // namespace gbm3d

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
  if (image < 0 || image >= (int)c->l_spins.size()) return fail(GBM3D_ERR_ARG, "shrink stats: no such image");
  const int n1 = c->l_spins[image];
  double h[GBM3D_MAX_SPINS * 3];
  GBM3D_HIP(hipStreamSynchronize(c->stream));
  GBM3D_HIP(hipMemcpy(h, learned_slot(c, image).res, sizeof(double) * 3 * n1, hipMemcpyDeviceToHost));
  double a = 0.0, b = 0.0;
  int f = 0;
  for (int s = 0; s < n1; ++s) {
    a += h[3 * s] / n1;
    b += h[3 * s + 1] / n1;
    f += h[3 * s + 2] != 0.0;
  }
  if (est_learned) *est_learned = a;
  if (est_ht) *est_ht = b;
  if (fell_back) *fell_back = f;
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
                          const uint32_t* dSel, int nsel, int target, float* dX, float* dT, float* dWt) {
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
  if (target != GBM3D_SHRINK_TARGET_WIENER && target != GBM3D_SHRINK_TARGET_ENERGY)
    return fail(GBM3D_ERR_ARG, "shrink features: unknown target");
  return launch_shrink_features(pc, pn, S, m, d2, dSel, nsel, sigma, target, dX, dT, dWt, c->stream);
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

