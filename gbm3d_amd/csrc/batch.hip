// batch.hip -- [U_i, out_i] = GBM3D(I_i, sigma_i, opts) for many images of
// uneven sizes and sigmas in one call (gbm3d_denoise_batch_device): images
// largest first onto the least-loaded lane stream, each lane with its own
// workspace, the whole issue captured once into a hipGraph and replayed; an
// optional tuner times lane counts and matcher launch shapes on the batch.
// No reference counterpart beyond the per-image call (source/GBM3D.m:1).
#include <algorithm>

#include "ctx.hpp"

using namespace gbm3d;

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
    const LearnedSlot slot = c->ext.learned ? learned_slot(c, i) : LearnedSlot{};
    st = run_pipeline_on(c, j.dI[i], j.plans[i], j.sigma[i], o, j.dU[i], j.dU1 ? j.dU1[i] : nullptr, c->bstreams[s],
                         c->arena + off[s], c->ext.learned ? &slot : nullptr);
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
  if (o.numMin != o.numMax)
    return fail(GBM3D_ERR_UNSUPPORTED, "batch: the tolerance branch (numMax ~= numMin) is single-image only");
  BatchJob j{count, dI, nullptr, m, n, sigma, dU, dU1, {}};
  j.plans.resize(count);
  for (int i = 0; i < count; ++i) {
    if (!dI[i] || !dU[i]) return fail(GBM3D_ERR_ARG, "batch: NULL image " + std::to_string(i));
    GBM3D_TRY(make_plan(o, m[i], n[i], &j.plans[i]));
  }
  GBM3D_TRY(learned_setup(c, o, j.plans));
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
  key.push_back((int64_t)(intptr_t)c->lslots);
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

