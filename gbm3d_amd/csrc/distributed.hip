// distributed.hip -- [U, out] = GBM3D_distributed(I, sigma, opts)
// (source/GBM3D_distributed.m): the image is cut into overlapping tiles of
// 256 + blockSize, each tile and match spin ("case") is denoised on its own,
// and the unnormalised U, W of the cases are stitched with ramp windows.  The
// cases are independent and run on the context's lane streams; one gather
// kernel per step stitches them in case order (deterministic).
#include <algorithm>

#include "ctx.hpp"
#include "wave.hpp"

using namespace gbm3d;

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
  if (o.numMin != o.numMax)
    return fail(GBM3D_ERR_UNSUPPORTED, "distributed: the tolerance branch (numMax ~= numMin) is single-image only");
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
  GBM3D_TRY(on_lanes(c, (int)c1.size(), l1, st, [&](int i) {
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
    bool wdone = false;  // the step-1 weights written with Vavg
    if (o.filtType == GBM3D_FILT_MEDIAN) {
      StageTimer tm(c, GBM3D_STAGE_HT_SYN0);
      GBM3D_TRY(launch_median3(b.vol, th, tw, o.numMax, k, b.vavg, ls));
    } else {
      HtWork hw;
      ht_work_bind(b.htw, th, tw, o.numMax, o.levels, o.cycleSpin, &hw);
      GBM3D_TRY(run_ht(b.vol, th, tw, o.numMax, o.levels, o.cycleSpin, tau.data(), hw, b.vavg, k, ls, ht_hook, c,
                       nullptr, b.w1, &wdone));
    }
    if (!wdone) {
      StageTimer tm(c, GBM3D_STAGE_AGG1_PREP);
      GBM3D_TRY(launch_agg1_prep(b.vavg, th, tw, o.numMax, k, b.w1, ls));
    }
    StageTimer tm(c, GBM3D_STAGE_AGG1);
    return launch_aggregate_m(1, b.vavg, b.w1, b.S1, masks1, th, tw, k, o.matchSize, o.numMax, 0, Uc + stride * i,
                              Wc + stride * i, ls);
  }));
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
  GBM3D_TRY(on_lanes(c, (int)c2.size(), l2, st, [&](int i) {
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
    return launch_aggregate_m(2, b.V2f, b.g2, b.S2, masks2, th, tw, bw, o.matchSize, o.numMax, 0, Uc + stride * i,
                              Wc + stride * i, ls);
  }));
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

