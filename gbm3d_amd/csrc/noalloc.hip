// noalloc.hip -- gbm3d_denoise_device_noalloc: GBM3D(I, sigma, opts) on
// device buffers with no device allocation at all, for callers that cannot
// call hipMalloc (code under stream capture, tight serving loops).
//
// The pipeline's workspace lives in a file-scope device array, `scratch`, of
// GBM3D_NOALLOC_SCRATCH_BYTES per GPU (the code object holds it from load to
// process end).  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in symbol): its device address is looked up once
// per device with hipGetSymbolAddress, at context creation with that device
// current, and kept in a per-device table.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (two streams, two contexts)
// race; calls in stream order on one stream do not.  No buffer of the
// workspace needs to start at zero: every one is written before it is read.
#include <mutex>

#include "ctx.hpp"

namespace {

static __device__ __align__(256) char scratch[GBM3D_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace gbm3d {

int noalloc_scratch_init(int device, char** ptr) {
  if (device < 0 || device >= kMaxDevices) return fail(GBM3D_ERR_UNSUPPORTED, "noalloc: device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // the caller made `device` current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(GBM3D_ERR_HIP, std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *ptr = g_scratch_ptr[device];
  return GBM3D_OK;
}

}  // namespace gbm3d

using namespace gbm3d;

extern "C" {

int gbm3d_denoise_device_noalloc(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts* opts,
                                 float* dU, float* dU1) {
  GBM3D_TRY(set_dev(c));
  if (!dI || !dU) return fail(GBM3D_ERR_ARG, "NULL image");
  if (!c->noalloc_ws) return fail(GBM3D_ERR_ARG, "noalloc: the scratch array of this device was not looked up");
  if (c->ext.learned)
    return fail(GBM3D_ERR_UNSUPPORTED, "noalloc: the learned form needs its per-image device state");
  gbm3d_opts o;
  GBM3D_TRY(resolve_checked(opts, &o));
  if (o.numMin != o.numMax)
    return fail(GBM3D_ERR_UNSUPPORTED, "noalloc: the tolerance branch (numMax ~= numMin) waits for the stream");
  Plan p;
  GBM3D_TRY(make_plan(o, m, n, &p));
  const size_t need = pipeline_bytes(p, o, 1);
  if (need > (size_t)GBM3D_NOALLOC_SCRATCH_BYTES)
    return fail(GBM3D_ERR_ARG, "noalloc: this image needs " + std::to_string(need) +
                                   " bytes of workspace, above the scratch array's " +
                                   std::to_string((size_t)GBM3D_NOALLOC_SCRATCH_BYTES));
  // stage timers would create events: off for this call
  const bool timing = c->timing;
  c->timing = false;
  const int st = run_pipeline_on(c, dI, p, sigma, o, dU, dU1, c->stream, c->noalloc_ws, nullptr, 1);
  c->timing = timing;
  return st;
}

}  // extern "C"
