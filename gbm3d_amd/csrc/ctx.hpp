// ctx.hpp -- the context object of libgbm3d and the pieces of the two-step
// pipeline (source/GBM3D.m:21-83) shared by the entry-point files:
//   api.hip          contexts, stage entry points
//   pipeline.hip     option validation, workspace layout, run_pipeline_on,
//                    gbm3d_denoise*, coloured noise
//   batch.hip        many images per call (graph capture, launch-shape tuner)
//   learned.hip      the learned-shrinkage form and its training entry points
//   distributed.hip  GBM3D_distributed (tiles, ramp-window stitching)
//   noalloc.hip      the entry point that allocates nothing
#pragma once

#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

struct gbm3d_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  char* arena = nullptr;
  size_t arena_bytes = 0;
  char* io = nullptr;  // staging for the host-buffer entry points
  size_t io_bytes = 0;
  float* masks = nullptr;  // [4 block sizes][2][32*32]
  char* noalloc_ws = nullptr;  // this device's static scratch array (noalloc.hip)
  gbm3d_ext ext{};
  uint32_t* bf16_stats = nullptr;  // BF16 matcher path counters (gbm3d_bf16_match_stats)
  bool timing = false;
  bool in_batch = false;  // stage timers off while a batch runs on several streams
  // lane streams: match spins / batch images / tile cases in flight at once
  std::vector<hipStream_t> bstreams;
  std::vector<hipEvent_t> bjoin;
  hipEvent_t bfork = nullptr;
  hipGraphExec_t gexec = nullptr;  // batch graph
  std::vector<int64_t> gkey;
  hipGraphExec_t dexec = nullptr;  // distributed driver
  std::vector<int64_t> dkey;
  int tuned_streams = 0, tuned_match_waves = 0;
  double tuned_ms = 0.0;
  std::vector<int64_t> tkey;
  int match_waves = 4;  // reference tiles per matching workgroup (single-image path)
  // coloured noise of the current call (gbm3d_denoise_colored)
  bool colored = false;
  double tol_sigma = 0.0;       // coloured noise: mean(sqrt(PSD)), the tolerance branch's sigma
  int* tolflag = nullptr;       // device: a tile of the tolerance branch came out short
  std::vector<double> ctau;   // levels x 7
  std::vector<float> wie_s2;  // blockSizeWie^2: sigmaWie.^2
  // learned shrinkage
  uint16_t* shrink_wb = nullptr;  // BF16 weights [4][64][192]
  float* shrink_fb = nullptr;     // FP32 biases and sigma-channel weights
  char* lslots = nullptr;         // per image: gbm3d::LearnedSlot's device arrays
  size_t lslots_images = 0;
  std::vector<int> l_spins;       // last call: match spins per image
  struct Mark {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  size_t pool_used = 0;
};

namespace gbm3d {

constexpr int kMaskStride = 2 * 32 * 32;
inline int mask_slot(int d) { return d == 4 ? 0 : d == 8 ? 1 : d == 16 ? 2 : d == 32 ? 3 : -1; }
constexpr int kMaxLanes = 8;  // match spins (or batch images) in flight at once, one stream each

// MATLAB linspace(a,b,n) with the last element pinned to b.
std::vector<double> linspace(double a, double b, int n);

// Bump allocator over a workspace; a first pass with base == nullptr
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

// Grow-only device buffer (hipMalloc): *buf holds at least `need` bytes.
int ensure(char** buf, size_t* have, size_t need);

// Stage timers: HIP events on the context stream (when timing is enabled).
void mark_begin(gbm3d_ctx* c, int stage);
void mark_end(gbm3d_ctx* c, int stage);
void ht_hook(void* user, int stage, int begin);
struct StageTimer {
  gbm3d_ctx* c;
  int stage;
  StageTimer(gbm3d_ctx* ctx, int s) : c(ctx), stage(s) { mark_begin(c, s); }
  ~StageTimer() { mark_end(c, stage); }
};

// One image's geometry under resolved options.
struct Plan {
  int m, n, M, N, bs, bw, d2, K, n1, n2;
  size_t MN() const { return (size_t)M * N; }
};

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

// Every option combination the kernels are built for, checked before any
// launch: an unsupported one is GBM3D_ERR_UNSUPPORTED with nothing issued.
int validate(const gbm3d_opts& o, int m, int n, int* M, int* N);
int make_plan(const gbm3d_opts& o, int m, int n, Plan* p);
void carve_spin(Carve& cv, const Plan& p, const gbm3d_opts& o, SpinBufs* b);
void carve(Carve& cv, const Plan& p, const gbm3d_opts& o, int lanes, Bufs* b);
// Device workspace bytes of one image's pipeline.
size_t pipeline_bytes(const Plan& p, const gbm3d_opts& o, int lanes = 1);
int spin_lanes(const gbm3d_ctx* c, const Plan& p);

// Block matching on a padded image with the context's distance variant.
// bn_ready: bn already holds the block norms of pad (launch_search_pad_norm).
int do_match(gbm3d_ctx* c, const float* pad, float* bn, float* smap, int M, int N, int d, int d2, int K, uint32_t* S,
             float* vol, hipStream_t st, int stage, const TolArgs* ta = nullptr, bool bn_ready = false);

// Fork `lanes` lane streams from st and join them back (no-ops for lanes <= 1).
// A fork is always joined, also after an error between the two, so that a
// stream capture stays well formed.
int lanes_fork(gbm3d_ctx* c, int lanes, hipStream_t st);
int lanes_join(gbm3d_ctx* c, int lanes, hipStream_t st);
// Runs body(i) for i = 0 .. count-1 between a fork of `nl` lanes from st and
// their join; stops at the first error and joins the lanes whatever body
// returned.
template <class F>
int on_lanes(gbm3d_ctx* c, int count, int nl, hipStream_t st, F&& body) {
  int status = lanes_fork(c, nl, st);
  for (int i = 0; i < count && status == GBM3D_OK; ++i) status = body(i);
  const int sj = lanes_join(c, nl, st);
  return status != GBM3D_OK ? status : sj;
}

// One image through both steps on stream st, workspace at ws
// (pipeline_bytes(p, o, lanes)); with lanes > 1 the match spins of each step
// run on the context's lane streams.
// The learned form's device state of one image: per match spin the shrink
// kernel's residual partials (stats[spin][kShrinkGrid][2]), the gate of the
// reference filter (run[spin], 1 = fall back) and the judge's figures
// (res[spin][3]: removed / noise energy of the learned gains and of a 2.7
// sigma hard threshold, the gate).
struct LearnedSlot {
  float* stats;
  int* run;
  double* res;
};
int run_pipeline_on(gbm3d_ctx* c, const float* dI, const Plan& p, float sigma, const gbm3d_opts& o, float* dU,
                    float* dU1, hipStream_t st, char* ws, const LearnedSlot* ls, int lanes = 1);
// One image on the context stream with the context arena.
int run_pipeline(gbm3d_ctx* c, const float* dI, int m, int n, float sigma, const gbm3d_opts& o, float* dU,
                 float* dU1);

// learned.hip: the learned form's prerequisites and device state.  setup
// records each image's match spin count for gbm3d_shrink_stats.
int learned_setup(gbm3d_ctx* c, const gbm3d_opts& o, const std::vector<Plan>& plans);
LearnedSlot learned_slot(gbm3d_ctx* c, int image);

// noalloc.hip: the device address of this device's scratch array, looked up
// once per device (the device must be current).
int noalloc_scratch_init(int device, char** ptr);

int set_dev(gbm3d_ctx* c);
int resolve_checked(const gbm3d_opts* in, gbm3d_opts* out);

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

}  // namespace gbm3d
