// k_learn.hip -- learned shrinkage of step-1 groups (a variant beside the
// reference path, which uses the 3-D wavelet hard threshold of
// source/denoiseAndAggregateGPU.m:19-47 here).
//
// Per group (8 x 8 x 16 stack of matched blocks): coefficients c(t, k) of the
// orthonormal 8x8 DCT-II of every slice (t = u + 8v) followed by the 3-level
// Haar transform along the 16 slices (k), the transform of the step-2 Wiener
// filter (source/denoiseAndAggregateWieFast.m:33-47).  A 4-layer 1-D
// convolutional network runs along k with the 64 coefficients as channels:
//
//   h1 = relu(conv3(c / sigma; W0) + b0 + (sigma / 50) * conv3(1; wsig))
//   h2 = relu(conv3(h1; W1) + b1),  h3 = relu(conv3(h2; W2) + b2)
//   g  = sigmoid(conv3(h3; W3) + b3)
//
// (conv3: width 3, zero padding; the constant sigma / 50 input channel is the
// wsig term).  g(0, 14..15) is fixed to 1 as in the Wiener filter
// (denoiseAndAggregateWieFast.m:50); c(t, k) is multiplied by g(t, k).  The filtered group is
// the inverse transform clipped at 0, written group-contiguous for the step-1
// aggregation.  Matrix products are BF16 MFMA (16x16x32) with FP32
// accumulation; operands are rounded to BF16 to nearest even.
//
// Training targets are the oracle Wiener gains g* = |c_clean|^2 / (|c_clean|^2
// + sigma^2) (the gain formula of denoiseAndAggregateWieFast.m:49 evaluated on
// the clean group), with loss sum w (g - g*)^2, w = min(1 + |c_clean|^2 /
// sigma^2, 100) (the excess squared error of gain g over g* in units of
// sigma^2, capped), weight 0 on the fixed coefficients.
//
// Parameter vector (FP32, 49600 values):
//   W0 [3][64][64] | wsig [3][64] | b0 [64] | (W_l [3][64][64] | b_l [64]) l = 1..3
// W_l[tap][cin][cout]; tap 0 reads position k - 1.
#include "group_xform.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

using namespace shrink;

__host__ __device__ inline uint16_t bf16_rne(float x) {
  uint32_t b = __builtin_bit_cast(uint32_t, x);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

__device__ __forceinline__ bf16x8 ld_frag(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

__device__ __forceinline__ f32x4 mfma16x16x32(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Sum of the sigma-channel weights over the taps that fall inside the group
// at position pos: tap 0 reads pos - 1, tap 2 reads pos + 1.
__device__ __forceinline__ float sig_taps(const float* wsig, int pos, int ch) {
  float a = wsig[CH + ch];
  if (pos >= 1) a += wsig[ch];
  if (pos <= POS - 2) a += wsig[2 * CH + ch];
  return a;
}

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + __expf(-z)); }

// -------------------------------------------------------- fused inference
constexpr int KS = KW + 8;   // LDS row stride of a weight row (bf16)
constexpr int AS = CH + 8;   // LDS row stride of an activation row (bf16)
constexpr int AR = POS + 2;  // activation rows: positions -1 .. 16
constexpr int GPW = 1;       // groups per wave and pass
constexpr int WAVES = 12;    // 768 threads share one LDS copy of the weights
// Per-wave LDS region: the transform scratch (16 x 64 floats) of the wave's
// group, later its activation rows (they alias: the activations are written
// once the transform is done, and the gains overwrite them once the last layer
// has read them).
constexpr int REGION = POS * 64;  // floats
static_assert(2 * REGION >= GPW * AR * AS, "region holds the activation rows");

// Coefficients of group g (clamped to G - 1) for lane t = u + 8v: the
// group's 16 blocks, 2-D DCT of every slice, Haar along the slices.
template <int MODE>
__device__ __forceinline__ void load_group(const float* __restrict__ src, const uint32_t* __restrict__ S, int M,
                                           int d2, long G, long g, int lane, float* X, const float* T,
                                           float (&x)[POS]) {
  if (g >= G) g = G - 1;
  const int u = lane & 7, v = lane >> 3;
  if constexpr (MODE == 0) {
    const int Mt = M / 8;
    const size_t P = (size_t)M + d2;
    const int ti = (int)(g % Mt), tj = (int)(g / Mt);
    // lane k < POS: slot k's block origin (its window index divided by d2
    // once, all slots in parallel), then each slot's origin from lane k
    size_t org = 0;
    if (lane < POS) {
      const uint32_t idx = S[(size_t)POS * g + lane] - 1u;
      const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
      org = (size_t)(ti * 8 + a) + P * (size_t)(tj * 8 + b);
    }
    const size_t off = (size_t)u + P * (size_t)v;
#pragma unroll
    for (int k = 0; k < POS; ++k) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)org, k);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(org >> 32), k);
      x[k] = src[(((size_t)hi << 32) | lo) + off];
    }
  } else {
#pragma unroll
    for (int k = 0; k < POS; ++k) x[k] = src[(size_t)lane + 64 * ((size_t)k + POS * (size_t)g)];
  }
  page_transform<8, POS, false>(x, X, T, u, v);
  haar_dec<POS, 3>(x);
}

// Network input rows of group j: c / sigma rounded to BF16, positions 0..15
// at rows 1..16.
__device__ __forceinline__ void put_act(uint16_t* A, int j, int lane, const float (&x)[POS], float inv_s) {
#pragma unroll
  for (int k = 0; k < POS; ++k) A[(j * AR + k + 1) * AS + lane] = bf16_rne(x[k] * inv_s);
}

// Gains from the last layer's accumulators, shrinkage, removed-energy sums,
// inverse transform, clip at 0 and the store of group g (when g < G).
__device__ __forceinline__ void finish_group(const f32x4 (&acc)[4], const float (&c)[POS], long g, long G,
                                             int lane, float* X, const float* T, const float* b3, float tau_h, float s2, float* __restrict__ vout,
                                             float* __restrict__ gout, float* e_l, float* e_h) {
#pragma unroll
  for (int ct = 0; ct < 4; ++ct)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int pos = 4 * (lane >> 4) + i, ch = ct * 16 + (lane & 15);
      X[pos * 64 + ch] = sigmoidf_(acc[ct][i] + b3[ch]);
    }
  wave_sync();
  const bool live = g < G;
  float* go = (gout && live) ? gout + (size_t)lane + 64 * POS * (size_t)g : nullptr;
  float y[POS];
#pragma unroll
  for (int k = 0; k < POS; ++k) {
    const float a = (lane == 0 && k >= POS - 2) ? 1.f : X[k * 64 + lane];
    const float cc = c[k], c2 = cc * cc;
    if (live) {
      // energy removed from the noisy coefficients by the learned gains and
      // by a hard threshold at tau_h
      *e_l += (1.f - a) * (1.f - a) * c2;
      *e_h += fabsf(cc) > tau_h ? 0.f : c2;
    }
    if (go) go[64 * k] = a;
    y[k] = cc * a;
  }
  wave_sync();
  haar_rec<POS, 3>(y);
  page_transform<8, POS, true>(y, X, T, lane & 7, lane >> 3);
  if (live) {
    float* vo = vout + (size_t)lane + 64 * POS * (size_t)g;
#pragma unroll
    for (int k = 0; k < POS; ++k) vo[64 * k] = y[k] > 0.f ? y[k] : 0.f;
  }
}

}  // namespace

// MODE 0: groups gathered from the padded search image through S (pipeline);
// MODE 1: groups read from V (d x d x K x G, stage entry point).
// Persistent grid: every block loads the BF16 weights into LDS once.
template <int MODE>
__global__ __launch_bounds__(768) void shrink_kernel(const float* __restrict__ src, const uint32_t* __restrict__ S,
                                                     int M, int d2, long G, float sigma, float tau_h,
                                                     const uint16_t* __restrict__ wb, const float* __restrict__ fb,
                                                     DctTable<8> tab, float* __restrict__ vout,
                                                     float* __restrict__ gout, float* __restrict__ stats) {
  __shared__ __align__(16) uint16_t wl[NL * CH * KS];
  __shared__ __align__(16) float region[WAVES][REGION];
  __shared__ float bias[NL * CH];
  __shared__ float sb0[3 * CH];
  __shared__ float T[64];
  __shared__ float red[WAVES][2];

  // weights: global [L][cout][KW] -> LDS [L][cout][KS], 16-byte chunks
  for (int e = threadIdx.x; e < NL * CH * (KW / 8); e += blockDim.x) {
    const int row = e / (KW / 8), kk = (e % (KW / 8)) * 8;
    *reinterpret_cast<u32x4*>(wl + row * KS + kk) = *reinterpret_cast<const u32x4*>(wb + (size_t)row * KW + kk);
  }
  const float s = sigma * 0.02f;
  for (int e = threadIdx.x; e < NL * CH; e += blockDim.x) bias[e] = fb[e];
  for (int e = threadIdx.x; e < 3 * CH; e += blockDim.x) {
    const int cls = e / CH, ch = e % CH;
    const int pos = cls == 0 ? 0 : (cls == 1 ? 1 : POS - 1);
    sb0[e] = fb[ch] + s * sig_taps(fb + NL * CH, pos, ch);
  }
  if (threadIdx.x < 64) T[threadIdx.x] = tab.t[threadIdx.x];
  __syncthreads();

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float* X = region[wave];
  uint16_t* A = reinterpret_cast<uint16_t*>(region[wave]);
  const float inv_s = 1.0f / sigma, s2 = sigma * sigma;
  float e_l = 0.f, e_h = 0.f;
  const long stride = (long)gridDim.x * WAVES * GPW;
  static_assert(GPW == 1, "one group per wave and pass");
  for (long base = (long)blockIdx.x * WAVES + wave; base < G; base += stride) {
    float c0[POS];  // the group's coefficients, in registers
    load_group<MODE>(src, S, M, d2, G, base, lane, X, T, c0);
    // the transform scratch is free now (page_transform ends with a wave sync)
    A[lane] = 0;  // rows -1 and 16: the conv zero padding
    A[(AR - 1) * AS + lane] = 0;
    put_act(A, 0, lane, c0, inv_s);
    wave_sync();
    f32x4 acc[GPW][4];
#pragma unroll 1
    for (int L = 0; L < NL; ++L) {
#pragma unroll
      for (int j = 0; j < GPW; ++j)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[j][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
      const uint16_t* WL = wl + L * CH * KS;
#pragma unroll
      for (int kc = 0; kc < KW / 32; ++kc) {
        const int tap = kc >> 1, cin = (kc & 1) * 32 + 8 * (lane >> 4);
        bf16x8 b[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) b[ct] = ld_frag(WL + (ct * 16 + (lane & 15)) * KS + kc * 32 + 8 * (lane >> 4));
#pragma unroll
        for (int j = 0; j < GPW; ++j) {
          const bf16x8 a = ld_frag(A + (j * AR + (lane & 15) + tap) * AS + cin);
#pragma unroll
          for (int ct = 0; ct < 4; ++ct) acc[j][ct] = mfma16x16x32(a, b[ct], acc[j][ct]);
        }
      }
      if (L < NL - 1) {
        wave_sync();  // every lane has read this layer's input
        const float* add = L == 0 ? sb0 : bias + L * CH;
#pragma unroll
        for (int j = 0; j < GPW; ++j)
#pragma unroll
          for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int pos = 4 * (lane >> 4) + i, ch = ct * 16 + (lane & 15);
              const int row = L == 0 ? (pos == 0 ? 0 : (pos == POS - 1 ? 2 : 1)) : 0;
              const float z = acc[j][ct][i] + add[row * CH + ch];
              A[(j * AR + pos + 1) * AS + ch] = bf16_rne(z > 0.f ? z : 0.f);
            }
        wave_sync();
      }
    }
    wave_sync();  // the last layer has read the activations: X may overwrite them
    const float* b3 = bias + (NL - 1) * CH;
    finish_group(acc[0], c0, base, G, lane, X, T, b3, tau_h, s2, vout, gout, &e_l, &e_h);
  }
  e_l = wave_sum_f32(e_l);
  e_h = wave_sum_f32(e_h);
  if (lane == 0) {
    red[wave][0] = e_l;
    red[wave][1] = e_h;
  }
  __syncthreads();
  if (threadIdx.x == 0 && stats) {
    float a = 0.f, b = 0.f;
    for (int w = 0; w < WAVES; ++w) {
      a += red[w][0];
      b += red[w][1];
    }
    stats[2 * blockIdx.x] = a;
    stats[2 * blockIdx.x + 1] = b;
  }
}

int shrink_grid(long G) {
  const long passes = (G + WAVES * GPW - 1) / (WAVES * GPW);
  return (int)(passes < kShrinkGrid ? passes : kShrinkGrid);
}

void shrink_pack_weights(const float* p, uint16_t* wb, float* fb) {
  for (int L = 0; L < NL; ++L)
    for (int co = 0; co < CH; ++co)
      for (int k = 0; k < KW; ++k) wb[((size_t)L * CH + co) * KW + k] = bf16_rne(p[off_w(L) + k * CH + co]);
  for (int L = 0; L < NL; ++L)
    for (int ch = 0; ch < CH; ++ch) fb[L * CH + ch] = p[off_b(L) + ch];
  for (int e = 0; e < 3 * CH; ++e) fb[NL * CH + e] = p[OFF_WSIG + e];
}

int launch_shrink(const float* src, const uint32_t* S, int M, int d2, long G, float sigma, const uint16_t* wb,
                  const float* fb, float* vout, float* gout, float* stats, int* grid_used, hipStream_t st) {
  if (G <= 0) {
    if (grid_used) *grid_used = 0;
    return GBM3D_OK;
  }
  if (S && (M % 8 || d2 <= 0)) return fail(GBM3D_ERR_ARG, "shrink: geometry");
  const int grid = shrink_grid(G);
  if (grid_used) *grid_used = grid;
  const DctTable<8> tab = make_dct_table<8>();
  const float tau_h = 2.7f * sigma;
  if (S)
    hipLaunchKernelGGL(shrink_kernel<0>, dim3(grid), dim3(64 * WAVES), 0, st, src, S, M, d2, G, sigma, tau_h, wb, fb, tab,
                       vout, gout, stats);
  else
    hipLaunchKernelGGL(shrink_kernel<1>, dim3(grid), dim3(64 * WAVES), 0, st, src, S, M, d2, G, sigma, tau_h, wb, fb, tab,
                       vout, gout, stats);
  return check_launch("shrink");
}

// ------------------------------------------------------------ the judge
// One block after the shrink kernel of one match spin: r = (energy the
// learned gains removed) / noise and r_ht, the same for the 2.7 sigma hard
// threshold, each the sum of the kernel's per-block partials in a fixed
// order; run[0] = 1 (fall back) when !(r <= bound), else 0; res = {r, r_ht,
// run}.  The gated launches that follow read run on the device: no host
// round trip.
__global__ __launch_bounds__(256) void learned_judge_kernel(const float* __restrict__ stats, int grid, double noise,
                                                            double bound, int* __restrict__ run,
                                                            double* __restrict__ res) {
  __shared__ double sa[256], sb[256];
  const int t = threadIdx.x;
  sa[t] = t < grid ? (double)stats[2 * t] : 0.0;
  sb[t] = t < grid ? (double)stats[2 * t + 1] : 0.0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) {
      sa[t] += sa[t + o];
      sb[t] += sb[t + o];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double r = sa[0] / noise, rh = sb[0] / noise;
    const int f = (r <= bound) ? 0 : 1;
    run[0] = f;
    res[0] = r;
    res[1] = rh;
    res[2] = (double)f;
  }
}

int launch_learned_judge(const float* stats, int grid, double noise, double bound, int* run, double* res,
                         hipStream_t st) {
  if (grid < 0 || grid > kShrinkGrid) return fail(GBM3D_ERR_ARG, "learned judge: grid");
  hipLaunchKernelGGL(learned_judge_kernel, dim3(1), dim3(256), 0, st, stats, grid, noise, bound, run, res);
  return check_launch("learned_judge");
}

// ------------------------------------------------------ training features
// One wave per selected group: coefficients of the noisy and clean groups at
// the same matched positions; X = c_noisy / sigma, T = g*, Wt = loss weight.
// Rows r = 16 * i + k (selected group i, position k), columns t.
__global__ __launch_bounds__(256) void shrink_features_kernel(const float* __restrict__ padC,
                                                              const float* __restrict__ padN,
                                                              const uint32_t* __restrict__ S, int M, int d2,
                                                              const uint32_t* __restrict__ sel, int nsel,
                                                              float sigma, int target, DctTable<8> tab,
                                                              float* __restrict__ Xf, float* __restrict__ Tg,
                                                              float* __restrict__ Wt) {
  __shared__ float sm[4][POS * 64];
  __shared__ float T[64];
  if (threadIdx.x < 64) T[threadIdx.x] = tab.t[threadIdx.x];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= nsel) return;
  const int u = lane & 7, v = lane >> 3;
  const uint32_t g = sel[i];
  const int Mt = M / 8;
  const int ti = (int)(g % (uint32_t)Mt), tj = (int)(g / (uint32_t)Mt);
  const size_t P = (size_t)M + d2;
  float xc[POS], xn[POS];
#pragma unroll
  for (int k = 0; k < POS; ++k) {
    const uint32_t idx = S[(size_t)POS * g + k] - 1u;
    const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
    const size_t o = (size_t)(ti * 8 + a + u) + P * (size_t)(tj * 8 + b + v);
    xc[k] = padC[o];
    xn[k] = padN[o];
  }
  page_transform<8, POS, false>(xc, sm[wave], T, u, v);
  haar_dec<POS, 3>(xc);
  page_transform<8, POS, false>(xn, sm[wave], T, u, v);
  haar_dec<POS, 3>(xn);
  const float inv_s = 1.0f / sigma, s2 = sigma * sigma;
#pragma unroll
  for (int k = 0; k < POS; ++k) {
    const size_t o = ((size_t)i * POS + k) * CH + lane;
    const float e2 = xc[k] * xc[k];
    const float x = xn[k] * inv_s;
    Xf[o] = x;
    const bool dc = lane == 0 && k >= POS - 2;  // the fixed DC approximation: no loss
    if (target == GBM3D_SHRINK_TARGET_ENERGY) {
      // the gain whose pilot coefficient has the clean coefficient's magnitude,
      // |g y| = |c| (then step 2's gain |P|^2/(|P|^2 + sigma^2) is the oracle
      // Wiener gain), weighted by the pilot error it measures: (g - T)^2 y^2/sigma^2
      const float ay = fabsf(xn[k]);
      Tg[o] = ay > 0.f ? fminf(fabsf(xc[k]) / ay, 1.f) : 1.f;
      Wt[o] = dc ? 0.f : fminf(x * x, 100.f);
    } else {
      Tg[o] = e2 / (e2 + s2);
      Wt[o] = dc ? 0.f : fminf(1.f + e2 / s2, 100.f);
    }
  }
}

int launch_shrink_features(const float* padC, const float* padN, const uint32_t* S, int M, int d2,
                           const uint32_t* sel, int nsel, float sigma, int target, float* X, float* Tg, float* Wt,
                           hipStream_t st) {
  if (nsel <= 0) return GBM3D_OK;
  hipLaunchKernelGGL(shrink_features_kernel, dim3((unsigned)((nsel + 3) / 4)), dim3(256), 0, st, padC, padN, S, M,
                     d2, sel, nsel, sigma, target, make_dct_table<8>(), X, Tg, Wt);
  return check_launch("shrink_features");
}

// ------------------------------------------------------ training GEMMs
// One kernel for the three products of a conv layer over rows r = 16 g + k:
//   forward      Z[r][n]  = sum_k A(r, k) W(k, n) + bias     (k = tap*64 + cin)
//   input grad   dX[r][n] = sum_k dZ(r - tap + 1, co) W(tap, n, co) * relu'(Zprev)
//   weight grad  dW[m][n] = sum_r A'(m, r) dZ(r, n), m = tap*64 + cin, plus
//                the bias row (m = 192) and the sigma-channel rows (193..195)
// 64 x 64 output tile per block, 4 waves of 2 x 2 MFMA tiles, K in steps of 32
// through LDS; operands rounded to BF16, FP32 accumulation.
enum { kFwd = 0, kBwdData = 1, kBwdW = 2 };

struct GemmArgs {
  int mode, layer, rows, chunk;
  const float* in;     // layer input (kFwd, kBwdW), rows x 64
  const float* par;    // parameter vector
  const float* dZ;     // kBwdData, kBwdW
  const float* Zprev;  // kBwdData: pre-activation of the previous layer
  const float* sig;    // sigma per group
  float* Z;            // kFwd
  float* H;            // kFwd
  float* dX;           // kBwdData
  float* part;         // kBwdW: [chunks][256][64]
};

namespace {

__device__ __forceinline__ float in_at(const float* in, int rows, int r, int tap, int ci) {
  const int p = (r & (POS - 1)) + tap - 1;
  if (r >= rows || p < 0 || p >= POS) return 0.f;
  return in[(size_t)((r & ~(POS - 1)) + p) * CH + ci];
}

__device__ __forceinline__ float a_elem(const GemmArgs& g, int m, int k) {
  if (g.mode == kFwd) return in_at(g.in, g.rows, m, k >> 6, k & 63);
  if (g.mode == kBwdData) {
    const int tap = k >> 6, co = k & 63;
    const int p = (m & (POS - 1)) - tap + 1;
    if (m >= g.rows || p < 0 || p >= POS) return 0.f;
    return g.dZ[(size_t)((m & ~(POS - 1)) + p) * CH + co];
  }
  // kBwdW: m indexes dW rows, k indexes data rows
  if (m < KW) return in_at(g.in, g.rows, k, m >> 6, m & 63);
  if (k >= g.rows) return 0.f;
  if (m == KW) return 1.f;
  if (g.layer == 0 && m <= KW + 3) {
    const int p = (k & (POS - 1)) + (m - KW - 1) - 1;
    return (p >= 0 && p < POS) ? g.sig[k >> 4] * 0.02f : 0.f;
  }
  return 0.f;
}

__device__ __forceinline__ float b_elem(const GemmArgs& g, int k, int n) {
  if (g.mode == kFwd) return g.par[off_w(g.layer) + k * CH + n];
  if (g.mode == kBwdData) return g.par[off_w(g.layer) + ((k >> 6) * CH + n) * CH + (k & 63)];
  return k < g.rows ? g.dZ[(size_t)k * CH + n] : 0.f;
}

}  // namespace

__global__ __launch_bounds__(256) void shrink_gemm_kernel(GemmArgs g) {
  constexpr int LS = 40;  // LDS row stride (bf16)
  __shared__ __align__(16) uint16_t As[64 * LS];
  __shared__ __align__(16) uint16_t Bs[64 * LS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave & 1, wn = wave >> 1;
  int m0, k0, k1;
  if (g.mode == kBwdW) {
    m0 = blockIdx.x * 64;
    k0 = blockIdx.y * g.chunk;
    k1 = k0 + g.chunk;
  } else {
    m0 = blockIdx.x * 64;
    k0 = 0;
    k1 = KW;
  }
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  for (int kb = k0; kb < k1; kb += 32) {
    uint16_t av[8], bv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      av[i] = bf16_rne(a_elem(g, m0 + lr, kb + lk + i));
      bv[i] = bf16_rne(b_elem(g, kb + lk + i, lr));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      As[lr * LS + lk + i] = av[i];
      Bs[lr * LS + lk + i] = bv[i];
    }
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) {
      const bf16x8 a = ld_frag(As + (wm * 32 + mi * 16 + (lane & 15)) * LS + 8 * (lane >> 4));
#pragma unroll
      for (int ni = 0; ni < 2; ++ni) {
        const bf16x8 b = ld_frag(Bs + (wn * 32 + ni * 16 + (lane & 15)) * LS + 8 * (lane >> 4));
        acc[mi][ni] = mfma16x16x32(a, b, acc[mi][ni]);
      }
    }
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * 32 + mi * 16 + 4 * (lane >> 4) + i;
        const int n = wn * 32 + ni * 16 + (lane & 15);
        const float a = acc[mi][ni][i];
        if (g.mode == kFwd) {
          if (m >= g.rows) continue;
          const int pos = m & (POS - 1);
          const float add = g.layer == 0
                                ? g.par[OFF_B0 + n] + (g.sig[m >> 4] * 0.02f) * sig_taps(g.par + OFF_WSIG, pos, n)
                                : g.par[off_b(g.layer) + n];
          const float z = a + add;
          g.Z[(size_t)m * CH + n] = z;
          g.H[(size_t)m * CH + n] = g.layer < NL - 1 ? (z > 0.f ? z : 0.f) : sigmoidf_(z);
        } else if (g.mode == kBwdData) {
          if (m >= g.rows) continue;
          g.dX[(size_t)m * CH + n] = g.Zprev[(size_t)m * CH + n] > 0.f ? a : 0.f;
        } else {
          g.part[((size_t)blockIdx.y * 256 + m) * CH + n] = a;
        }
      }
}

// Loss sum w (g - t)^2 / ngroups and its gradient at the last pre-activation.
__global__ __launch_bounds__(256) void shrink_loss_kernel(const float* __restrict__ Hl, const float* __restrict__ Tg,
                                                          const float* __restrict__ Wt, size_t n, float scale,
                                                          float* __restrict__ dZ, float* __restrict__ part) {
  __shared__ float red[4];
  float acc = 0.f;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const float gv = Hl[e], d = gv - Tg[e], w = Wt[e];
    acc += w * d * d;
    dZ[e] = 2.f * scale * w * d * gv * (1.f - gv);
  }
  acc = wave_sum_f32(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// grad[layer] from the per-chunk weight-gradient partials, summed in chunk order.
__global__ void shrink_wgrad_reduce_kernel(const float* __restrict__ part, int chunks, int layer,
                                           float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (KW + 4) * CH) return;
  const int m = e / CH, n = e % CH;
  if (m > KW && layer != 0) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += part[((size_t)c * 256 + m) * CH + n];
  if (m < KW)
    grad[off_w(layer) + m * CH + n] = s;
  else if (m == KW)
    grad[off_b(layer) + n] = s;
  else
    grad[OFF_WSIG + (m - KW - 1) * CH + n] = s;
}

__global__ void shrink_adam_kernel(float* __restrict__ p, float* __restrict__ mo, float* __restrict__ ve,
                                   const float* __restrict__ g, int n, float lr_t, float b1, float b2, float eps) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gv = g[e];
  const float m = b1 * mo[e] + (1.f - b1) * gv;
  const float v = b2 * ve[e] + (1.f - b2) * gv * gv;
  mo[e] = m;
  ve[e] = v;
  p[e] -= lr_t * m / (sqrtf(v) + eps);
}

size_t shrink_train_floats(int ngroups) {
  const size_t rows = (size_t)ngroups * POS;
  const size_t chunks = (rows + kShrinkChunk - 1) / kShrinkChunk;
  return 10 * rows * CH + chunks * 256 * CH + 1024;
}

static void bind(float* ws, int ngroups, float** Z, float** H, float** dA, float** dB, float** part,
                 float** lpart) {
  const size_t rc = (size_t)ngroups * POS * CH;
  for (int l = 0; l < NL; ++l) {
    Z[l] = ws + rc * l;
    H[l] = ws + rc * (NL + l);
  }
  *dA = ws + rc * 8;
  *dB = ws + rc * 9;
  *part = ws + rc * 10;
  const size_t rows = (size_t)ngroups * POS;
  const size_t chunks = (rows + kShrinkChunk - 1) / kShrinkChunk;
  *lpart = *part + chunks * 256 * CH;
}

static int run_forward(const float* par, const float* X, const float* sig, int ngroups, float* const* Z,
                       float* const* H, hipStream_t st) {
  const int rows = ngroups * POS;
  for (int l = 0; l < NL; ++l) {
    GemmArgs a{};
    a.mode = kFwd;
    a.layer = l;
    a.rows = rows;
    a.in = l == 0 ? X : H[l - 1];
    a.par = par;
    a.sig = sig;
    a.Z = Z[l];
    a.H = H[l];
    hipLaunchKernelGGL(shrink_gemm_kernel, dim3((unsigned)((rows + 63) / 64)), dim3(256), 0, st, a);
    GBM3D_TRY(check_launch("shrink_gemm fwd"));
  }
  return GBM3D_OK;
}

int shrink_forward(const float* par, const float* X, const float* sig, int ngroups, float* ws, float* gains,
                   hipStream_t st) {
  if (ngroups <= 0) return GBM3D_OK;
  float *Z[NL], *H[NL], *dA, *dB, *part, *lpart;
  bind(ws, ngroups, Z, H, &dA, &dB, &part, &lpart);
  H[NL - 1] = gains;
  return run_forward(par, X, sig, ngroups, Z, H, st);
}

int shrink_grad(const float* par, const float* X, const float* Tg, const float* Wt, const float* sig, int ngroups,
                float* ws, float* grad, float* loss_part, int* loss_blocks, hipStream_t st) {
  if (ngroups <= 0) return fail(GBM3D_ERR_ARG, "shrink_grad: empty batch");
  float *Z[NL], *H[NL], *dA, *dB, *part, *lpart;
  bind(ws, ngroups, Z, H, &dA, &dB, &part, &lpart);
  if (!loss_part) loss_part = lpart;
  GBM3D_TRY(run_forward(par, X, sig, ngroups, Z, H, st));
  const int rows = ngroups * POS;
  const size_t n = (size_t)rows * CH;
  const int lb = 1024;
  *loss_blocks = lb;
  hipLaunchKernelGGL(shrink_loss_kernel, dim3(lb), dim3(256), 0, st, H[NL - 1], Tg, Wt, n, 1.0f / (float)ngroups,
                     dA, loss_part);
  GBM3D_TRY(check_launch("shrink_loss"));
  const int chunks = (rows + kShrinkChunk - 1) / kShrinkChunk;
  float* dcur = dA;
  float* dnext = dB;
  for (int l = NL - 1; l >= 0; --l) {
    GemmArgs w{};
    w.mode = kBwdW;
    w.layer = l;
    w.rows = rows;
    w.chunk = kShrinkChunk;
    w.in = l == 0 ? X : H[l - 1];
    w.dZ = dcur;
    w.sig = sig;
    w.part = part;
    hipLaunchKernelGGL(shrink_gemm_kernel, dim3(4, (unsigned)chunks), dim3(256), 0, st, w);
    GBM3D_TRY(check_launch("shrink_gemm wgrad"));
    hipLaunchKernelGGL(shrink_wgrad_reduce_kernel, dim3(((KW + 4) * CH + 255) / 256), dim3(256), 0, st, part, chunks,
                       l, grad);
    GBM3D_TRY(check_launch("shrink_wgrad_reduce"));
    if (l == 0) break;
    GemmArgs d{};
    d.mode = kBwdData;
    d.layer = l;
    d.rows = rows;
    d.par = par;
    d.dZ = dcur;
    d.Zprev = Z[l - 1];
    d.dX = dnext;
    hipLaunchKernelGGL(shrink_gemm_kernel, dim3((unsigned)((rows + 63) / 64)), dim3(256), 0, st, d);
    GBM3D_TRY(check_launch("shrink_gemm dgrad"));
    float* t = dcur;
    dcur = dnext;
    dnext = t;
  }
  return GBM3D_OK;
}

int launch_shrink_adam(float* p, float* m, float* v, const float* g, int n, float lr_t, float b1, float b2, float eps,
                       hipStream_t st) {
  hipLaunchKernelGGL(shrink_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, m, v, g, n, lr_t, b1, b2, eps);
  return check_launch("shrink_adam");
}

}  // namespace gbm3d
