// k_ht.hip -- the step-1 filter: cycle-spun, 3-level bior1.5 (x,y) x Haar (z)
// hard threshold on the whole mosaic volume with circular boundaries
// (source/denoiseAndAggregateGPU.m:27-47, waveTransforms3D/myWavDec3DFFT.m,
// waveTransforms3D/myWavRec3DGPUfast.m), computed as direct separable
// filters instead of full-volume FFTs.
//
// Conventions (0-based, per axis of length L):
//   analysis   c[t] = sum_{k<n} h[k] x[(2t - k) mod L]           (conv, keep even)
//   synthesis  y[x] = sum_{k<n, x+k even} g[k] c[((x+k) mod L)/2] (upsample, corr)
// n = min(taps, L) reproduces fftn's truncation of a filter longer than the
// axis.  Along z the pair (2t-1, 2t) forms coefficient t and is rebuilt from it
// alone, so one z-pair of the input maps to one z-pair of the output.
//
// Levels: lo-kernel X_l -> X_{l+1} (approximation only); syn-kernel rebuilds
// X_l from its thresholded details (recomputed from X_l) plus the rebuilt
// X_{l+1}.  At level 0 cycle spin s reads the volume at (x - s) and writes
// the rebuilt voxel back at (x - s), which is circshift(., [s s s]) and its
// inverse.
#include <cmath>

#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {
constexpr float pCa = (float)(kCa / kCc), pCb = (float)(kCb / kCc);

// analysis lo / hi and synthesis lo / hi (getWaveFilters3Dbior.m:21-23,49-50)
// in the primed scaling of the register-blocked kernels below: analysis taps
// divided by Cc, synthesis taps times Cc and a gain g (1, or 1/2 to restore
// the z axis's factor as well)
__device__ __forceinline__ float a_lo(int k) {
  const float t[10] = {pCa, -pCa, -pCb, pCb, 1.f, 1.f, pCb, -pCb, -pCa, pCa};
  return t[k];
}
__device__ __forceinline__ float a_hi(int k) { return k == 4 ? -1.f : (k == 5 ? 1.f : 0.f); }
template <int GS>
__device__ __forceinline__ float s_lo(int k) {
  constexpr float c = GS ? 0.25f : 0.5f;
  return (k == 4 || k == 5) ? c : 0.f;
}
template <int GS>
__device__ __forceinline__ float s_hi(int k) {
  constexpr double g = GS ? 0.5 : 1.0;
  constexpr float a = (float)(kCc * kCa * g), b = (float)(kCc * kCb * g), c = (float)(0.5 * g);
  const float t[10] = {-a, -a, b, b, -c, c, -b, -b, a, a};
  return t[k];
}
}  // namespace

// z analysis of the pair (a = plane 2t, b = plane 2t-1) and its synthesis,
// unscaled (the synthesis factor Cc^2 sits in the in-plane taps, G = 1/2)
__device__ __forceinline__ float z_sum(float a, float b) { return a + b; }
__device__ __forceinline__ float z_dif(float a, float b) { return b - a; }
__device__ __forceinline__ float z_even(float lo, float hi) { return lo - hi; }  // plane 2t
__device__ __forceinline__ float z_odd(float lo, float hi) { return lo + hi; }   // plane 2t-1

struct Tau7 {
  float t[8];  // t[j] for subband j = 1..7 (t[0] unused)
};

// subband bits b = dim1_hi | dim2_hi<<1 | z_hi<<2 -> column j of C{i,j}
__device__ __forceinline__ int band_j(int b) { return b == 3 ? 4 : (b == 4 ? 3 : b); }

// tau.t[band_j(b)] selected over constant indices: the argument block stays in
// scalar registers (an indexed read of it is a memory load, and a wait for it
// would also wait for every load issued before it).
__device__ __forceinline__ float tau_of(const Tau7& tau, int b) {
  const int j = band_j(b);
  float t = tau.t[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) t = j == i ? tau.t[i] : t;
  return t;
}

// ---------------------------------------------------------------- lo kernel
constexpr int LT = 32;           // coefficients per tile edge
constexpr int LR = 2 * LT + 8;   // input rows/cols per tile
constexpr int LRP = LR + 1;

__global__ __launch_bounds__(256) void ht_lo_kernel(const float* __restrict__ X, int p, int q, int r, int sh,
                                                    float* __restrict__ Y, const int* __restrict__ run) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  __shared__ float Z[LR * LRP];   // z-lowpass plane, [col][row]
  __shared__ float T2[LT * LRP];  // after dim-2 filtering, [t2][row]
  const int t1b = blockIdx.x * LT, t2b = blockIdx.y * LT, t3 = blockIdx.z;
  const int rb = 2 * t1b - 9 - sh, cb = 2 * t2b - 9 - sh;
  const size_t pq = (size_t)p * q;
  const float* X0 = X + pq * pmod_i(2 * t3 - sh, r);
  const float* X1 = X + pq * pmod_i(2 * t3 - 1 - sh, r);
  for (int e = threadIdx.x; e < LR * LR; e += 256) {
    const int rr = e % LR, cc = e / LR;
    const size_t o = (size_t)pmod_i(rb + rr, p) + (size_t)p * pmod_i(cb + cc, q);
    Z[cc * LRP + rr] = X0[o] + X1[o];
  }
  __syncthreads();
  const int n1 = min(10, p), n2 = min(10, q);
  for (int e = threadIdx.x; e < LT * LR; e += 256) {
    const int rr = e % LR, t2l = e / LR;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 10; ++k)
      if (k < n2) acc += a_lo(k) * Z[(2 * t2l + 9 - k) * LRP + rr];
    T2[t2l * LRP + rr] = acc;
  }
  __syncthreads();
  const int ph = p / 2, qh = q / 2;
  for (int e = threadIdx.x; e < LT * LT; e += 256) {
    const int t1l = e % LT, t2l = e / LT;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 10; ++k)
      if (k < n1) acc += a_lo(k) * T2[t2l * LRP + 2 * t1l + 9 - k];
    const int t1 = t1b + t1l, t2 = t2b + t2l;
    if (t1 < ph && t2 < qh) Y[t1 + (size_t)ph * (t2 + (size_t)qh * t3)] = acc;
  }
}

// --------------------------------------------------------------- syn kernel
constexpr int SP = 32;            // output voxels per tile edge
constexpr int CP = SP / 2 + 5;    // coefficients per tile edge
constexpr int SR = 2 * CP + 8;    // input rows/cols per tile
constexpr int SRP = SR + 1;
constexpr int CPP = CP + 1;
constexpr int SPP = SP + 1;

// Store modes of the finest level (cycle spins accumulate into one volume):
// kStore  Y = v;  kAccum  Y = Y + v;  kFinal  Y = max((Y + v)/cs, 0), with Y
// read from acc (acc == nullptr reads 0).  Coarser levels always kStore.
enum { kStore = 0, kAccum = 1, kFinal = 2 };

// Offset of voxel (pr, pc, z) of a p x q x r volume: mosaic (column-major) for
// gd == 0, else group-contiguous with gd x gd blocks ([group][z][col][row]),
// the layout the aggregation reads.
__device__ __forceinline__ size_t vol_off(int pr, int pc, int z, int p, int q, int r, int gd) {
  if (gd == 0) return (size_t)pr + (size_t)p * ((size_t)pc + (size_t)q * z);
  const int g = pr / gd + (p / gd) * (pc / gd);
  return (size_t)(pr % gd) + (size_t)gd * ((size_t)(pc % gd) + (size_t)gd * ((size_t)z + (size_t)r * g));
}

// Store of one rebuilt voxel; the accumulator is always in mosaic layout.
__device__ __forceinline__ void put(float* Y, const float* acc, int pr, int pc, int z, float v, int p, int q, int r,
                                   int mode, float cs, int gd) {
  const size_t om = vol_off(pr, pc, z, p, q, r, 0);
  if (mode == kStore) {
    Y[om] = v;
    return;
  }
  const float a = acc ? acc[om] : 0.f;
  const float t = a + v;
  if (mode == kAccum) {
    Y[om] = t;
    return;
  }
  const float mean = t / cs;
  Y[vol_off(pr, pc, z, p, q, r, gd)] = mean > 0.f ? mean : 0.f;  // max(mean(Vrec,4), 0)
}

__global__ __launch_bounds__(256) void ht_syn_kernel(const float* __restrict__ X, const float* __restrict__ Ylo,
                                                     int p, int q, int r, int sh, Tau7 tau,
                                                     float* Y, const float* acc, int mode, float cs, int gd, const int* __restrict__ run) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  __shared__ float Z[2][SR * SRP];   // z-lo / z-hi planes, [col][row]
  __shared__ float A[2][CP * SRP];   // dim-2 lo / hi, [c][row]
  __shared__ float B[4][CP * CPP];   // bands b1 | b2<<1, [c][t]
  __shared__ float E[2][CP * SPP];   // dim-1 synthesis per dim-2 band, [c][x]
  const int x0 = blockIdx.x * SP, y0 = blockIdx.y * SP, t3 = blockIdx.z;
  const int rb = x0 - 9 - sh, cb = y0 - 9 - sh;
  const size_t pq = (size_t)p * q;
  const float* X0 = X + pq * pmod_i(2 * t3 - sh, r);
  const float* X1 = X + pq * pmod_i(2 * t3 - 1 - sh, r);
  for (int e = threadIdx.x; e < SR * SR; e += 256) {
    const int rr = e % SR, cc = e / SR;
    const size_t o = (size_t)pmod_i(rb + rr, p) + (size_t)p * pmod_i(cb + cc, q);
    const float a = X0[o], b = X1[o];
    Z[0][cc * SRP + rr] = z_sum(a, b);
    Z[1][cc * SRP + rr] = z_dif(a, b);
  }
  __syncthreads();
  const int n1 = min(10, p), n2 = min(10, q);
  const int ph = p / 2, qh = q / 2;
  float out[2][4];
  for (int zb = 0; zb < 2; ++zb) {
    // dim-2 analysis (lo, hi)
    for (int e = threadIdx.x; e < 2 * CP * SR; e += 256) {
      const int rr = e % SR, c = (e / SR) % CP, b2 = e / (SR * CP);
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 10; ++k)
        if (k < n2) acc += (b2 ? a_hi(k) : a_lo(k)) * Z[zb][(2 * c + 9 - k) * SRP + rr];
      A[b2][c * SRP + rr] = acc;
    }
    __syncthreads();
    // dim-1 analysis, threshold / approximation replacement
    for (int e = threadIdx.x; e < 4 * CP * CP; e += 256) {
      const int t = e % CP, c = (e / CP) % CP, bb = e / (CP * CP);
      const int b1 = bb & 1, b2 = bb >> 1;
      const int bits = b1 | (b2 << 1) | (zb << 2);
      float v;
      if (bits == 0) {
        v = Ylo[(size_t)pmod_i(x0 / 2 + t, ph) + (size_t)ph * ((size_t)pmod_i(y0 / 2 + c, qh) + (size_t)qh * t3)];
      } else {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < 10; ++k)
          if (k < n1) acc += (b1 ? a_hi(k) : a_lo(k)) * A[b2][c * SRP + 2 * t + 9 - k];
        // C .* (abs(C) > tau)  (denoiseAndAggregateGPU.m:37)
        v = acc * (fabsf(acc) > tau.t[band_j(bits)] ? 1.f : 0.f);
      }
      B[bb][c * CPP + t] = v;
    }
    __syncthreads();
    // dim-1 synthesis
    for (int e = threadIdx.x; e < 2 * CP * SP; e += 256) {
      const int x = e % SP, c = (e / SP) % CP, b2 = e / (SP * CP);
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        if (k < n1 && ((x + k) & 1) == 0) {
          const int t = (x + k) >> 1;
          acc += s_lo<0>(k) * B[b2 << 1][c * CPP + t] + s_hi<0>(k) * B[1 | (b2 << 1)][c * CPP + t];
        }
      }
      E[b2][c * SPP + x] = acc;
    }
    __syncthreads();
    // dim-2 synthesis -> this thread's 4 voxels of the z-band plane
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = threadIdx.x + 256 * i;
      const int x = e % SP, y = e / SP;
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        if (k < n2 && ((y + k) & 1) == 0) {
          const int c = (y + k) >> 1;
          acc += s_lo<1>(k) * E[0][c * SPP + x] + s_hi<1>(k) * E[1][c * SPP + x];
        }
      }
      out[zb][i] = acc;
    }
    __syncthreads();
  }
  // z synthesis: plane 2t3 <- g[0], plane 2t3-1 <- g[1]; lo g = [c c], hi g = [-c c]
  const int ze = pmod_i(2 * t3 - sh, r), zo = pmod_i(2 * t3 - 1 - sh, r);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int x = e % SP, y = e / SP;
    if (x0 + x >= p || y0 + y >= q) continue;
    const int pr = pmod_i(x0 + x - sh, p), pc = pmod_i(y0 + y - sh, q);
    put(Y, acc, pr, pc, ze, z_even(out[0][i], out[1][i]), p, q, r, mode, cs, gd);
    put(Y, acc, pr, pc, zo, z_odd(out[0][i], out[1][i]), p, q, r, mode, cs, gd);
  }
}

// ------------------------------------------------ register-blocked kernels
// The same level operators for axes of at least 10 samples (no filter
// truncation).  Each thread filters a run of outputs from one register
// window, computing the low and the high band from the same loads; zero taps
// are skipped at compile time; circular indices come from per-tile tables.

// The filters in lifting form, each axis's analysis without its factor Cc.
// With the input pairs D_j = v[2j] - v[2j+1], S_j = v[2j] + v[2j+1] (a_lo and
// s_hi are symmetric up to sign) the reference's coefficients are
//   lo[i] = Ca (D_i - D_{i+4}) + Cb (D_{i+3} - D_{i+1}) + Cc S_{i+2},
//   hi[i] = Cc D_{i+2};
// the kernels compute lo' = lo / Cc and hi' = hi / Cc (taps Ca/Cc, Cb/Cc, 1:
// 4 operations per low coefficient, none per high one), and along z the sum
// a + b and difference b - a of the Haar pair.  A level's coefficients are
// then (1/Cc)^3 = 2 sqrt(2) times the reference's, and its approximation,
// the next level's input, too: level l (0 = finest) carries (2 sqrt 2)^(l+1),
// and run_ht scales its thresholds by that.  With Q_j = Ca (hi[j+4] - hi[j])
// + Cb (hi[j+1] - hi[j+3]), the synthesis outputs of coefficient window
// position j are Q_j + Cc (lo[j+2] - hi[j+2]) (even output 2j) and Q_j + Cc
// (lo[j+2] + hi[j+2]) (odd output 2j-1); on the primed coefficients the taps
// are Cc Ca, Cc Cb and Cc Cc = 1/2, which restore the axis's own factor.  The
// in-plane synthesis the z synthesis follows takes its taps times a gain G:
// G = 1/2 restores z's factor Cc^2 as well (the z steps are then one addition
// or subtraction each), and G = 1/4 also takes the mean over two cycle spins
// (a power of two: the results are exactly those of G = 1/2, halved).  Every
// value is a fixed function of its own inputs, whatever the run it is
// computed in, so kernels that tile the axes differently agree bitwise.
// synthesis taps on primed coefficients, times the gain 2^-GS
template <int GS>
struct STaps {
  static constexpr double g = GS == 0 ? 1.0 : (GS == 1 ? 0.5 : 0.25);
  static constexpr float a = (float)(kCc * kCa * g), b = (float)(kCc * kCb * g), c = (float)(0.5 * g);
};

// lo'[i] = sum_k a_lo(k) v[2i+9-k] / Cc, hi'[i] = sum_k a_hi(k) v[2i+9-k] / Cc, i < R
template <int R, int NV>
__device__ __forceinline__ void ana_lo_hi(const float (&v)[NV], float (&lo)[R], float (&hi)[R]) {
  static_assert(NV >= 2 * R + 8, "analysis window");
  float D[R + 4], S[R + 4];
#pragma unroll
  for (int j = 0; j < R + 4; ++j) {
    D[j] = v[2 * j] - v[2 * j + 1];
    S[j] = v[2 * j] + v[2 * j + 1];
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    lo[i] = fmaf(pCa, D[i] - D[i + 4], fmaf(pCb, D[i + 3] - D[i + 1], S[i + 2]));
    hi[i] = D[i + 2];
  }
}

template <int R, int NV>
__device__ __forceinline__ void ana_lo(const float (&v)[NV], float (&lo)[R]) {
  static_assert(NV >= 2 * R + 8, "analysis window");
  float D[R + 4], S[R + 4];
#pragma unroll
  for (int j = 0; j < R + 4; ++j) {
    D[j] = v[2 * j] - v[2 * j + 1];
    S[j] = v[2 * j] + v[2 * j + 1];
  }
#pragma unroll
  for (int i = 0; i < R; ++i) lo[i] = fmaf(pCa, D[i] - D[i + 4], fmaf(pCb, D[i + 3] - D[i + 1], S[i + 2]));
}

// the high band of ana_lo_hi alone (reads v[4 .. 2R+3])
template <int R, int NV>
__device__ __forceinline__ void ana_hi(const float (&v)[NV], float (&hi)[R]) {
  static_assert(NV >= 2 * R + 8, "analysis window");
#pragma unroll
  for (int i = 0; i < R; ++i) hi[i] = v[2 * i + 4] - v[2 * i + 5];
}

template <int GS>
__device__ __forceinline__ float syn_q(const float* hi, int j) {
  using T = STaps<GS>;
  return fmaf(T::a, hi[j + 4] - hi[j], T::b * (hi[j + 1] - hi[j + 3]));
}

// out[i] = sum_{k: i+k even} s_lo(k) lo[(i+k)/2] + s_hi(k) hi[(i+k)/2], i < R2
// (the window starts at an even output)
template <int R2, int GS, int NC>
__device__ __forceinline__ void syn_lo_hi(const float (&lo)[NC], const float (&hi)[NC], float (&out)[R2]) {
  static_assert(NC >= R2 / 2 + 5 && R2 % 2 == 0, "synthesis window");
  using T = STaps<GS>;
#pragma unroll
  for (int j = 0; j <= R2 / 2; ++j) {
    const float q = syn_q<GS>(hi, j);
    if (j < R2 / 2) out[2 * j] = fmaf(T::c, lo[j + 2] - hi[j + 2], q);
    if (j >= 1) out[2 * j - 1] = fmaf(T::c, lo[j + 2] + hi[j + 2], q);
  }
}

// the same synthesis for a window that starts at an odd output 2m-1, its
// coefficient window starting at m: out[i] = sum_{k: i+k odd} s_lo(k)
// lo[(i+k-1)/2] + s_hi(k) hi[(i+k-1)/2]
template <int R2, int GS, int NC>
__device__ __forceinline__ void syn_lo_hi_odd(const float (&lo)[NC], const float (&hi)[NC], float (&out)[R2]) {
  static_assert(NC >= R2 / 2 + 4 && R2 % 2 == 0, "synthesis window");
  using T = STaps<GS>;
#pragma unroll
  for (int j = 0; j < R2 / 2; ++j) {
    const float q = syn_q<GS>(hi, j);
    out[2 * j] = fmaf(T::c, lo[j + 2] + hi[j + 2], q);
    out[2 * j + 1] = fmaf(T::c, lo[j + 2] - hi[j + 2], q);
  }
}

// C .* (abs(C) > tau) (denoiseAndAggregateGPU.m:37) as one compare and one
// select; a NaN coefficient is kept, as the product keeps it (a coefficient
// at or below tau becomes +0 instead of a signed zero, which no later sum or
// the final max(., 0) can tell apart).  A NaN tau arrives as +inf (run_ht).
__device__ __forceinline__ float hard_thr(float c, float t) { return fabsf(c) <= t ? 0.f : c; }

constexpr int kSynPairsPerBlock = 4;  // z-pairs per block of the register-blocked synthesis

namespace fs {
constexpr int P = 32, C = 21, R = 50;    // output tile, coefficients, input rows per edge
constexpr int ZS = R + 1, AS = R + 1, BS = C + 1, ES = P + 16;  // ES: the two half-waves of phase 5 hit disjoint banks
constexpr int REG0 = 2 * R * ZS;  // Z (later B)
constexpr int REG1 = 4 * C * AS;  // A (later E)
static_assert(8 * C * BS <= REG0, "B fits in Z");
static_assert(4 * C * ES <= REG1, "E fits in A");
}  // namespace fs

// One block per 32 x 32 output tile and a run of `zpb` z-pairs: the loads of
// the next pair (input planes and the coarser reconstruction) are issued into
// registers before the current pair's filter phases run, so their latency
// overlaps the LDS work.
// nsp = 2: one launch for both cycle spins of a coarse level (shift 0), the
// second spin's workgroups after the first's, reading X2 / Ylo2, writing Y2.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ht_syn_fast_kernel(
    const float* __restrict__ X, const float* __restrict__ Ylo, int p, int q, int r, int sh, Tau7 tau, float* Y,
    const float* acc, int mode, float cs, int gd, int zpb, int nsp, const float* __restrict__ X2,
    const float* __restrict__ Ylo2, float* Y2, const int* __restrict__ run) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  using namespace fs;
  __shared__ float lds[REG0 + REG1];
  __shared__ int ridx[R], cidx[R], yrow[C];
  float* Z = lds;
  float* B = lds;
  float* A = lds + REG0;
  float* E = lds + REG0;
  const int tid = threadIdx.x;
  __builtin_assume(tid < 256);  // 256-thread blocks: drops the guards of full element rounds
  const int nbx = gridDim.x, nby = gridDim.y;
  const int lb = xcd_remap(blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z), nbx * nby * gridDim.z);
  const int x0 = (lb % nbx) * P, y0 = ((lb / nbx) % nby) * P;
  const int nzc = gridDim.z / nsp, zc = (lb / (nbx * nby)) % nzc;
  if (lb / (nbx * nby) >= nzc) {  // second spin
    X = X2;
    Ylo = Ylo2;
    Y = Y2;
  }
  const int tz0 = zc * zpb, tz1 = min(r / 2, tz0 + zpb);
  const int ph = p >> 1, qh = q >> 1;
  if (tid < R)
    ridx[tid] = pmod_i(x0 - 9 - sh + tid, p);
  else if (tid < 2 * R)
    cidx[tid - R] = pmod_i(y0 - 9 - sh + tid - R, q);
  else if (tid < 2 * R + C)
    yrow[tid - 2 * R] = pmod_i(x0 / 2 + tid - 2 * R, ph);  // coarse rows of the tile
  __syncthreads();
  const size_t pq = (size_t)p * q;
  constexpr int NIT = (R * R + 255) / 256;
  // in-plane byte offsets of this thread's window elements, fixed over the z
  // loop (planes of fewer than 2^30 elements, read through buffer loads;
  // larger planes take 64-bit addresses)
  const bool buf = pq < ((size_t)1 << 30);
  uint32_t offb[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int e = tid + 256 * it;
    offb[it] = e < R * R ? 4u * ((uint32_t)ridx[e % R] + (uint32_t)p * (uint32_t)cidx[e / R]) : 0u;
  }
  // phase-3 roles (fixed per thread): coefficient column, run, dim-2 band, z
  // band.  The 6C tasks of the three detail (z band, dim-2 band) pairs run on
  // threads 0 .. 6C-1 (waves 0 and 1), the 2C tasks of the approximation pair,
  // whose low band comes from Ylo (no low-pass analysis), on threads
  // 128 .. 128+2C-1 (wave 2): no wave runs both kinds
  static_assert(6 * C <= 128 && 128 + 2 * C <= 256, "phase-3 roles");
  const int t3r = tid < 128 ? (tid < 6 * C ? tid + 2 * C : -1) : (tid < 128 + 2 * C ? tid - 128 : -1);
  const int c3 = t3r % C, run3 = (t3r / C) & 1, b23 = (t3r / (2 * C)) & 1, zb3 = t3r / (4 * C);
  const bool coarse = t3r >= 0 && t3r < 2 * C;
  const float tl = tau_of(tau, (b23 << 1) | (zb3 << 2)), th = tau_of(tau, (b23 << 1) | (zb3 << 2) | 1);
  float a[NIT], b[NIT], yl[11];
  const size_t ycol = coarse ? (size_t)ph * pmod_i(y0 / 2 + c3, qh) : 0;  // this thread's coarse column
  auto fetch = [&](int t3) {
    const float* X0 = X + pq * pmod_i(2 * t3 - sh, r);
    const float* X1 = X + pq * pmod_i(2 * t3 - 1 - sh, r);
    if (buf) {
      const __amdgpu_buffer_rsrc_t r0 = buf_rsrc(X0, (uint32_t)(4 * pq)), r1 = buf_rsrc(X1, (uint32_t)(4 * pq));
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        a[it] = b[it] = 0.f;
        if (tid + 256 * it < R * R) {
          a[it] = buf_load(r0, offb[it]);
          b[it] = buf_load(r1, offb[it]);
        }
      }
    } else {
#pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int e = tid + 256 * it;
        a[it] = b[it] = 0.f;
        if (e < R * R) {
          const size_t o = (size_t)ridx[e % R] + (size_t)p * cidx[e / R];
          a[it] = X0[o];
          b[it] = X1[o];
        }
      }
    }
    if (coarse) {
      const float* Yl = Ylo + (size_t)ph * qh * t3 + ycol;
#pragma unroll
      for (int i = 0; i < 11; ++i) yl[i] = Yl[yrow[run3 * 10 + i]];
    }
  };
  if (tz0 < tz1) fetch(tz0);
  for (int t3 = tz0; t3 < tz1; ++t3) {
    // z analysis of the pair (2t3-1, 2t3) from the prefetched registers
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = tid + 256 * it;
      if (e < R * R) {
        const int rr = e % R, cc = e / R;
        Z[cc * ZS + rr] = z_sum(a[it], b[it]);
        Z[R * ZS + cc * ZS + rr] = z_dif(a[it], b[it]);
      }
    }
    float ycur[11];
#pragma unroll
    for (int i = 0; i < 11; ++i) ycur[i] = yl[i];
    if (t3 + 1 < tz1) fetch(t3 + 1);  // in flight during the phases below
    __syncthreads();
    // dim-2 analysis: two runs of 11 coefficients, [0,11) and [10,21), per (z band, row)
    if (tid < 2 * 2 * R) {
      const int rr = tid % R, run = (tid / R) & 1, zb = tid / (2 * R);
      const int c0 = run * 10;
      float v[30];
#pragma unroll
      for (int j = 0; j < 30; ++j) v[j] = Z[zb * R * ZS + (2 * c0 + j) * ZS + rr];
      float lo[11], hi[11];
      ana_lo_hi<11>(v, lo, hi);
#pragma unroll
      for (int i = 0; i < 11; ++i) {
        A[((zb * 2 + 0) * C + c0 + i) * AS + rr] = lo[i];
        A[((zb * 2 + 1) * C + c0 + i) * AS + rr] = hi[i];
      }
    }
    __syncthreads();
    // dim-1 analysis + hard threshold (the approximation is the rebuilt coarser level)
    if (t3r >= 0) {
      const int t0 = run3 * 10;
      const float* Ac = A + ((zb3 * 2 + b23) * C + c3) * AS + 2 * t0;
      float* Bl = B + ((zb3 * 4 + (b23 << 1)) * C + c3) * BS + t0;
      float* Bh = Bl + C * BS;
      float v[30], hi[11];
      if (coarse) {
#pragma unroll
        for (int j = 4; j < 26; ++j) v[j] = Ac[j];
        ana_hi<11>(v, hi);
#pragma unroll
        for (int i = 0; i < 11; ++i) Bl[i] = ycur[i];
      } else {
#pragma unroll
        for (int j = 0; j < 30; ++j) v[j] = Ac[j];
        float lo[11];
        ana_lo_hi<11>(v, lo, hi);
#pragma unroll
        for (int i = 0; i < 11; ++i) Bl[i] = hard_thr(lo[i], tl);
      }
#pragma unroll
      for (int i = 0; i < 11; ++i) Bh[i] = hard_thr(hi[i], th);
    }
    __syncthreads();
    // dim-1 synthesis: runs of 16 outputs per (z band, dim-2 band, coefficient column)
    if (tid < 2 * 2 * 2 * C) {
      const int c = tid % C, xr = (tid / C) & 1, b2 = (tid / (2 * C)) & 1, zb = tid / (4 * C);
      const int xl = xr * 16;
      const float* Bl = B + ((zb * 4 + (b2 << 1)) * C + c) * BS + xl / 2;
      const float* Bh = Bl + C * BS;
      float lo[13], hi[13], out[16];
#pragma unroll
      for (int j = 0; j < 13; ++j) {
        lo[j] = Bl[j];
        hi[j] = Bh[j];
      }
      syn_lo_hi<16, 0>(lo, hi, out);
#pragma unroll
      for (int i = 0; i < 16; ++i) E[((zb * 2 + b2) * C + c) * ES + xl + i] = out[i];
    }
    __syncthreads();
    // dim-2 synthesis (runs of 4), then z synthesis and the store
    {
      const int x = tid % P, yq = (tid / P) * 4;
      float o[2][4];
#pragma unroll
      for (int zb = 0; zb < 2; ++zb) {
        float lo[7], hi[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) {
          lo[j] = E[((zb * 2 + 0) * C + yq / 2 + j) * ES + x];
          hi[j] = E[((zb * 2 + 1) * C + yq / 2 + j) * ES + x];
        }
        syn_lo_hi<4, 1>(lo, hi, o[zb]);
      }
      const int ze = pmod_i(2 * t3 - sh, r), zo = pmod_i(2 * t3 - 1 - sh, r);
      if (x0 + x < p) {
        const int pr = ridx[x + 9];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (y0 + yq + i >= q) continue;
          const int pc = cidx[yq + i + 9];
          put(Y, acc, pr, pc, ze, z_even(o[0][i], o[1][i]), p, q, r, mode, cs, gd);
          put(Y, acc, pr, pc, zo, z_odd(o[0][i], o[1][i]), p, q, r, mode, cs, gd);
        }
      }
    }
    // No barrier here: the next pair first writes Z (over B, last read in phase 4,
    // behind the barrier above) and writes A (over E, read in phase 5) only
    // after its own first barrier, which every thread reaches after phase 5.
  }
}

// ------------------------------------- finest level, both cycle spins fused
// Cycle spin s (of 2) in original coordinates: circshift(V,[s s s]), the
// decimated transform and the inverse shift amount to the same analysis and
// synthesis with the coefficient anchors at parity s in x, y and z:
//   C(a) = sum_k h[k] V(a - k),  a = s (mod 2);  Vrec_s(o) = sum_k g[k] C(o + k).
// One workgroup owns a 64 x 32 output column (rows x0.., columns y0..) over
// all r planes and walks the z anchors az = 0..r-1 (spin az & 1, planes az-1
// and az); each output plane is complete after two consecutive anchors and
// Vavg = max((Vrec_0 + Vrec_1) / 2, 0) is stored once.  Spin s's x coefficient
// j has its anchor at x0 + s + 2j and reads window rows s + 2j .. s + 2j + 9
// (window row w = image row x0 - 9 + w, 83 rows); y likewise (51 columns).
// Per anchor, three phases and two barriers:
//   A  (thread = window row w, run of 7 y coefficients): the thread's 23
//      window columns of both planes are in registers; z analysis, dim-2
//      analysis of both z bands -> A[zb][b2][c][w];
//   B  (thread = (zb, b2, c), run of 22 output rows): dim-1 analysis of 16
//      coefficients, hard threshold (or, for the approximation band, the
//      rebuilt coarser level), dim-1 synthesis -> E[zb][b2][c][x];
//   C  (thread = output row x, 8 output columns): dim-2 synthesis of both z
//      bands, z synthesis, spin combine, store.
// The planes of the next anchor and the coarse values it needs are loaded
// during the current one.  The per-coefficient and per-output arithmetic is
// that of the per-spin kernels (lifting helpers), so the two agree bitwise.
namespace f2 {
constexpr int TX = 64, TY = 32;            // output tile: rows x columns
constexpr int WX = TX + 19, WY = TY + 19;  // window rows, columns (both spins)
constexpr int C = TY / 2 + 5;              // y coefficients per spin (21)
constexpr int JX = TX / 2 + 5;             // x coefficients per spin (37)
constexpr int NJ = 23;                     // window columns per phase-A thread (spin offset + 22)
constexpr int NA = 3 * WX;                 // phase-A tasks: 83 rows x 3 runs of 7 coefficients
constexpr int ND = 3 * 3 * C;              // phase-B detail tasks: 63 (zb, b2, c) x 3 runs
constexpr int NCO = 3 * C;                 // phase-B approximation-band tasks (wave 3)
constexpr int RB = 22, NB = 16;            // phase-B run: output rows, coefficients
// A's row stride: odd (consecutive c in distinct banks), and 7 AS = 19 (mod 64)
// so that the two runs a phase-A wave straddles fill disjoint banks
constexpr int AS = 85, ES = TX + 1;
constexpr int NYC = C * JX;                // coarse values staged per anchor
constexpr int NYT = (NYC + 255) / 256;
static_assert(NA <= 256 && ND <= 192 && 192 + NCO <= 256, "thread roles");
static_assert(42 + 1 + 2 * (NB - 1) + 9 < WX, "phase-B rows inside the window");
static_assert(14 * 2 + NJ <= WY, "phase-A columns inside the window");
// LDS-staged planes (p a multiple of 4): the window plane from row x0 - 12 in
// NQ aligned quads of 4 rows per column, one 128-bit direct-to-LDS load per
// quad; window row w at staged row w + 3
constexpr int NQ = 22, PS = 4 * NQ;        // quads per column, staged column stride
constexpr int NPQ = NQ * WY;               // quads per plane (1,122)
constexpr int NPC = (NPQ + 63) / 64;       // wave chunks of 64 quads (18)
static_assert(PS >= WX + 3 + 2 && NPC <= 20, "staged plane");
// and (ph a multiple of 4) the coarse values of an anchor from coarse row
// x0/2 in YQ quads per coarse column, one direct-to-LDS load per wave
constexpr int YQ = 10, YS = 4 * YQ;        // quads per coarse column, staged stride
static_assert(YS >= JX + 1 && C * YQ <= 256, "staged coarse values");
}  // namespace f2

// CS = 2: the finest level, both cycle spins fused (above).  CS = 1: one
// level of one cycle spin with no shift (the coarse levels), every anchor at
// parity 0 (z pairs (2t-1, 2t)), both rebuilt planes stored as they are
// (Y = Vrec, mosaic layout); grid z selects the spin (X/Ylo0/Y or
// X1/Ylo1/Y1).  The arithmetic per coefficient and per output is that of
// ht_syn_fast_kernel, so the two agree bitwise.
// LP (CS = 2, p a multiple of 8): each plane is staged through LDS by 128-bit
// direct-to-LDS loads (4.4 per thread instead of 23 32-bit loads, each
// column's window rows read once instead of by the up to three runs whose
// coefficients use it); phase A reads its 23 columns from there.  The staged
// plane is reloaded after the first barrier and has landed at the second, and
// so have the coarse values of the next anchor (two buffers, one per spin
// parity).  The staging's LDS (24 KB) leaves room for two workgroups per CU.
// The values are the same, so both forms agree bitwise.
// WG (LP, gd = 8): the kernel also writes the step-1 weight of every 8x8
// group of its tile, wgt[g] = 1/(D + eps) with D the sums of
// getBM3DAggWeights(3,1) over the group's stack (the output tile holds whole
// groups over all planes): per pixel the terms are added plane by plane as
// the planes are stored, then summed over the group in agg1_prep's order, so
// the weights are those of agg1_prep_wave_kernel bit for bit.
template <int CS, bool LP = false, bool WG = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LP ? 2 : 3))) void ht_fine_kernel(
    const float* __restrict__ X, const float* __restrict__ X1, const float* __restrict__ Ylo0,
    const float* __restrict__ Ylo1, int p, int q, int r, Tau7 tau, float* __restrict__ Y, float* __restrict__ Y1,
    int gd, const int* __restrict__ run, float* __restrict__ wgt) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  using namespace f2;
  static_assert(CS == 1 || CS == 2, "cycle spins per launch");
  static_assert(!LP || CS == 2, "LDS-staged planes: the fused finest level");
  static_assert(!WG || LP, "step-1 weights: the LDS-staged fused finest level (gd = 8)");
  if (CS == 1 && blockIdx.z) {  // the second spin of a coarse level
    X = X1;
    Ylo0 = Ylo1;
    Y = Y1;
  }
  __shared__ float A[4 * C * AS];
  __shared__ float E[4 * C * ES];
  __shared__ float Yc[LP ? 4 : NYC];
  __shared__ int cidxp[WY];             // p x window column (circular)
  __shared__ int yrow[2][JX], ycol[2][C];  // per spin: coarse rows, ph x coarse columns (circular)
  __shared__ __align__(16) float PB[LP ? WY * PS : 4];  // the staged plane [column][staged row]
  __shared__ int qrow[NQ];                               // first image row of each staged quad
  __shared__ __align__(16) float YcL[LP ? 2 * C * YS : 4];  // staged coarse values [az & 1][column][coarse row - x0/2]
  __shared__ int yqrow[YQ];                              // first coarse row of each staged quad
  const int tid = threadIdx.x;
  __builtin_assume(tid < 256);
  const int nbx = gridDim.x, nby = gridDim.y;
  const int lb = xcd_remap(blockIdx.x + nbx * blockIdx.y, nbx * nby);
  int bx, by;
  tile_order8(lb, nbx, nby, bx, by);
  const int x0 = bx * TX, y0 = by * TY;
  const int ph = p >> 1, qh = q >> 1, rh = r >> 1;
  if (tid < WY) {
    cidxp[tid] = p * pmod_i(y0 - 9 + tid, q);
  } else if (tid < WY + 2 * JX) {
    const int k = tid - WY, sp = k / JX;
    yrow[sp][k % JX] = pmod_i((x0 >> 1) + sp + k % JX, ph);
  } else if (tid < WY + 2 * JX + 2 * C) {
    const int k = tid - WY - 2 * JX, sp = k / C;
    ycol[sp][k % C] = ph * pmod_i((y0 >> 1) + sp + k % C, qh);
  } else if (LP && tid < WY + 2 * JX + 2 * C + NQ) {
    // p % 4 == 0: the quad lies inside one column whatever the circular shift
    const int k = tid - WY - 2 * JX - 2 * C;
    qrow[k] = pmod_i(x0 - 12 + 4 * k, p);
  } else if (LP && tid < WY + 2 * JX + 2 * C + NQ + YQ) {
    const int k = tid - WY - 2 * JX - 2 * C - NQ;
    yqrow[k] = pmod_i((x0 >> 1) + 4 * k, ph);
  }
  __syncthreads();
  const size_t pq = (size_t)p * q;
  // ---- phase-A role: window row wa, first y coefficient 7 ra
  const int wa = tid % WX, ra = tid / WX;
  const bool actA = tid < NA;
  const int rowA = pmod_i(x0 - 9 + wa, p);
  // interior column range (no wrap): column offsets p * j on a per-thread base
  const bool colin = y0 - 9 >= 0 && y0 + TY + 9 < q;
  const uint32_t voffA = 4u * (uint32_t)(rowA + (colin ? p * (y0 - 9 + 14 * ra) : 0));
  const uint32_t pbytes = (uint32_t)(4 * pq), cstride = (uint32_t)(4 * p);
  auto fetch_plane = [&](float (&dst)[NJ], int z) {
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(X + pq * z, pbytes);
    if (!actA) return;
    if (colin) {
      // column j at scalar offset 4 p j, stepped by an opaque scalar add so
      // that the 23 offsets are not all kept live across the anchor loop
      uint32_t so = 0;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        dst[j] = buf_load(rs, voffA, so);
        so = sadd_u32(so, cstride);
      }
    } else {
#pragma unroll
      for (int j = 0; j < NJ; ++j) dst[j] = buf_load(rs, 4u * (uint32_t)(rowA + cidxp[14 * ra + j]));
    }
  };
  // ---- phase-B role: band row kb = (zb * 2 + b2) * C + c, output rows xb .. xb + 21
  // detail bands on threads 0 .. 188 (waves 0-2), the approximation band on
  // threads 192 .. 254 (wave 3): no wave runs both kinds
  int kb = -1, xr = 0;
  if (tid < ND) {
    kb = C + tid % (3 * C);
    xr = tid / (3 * C);
  } else if (tid >= 192 && tid < 192 + NCO) {
    kb = (tid - 192) % C;
    xr = (tid - 192) / C;
  }
  const bool actB = kb >= 0, coarseB = kb >= 0 && kb < C;
  const int xb = xr == 0 ? 0 : (xr == 1 ? RB : TX - RB);  // runs at 0, 22, 42 (42, 43 stored by the middle run)
  const int kbc = kb < 0 ? 0 : kb;
  const int zbB = kbc / (2 * C), b2B = (kbc / C) & 1, cB = kbc % C;
  const float tl = tau_of(tau, (b2B << 1) | (zbB << 2)), th = tau_of(tau, (b2B << 1) | (zbB << 2) | 1);
  // ---- phase-C role: output row xo, output columns 8 yc .. 8 yc + 7 (yc = wave index)
  const int xo = tid & (TX - 1);
  const int yc = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool rowC = x0 + xo < p;
  // output offsets: voffset ob0 (per lane) + a uniform part per column i
  // (mosaic: p i; group layout gd x gd: gd (i % 8) within a group column for
  // gd >= 8, and for gd = 4 the next group column after 4)
  const int pc0 = y0 + 8 * yc;
  int ob0 = 0, zstride, cst, gst = 0, cm = 7, csh = 3;
  if (gd == 0) {
    ob0 = x0 + xo + p * pc0;
    zstride = p * q;
    cst = p;
  } else {
    const int pr = x0 + xo;
    const int g = pr / gd + (p / gd) * (pc0 / gd);
    ob0 = pr % gd + gd * (pc0 % gd + gd * r * g);
    zstride = gd * gd;
    cst = gd;
    if (gd == 4) {
      cm = 3;
      csh = 2;
      gst = gd * gd * r * (p / gd);
    }
  }
  const __amdgpu_buffer_rsrc_t ys = buf_rsrc(Y, (uint32_t)(4 * pq * r));
  // CS = 1: the rebuilt plane as it is
  auto store_vrec = [&](int z, const float (&v)[8]) {
    if (!rowC) return;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (pc0 + i >= q) continue;
      const int so = (i & cm) * cst + (i >> csh) * gst + z * zstride;
      buf_store(ys, 4u * (uint32_t)ob0, v[i], 4u * (uint32_t)so);
    }
  };
  // WG: per pixel (row xo & 7 of its group, column i), the weight terms of
  // the planes stored so far: |third diff| down the rows, along the columns,
  // |diff| across the planes (agg1_prep_wave_kernel's s1, s2, s3)
  float ws1[8], ws2[8], ws3[8], wprev[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) ws1[i] = ws2[i] = ws3[i] = wprev[i] = 0.f;
  // the value of the lane below in the same 16-lane row (0 past its end)
  auto below = [](float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xF, 0xF, true));
  };
  auto weigh_plane = [&](int z, const float (&x)[8]) {
    const bool down = (xo & 7) + 3 < 8;  // the row's group holds rows +1 .. +3
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (z > 0) ws3[i] += fabsf(x[i] - wprev[i]);
      wprev[i] = x[i];
      if (i + 3 < 8) ws2[i] += fabsf(third_diff(x[i], x[i + 1], x[i + 2], x[i + 3]));
      // down the rows: the differences of the differences, each from the lane
      // below (a group's 8 rows are 8 consecutive lanes of a 16-lane DPP row,
      // and row u <= 4 reaches rows up to u + 3): the steps of third_diff on
      // rows u .. u + 3, the same operations on the same values
      const float d = below(x[i]) - x[i];
      const float e = below(d) - d;
      const float f = below(e) - e;
      if (down) ws1[i] += fabsf(f);
    }
  };
  // 16-byte stores: with p a multiple of 4 the four rows of a lane quad are
  // all inside the volume or all outside, never straddle a group, and one
  // column's four are contiguous and 16-byte aligned in every layout once Y is
  const bool st16 = (p & 3) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
  // the value of lane ql ^ 1 (quad_perm [1,0,3,2]) / ql ^ 2 ([2,3,0,1]) of the quad
  auto xdpp = [](float x, int b) {
    return __int_as_float(b ? __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, true)
                            : __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, true));
  };
  auto store_plane = [&](int z, const float (&u)[8], const float (&v)[8]) {
    if (!rowC) return;
    float val[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      // mean(Vrec,4), then max(.,0) (denoiseAndAggregateGPU.m:45): u and v
      // are the two spins' planes already halved by the synthesis taps
      const float mean = u[i] + v[i];
      val[i] = mean > 0.f ? mean : 0.f;
    }
    if constexpr (WG) weigh_plane(z, val);
    if (st16) {
      // the quad's 4 x 8 values transposed by column pairs (two butterfly
      // exchanges within the quad), so that lane ql holds columns 2 ql and
      // 2 ql + 1 of the quad's 4 rows: t[2 r + h] = row 4k + r, column 2 ql + h
      const int ql = xo & 3;
      float t[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) t[i] = val[i];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const bool hi = (ql >> b) & 1;
#pragma unroll
        for (int c0 = 0; c0 < 4; ++c0) {
          if ((c0 >> b) & 1) continue;
          const int c1 = c0 | (1 << b);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const float a0 = t[2 * c0 + h], a1 = t[2 * c1 + h];
            const float from1 = xdpp(a1, b), from0 = xdpp(a0, b);
            t[2 * c0 + h] = hi ? from1 : a0;
            t[2 * c1 + h] = hi ? a1 : from0;
          }
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * ql + h;
        if (pc0 + i >= q) continue;
        const int so = (i & cm) * cst + (i >> csh) * gst + z * zstride;
        typedef __attribute__((ext_vector_type(4))) unsigned u32x4;
        const u32x4 w = {__float_as_uint(t[h]), __float_as_uint(t[2 + h]), __float_as_uint(t[4 + h]),
                         __float_as_uint(t[6 + h])};
        __builtin_amdgcn_raw_buffer_store_b128(w, ys, 4u * (uint32_t)(ob0 - ql + so), 0, 0);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (pc0 + i >= q) continue;
        const int so = (i & cm) * cst + (i >> csh) * gst + z * zstride;
        buf_store(ys, 4u * (uint32_t)ob0, val[i], 4u * (uint32_t)so);
      }
    }
  };
  // ---- coarse (rebuilt level-1) approximation of anchor az, staged through registers
  float yv[NYT];
  auto fetch_coarse = [&](int az) {
    const int s = az & 1;
    const int t3 = ((az + s) >> 1) % rh;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(s ? Ylo1 : Ylo0, (uint32_t)(4 * (size_t)ph * qh * rh));
    const uint32_t so = 4u * (uint32_t)(ph * qh * t3);
#pragma unroll
    for (int i = 0; i < NYT; ++i) {
      const int e = tid + 256 * i;
      if (e < NYC) yv[i] = buf_load(rs, 4u * (uint32_t)(yrow[s][e % JX] + ycol[s][e / JX]), so);
    }
  };
  // LP: plane z into PB, quad g = column * NQ + quad at PB + 4 g (the lanes of
  // a direct-to-LDS load write 16 consecutive bytes each from the wave's base)
  auto fetch_pb = [&](int z) {
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(X + pq * z, pbytes);
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), ln = tid & 63;
#pragma unroll
    for (int it = 0; it < (NPC + 3) / 4; ++it) {
      const int c = 4 * it + wv;
      if (c >= NPC) break;
      const int g = 64 * c + ln;
      if (g < NPQ) {
        const uint32_t off = 4u * (uint32_t)(qrow[g % NQ] + cidxp[g / NQ]);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(PB + 256 * c), 16, off,
                                                 0, 0, 0);
      }
    }
  };
  // LP: the coarse values of anchor az (spin s = az & 1) into buffer s of YcL,
  // quad g = column * YQ + quad at 4 g
  auto fetch_yc = [&](int az) {
    const int s = az & 1;
    const int t3 = ((az + s) >> 1) % rh;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(s ? Ylo1 : Ylo0, (uint32_t)(4 * (size_t)ph * qh * rh));
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), g = 64 * wv + (tid & 63);
    if (g < C * YQ)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)(YcL + s * C * YS + 256 * wv), 16,
                                               4u * (uint32_t)(yqrow[g % YQ] + ycol[s][g / YQ]),
                                               4u * (uint32_t)(ph * qh * t3), 0, 0);
  };
  // LP: the thread's 23 window columns of the staged plane
  auto read_pb = [&](float (&dst)[NJ]) {
    const float* Pr = PB + 14 * ra * PS + wa + 3;
#pragma unroll
    for (int j = 0; j < NJ; ++j) dst[j] = Pr[j * PS];
  };
  float P0[NJ], P1[NJ];  // planes of even / odd index
#pragma unroll
  for (int j = 0; j < NJ; ++j) P0[j] = P1[j] = 0.f;
  fetch_plane(P1, r - 1);
  if constexpr (LP) {
    fetch_pb(0);
    fetch_yc(0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  } else {
    fetch_plane(P0, 0);
    fetch_coarse(0);
  }
  float first[8], pend[8];
  auto anchor = [&](auto sc, int az, float (&cur)[NJ], float (&prev)[NJ]) {
    constexpr int s = decltype(sc)::value;
    // phase A: z analysis of the pair (az-1, az), dim-2 analysis of both z bands
    if (actA) {
      if constexpr (LP) read_pb(cur);
      float zl[NJ - 1], zh[NJ - 1];
#pragma unroll
      for (int i = 0; i < NJ - 1; ++i) {
        zl[i] = z_sum(cur[i + s], prev[i + s]);
        zh[i] = z_dif(cur[i + s], prev[i + s]);
      }
      // in flight during the phases below: CS = 2, the next plane into the
      // registers of plane az-1; CS = 1, the next pair into both
      if constexpr (CS == 2) {
        if (!LP && az + 1 < r) fetch_plane(prev, az + 1);
      } else if (az + 2 < r) {
        fetch_plane(prev, az + 1);
        fetch_plane(cur, az + 2);
      }
      float lo[7], hi[7];
      float* Aw = A + 7 * ra * AS + wa;
      ana_lo_hi<7>(zl, lo, hi);
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        Aw[(0 * C + i) * AS] = lo[i];
        Aw[(1 * C + i) * AS] = hi[i];
      }
      ana_lo_hi<7>(zh, lo, hi);
#pragma unroll
      for (int i = 0; i < 7; ++i) {
        Aw[(2 * C + i) * AS] = lo[i];
        Aw[(3 * C + i) * AS] = hi[i];
      }
    }
    if constexpr (!LP) {
#pragma unroll
      for (int i = 0; i < NYT; ++i) {
        const int e = tid + 256 * i;
        if (e < NYC) Yc[e] = yv[i];
      }
      constexpr int step = CS == 2 ? 1 : 2;  // the next anchor: both spins alternate, or one spin's pairs
      if (az + step < r) fetch_coarse(az + step);
    }
    __syncthreads();
    // LP: every thread has read PB; the next plane lands there during phase B,
    // the next anchor's coarse values in the other buffer (last read in the
    // previous anchor's phase B)
    if constexpr (LP) {
      if (az + 1 < r) {
        fetch_pb(az + 1);
        fetch_yc(az + 1);
      }
    }
    // phase B: dim-1 analysis, threshold, dim-1 synthesis
    if (actB) {
      const float* Ar = A + kbc * AS + xb + s;
      float v[2 * NB + 8];
      float lo[NB], hi[NB];
      if (coarseB) {
#pragma unroll
        for (int j = 4; j < 2 * NB + 4; ++j) v[j] = Ar[j];
        ana_hi<NB>(v, hi);
        const float* Yr = LP ? YcL + (s * C + cB) * YS + s + (xb >> 1) : Yc + cB * JX + (xb >> 1);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          lo[i] = Yr[i];
          hi[i] = hard_thr(hi[i], th);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 2 * NB + 8; ++j) v[j] = Ar[j];
        ana_lo_hi<NB>(v, lo, hi);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          lo[i] = hard_thr(lo[i], tl);
          hi[i] = hard_thr(hi[i], th);
        }
      }
      float out[RB];
      if constexpr (s == 0) {
        syn_lo_hi<RB, 0>(lo, hi, out);
      } else {
        syn_lo_hi_odd<RB, 0>(lo, hi, out);
      }
      float* Er = E + kbc * ES + xb;
#pragma unroll
      for (int i = 0; i < RB; ++i)
        if (i >= 2 || xr != 2) Er[i] = out[i];
    }
    if constexpr (LP) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this thread's share of the next plane
    __syncthreads();
    // phase C: dim-2 synthesis, z synthesis, spin combine, store
    {
      constexpr int NCC = s == 0 ? 9 : 8;
      constexpr int GC = CS == 2 ? 2 : 1;  // taps x 1/2 (z), x 1/4 (z and the mean over both spins)
      float o[2][8];
#pragma unroll
      for (int zb = 0; zb < 2; ++zb) {
        const float* El = E + ((zb * 2 + 0) * C + 4 * yc) * ES + xo;
        const float* Eh = El + C * ES;
        float lo[NCC], hi[NCC];
#pragma unroll
        for (int j = 0; j < NCC; ++j) {
          lo[j] = El[j * ES];
          hi[j] = Eh[j * ES];
        }
        if constexpr (s == 0) {
          syn_lo_hi<8, GC>(lo, hi, o[zb]);
        } else {
          syn_lo_hi_odd<8, GC>(lo, hi, o[zb]);
        }
      }
      float ve[8], vo[8];  // plane az (even tap g[0]) and plane az-1 (odd tap g[1])
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        ve[i] = z_even(o[0][i], o[1][i]);
        vo[i] = z_odd(o[0][i], o[1][i]);
      }
      if constexpr (CS == 1) {
        store_vrec(az, ve);
        store_vrec(az == 0 ? r - 1 : az - 1, vo);
      } else {
        if (az == 0) {
#pragma unroll
          for (int i = 0; i < 8; ++i) first[i] = vo[i];  // plane r-1, completed by anchor r-1
        } else {
          store_plane(az - 1, pend, vo);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) pend[i] = ve[i];
        if (az == r - 1) store_plane(r - 1, first, pend);
      }
    }
    // No barrier here: the next anchor writes A (read in phase B, behind the
    // barrier above) and the coarse staging, and writes E only after its own
    // first barrier, which every thread reaches after this phase.
  };
  for (int az = 0; az < r; az += 2) {  // r is even: anchors in spin pairs, plane az in P0, az+1 in P1
    anchor(std::integral_constant<int, 0>{}, az, P0, P1);
    if constexpr (CS == 2) anchor(std::integral_constant<int, 1>{}, az + 1, P1, P0);
  }
  if constexpr (WG) {
    // sums over the group in the order of agg1_prep's wave sum over lanes
    // u + 8 v (xor 32, 16, 8 pair the columns v, xor 4, 2, 1 the rows u)
    auto cols = [](const float (&a)[8]) { return ((a[0] + a[4]) + (a[2] + a[6])) + ((a[1] + a[5]) + (a[3] + a[7])); };
    float s1 = cols(ws1), s2 = cols(ws2), s3 = cols(ws3);
#pragma unroll
    for (int o = 4; o >= 1; o >>= 1) {
      s1 += __uint_as_float(shfl_xor_u32(__float_as_uint(s1), o));
      s2 += __uint_as_float(shfl_xor_u32(__float_as_uint(s2), o));
      s3 += __uint_as_float(shfl_xor_u32(__float_as_uint(s3), o));
    }
    if ((xo & 7) == 0 && rowC && pc0 < q) {
      float Dv = s2 * 0.25f;
      Dv = Dv + s1 * 0.25f;
      Dv = Dv + s3;
      wgt[(x0 + xo) / 8 + (p / 8) * (pc0 / 8)] = 1.0f / (Dv + 2.220446049250313e-16f);  // 1/(D(blocks)+eps)
    }
  }
}

namespace fl {
constexpr int P = 32, R = 2 * P + 8, ZS = R + 1, OS = P + 1;
}

// nsp = 2: both cycle spins of a coarse level in one launch (see ht_syn_fast_kernel).
__global__ __launch_bounds__(256) void ht_lo_fast_kernel(const float* __restrict__ X, int p, int q, int r, int sh,
                                                         float* __restrict__ Y, int nsp, const float* __restrict__ X2,
                                                         float* __restrict__ Y2, const int* __restrict__ run) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  using namespace fl;
  __shared__ float Z[R * ZS];   // z-lowpass plane [col][row]; later the output tile [c][t]
  __shared__ float T2[P * ZS];  // dim-2 lowpass [c][row]
  __shared__ int ridx[R], cidx[R];
  const int tid = threadIdx.x;
  const int nbx = gridDim.x, nby = gridDim.y;
  const int lb = xcd_remap(blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z), nbx * nby * gridDim.z);
  const int nz = gridDim.z / nsp;
  const int t1b = (lb % nbx) * P, t2b = ((lb / nbx) % nby) * P, t3 = (lb / (nbx * nby)) % nz;
  if (lb / (nbx * nby) >= nz) {  // second spin
    X = X2;
    Y = Y2;
  }
  if (tid < R)
    ridx[tid] = pmod_i(2 * t1b - 9 - sh + tid, p);
  else if (tid < 2 * R)
    cidx[tid - R] = pmod_i(2 * t2b - 9 - sh + tid - R, q);
  const size_t pq = (size_t)p * q;
  const float* X0 = X + pq * pmod_i(2 * t3 - sh, r);
  const float* X1 = X + pq * pmod_i(2 * t3 - 1 - sh, r);
  __syncthreads();
  {
    constexpr int NIT = (R * R + 255) / 256;
    float a[NIT], b[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = tid + 256 * it;
      a[it] = b[it] = 0.f;
      if (e < R * R) {
        const size_t o = (size_t)ridx[e % R] + (size_t)p * cidx[e / R];
        a[it] = X0[o];
        b[it] = X1[o];
      }
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = tid + 256 * it;
      if (e < R * R) Z[(e / R) * ZS + e % R] = z_sum(a[it], b[it]);
    }
  }
  __syncthreads();
  for (int e = tid; e < 4 * R; e += 256) {
    const int rr = e % R, c0 = (e / R) * 8;
    float v[24], lo[8];
#pragma unroll
    for (int j = 0; j < 24; ++j) v[j] = Z[(2 * c0 + j) * ZS + rr];
    ana_lo<8>(v, lo);
#pragma unroll
    for (int i = 0; i < 8; ++i) T2[(c0 + i) * ZS + rr] = lo[i];
  }
  __syncthreads();
  float* Out = Z;
  if (tid < 4 * P) {
    const int c = tid % P, t0 = (tid / P) * 8;
    float v[24], lo[8];
#pragma unroll
    for (int j = 0; j < 24; ++j) v[j] = T2[c * ZS + 2 * t0 + j];
    ana_lo<8>(v, lo);
#pragma unroll
    for (int i = 0; i < 8; ++i) Out[c * OS + t0 + i] = lo[i];
  }
  __syncthreads();
  const int ph = p / 2, qh = q / 2;
  for (int e = tid; e < P * P; e += 256) {
    const int t = e % P, c = e / P;
    const int t1 = t1b + t, t2 = t2b + c;
    if (t1 < ph && t2 < qh) Y[t1 + (size_t)ph * (t2 + (size_t)qh * t3)] = Out[c * OS + t];
  }
}

// Level-0 approximation of both cycle spins from one read of the volume.
// Spin s's coefficient (t1, t2, t3) combines planes (2t3 - s, 2t3 - 1 - s) and
// the window rows starting at 2 t1 - 9 - s, so a 73 x 73 window over planes
// 2t3-2 .. 2t3 serves both: spin 1 uses planes (2t3-1, 2t3-2) and the
// sub-window at (0, 0), spin 0 planes (2t3, 2t3-1) and the sub-window at
// (1, 1).  Per spin the arithmetic is that of ht_lo_fast_kernel.  Workgroups
// take the z pairs of one tile consecutively, so the plane two neighbours share
// is read from the L2.
namespace fl2 {
constexpr int P = fl::P, R = fl::R, U = R + 1, ZS = U, OS = P + 1;
constexpr int NIT = (U * U + 255) / 256;
}  // namespace fl2

__global__ __launch_bounds__(256) void ht_lo0_cs2_kernel(const float* __restrict__ X, int p, int q, int r,
                                                         float* __restrict__ Y0, float* __restrict__ Y1, const int* __restrict__ run) {
  if (run && *run == 0) return;  // gated launch (the learned form's fallback)
  using namespace fl2;
  __shared__ float Z[U * ZS];    // z-lowpass plane [col][row] of the union window
  __shared__ float T2[P * ZS];   // dim-2 lowpass [c][row]
  __shared__ float Out[P * OS];  // output tile [c][t]
  __shared__ int ridx[U], cidxp[U];
  const int tid = threadIdx.x;
  const int ph = p >> 1, qh = q >> 1, rh = r >> 1;
  const int nbx = (ph + P - 1) / P;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int t3 = lb % rh, tile = lb / rh;
  const int t1b = (tile % nbx) * P, t2b = (tile / nbx) * P;
  if (tid < U)
    ridx[tid] = pmod_i(2 * t1b - 10 + tid, p);
  else if (tid < 2 * U)
    cidxp[tid - U] = p * pmod_i(2 * t2b - 10 + tid - U, q);
  __syncthreads();
  const size_t pq = (size_t)p * q;
  const float* Xa = X + pq * pmod_i(2 * t3 - 2, r);
  const float* Xb = X + pq * pmod_i(2 * t3 - 1, r);
  const float* Xc = X + pq * pmod_i(2 * t3, r);
  float pa[NIT], pb[NIT], pc[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int e = tid + 256 * it;
    pa[it] = pb[it] = pc[it] = 0.f;
    if (e < U * U) {
      const int o = ridx[e % U] + cidxp[e / U];
      pa[it] = Xa[o];
      pb[it] = Xb[o];
      pc[it] = Xc[o];
    }
  }
  // s = 1: planes (2t3-1, 2t3-2), offset 0; s = 0: planes (2t3, 2t3-1), offset 1
#pragma unroll 1
  for (int s = 1; s >= 0; --s) {
    const int off = 1 - s;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = tid + 256 * it;  // = column * U + row
      if (e < U * U) Z[e] = s ? z_sum(pb[it], pa[it]) : z_sum(pc[it], pb[it]);
    }
    __syncthreads();
    for (int e = tid; e < 4 * R; e += 256) {
      const int rr = e % R, c0 = (e / R) * 8;
      float v[24], lo[8];
#pragma unroll
      for (int j = 0; j < 24; ++j) v[j] = Z[(2 * c0 + j + off) * ZS + rr + off];
      ana_lo<8>(v, lo);
#pragma unroll
      for (int i = 0; i < 8; ++i) T2[(c0 + i) * ZS + rr] = lo[i];
    }
    __syncthreads();
    if (tid < 4 * P) {
      const int c = tid % P, t0 = (tid / P) * 8;
      float v[24], lo[8];
#pragma unroll
      for (int j = 0; j < 24; ++j) v[j] = T2[c * ZS + 2 * t0 + j];
      ana_lo<8>(v, lo);
#pragma unroll
      for (int i = 0; i < 8; ++i) Out[c * OS + t0 + i] = lo[i];
    }
    __syncthreads();
    float* Y = s ? Y1 : Y0;
    for (int e = tid; e < P * P; e += 256) {
      const int t = e % P, c = e / P;
      const int t1 = t1b + t, t2 = t2b + c;
      if (t1 < ph && t2 < qh) Y[t1 + (size_t)ph * (t2 + (size_t)qh * t3)] = Out[c * OS + t];
    }
    // the next spin rewrites Z only after every thread has passed the barrier
    // above, and T2 / Out only after the next two barriers
  }
}

// --------------------------------------------------------------- driver
// Force the generic (per-tap predicated) kernels everywhere; set through
// GBM3D_HT_GENERIC=1 for cross-checking the two implementations.
static const bool g_ht_generic = [] {
  const char* e = getenv("GBM3D_HT_GENERIC");
  return e && e[0] == '1';
}();

size_t ht_work_floats(int M, int N, int K, int levels, int cycleSpin) {
  size_t vol = (size_t)M * N * K, tot = cycleSpin > 1 ? vol : 0;
  for (int l = 1; l <= levels; ++l) tot += 2 * (vol >> (3 * l)) * cycleSpin;
  return tot;
}

void ht_work_bind(float* base, int M, int N, int K, int levels, int cycleSpin, HtWork* w) {
  size_t vol = (size_t)M * N * K;
  w->vrec = cycleSpin > 1 ? base : nullptr;
  float* p = base + (cycleSpin > 1 ? vol : 0);
  for (int l = 0; l < 4; ++l) {
    w->lo[l].assign(cycleSpin, nullptr);
    w->rec[l].assign(cycleSpin, nullptr);
  }
  for (int l = 0; l < levels; ++l) {
    size_t sz = vol >> (3 * (l + 1));
    for (int s = 0; s < cycleSpin; ++s) {
      w->lo[l][s] = p;
      p += sz;
      w->rec[l][s] = p;
      p += sz;
    }
  }
}

int run_ht(const float* vol, int M, int N, int K, int levels, int cycleSpin, const double* tau, const HtWork& w,
           float* vavg, int gd, hipStream_t st, StageHook hook, void* user, const int* run, float* wgt, bool* wgt_done) {
  if (wgt_done) *wgt_done = false;
  if (gd != 0 && (M % gd || N % gd)) return fail(GBM3D_ERR_ARG, "ht: group layout block size");
  auto mark = [&](int stage, int begin) {
    if (hook) hook(user, stage, begin);
  };
  if (levels < 1 || levels > 4) return fail(GBM3D_ERR_UNSUPPORTED, "levels must be 1..4");
  const int f = 1 << levels;
  if (M % f || N % f || K % f) return fail(GBM3D_ERR_UNSUPPORTED, "image size and numMax must be multiples of 2^levels");
  if (cycleSpin < 1) return fail(GBM3D_ERR_ARG, "cycleSpin must be >= 1");
  std::vector<Tau7> tl(levels);
  for (int l = 0; l < levels; ++l) {
    tl[l].t[0] = 0.f;
    // the kernels' level-l coefficients are (1/Cc)^(3(l+1)) = (2 sqrt 2)^(l+1)
    // times the reference's (analysis without the factor Cc per axis)
    const double scale = std::pow(kCc, -3.0 * (l + 1));
    // a NaN threshold (getColoredTau on a level coarser than its PSD) zeroes
    // the band: C .* (abs(C) > NaN) is 0 for every finite C and NaN for a NaN
    // C, which is what +inf gives with the kernels' compare-and-select
    for (int j = 1; j <= 7; ++j) {
      const double t = tau[l + levels * (j - 1)];
      tl[l].t[j] = std::isnan(t) ? INFINITY : (float)(t * scale);
    }
  }
  // Two cycle spins (every profile): the finest level of both spins in one
  // pass over the volume.  GBM3D_HT_UNFUSED=1 keeps the per-spin kernels, for
  // cross-checking (the two give bitwise equal results).
  const char* unf = getenv("GBM3D_HT_UNFUSED");
  // (the fused kernel addresses the volume through 32-bit byte offsets: < 2^30 elements)
  const bool fused = cycleSpin == 2 && M >= 10 && N >= 10 && (size_t)M * N * K < ((size_t)1 << 30) &&
                     (gd == 0 || gd == 4 || gd == 8 || gd == 16 || gd == 32) &&
                     !g_ht_generic && !(unf && unf[0] == '1');
  if (fused) {
    // level-0 approximation of both spins from one read; coarser levels of
    // both spins in one launch each (no shift below level 0)
    mark(GBM3D_STAGE_HT_LO, 1);
    const int nb = ((M / 2 + fl::P - 1) / fl::P) * ((N / 2 + fl::P - 1) / fl::P) * (K / 2);
    hipLaunchKernelGGL(ht_lo0_cs2_kernel, dim3(nb), dim3(256), 0, st, vol, M, N, K, w.lo[0][0], w.lo[0][1], run);
    GBM3D_TRY(check_launch("ht_lo0_cs2"));
    for (int l = 1; l < levels; ++l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      dim3 grid((p / 2 + LT - 1) / LT, (q / 2 + LT - 1) / LT, 2 * (r / 2));
      if (p >= 10 && q >= 10) {
        hipLaunchKernelGGL(ht_lo_fast_kernel, grid, dim3(256), 0, st, w.lo[l - 1][0], p, q, r, 0, w.lo[l][0], 2,
                           w.lo[l - 1][1], w.lo[l][1], run);
      } else {
        for (int s = 0; s < 2; ++s) {
          dim3 g1(grid.x, grid.y, r / 2);
          hipLaunchKernelGGL(ht_lo_kernel, g1, dim3(256), 0, st, w.lo[l - 1][s], p, q, r, 0, w.lo[l][s], run);
        }
      }
      GBM3D_TRY(check_launch("ht_lo"));
    }
    mark(GBM3D_STAGE_HT_LO, 0);
    for (int l = levels - 1; l >= 1; --l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      const float* ylo0 = (l == levels - 1) ? w.lo[l][0] : w.rec[l][0];
      const float* ylo1 = (l == levels - 1) ? w.lo[l][1] : w.rec[l][1];
      mark(GBM3D_STAGE_HT_SYN, 1);
      if (p >= 10 && q >= 10) {
        // both spins of the level in one launch (grid z)
        dim3 gf((p + f2::TX - 1) / f2::TX, (q + f2::TY - 1) / f2::TY, 2);
        hipLaunchKernelGGL(ht_fine_kernel<1>, gf, dim3(256), 0, st, w.lo[l - 1][0], w.lo[l - 1][1], ylo0, ylo1, p, q,
                           r, tl[l], w.rec[l - 1][0], w.rec[l - 1][1], 0, run, nullptr);
      } else {
        dim3 grid((p + SP - 1) / SP, (q + SP - 1) / SP, r / 2);
        for (int s = 0; s < 2; ++s)
          hipLaunchKernelGGL(ht_syn_kernel, grid, dim3(256), 0, st, w.lo[l - 1][s], s ? ylo1 : ylo0, p, q, r, 0,
                             tl[l], w.rec[l - 1][s], nullptr, kStore, 2.f, 0, run);
      }
      GBM3D_TRY(check_launch("ht_syn"));
      mark(GBM3D_STAGE_HT_SYN, 0);
    }
  }
  for (int s = 0; s < (fused ? 0 : cycleSpin); ++s) {  // per-spin chains (cycleSpin != 2, cross-checks)
    // analysis chain (approximations only)
    mark(GBM3D_STAGE_HT_LO, 1);
    for (int l = 0; l < levels; ++l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      const float* in = l == 0 ? vol : w.lo[l - 1][s];
      dim3 grid((p / 2 + LT - 1) / LT, (q / 2 + LT - 1) / LT, r / 2);
      if (p >= 10 && q >= 10 && !g_ht_generic)
        hipLaunchKernelGGL(ht_lo_fast_kernel, grid, dim3(256), 0, st, in, p, q, r, l == 0 ? s : 0, w.lo[l][s], 1,
                           nullptr, nullptr, run);
      else
        hipLaunchKernelGGL(ht_lo_kernel, grid, dim3(256), 0, st, in, p, q, r, l == 0 ? s : 0, w.lo[l][s], run);
      GBM3D_TRY(check_launch("ht_lo"));
    }
    mark(GBM3D_STAGE_HT_LO, 0);
    // synthesis chain, coarse to fine
    for (int l = levels - 1; l >= 0; --l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      const float* in = l == 0 ? vol : w.lo[l - 1][s];
      const float* ylo = (l == levels - 1) ? w.lo[l][s] : w.rec[l][s];
      // finest level: spins 0..cs-2 accumulate in w.vrec, the last one writes
      // Vavg = max(sum/cs, 0) (denoiseAndAggregateGPU.m:45)
      const bool last = s == cycleSpin - 1;
      float* out = l == 0 ? (last ? vavg : w.vrec) : w.rec[l - 1][s];
      const float* acc = (l == 0 && s > 0) ? w.vrec : nullptr;
      const int mode = l > 0 ? kStore : (last ? kFinal : (s == 0 ? kStore : kAccum));
      const float cs = (float)cycleSpin;
      const int lgd = (l == 0 && last) ? gd : 0;
      dim3 grid((p + SP - 1) / SP, (q + SP - 1) / SP, r / 2);
      const int stage = l == 0 ? GBM3D_STAGE_HT_SYN0 : GBM3D_STAGE_HT_SYN;
      mark(stage, 1);
      if (p >= 10 && q >= 10 && !g_ht_generic) {
        const int zpb = kSynPairsPerBlock;
        dim3 gf((p + fs::P - 1) / fs::P, (q + fs::P - 1) / fs::P, (r / 2 + zpb - 1) / zpb);
        hipLaunchKernelGGL(ht_syn_fast_kernel, gf, dim3(256), 0, st, in, ylo, p, q, r, l == 0 ? s : 0, tl[l], out,
                           acc, mode, cs, lgd, zpb, 1, nullptr, nullptr, nullptr, run);
      } else
        hipLaunchKernelGGL(ht_syn_kernel, grid, dim3(256), 0, st, in, ylo, p, q, r, l == 0 ? s : 0, tl[l], out, acc,
                           mode, cs, lgd, run);
      GBM3D_TRY(check_launch("ht_syn"));
      mark(stage, 0);
    }
  }
  if (fused) {
    const float* y0 = levels == 1 ? w.lo[0][0] : w.rec[0][0];
    const float* y1 = levels == 1 ? w.lo[0][1] : w.rec[0][1];
    dim3 grid((M + f2::TX - 1) / f2::TX, (N + f2::TY - 1) / f2::TY, 1);
    mark(GBM3D_STAGE_HT_SYN0, 1);
    if (wgt && gd == 8) {  // M % 8 == 0
      hipLaunchKernelGGL((ht_fine_kernel<2, true, true>), grid, dim3(256), 0, st, vol, nullptr, y0, y1, M, N, K, tl[0],
                         vavg, nullptr, gd, run, wgt);
      if (wgt_done) *wgt_done = true;
    } else if (M % 8 == 0) {
      hipLaunchKernelGGL((ht_fine_kernel<2, true>), grid, dim3(256), 0, st, vol, nullptr, y0, y1, M, N, K, tl[0], vavg,
                         nullptr, gd, run, nullptr);
    } else {
      hipLaunchKernelGGL((ht_fine_kernel<2, false>), grid, dim3(256), 0, st, vol, nullptr, y0, y1, M, N, K, tl[0], vavg,
                         nullptr, gd, run, nullptr);
    }
    GBM3D_TRY(check_launch("ht_fine (finest level, both spins)"));
    mark(GBM3D_STAGE_HT_SYN0, 0);
  }
  return GBM3D_OK;
}

}  // namespace gbm3d
