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
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {
constexpr float fCa = (float)kCa, fCb = (float)kCb, fCc = (float)kCc;

// analysis lo / hi and synthesis lo / hi (getWaveFilters3Dbior.m:21-23,49-50)
__device__ __forceinline__ float a_lo(int k) {
  const float t[10] = {fCa, -fCa, -fCb, fCb, fCc, fCc, fCb, -fCb, -fCa, fCa};
  return t[k];
}
__device__ __forceinline__ float a_hi(int k) { return k == 4 ? -fCc : (k == 5 ? fCc : 0.f); }
__device__ __forceinline__ float s_lo(int k) { return (k == 4 || k == 5) ? fCc : 0.f; }
__device__ __forceinline__ float s_hi(int k) {
  const float t[10] = {-fCa, -fCa, fCb, fCb, -fCc, fCc, -fCb, -fCb, fCa, fCa};
  return t[k];
}
}  // namespace

struct Tau7 {
  float t[8];  // t[j] for subband j = 1..7 (t[0] unused)
};

// subband bits b = dim1_hi | dim2_hi<<1 | z_hi<<2 -> column j of C{i,j}
__device__ __forceinline__ int band_j(int b) { return b == 3 ? 4 : (b == 4 ? 3 : b); }

// ---------------------------------------------------------------- lo kernel
constexpr int LT = 32;           // coefficients per tile edge
constexpr int LR = 2 * LT + 8;   // input rows/cols per tile
constexpr int LRP = LR + 1;

__global__ __launch_bounds__(256) void ht_lo_kernel(const float* __restrict__ X, int p, int q, int r, int sh,
                                                    float* __restrict__ Y) {
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
    Z[cc * LRP + rr] = fCc * X0[o] + fCc * X1[o];
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

__global__ __launch_bounds__(256) void ht_syn_kernel(const float* __restrict__ X, const float* __restrict__ Ylo,
                                                     int p, int q, int r, int sh, Tau7 tau,
                                                     float* __restrict__ Y) {
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
    Z[0][cc * SRP + rr] = fCc * a + fCc * b;
    Z[1][cc * SRP + rr] = -fCc * a + fCc * b;
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
          acc += s_lo(k) * B[b2 << 1][c * CPP + t] + s_hi(k) * B[1 | (b2 << 1)][c * CPP + t];
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
          acc += s_lo(k) * E[0][c * SPP + x] + s_hi(k) * E[1][c * SPP + x];
        }
      }
      out[zb][i] = acc;
    }
    __syncthreads();
  }
  // z synthesis: plane 2t3 <- g[0], plane 2t3-1 <- g[1]; lo g = [c c], hi g = [-c c]
  float* Ye = Y + pq * pmod_i(2 * t3 - sh, r);
  float* Yo = Y + pq * pmod_i(2 * t3 - 1 - sh, r);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int x = e % SP, y = e / SP;
    if (x0 + x >= p || y0 + y >= q) continue;
    const size_t o = (size_t)pmod_i(x0 + x - sh, p) + (size_t)p * pmod_i(y0 + y - sh, q);
    Ye[o] = fCc * out[0][i] - fCc * out[1][i];
    Yo[o] = fCc * out[0][i] + fCc * out[1][i];
  }
}

// --------------------------------------------------------------- driver
size_t ht_work_floats(int M, int N, int K, int levels, int cycleSpin) {
  size_t vol = (size_t)M * N * K, tot = vol * cycleSpin;
  for (int l = 1; l <= levels; ++l) tot += 2 * (vol >> (3 * l)) * cycleSpin;
  return tot;
}

void ht_work_bind(float* base, int M, int N, int K, int levels, int cycleSpin, HtWork* w) {
  size_t vol = (size_t)M * N * K;
  w->vrec = base;
  float* p = base + vol * cycleSpin;
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
           hipStream_t st) {
  if (levels < 1 || levels > 4) return fail(GBM3D_ERR_UNSUPPORTED, "levels must be 1..4");
  const int f = 1 << levels;
  if (M % f || N % f || K % f) return fail(GBM3D_ERR_UNSUPPORTED, "image size and numMax must be multiples of 2^levels");
  if (cycleSpin < 1) return fail(GBM3D_ERR_ARG, "cycleSpin must be >= 1");
  std::vector<Tau7> tl(levels);
  for (int l = 0; l < levels; ++l) {
    tl[l].t[0] = 0.f;
    for (int j = 1; j <= 7; ++j) tl[l].t[j] = (float)tau[l + levels * (j - 1)];
  }
  for (int s = 0; s < cycleSpin; ++s) {
    // analysis chain (approximations only)
    for (int l = 0; l < levels; ++l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      const float* in = l == 0 ? vol : w.lo[l - 1][s];
      dim3 grid((p / 2 + LT - 1) / LT, (q / 2 + LT - 1) / LT, r / 2);
      hipLaunchKernelGGL(ht_lo_kernel, grid, dim3(256), 0, st, in, p, q, r, l == 0 ? s : 0, w.lo[l][s]);
      GBM3D_TRY(check_launch("ht_lo"));
    }
    // synthesis chain, coarse to fine
    for (int l = levels - 1; l >= 0; --l) {
      const int p = M >> l, q = N >> l, r = K >> l;
      const float* in = l == 0 ? vol : w.lo[l - 1][s];
      const float* ylo = (l == levels - 1) ? w.lo[l][s] : w.rec[l][s];
      float* out = l == 0 ? w.vrec + (size_t)M * N * K * s : w.rec[l - 1][s];
      dim3 grid((p + SP - 1) / SP, (q + SP - 1) / SP, r / 2);
      hipLaunchKernelGGL(ht_syn_kernel, grid, dim3(256), 0, st, in, ylo, p, q, r, l == 0 ? s : 0, tl[l], out);
      GBM3D_TRY(check_launch("ht_syn"));
    }
  }
  return GBM3D_OK;
}

}  // namespace gbm3d
