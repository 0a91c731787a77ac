// k_wiener.hip -- step-2 empirical Wiener filter on matched groups.
//
// Reference: source/denoiseAndAggregateWieFast.m:19-66.  Per group (d x d x K
// stack): orthonormal 2-D DCT-II of every slice (T*B*T', T = dct(eye(d))),
// 1-D Haar along the K slices (waveTransforms3D/myWavDec1D.m: pairs (2t-1, 2t),
// layout [details L1 | L2 | ... | approximation]), gains g = |V1|^2/(|V1|^2 +
// sigma^2) from the pilot group with g(1,1,15:end) = 1, applied to the noisy
// group, inverse Haar (myWavRec1D.m), inverse DCT (T'*X*T), clip at 0.
//
// One thread per (u, v) pixel position holds the K values of its column in
// registers, so the Haar transform along K is register-local; the two DCT
// passes of every slice go through LDS.  The groups are gathered straight from
// the padded pilot / noisy search images through S (no V1/V2 arrays), or read
// from V1/V2 for the stage entry point.
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {

template <int D>
struct DctTable {
  float t[D * D];  // t[k + D*n] = c_k cos(pi (2n+1) k / (2D))
};

// Haar analysis along the slices, in place (K compile-time, L levels).
template <int K, int L>
__device__ __forceinline__ void haar_dec(float (&x)[K]) {
  constexpr float c = (float)kCc;
  float out[K];
  float cur[K];
#pragma unroll
  for (int i = 0; i < K; ++i) cur[i] = x[i];
  int len = K, cnt = 0;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    float lo[K];
#pragma unroll
    for (int t = 0; t < K / 2; ++t) {
      if (t < len / 2) {
        const float a = cur[2 * t];
        const float b = cur[(2 * t - 1 + len) % len];
        out[cnt + t] = -c * a + c * b;
        lo[t] = c * a + c * b;
      }
    }
#pragma unroll
    for (int t = 0; t < K / 2; ++t)
      if (t < len / 2) cur[t] = lo[t];
    cnt += len / 2;
    len /= 2;
  }
#pragma unroll
  for (int t = 0; t < K; ++t)
    if (t < len) out[cnt + t] = cur[t];
#pragma unroll
  for (int i = 0; i < K; ++i) x[i] = out[i];
}

// Haar synthesis: y[2t] = g0 c[t], y[2t-1] = g1 c[t] for the detail (g = [-c c])
// and approximation (g = [c c]) paths of each level (myWavRec1D.m:10-26).
template <int K, int L>
__device__ __forceinline__ void haar_rec(float (&x)[K]) {
  constexpr float c = (float)kCc;
  float cur[K];
  int len = K >> L;
  int cnt = K - len;
#pragma unroll
  for (int t = 0; t < K; ++t)
    if (t < len) cur[t] = x[cnt + t];
#pragma unroll
  for (int l = L - 1; l >= 0; --l) {
    const int out_len = K >> l;
    cnt -= out_len / 2;
    float nxt[K];
#pragma unroll
    for (int t = 0; t < K; ++t) nxt[t] = 0.f;
#pragma unroll
    for (int t = 0; t < K / 2; ++t) {
      if (t < out_len / 2) {
        const float dt = x[cnt + t];
        nxt[2 * t] += -c * dt;
        nxt[(2 * t - 1 + out_len) % out_len] += c * dt;
      }
    }
#pragma unroll
    for (int t = 0; t < K / 2; ++t) {
      if (t < out_len / 2) {
        nxt[2 * t] += c * cur[t];
        nxt[(2 * t - 1 + out_len) % out_len] += c * cur[t];
      }
    }
#pragma unroll
    for (int t = 0; t < K; ++t) cur[t] = nxt[t];
  }
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = cur[t];
}

template <int TPG>
__device__ __forceinline__ void group_sync() {
  if constexpr (TPG <= 64)
    wave_sync();
  else
    __syncthreads();
}

// 2-D transform of all K slices held one value per thread:
// forward  X <- T X T'   (thread t = u + D v holds X(u, v));
// inverse  X <- T' X T.
// Pass 1 combines along u (column of the slice), pass 2 along v.
template <int D, int K, bool INV>
__device__ __forceinline__ void page_transform(float (&x)[K], float* sm, const float* __restrict__ T, int u, int v) {
  constexpr int TPG = D * D;
  // stage slices as sm[k][v][u]
#pragma unroll
  for (int k = 0; k < K; ++k) sm[k * TPG + v * D + u] = x[k];
  group_sync<TPG>();
  // thread (p = u, v): tmp(p, v) = sum_u' L(p, u') X(u', v)
  float Lrow[D];
#pragma unroll
  for (int j = 0; j < D; ++j) Lrow[j] = INV ? T[j + D * u] : T[u + D * j];
  float tmp[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < D; ++j) acc += Lrow[j] * sm[k * TPG + v * D + j];
    tmp[k] = acc;
  }
  group_sync<TPG>();
  // store tmp as sm[k][p][v] so pass 2 reads a row
#pragma unroll
  for (int k = 0; k < K; ++k) sm[k * TPG + u * D + v] = tmp[k];
  group_sync<TPG>();
  // thread (p = u, q = v): out(p, q) = sum_v' tmp(p, v') R(q, v')
  float Rrow[D];
#pragma unroll
  for (int j = 0; j < D; ++j) Rrow[j] = INV ? T[j + D * v] : T[v + D * j];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int j = 0; j < D; ++j) acc += sm[k * TPG + u * D + j] * Rrow[j];
    x[k] = acc;
  }
  group_sync<TPG>();
}

}  // namespace

template <int D, int K, bool GATHER>
__global__ __launch_bounds__(256) void wiener_kernel(const float* __restrict__ padP, const float* __restrict__ padN,
                                                     const uint32_t* __restrict__ S, const float* __restrict__ V1,
                                                     const float* __restrict__ V2, int M, int N, int d2, long G,
                                                     float sigma2, DctTable<D> tab, float* __restrict__ V2f,
                                                     float* __restrict__ gain2) {
  constexpr int TPG = D * D, GPB = 256 / TPG;
  constexpr int L = K < 16 ? 2 : 3;
  __shared__ float smem[GPB][K * TPG];
  __shared__ float T[D * D];
  __shared__ float red[GPB][TPG > 64 ? TPG / 64 : 1];
  for (int e = threadIdx.x; e < D * D; e += blockDim.x) T[e] = tab.t[e];
  __syncthreads();
  const int lg = threadIdx.x / TPG, t = threadIdx.x % TPG;
  const int u = t % D, v = t / D;
  long g = (long)blockIdx.x * GPB + lg;
  const bool active = g < G;
  if (!active) g = G - 1;
  float* sm = smem[lg];
  float c1[K], c2[K];
  if constexpr (GATHER) {
    const int Mt = M / D;
    const int ti = (int)(g % Mt), tj = (int)(g / Mt);
    const size_t P = (size_t)M + d2;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t idx = S[(size_t)K * g + k] - 1u;
      const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
      const size_t o = (size_t)(ti * D + a + u) + P * (size_t)(tj * D + b + v);
      c1[k] = padP[o];
      c2[k] = padN[o];
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const size_t o = (size_t)t + (size_t)TPG * (k + (size_t)K * g);
      c1[k] = V1[o];
      c2[k] = V2[o];
    }
  }
  page_transform<D, K, false>(c1, sm, T, u, v);
  haar_dec<K, L>(c1);
  page_transform<D, K, false>(c2, sm, T, u, v);
  haar_dec<K, L>(c2);
  float n2 = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float a2 = c1[k] * c1[k];
    float f = a2 / (a2 + sigma2);              // abs(V1).^2./(abs(V1).^2 + sigma.^2)
    if (t == 0 && k >= 14) f = 1.f;            // filt(1,1,15:end) = 1
    c2[k] *= f;
    n2 += f * f;
  }
  haar_rec<K, L>(c2);
  page_transform<D, K, true>(c2, sm, T, u, v);
  if (active) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float y = c2[k];
      V2f[(size_t)t + (size_t)TPG * (k + (size_t)K * g)] = y > 0.f ? y : 0.f;  // max(V2, 0)
    }
  }
  // |filt|_F^2 of the group
  if constexpr (TPG == 64) {
    n2 = wave_sum_f32(n2);
    if (t == 0 && active) gain2[g] = n2;
  } else if constexpr (TPG < 64) {
#pragma unroll
    for (int o = TPG / 2; o >= 1; o >>= 1) n2 += __shfl_xor(n2, o, 64);
    if (t == 0 && active) gain2[g] = n2;
  } else {
    n2 = wave_sum_f32(n2);
    if ((threadIdx.x & 63) == 0) red[lg][(t >> 6)] = n2;
    __syncthreads();
    if (t == 0 && active) {
      float s = 0.f;
      for (int w = 0; w < TPG / 64; ++w) s += red[lg][w];
      gain2[g] = s;
    }
  }
}

template <int D, int K>
static int launch_wiener_t(const float* padP, const float* padN, const uint32_t* S, const float* V1, const float* V2,
                           int M, int N, int d2, long G, float sigma, float* V2f, float* gain2, hipStream_t st) {
  DctTable<D> tab;
  for (int k = 0; k < D; ++k)
    for (int n = 0; n < D; ++n) {
      const double ck = k == 0 ? std::sqrt(1.0 / D) : std::sqrt(2.0 / D);
      tab.t[k + D * n] = (float)(ck * std::cos(M_PI * (2 * n + 1) * k / (2.0 * D)));
    }
  constexpr int GPB = 256 / (D * D);
  dim3 grid((unsigned)((G + GPB - 1) / GPB));
  const float s2 = sigma * sigma;
  if (padP)
    hipLaunchKernelGGL((wiener_kernel<D, K, true>), grid, dim3(256), 0, st, padP, padN, S, V1, V2, M, N, d2, G, s2,
                       tab, V2f, gain2);
  else
    hipLaunchKernelGGL((wiener_kernel<D, K, false>), grid, dim3(256), 0, st, padP, padN, S, V1, V2, M, N, d2, G, s2,
                       tab, V2f, gain2);
  return check_launch("wiener");
}

int launch_wiener(const float* pad_pilot, const float* pad_noisy, const uint32_t* S, const float* V1, const float* V2,
                  int M, int N, int d, int d2, int K, long G, float sigma, float* V2f, float* gain2, hipStream_t st) {
  if (G <= 0) return GBM3D_OK;
#define GBM3D_WIE_CASE(DD, KK)                                                                          \
  if (d == DD && K == KK)                                                                               \
    return launch_wiener_t<DD, KK>(pad_pilot, pad_noisy, S, V1, V2, M, N, d2, G, sigma, V2f, gain2, st);
  GBM3D_WIE_CASE(8, 16)
  GBM3D_WIE_CASE(8, 8)
  GBM3D_WIE_CASE(8, 32)
  GBM3D_WIE_CASE(4, 16)
  GBM3D_WIE_CASE(16, 16)
#undef GBM3D_WIE_CASE
  return fail(GBM3D_ERR_UNSUPPORTED, "wiener: (blockSizeWie, numMax) must be (8, 8|16|32), (4, 16) or (16, 16)");
}

}  // namespace gbm3d
