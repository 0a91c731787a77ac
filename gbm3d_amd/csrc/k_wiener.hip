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
#include "group_xform.hpp"
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

// Noise variance per DCT coefficient (u, v): sigma^2 everywhere for white
// noise, sigmaWie(u, v)^2 for coloured noise (the implicit expansion of the
// d x d sigma in denoiseAndAggregateWieFast.m:49).
template <int D>
struct Sig2Tab {
  float s[D * D];
};

template <int D, int K, bool GATHER>
__global__ __launch_bounds__(256) void wiener_kernel(const float* __restrict__ padP, const float* __restrict__ padN,
                                                     const uint32_t* __restrict__ S, const float* __restrict__ V1,
                                                     const float* __restrict__ V2, int M, int N, int d2, long G,
                                                     Sig2Tab<D> sig2, DctTable<D> tab, float hi,
                                                     float* __restrict__ V2f, float* __restrict__ gain2) {
  constexpr int TPG = D * D, GPB = 256 / TPG;
  constexpr int L = K < 16 ? 2 : 3;
  __shared__ float smem[GPB][D == 8 ? K * kDct8Slice : K * TPG];
  __shared__ float T[D * D];
  __shared__ float red[GPB][TPG > 64 ? TPG / 64 : 1];
  for (int e = threadIdx.x; e < D * D; e += blockDim.x) T[e] = tab.t[e];
  __syncthreads();
  const int lg = threadIdx.x / TPG, t = threadIdx.x % TPG;
  const int u = t % D, v = t / D;
  // one group per wave (D = 8): g is wave-uniform, so its slot indices are scalar loads
  long g = (long)blockIdx.x * GPB + lg;
  if constexpr (TPG == 64) g = (long)__builtin_amdgcn_readfirstlane((int)g);
  const bool active = g < G;
  if (!active) g = G - 1;
  float* sm = smem[lg];
  float c1[K], c2[K];
  if constexpr (GATHER) {
    const int Mt = M / D;
    const int ti = (int)(g % Mt), tj = (int)(g / Mt);
    const size_t P = (size_t)M + d2;
    if (P * ((size_t)N + d2) < ((size_t)1 << 30)) {
      // buffer loads: this lane's pixel offset in a vector register, the
      // slice's block origin (uniform: one group per wave at D = 8) in a
      // scalar register
      const uint32_t pbytes = (uint32_t)(4 * P * ((size_t)N + d2));
      const __amdgpu_buffer_rsrc_t rP = buf_rsrc(padP, pbytes), rN = buf_rsrc(padN, pbytes);
      const uint32_t lo = 4u * (uint32_t)(u + (int)P * v);
      if constexpr (TPG == 64) {
        // lane k < K: slot k's block origin (its window index divided by d2
        // once, all slots in parallel), then each slot's offset from lane k
        uint32_t sol = 0;
        if (t < K) {
          const uint32_t idx = S[(size_t)K * g + t] - 1u;
          const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
          sol = 4u * (uint32_t)((ti * D + a) + (int)P * (tj * D + b));
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const uint32_t so = (uint32_t)__builtin_amdgcn_readlane((int)sol, k);
          c1[k] = buf_load(rP, lo, so);
          c2[k] = buf_load(rN, lo, so);
        }
      } else {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const uint32_t idx = S[(size_t)K * g + k] - 1u;
          const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
          const uint32_t so = 4u * (uint32_t)((ti * D + a) + (int)P * (tj * D + b));
          c1[k] = buf_load(rP, lo, so);
          c2[k] = buf_load(rN, lo, so);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint32_t idx = S[(size_t)K * g + k] - 1u;
        const int a = (int)(idx % (uint32_t)d2), b = (int)(idx / (uint32_t)d2);
        const size_t o = (size_t)(ti * D + a + u) + P * (size_t)(tj * D + b + v);
        c1[k] = padP[o];
        c2[k] = padN[o];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const size_t o = (size_t)t + (size_t)TPG * (k + (size_t)K * g);
      c1[k] = V1[o];
      c2[k] = V2[o];
    }
  }
  if constexpr (D == 8) {
    page_transform_dct8<K, false>(c1, sm, tab, u, v);
    haar_dec<K, L>(c1);
    page_transform_dct8<K, false>(c2, sm, tab, u, v);
  } else {
    page_transform<D, K, false, true>(c1, sm, T, u, v);
    haar_dec<K, L>(c1);
    page_transform<D, K, false, true>(c2, sm, T, u, v);
  }
  haar_dec<K, L>(c2);
  float n2 = 0.f;
  const float sigma2 = sig2.s[t];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float a2 = c1[k] * c1[k];
    // abs(V1).^2./(abs(V1).^2 + sigma.^2), with the hardware reciprocal (1 ulp)
    float f = a2 * __builtin_amdgcn_rcpf(a2 + sigma2);
    if (t == 0 && k >= 14) f = 1.f;            // filt(1,1,15:end) = 1
    c2[k] *= f;
    n2 += f * f;
  }
  haar_rec<K, L>(c2);
  if constexpr (D == 8)
    page_transform_dct8<K, true>(c2, sm, tab, u, v);
  else
    page_transform<D, K, true, true>(c2, sm, T, u, v);
  if (active) {
    float* dst = V2f + (size_t)TPG * K * g;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float y = c2[k];
      // max(V2, 0) (WieFast :64), or min(max(V2, 0), 1) (the dct() twin :55; hi = 1)
      dst[t + TPG * k] = fminf(y > 0.f ? y : 0.f, hi);
    }
  }
  // |filt|_F^2 of the group
  if constexpr (TPG == 64) {
    n2 = wave_sum_f32(n2);
    if (t == 0 && active) gain2[g] = n2;
  } else if constexpr (TPG < 64) {
#pragma unroll
    for (int o = TPG / 2; o >= 1; o >>= 1) n2 += __uint_as_float(shfl_xor_u32(__float_as_uint(n2), o));
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
                           int M, int N, int d2, long G, float sigma, const float* sig2tab, float hi, float* V2f,
                           float* gain2, hipStream_t st) {
  const DctTable<D> tab = make_dct_table<D>();
  constexpr int GPB = 256 / (D * D);
  dim3 grid((unsigned)((G + GPB - 1) / GPB));
  Sig2Tab<D> s2;
  for (int e = 0; e < D * D; ++e) s2.s[e] = sig2tab ? sig2tab[e] : sigma * sigma;
  if (padP)
    hipLaunchKernelGGL((wiener_kernel<D, K, true>), grid, dim3(256), 0, st, padP, padN, S, V1, V2, M, N, d2, G, s2,
                       tab, hi, V2f, gain2);
  else
    hipLaunchKernelGGL((wiener_kernel<D, K, false>), grid, dim3(256), 0, st, padP, padN, S, V1, V2, M, N, d2, G, s2,
                       tab, hi, V2f, gain2);
  return check_launch("wiener");
}

int launch_wiener(const float* pad_pilot, const float* pad_noisy, const uint32_t* S, const float* V1, const float* V2,
                  int M, int N, int d, int d2, int K, long G, float sigma, float* V2f, float* gain2, hipStream_t st,
                  const float* sig2tab, bool clip01) {
  const float hi = clip01 ? 1.f : __builtin_huge_valf();
  if (G <= 0) return GBM3D_OK;
#define GBM3D_WIE_CASE(DD, KK)                                                                                  \
  if (d == DD && K == KK)                                                                                       \
    return launch_wiener_t<DD, KK>(pad_pilot, pad_noisy, S, V1, V2, M, N, d2, G, sigma, sig2tab, hi, V2f, gain2, st);
  GBM3D_WIE_CASE(8, 16)
  GBM3D_WIE_CASE(8, 8)
  GBM3D_WIE_CASE(8, 32)
  GBM3D_WIE_CASE(4, 8)
  GBM3D_WIE_CASE(4, 16)
  GBM3D_WIE_CASE(4, 32)
  GBM3D_WIE_CASE(16, 8)
  GBM3D_WIE_CASE(16, 16)
  GBM3D_WIE_CASE(16, 32)
#undef GBM3D_WIE_CASE
  return fail(GBM3D_ERR_UNSUPPORTED, "wiener: (blockSizeWie, numMax) must be (4|8|16, 8|16|32)");
}

bool wiener_supported(int d, int K) { return (d == 4 || d == 8 || d == 16) && (K == 8 || K == 16 || K == 32); }

}  // namespace gbm3d
