// group_xform.hpp -- per-group transforms shared by the step-2 Wiener filter
// and the learned shrinkage: orthonormal 2-D DCT-II of every d x d slice
// (T = dct(eye(d)), source/denoiseAndAggregateWieFast.m:33) and the 1-D Haar
// transform along the K slices (waveTransforms3D/myWavDec1D.m:10-26,
// myWavRec1D.m:5-26; pairs (2t-1, 2t) circularly, coefficient layout
// [details L1 | L2 | ... | approximation]).
#pragma once

#include <cmath>

#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

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
// Pass 1 combines along u (column of the slice), pass 2 along v.  FMA: each
// dot product as one fused multiply-add chain (else multiply, then add).
template <int D, int K, bool INV, bool FMA = false>
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
    for (int j = 0; j < D; ++j) acc = FMA ? fmaf(Lrow[j], sm[k * TPG + v * D + j], acc) : acc + Lrow[j] * sm[k * TPG + v * D + j];
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
    for (int j = 0; j < D; ++j) acc = FMA ? fmaf(sm[k * TPG + u * D + j], Rrow[j], acc) : acc + sm[k * TPG + u * D + j] * Rrow[j];
    x[k] = acc;
  }
  group_sync<TPG>();
}


// T(k, n) = c_k cos(pi (2n + 1) k / (2D)), c_0 = sqrt(1/D), c_k = sqrt(2/D).
template <int D>
inline DctTable<D> make_dct_table() {
  DctTable<D> tab;
  for (int k = 0; k < D; ++k)
    for (int n = 0; n < D; ++n) {
      const double ck = k == 0 ? std::sqrt(1.0 / D) : std::sqrt(2.0 / D);
      tab.t[k + D * n] = (float)(ck * std::cos(M_PI * (2 * n + 1) * k / (2.0 * D)));
    }
  return tab;
}

}  // namespace gbm3d
