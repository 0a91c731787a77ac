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


// 8-point DCT-II along one line of 8 values (stride ST in LDS), in place, from
// the table entries (uniform: scalar registers).  T(k, 7-n) = (-1)^k T(k, n),
// so with s_n = x_n + x_{7-n}, d_n = x_n - x_{7-n} (n < 4)
//   X_{2m} = sum_{n<4} T(2m, n) s_n,  X_{2m+1} = sum_{n<4} T(2m+1, n) d_n,
// and the inverse x_n = E_n + O_n, x_{7-n} = E_n - O_n with
//   E_n = sum_m T(2m, n) X_{2m},  O_n = sum_m T(2m+1, n) X_{2m+1}:
// 8 additions and 32 fused multiply-adds instead of 64.
template <bool INV, int ST>
__device__ __forceinline__ void dct8_line(float* p, const DctTable<8>& tab) {
  float x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = p[i * ST];
  if (!INV) {
    float sv[4], dv[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      sv[n] = x[n] + x[7 - n];
      dv[n] = x[n] - x[7 - n];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      float e = 0.f, o = 0.f;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        e = fmaf(tab.t[2 * m + 8 * n], sv[n], e);
        o = fmaf(tab.t[2 * m + 1 + 8 * n], dv[n], o);
      }
      p[(2 * m) * ST] = e;
      p[(2 * m + 1) * ST] = o;
    }
  } else {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      float e = 0.f, o = 0.f;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        e = fmaf(tab.t[2 * m + 8 * n], x[2 * m], e);
        o = fmaf(tab.t[2 * m + 1 + 8 * n], x[2 * m + 1], o);
      }
      p[n * ST] = e + o;
      p[(7 - n) * ST] = e - o;
    }
  }
}

// page_transform for D = 8 with the lines above: the 64 threads of a group
// take the 8 lines x K slices of each pass two or more at a time.  Slices are
// staged as sm[k][v][u] at k*72 + v*9 + u (kDct8Slice floats per slice): the
// line bases of one instruction, 9v + 8k (pass 1) and 8k + u (pass 2), are
// distinct modulo the 64 LDS banks.
constexpr int kDct8Row = 9, kDct8Slice = 72;

template <int K, bool INV>
__device__ __forceinline__ void page_transform_dct8(float (&x)[K], float* sm, const DctTable<8>& tab, int u, int v) {
  constexpr int TPG = 64;
  const int t = u + 8 * v;
#pragma unroll
  for (int k = 0; k < K; ++k) sm[k * kDct8Slice + v * kDct8Row + u] = x[k];
  group_sync<TPG>();
  // pass 1: along u for each (v, k)
#pragma unroll
  for (int r = 0; r < (8 * K + TPG - 1) / TPG; ++r) {
    const int task = t + TPG * r;
    if (task < 8 * K) dct8_line<INV, 1>(sm + (task >> 3) * kDct8Slice + (task & 7) * kDct8Row, tab);
  }
  group_sync<TPG>();
  // pass 2: along v for each (u, k)
#pragma unroll
  for (int r = 0; r < (8 * K + TPG - 1) / TPG; ++r) {
    const int task = t + TPG * r;
    if (task < 8 * K) dct8_line<INV, kDct8Row>(sm + (task >> 3) * kDct8Slice + (task & 7), tab);
  }
  group_sync<TPG>();
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = sm[k * kDct8Slice + v * kDct8Row + u];
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
