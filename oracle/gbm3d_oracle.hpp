// gbm3d_oracle.hpp -- CPU restatement of the GBM3D two-step hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library in
// gbm3d_amd/ is compared against; it is never linked into, loaded by or called
// from the product path.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it.
//
// Instantiated twice:
//   T = double : the FP64 oracle (every stage in double precision);
//   T = float  : the FP32 "CPU restatement" baseline (OpenMP), which follows
//                the reference's single-precision rules where they are known.
//
// All arrays are column-major (MATLAB layout).  0-based indices throughout;
// the reference's 1-based window indices appear only in S.
//
// Pinning (what checks each function independently of this code):
//   filters       : closed-form constants, archive/waveletFilters/bior15Filters.mat
//                   (tests/golden/bior15_taps.json), perfect reconstruction.
//   transform     : numpy FFT restatement of myWavDec3DFFT / myWavRec3DGPUfast
//                   (oracle/literal.py) on small volumes; PR; tau = inf.
//   matching      : numpy fft2 restatement of matchBlocksCPU / matchBlocksWie
//                   lines 57-88 plus brute force; stable argsort.
//   aggregation   : literal scatter loop (this file) == gather (this file),
//                   numpy literal loop (oracle/literal.py).
//   wiener        : scipy orthonormal DCT-II, numpy FFT Haar, PR of the 1-D Haar.
//   end to end    : numpy literal GBM3D on 64x64 inputs; scale equivariance.
//   whole-image   : sym4 taps == archive/waveletFilters/sym4Filters.mat
//   matching        (tests/golden/sym4_taps.json) and their orthonormality;
//                   prefilter PR without threshold; numpy fftn restatement of
//                   myWavDec2DFFT / myWavRec2DGPUfast / matchBlocksGPU (all three
//                   list branches); brute-force |x - y|^2 / |x|^2 ranking.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/gbm3d.h"

namespace gbm3d_oracle {

// ------------------------------------------------------------------ options
// utilities/setBM3Dopts.m:12-86
inline void set_opts(const char* profile, gbm3d_opts* o) {
  std::string p = profile ? profile : "default";
  if (p != "default" && p != "fast" && p != "superFast" && p != "accuracy")
    p = "default";  // :12-14
  std::memset(o, 0, sizeof(*o));
  o->numMax = 16;
  o->numMin = 16;
  o->levels = 3;
  o->cycleSpin = 2;
  o->tauMode = 1;
  o->blockSize = 16;
  o->blockSizeWie = 8;
  o->matchSize = 32;
  o->Wiener = 1;
  o->filtType = GBM3D_FILT_HT;
  auto spins = [](int32_t* dst, int32_t* n, std::initializer_list<int> v) {
    *n = 0;
    for (int x : v) dst[(*n)++] = x;
  };
  if (p == "default") {  // :18-34
    spins(o->matchSpins, &o->nMatchSpins, {0, 4});
    spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0, 4});
  } else if (p == "fast") {  // :35-51
    spins(o->matchSpins, &o->nMatchSpins, {0, 4, 6});
    spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0});
  } else if (p == "superFast") {  // :52-68
    spins(o->matchSpins, &o->nMatchSpins, {0, 4, 6});
    spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0});
    o->blockSize = 32;
    o->Wiener = 0;
  } else {  // accuracy :69-85
    spins(o->matchSpins, &o->nMatchSpins, {0, 4, 6});
    spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0, 4, 6});
  }
}

// utilities/checkBM3Dopts.m:10-24 (no default for matchSpinsWie)
inline void check_opts(gbm3d_opts* o) {
  auto def = [](int32_t& f, int v) {
    if (f == GBM3D_UNSET) f = v;
  };
  def(o->numMax, 16);
  def(o->numMin, 16);
  def(o->levels, 3);
  def(o->tauMode, 1);
  def(o->filtType, GBM3D_FILT_HT);
  def(o->blockSize, 16);
  def(o->blockSizeWie, 8);
  def(o->matchSize, std::max(32, (int)o->blockSize));
  def(o->cycleSpin, 2);
  if (o->nMatchSpins == GBM3D_UNSET) {
    o->nMatchSpins = 2;
    o->matchSpins[0] = 0;
    o->matchSpins[1] = std::min(o->blockSize, o->blockSizeWie) / 2;
  }
  def(o->Wiener, 1);
}

// source/GBM3D.m:15-19
inline void resolve_opts(const gbm3d_opts* in, gbm3d_opts* out) {
  if (!in) {
    set_opts("default", out);
  } else if (in->profile[0]) {
    set_opts(in->profile, out);
  } else {
    *out = *in;
  }
  check_opts(out);
}

// utilities/getTau.m:9-51; tau[i + levels*j]
inline void get_tau(int levels, int mode, double sigma, double* tau) {
  for (int j = 0; j < 7; ++j)
    for (int i = 0; i < levels; ++i) {
      double t = 0;
      double base = 3.3 - 0.3 * i;
      switch (mode) {
        case 1: t = base; break;
        case 2: t = (j < 3) ? base : (j < 6 ? base + 0.5 : base + 1.0); break;
        case 3: t = (j < 3) ? base : (j < 6 ? base + 0.25 : base + 1.5); break;
        case 4: t = (j < 3) ? 2.9 : (j < 6 ? 2.9 + 0.5 : 2.9 + 1.0); break;
        default: t = 0; break;
      }
      tau[i + levels * j] = sigma * t;
    }
}

// ------------------------------------------------------------ small helpers
inline int pmod(long long a, long long m) {
  long long r = a % m;
  return (int)(r < 0 ? r + m : r);
}

// utilities/myMirrorPadImage.m:3-8: rows m..M-1 repeat the edge
// (padded row m+j is source row m-1-j).
template <class T>
std::vector<T> mirror_pad(const T* I, int m, int n, int M, int N) {
  std::vector<T> out((size_t)M * N);
  for (int c = 0; c < N; ++c) {
    int sc = c < n ? c : 2 * n - 1 - c;
    for (int r = 0; r < M; ++r) {
      int sr = r < m ? r : 2 * m - 1 - r;
      out[r + (size_t)M * c] = I[sr + (size_t)m * sc];
    }
  }
  return out;
}

// circshift(X, [s s]) of an m x n image: out(r,c) = X(r-s, c-s) circularly.
template <class T>
std::vector<T> circshift2(const T* X, int m, int n, int s) {
  std::vector<T> out((size_t)m * n);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r)
      out[r + (size_t)m * c] = X[pmod(r - s, m) + (size_t)m * pmod(c - s, n)];
  return out;
}

// The padded search image of source/matchBlocksCPU.m:57-61 (and
// source/matchBlocksWie.m:42-52): 2x2 tiling of X truncated to
// (m+d2) x (n+d2), circshifted by d2/2 inside that size.  Row r of the
// result is X row (r < s ? r+s : (r-s) mod m), s = d2/2 (a shifted copy at
// the top, the periodic wrap at the bottom).  Needs m, n >= d2.
inline int padmap(int r, int m, int s) { return r < s ? r + s : (r - s) % m; }

template <class T>
std::vector<T> search_pad(const T* X, int m, int n, int d2) {
  int s = d2 / 2, P = m + d2, Q = n + d2;
  std::vector<T> out((size_t)P * Q);
  for (int c = 0; c < Q; ++c) {
    int sc = padmap(c, n, s);
    for (int r = 0; r < P; ++r) out[r + (size_t)P * c] = X[padmap(r, m, s) + (size_t)m * sc];
  }
  return out;
}

template <class T>
inline T madd(T a, T b, T c);
template <>
inline float madd<float>(float a, float b, float c) { return std::fmaf(a, b, c); }
template <>
inline double madd<double>(double a, double b, double c) { return a * b + c; }

// Orderable key for mink: NaN after everything, ties by linear index.
template <class T>
inline bool dist_less(T x, int ix, T y, int iy) {
  bool nx = std::isnan(x), ny = std::isnan(y);
  if (nx || ny) {
    if (nx && ny) return ix < iy;
    return ny;  // non-NaN < NaN
  }
  if (x < y) return true;
  if (y < x) return false;
  return ix < iy;
}

// ------------------------------------------------------------------ matching
// Box sums of Ipad^2 over d x d blocks (blockNormsPad, matchBlocksCPU.m:66),
// summed column by column (v outer, u inner) with fused multiply-adds.
template <class T>
std::vector<T> block_norm_map(const std::vector<T>& Ipad, int P, int Q, int d) {
  int Pm = P - d + 1, Qm = Q - d + 1;
  std::vector<T> out((size_t)Pm * Qm);
#pragma omp parallel for schedule(static)
  for (int q = 0; q < Qm; ++q)
    for (int p = 0; p < Pm; ++p) {
      T acc = 0;
      for (int v = 0; v < d; ++v)
        for (int u = 0; u < d; ++u) {
          T y = Ipad[(p + u) + (size_t)P * (q + v)];
          acc = madd<T>(y, y, acc);
        }
      out[p + (size_t)Pm * q] = acc;
    }
  return out;
}

// Windowed block matching shared by matchBlocksCPU (source/matchBlocksCPU.m:
// 56-103) and matchBlocksWie (source/matchBlocksWie.m:38-82).  For tile
// (i,j) and window offset (a,b) in [0,d2)^2:
//   dist = (|x|^2 + |y_ab|^2 - 2<x,y_ab>) / |x|^2          (:83-85)
// then the K smallest in ascending order (mink, :88; NaN last, ties by the
// lower linear index), then slot 1 := the window centre (:103).
// S[k + K*(i + Mt*j)] = 1-based window index; dist (nullable) receives all
// d2*d2 distances of every tile.
// Tolerance branch (Kmin != K, matchBlocksCPU.m:89-99): with `below` the
// number of window positions whose distance is < tol (the single value
// compared with the double tol exactly), below > K keeps the K smallest,
// below == K the K positions in index order (find), and below < K (mink of
// numMin, or find, returns fewer than numMax) is where the reference fails
// (Volume(.., 1:numel(S)) = blocks, :111, blocks has numMax slices): *short_out
// is set and the K smallest are kept.
template <class T>
void match_core(const T* I, int m, int n, int d, int d2, int K, uint32_t* S, T* dist_out,
                const std::vector<T>& Ipad, int Kmin = 0, double tol = 0.0, int* short_out = nullptr) {
  const int Mt = m / d, Nt = n / d, s = d2 / 2, P = m + d2, Q = n + d2;
  const int W2 = d2 * d2;
  std::vector<T> bn = block_norm_map<T>(Ipad, P, Q, d);
  const int Pm = P - d + 1;
#pragma omp parallel
  {
    std::vector<T> dist(W2), x((size_t)d * d);
    std::vector<int> order(W2);
#pragma omp for schedule(dynamic, 4) collapse(2)
    for (int j = 0; j < Nt; ++j)
      for (int i = 0; i < Mt; ++i) {
        for (int v = 0; v < d; ++v)
          for (int u = 0; u < d; ++u) x[u + d * v] = I[(i * d + u) + (size_t)m * (j * d + v)];
        T nx = 0;  // blockNorms at the tile origin (:65, :83)
        for (int v = 0; v < d; ++v)
          for (int u = 0; u < d; ++u) nx = madd<T>(x[u + d * v], x[u + d * v], nx);
        for (int b = 0; b < d2; ++b)
          for (int a = 0; a < d2; ++a) {
            const T* y = &Ipad[(i * d + a) + (size_t)P * (j * d + b)];
            T corr = 0;
            for (int v = 0; v < d; ++v)
              for (int u = 0; u < d; ++u) corr = madd<T>(x[u + d * v], y[u + (size_t)P * v], corr);
            T bnp = bn[(i * d + a) + (size_t)Pm * (j * d + b)];
            T num = (nx + bnp) - T(2) * corr;
            dist[a + d2 * b] = num / nx;
          }
        for (int t = 0; t < W2; ++t) order[t] = t;
        std::partial_sort(order.begin(), order.begin() + K, order.end(),
                          [&](int p, int q) { return dist_less<T>(dist[p], p, dist[q], q); });
        if (Kmin != 0 && Kmin != K) {
          int below = 0;
          for (int t = 0; t < W2; ++t) below += (double)dist[t] < tol;
          if (below == K) {
            std::sort(order.begin(), order.begin() + K);
          } else if (below < K && short_out) {
#pragma omp atomic write
            *short_out = 1;
          }
        }
        uint32_t* Sg = S + (size_t)K * (i + (size_t)Mt * j);
        for (int k = 0; k < K; ++k) Sg[k] = (uint32_t)order[k] + 1u;
        Sg[0] = (uint32_t)(s + d2 * s + 1);  // d2^2/2 + d2/2 + 1
        if (dist_out) std::memcpy(dist_out + (size_t)W2 * (i + (size_t)Mt * j), dist.data(), sizeof(T) * W2);
      }
  }
}

// [S, Volume] = matchBlocksCPU(I, sigma, opts); Kmin != K: the tolerance
// branch with tol = 0.005 sigma^2 d^2 (:21-28), returning GBM3D_ERR_ARG where
// the reference fails (see match_core).
template <class T>
int match_step1(const T* I, int m, int n, int d, int d2, int K, uint32_t* S, T* Volume, T* dist, int Kmin = 0,
                double sigma = 0.0) {
  std::vector<T> Ipad = search_pad<T>(I, m, n, d2);
  int short_tile = 0;
  match_core<T>(I, m, n, d, d2, K, S, dist, Ipad, Kmin, 0.005 * sigma * sigma * d * d, &short_tile);
  if (short_tile) return GBM3D_ERR_ARG;
  if (!Volume) return 0;
  const int Mt = m / d, Nt = n / d, P = m + d2;
  // Volume(tile, k) = Itmp(a:a+d-1, b:b+d-1)  (:107-111)
#pragma omp parallel for collapse(2)
  for (int j = 0; j < Nt; ++j)
    for (int i = 0; i < Mt; ++i)
      for (int k = 0; k < K; ++k) {
        uint32_t idx = S[k + (size_t)K * (i + (size_t)Mt * j)] - 1u;
        int a = idx % d2, b = idx / d2;
        for (int v = 0; v < d; ++v)
          for (int u = 0; u < d; ++u)
            Volume[(i * d + u) + (size_t)m * ((j * d + v) + (size_t)n * k)] =
                Ipad[(i * d + a + u) + (size_t)P * (j * d + b + v)];
      }
  return 0;
}

// [S, V1, V2] = matchBlocksWie(I, In, sigma, opts); V: d x d x K x Mt x Nt.
template <class T>
void match_step2(const T* I, const T* In, int m, int n, int d, int d2, int K, uint32_t* S, T* V1,
                 T* V2, T* dist) {
  std::vector<T> Ipad = search_pad<T>(I, m, n, d2);
  std::vector<T> Ipad2 = search_pad<T>(In, m, n, d2);
  match_core<T>(I, m, n, d, d2, K, S, dist, Ipad);
  const int Mt = m / d, Nt = n / d, P = m + d2;
#pragma omp parallel for collapse(2)
  for (int j = 0; j < Nt; ++j)
    for (int i = 0; i < Mt; ++i) {
      size_t g = i + (size_t)Mt * j;
      for (int k = 0; k < K; ++k) {
        uint32_t idx = S[k + K * g] - 1u;
        int a = idx % d2, b = idx / d2;
        for (int v = 0; v < d; ++v)
          for (int u = 0; u < d; ++u) {
            size_t src = (i * d + a + u) + (size_t)P * (j * d + b + v);
            size_t dst = u + (size_t)d * (v + (size_t)d * (k + (size_t)K * g));
            if (V1) V1[dst] = Ipad[src];
            if (V2) V2[dst] = Ipad2[src];
          }
      }
    }
}

// ---------------------------------------------------- 3-D wavelet transform
// waveTransforms3D/getWaveFilters3Dbior.m:16-27,49-50 (separable factors).
template <class T>
struct Bank {
  T alo[10], ahi[10], slo[10], shi[10], zlo[2], zhi[2];
  Bank() {
    const double Ca = 0.016572815184060, Cb = 0.121533978016438, Cc = std::sqrt(2.0) / 2.0;
    const double lo[10] = {Ca, -Ca, -Cb, Cb, Cc, Cc, Cb, -Cb, -Ca, Ca};
    const double hs[10] = {-Ca, -Ca, Cb, Cb, -Cc, Cc, -Cb, -Cb, Ca, Ca};
    for (int k = 0; k < 10; ++k) {
      alo[k] = (T)lo[k];
      ahi[k] = 0;
      slo[k] = 0;
      shi[k] = (T)hs[k];
    }
    ahi[4] = (T)-Cc;
    ahi[5] = (T)Cc;
    slo[4] = (T)Cc;
    slo[5] = (T)Cc;
    zlo[0] = (T)Cc;
    zlo[1] = (T)Cc;
    zhi[0] = (T)-Cc;
    zhi[1] = (T)Cc;
  }
};

template <class T>
struct Vol {
  int p = 0, q = 0, r = 0;
  std::vector<T> v;
  Vol() {}
  Vol(int p_, int q_, int r_) : p(p_), q(q_), r(r_), v((size_t)p_ * q_ * r_, T(0)) {}
  T& at(int i, int j, int k) { return v[i + (size_t)p * (j + (size_t)q * k)]; }
  const T& at(int i, int j, int k) const { return v[i + (size_t)p * (j + (size_t)q * k)]; }
  int dim(int a) const { return a == 0 ? p : (a == 1 ? q : r); }
};

// Analysis along one axis: out[t] = sum_{k<nt} h[k] in[(2t-k) mod L], i.e.
// circular convolution with the (zero-padded, fftn-truncated) filter kept at
// the even samples (myWavDec3DFFT.m:15-16).
template <class T>
Vol<T> ana_axis(const Vol<T>& in, int axis, const T* h, int ntaps) {
  int L = in.dim(axis);
  int nt = std::min(ntaps, L);
  Vol<T> out(axis == 0 ? in.p / 2 : in.p, axis == 1 ? in.q / 2 : in.q, axis == 2 ? in.r / 2 : in.r);
#pragma omp parallel for schedule(static) if (out.r * out.q > 64)
  for (int c = 0; c < out.r * out.q; ++c) {
    int j = c % out.q, kk = c / out.q;
    for (int i = 0; i < out.p; ++i) {
      T acc = 0;
      int t = axis == 0 ? i : (axis == 1 ? j : kk);
      for (int k = 0; k < nt; ++k) {
        int src = pmod(2 * t - k, L);
        T x = axis == 0 ? in.at(src, j, kk) : (axis == 1 ? in.at(i, src, kk) : in.at(i, j, src));
        acc += h[k] * x;
      }
      out.at(i, j, kk) = acc;
    }
  }
  return out;
}

// Synthesis along one axis: out[x] = sum_{k<nt, x+k even} g[k] in[((x+k) mod L)/2]
// (zero upsampling + circular correlation, myWavRec3DGPUfast.m:19-21,33-37).
template <class T>
Vol<T> syn_axis(const Vol<T>& in, int axis, const T* g, int ntaps) {
  int L = 2 * in.dim(axis);
  int nt = std::min(ntaps, L);
  Vol<T> out(axis == 0 ? in.p * 2 : in.p, axis == 1 ? in.q * 2 : in.q, axis == 2 ? in.r * 2 : in.r);
#pragma omp parallel for schedule(static) if (out.r * out.q > 64)
  for (int c = 0; c < out.r * out.q; ++c) {
    int j = c % out.q, kk = c / out.q;
    for (int i = 0; i < out.p; ++i) {
      T acc = 0;
      int x = axis == 0 ? i : (axis == 1 ? j : kk);
      for (int k = 0; k < nt; ++k) {
        int up = pmod(x + k, L);
        if (up & 1) continue;
        int src = up / 2;
        T v = axis == 0 ? in.at(src, j, kk) : (axis == 1 ? in.at(i, src, kk) : in.at(i, j, src));
        acc += g[k] * v;
      }
      out.at(i, j, kk) = acc;
    }
  }
  return out;
}

// Subband bits b = dim1_hi | dim2_hi<<1 | z_hi<<2 -> C{i,j} column j (1..7),
// Psi(:,:,:,j+1) of getWaveFilters3Dbior.m:39-46; b = 0 is the approximation.
inline int band_j(int b) {
  static const int jm[8] = {0, 1, 2, 4, 3, 5, 6, 7};
  return jm[b];
}

template <class T>
void analysis_level(const Vol<T>& X, Vol<T> C[8], const Bank<T>& F) {
  Vol<T> Z[2] = {ana_axis<T>(X, 2, F.zlo, 2), ana_axis<T>(X, 2, F.zhi, 2)};
  for (int zb = 0; zb < 2; ++zb) {
    Vol<T> A[2] = {ana_axis<T>(Z[zb], 1, F.alo, 10), ana_axis<T>(Z[zb], 1, F.ahi, 10)};
    for (int b2 = 0; b2 < 2; ++b2) {
      C[0 | (b2 << 1) | (zb << 2)] = ana_axis<T>(A[b2], 0, F.alo, 10);
      C[1 | (b2 << 1) | (zb << 2)] = ana_axis<T>(A[b2], 0, F.ahi, 10);
    }
  }
}

template <class T>
Vol<T> synthesis_level(const Vol<T> C[8], const Bank<T>& F) {
  Vol<T> Y;
  for (int b = 0; b < 8; ++b) {
    const T* g1 = (b & 1) ? F.shi : F.slo;
    const T* g2 = (b & 2) ? F.shi : F.slo;
    const T* gz = (b & 4) ? F.zhi : F.zlo;
    Vol<T> t = syn_axis<T>(syn_axis<T>(syn_axis<T>(C[b], 0, g1, 10), 1, g2, 10), 2, gz, 2);
    if (b == 0)
      Y = std::move(t);
    else
      for (size_t e = 0; e < Y.v.size(); ++e) Y.v[e] += t.v[e];
  }
  return Y;
}

// One cycle spin of denoiseAndAggregateGPU.m:29-42: decompose, hard
// threshold C.*(abs(C)>tau(i,j)) on the 7 details of every level, rebuild.
template <class T>
Vol<T> ht_one(const Vol<T>& X0, int levels, const double* tau, double* min_margin) {
  Bank<T> F;
  std::vector<Vol<T>> A(levels + 1);
  std::vector<std::vector<Vol<T>>> D(levels, std::vector<Vol<T>>(8));
  A[0] = X0;
  for (int l = 0; l < levels; ++l) {
    Vol<T> C[8];
    analysis_level<T>(A[l], C, F);
    for (int b = 1; b < 8; ++b) {
      double t = tau[l + levels * (band_j(b) - 1)];
      for (auto& c : C[b].v) {
        double ac = std::fabs((double)c);
        if (min_margin && t > 0) *min_margin = std::min(*min_margin, std::fabs(ac - t) / t);
        if (!(std::fabs(c) > (T)t)) c = 0;
      }
      D[l][b] = std::move(C[b]);
    }
    A[l + 1] = std::move(C[0]);
  }
  Vol<T> Y = A[levels];
  for (int l = levels - 1; l >= 0; --l) {
    D[l][0] = std::move(Y);
    Y = synthesis_level<T>(D[l].data(), F);
  }
  return Y;
}

// Vavg = max(mean over cycle spins of circshift-back(rec(thr(dec(circshift(V))))), 0)
// (source/denoiseAndAggregateGPU.m:27-47).
template <class T>
void ht_filter(const T* V, int m, int n, int K, int levels, int cycleSpin, const double* tau,
               T* Vavg, double* min_margin) {
  std::vector<T> acc((size_t)m * n * K, T(0));
  for (int sp = 0; sp < cycleSpin; ++sp) {
    Vol<T> X(m, n, K);
    for (int k = 0; k < K; ++k)
      for (int c = 0; c < n; ++c)
        for (int r = 0; r < m; ++r)
          X.at(r, c, k) = V[pmod(r - sp, m) + (size_t)m * (pmod(c - sp, n) + (size_t)n * pmod(k - sp, K))];
    Vol<T> Y = ht_one<T>(X, levels, tau, min_margin);
    for (int k = 0; k < K; ++k)
      for (int c = 0; c < n; ++c)
        for (int r = 0; r < m; ++r)
          acc[r + (size_t)m * (c + (size_t)n * k)] += Y.at(pmod(r + sp, m), pmod(c + sp, n), pmod(k + sp, K));
  }
  for (size_t e = 0; e < acc.size(); ++e) {
    T v = acc[e] / (T)cycleSpin;
    Vavg[e] = v > 0 ? v : T(0);
  }
}

// ------------------------------------------------- near-threshold masks
// Which outputs of ht_filter can move by more than round-off when the
// transform's operation order changes: every voxel in the synthesis support
// of a detail coefficient C with | |C| - tau | <= tol (an absolute margin on
// the coefficients' scale; the hard threshold C .* (abs(C) > tau) may decide
// either way there).  A level-l
// coefficient (l = 0 finest) at index s along an axis reaches the full-size
// positions [2^(l+1) s - (t - 1)(2^(l+1) - 1), 2^(l+1) s] (mod size), t the
// synthesis filter length (10 in x, y; 2 in z): each synthesis level maps
// index s to 2s - k, k in [0, t) (myWavRec3DGPUfast.m:19-21,33-37).  The box
// is taken in the cycle-spun frame and shifted back.  vmask: m x n x K, set
// to 1 where affected (never cleared).  Returns the number of such
// coefficients.
template <class T>
long ht_near_mask(const T* V, int m, int n, int K, int levels, int cycleSpin, const double* tau, double tol,
                  uint8_t* vmask) {
  Bank<T> F;
  long count = 0;
  auto mark_box = [&](int l, int si, int sj, int sk, int sp) {
    const int f = 1 << (l + 1);
    const int ri = 9 * (f - 1), rk = f - 1;
    for (int dk = 0; dk <= rk; ++dk) {
      const int k = pmod((long long)f * sk - dk - sp, K);
      for (int dj = 0; dj <= ri; ++dj) {
        const int c = pmod((long long)f * sj - dj - sp, n);
        for (int di = 0; di <= ri; ++di) {
          const int r = pmod((long long)f * si - di - sp, m);
          vmask[r + (size_t)m * (c + (size_t)n * k)] = 1;
        }
      }
    }
  };
  for (int sp = 0; sp < cycleSpin; ++sp) {
    Vol<T> X(m, n, K);
    for (int k = 0; k < K; ++k)
      for (int c = 0; c < n; ++c)
        for (int r = 0; r < m; ++r)
          X.at(r, c, k) = V[pmod(r - sp, m) + (size_t)m * (pmod(c - sp, n) + (size_t)n * pmod(k - sp, K))];
    Vol<T> A = X;
    for (int l = 0; l < levels; ++l) {
      Vol<T> C[8];
      analysis_level<T>(A, C, F);
      for (int b = 1; b < 8; ++b) {
        const double t = tau[l + levels * (band_j(b) - 1)];
        if (!(t > 0)) continue;
        const Vol<T>& Cb = C[b];
        for (int kk = 0; kk < Cb.r; ++kk)
          for (int j = 0; j < Cb.q; ++j)
            for (int i = 0; i < Cb.p; ++i)
              if (std::fabs(std::fabs((double)Cb.at(i, j, kk)) - t) <= tol) {
                ++count;
                mark_box(l, i, j, kk, sp);
              }
      }
      A = std::move(C[0]);
    }
  }
  return count;
}

// The pixels of GBM3D's first estimate U1 (m x n, after the crop) that a
// near-threshold coefficient of step 1 can move (ht_near_mask per match
// spin): a changed voxel changes its group's TV weight, so every pixel that
// any of the group's K blocks covers is marked (source/GBM3D.m:50-57,
// denoiseAndAggregateGPU.m:70-89).  Returns the number of near-threshold
// coefficients.
template <class T>
long u1_near_mask(const T* I0, int m, int n, double sigma, const gbm3d_opts& o, double tol, uint8_t* mask_out) {
  const int bs = o.blockSize, d2 = o.matchSize, K = o.numMax, s = d2 / 2;
  const int M = (m + bs - 1) / bs * bs, N = (n + bs - 1) / bs * bs;
  std::vector<T> I = mirror_pad<T>(I0, m, n, M, N);
  std::vector<double> tau((size_t)o.levels * 7);
  get_tau(o.levels, o.tauMode, sigma, tau.data());
  const size_t MN = (size_t)M * N;
  const int Mt = M / bs, Nt = N / bs;
  std::vector<uint8_t> pm(MN, 0);
  std::vector<uint32_t> S((size_t)K * Mt * Nt);
  std::vector<T> Vol(MN * K);
  long count = 0;
  for (int ii = 0; ii < o.nMatchSpins; ++ii) {
    const int sp = o.matchSpins[ii];
    std::vector<T> Is = circshift2<T>(I.data(), M, N, sp);
    match_step1<T>(Is.data(), M, N, bs, d2, K, S.data(), Vol.data(), nullptr);
    std::vector<uint8_t> vm(MN * K, 0);
    count += ht_near_mask<T>(Vol.data(), M, N, K, o.levels, o.cycleSpin, tau.data(), tol, vm.data());
    for (int x = 0; x < Nt; ++x)
      for (int y = 0; y < Mt; ++y) {
        bool hit = false;
        for (int k = 0; k < K && !hit; ++k)
          for (int v = 0; v < bs && !hit; ++v)
            for (int u = 0; u < bs && !hit; ++u) hit = vm[(y * bs + u) + (size_t)M * ((x * bs + v) + (size_t)N * k)];
        if (!hit) continue;
        const size_t g = y + (size_t)Mt * x;
        for (int k = 0; k < K; ++k) {
          const uint32_t idx = S[k + K * g] - 1u;
          const int r0 = y * bs + (int)(idx % d2) - s, c0 = x * bs + (int)(idx / d2) - s;
          for (int v = 0; v < bs; ++v)
            for (int u = 0; u < bs; ++u) {
              const int r = r0 + u, c = c0 + v;
              if (r < 0 || r >= M || c < 0 || c >= N) continue;  // crop-drop
              // this spin's frame is the image circshifted by sp
              pm[pmod(r - sp, M) + (size_t)M * pmod(c - sp, N)] = 1;
            }
        }
      }
  }
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r) mask_out[r + (size_t)m * c] = pm[r + (size_t)M * c];
  return count;
}

// ---------------------------------------------------------------- aggregation
// getBM3DAggWeights(3,1) (utilities/getBM3DAggWeights.m:15-19) on one group
// stack blk (d x d x K, column-major), summed in column-major order.
template <class T>
T agg_tv(const T* blk, int d, int K) {
  auto at = [&](int u, int v, int k) { return blk[u + d * (v + d * k)]; };
  auto d3 = [](T x0, T x1, T x2, T x3) {
    T a = x1 - x0, b = x2 - x1, c = x3 - x2;
    T e = b - a, f = c - b;
    return f - e;
  };
  T s2 = 0, s1 = 0, s3 = 0;
  for (int k = 0; k < K; ++k)
    for (int v = 0; v + 3 < d; ++v)
      for (int u = 0; u < d; ++u) s2 += std::fabs(d3(at(u, v, k), at(u, v + 1, k), at(u, v + 2, k), at(u, v + 3, k)));
  for (int k = 0; k < K; ++k)
    for (int v = 0; v < d; ++v)
      for (int u = 0; u + 3 < d; ++u) s1 += std::fabs(d3(at(u, v, k), at(u + 1, v, k), at(u + 2, v, k), at(u + 3, v, k)));
  for (int k = 0; k + 1 < K; ++k)
    for (int v = 0; v < d; ++v)
      for (int u = 0; u < d; ++u) s3 += std::fabs(at(u, v, k + 1) - at(u, v, k));
  T W = s2 * T(0.25);
  W = W + s1 * T(0.25);
  W = W + s3;
  return W;
}

// MATLAB linspace(a, b, n) (last element exactly b).
inline std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> y(n);
  if (n == 1) {
    y[0] = b;
    return y;
  }
  double step = (b - a) / (n - 1);
  for (int k = 0; k < n; ++k) y[k] = a + k * step;
  y[n - 1] = b;
  return y;
}

// mask = s' * s, s = sin(pi*linspace(lo, hi, d))  (denoiseAndAggregateGPU.m:65-67,
// denoiseAndAggregateWieFast.m:72-75); mask[u + d*v].
inline std::vector<double> sine_mask(double lo, double hi, int d) {
  std::vector<double> s = linspace(lo, hi, d), w((size_t)d * d);
  for (auto& x : s) x = std::sin(M_PI * x);
  for (int v = 0; v < d; ++v)
    for (int u = 0; u < d; ++u) w[u + d * v] = s[u] * s[v];
  return w;
}

// First tile row whose blocks can cover pixel row r: a block of tile y spans
// rows [y*d + a - s, y*d + a - s + d) with a in [0, d2), so y*d >= r + s - d2 - d + 2.
inline int first_tile(int r, int s, int d, int d2) {
  int x = r + s - d2 - d + 2;
  return x > 0 ? (x + d - 1) / d : 0;
}

// Step-1 contribution of one pixel of one block: U += w*b.*mask, W += w.*mask.
// FP32: single*single then .*mask in single (MATLAB mixed single/double
// arithmetic is single), accumulated into the canvas in single.
template <class T>
struct Agg1Math;
template <>
struct Agg1Math<double> {
  static void add(double& U, double& W, double w, double b, double mask) {
    U += w * b * mask;
    W += w * mask;
  }
};
template <>
struct Agg1Math<float> {
  static void add(float& U, float& W, float w, float b, float mask) {
    float p = w * b;
    p = p * mask;
    U = U + p;
    W = W + w * mask;
  }
};

// Group weights 1/(D(Vavg tile)+eps) of denoiseAndAggregateGPU.m:74-75.
template <class T>
std::vector<T> agg1_weights(const T* Vavg, int m, int n, int d, int K) {
  const int Mt = m / d, Nt = n / d;
  std::vector<T> w((size_t)Mt * Nt);
#pragma omp parallel
  {
    std::vector<T> blk((size_t)d * d * K);
#pragma omp for collapse(2)
    for (int j = 0; j < Nt; ++j)
      for (int i = 0; i < Mt; ++i) {
        for (int k = 0; k < K; ++k)
          for (int v = 0; v < d; ++v)
            for (int u = 0; u < d; ++u)
              blk[u + d * (v + d * k)] = Vavg[(i * d + u) + (size_t)m * ((j * d + v) + (size_t)n * k)];
        T D = agg_tv<T>(blk.data(), d, K);
        w[i + (size_t)Mt * j] = T(1) / (D + (T)std::numeric_limits<double>::epsilon());
      }
  }
  return w;
}

// Aggregation of denoiseAndAggregateGPU (:59-92).  Block k of tile (y,x) is
// placed at image origin (y*d + a - d2/2, x*d + b - d2/2); parts outside the
// m x n crop are dropped (canvas (m+d2)x(n+d2), crop [d2/2, m+d2/2)).
// scatter == the reference's loop order (y outer, x inner, k); gather visits,
// for every pixel, the same (y, x, k) sequence, so both give identical sums.
template <class T>
void aggregate_ht(const T* Vavg, const uint32_t* S, int m, int n, int d, int d2, int K, T* U, T* W,
                  bool gather) {
  const int Mt = m / d, Nt = n / d, s = d2 / 2;
  std::vector<T> w = agg1_weights<T>(Vavg, m, n, d, K);
  std::vector<double> maskd = sine_mask(0.1, 0.9, d);
  std::vector<T> mask(maskd.begin(), maskd.end());
  if (!gather) {
    std::fill(U, U + (size_t)m * n, T(0));
    std::fill(W, W + (size_t)m * n, T(0));
    for (int y = 0; y < Mt; ++y)
      for (int x = 0; x < Nt; ++x) {
        size_t g = y + (size_t)Mt * x;
        for (int k = 0; k < K; ++k) {
          uint32_t idx = S[k + K * g] - 1u;
          int r0 = y * d + (int)(idx % d2) - s, c0 = x * d + (int)(idx / d2) - s;
          for (int v = 0; v < d; ++v)
            for (int u = 0; u < d; ++u) {
              int r = r0 + u, c = c0 + v;
              if (r < 0 || r >= m || c < 0 || c >= n) continue;
              T b = Vavg[(y * d + u) + (size_t)m * ((x * d + v) + (size_t)n * k)];
              Agg1Math<T>::add(U[r + (size_t)m * c], W[r + (size_t)m * c], w[g], b, mask[u + d * v]);
            }
        }
      }
    return;
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r) {
      T Ua = 0, Wa = 0;
      int ylo = first_tile(r, s, d, d2), yhi = std::min(Mt - 1, (r + s) / d);
      int xlo = first_tile(c, s, d, d2), xhi = std::min(Nt - 1, (c + s) / d);
      for (int y = ylo; y <= yhi; ++y)
        for (int x = xlo; x <= xhi; ++x) {
          size_t g = y + (size_t)Mt * x;
          for (int k = 0; k < K; ++k) {
            uint32_t idx = S[k + K * g] - 1u;
            int u = r - (y * d + (int)(idx % d2) - s), v = c - (x * d + (int)(idx / d2) - s);
            if (u < 0 || u >= d || v < 0 || v >= d) continue;
            T b = Vavg[(y * d + u) + (size_t)m * ((x * d + v) + (size_t)n * k)];
            Agg1Math<T>::add(Ua, Wa, w[g], b, mask[u + d * v]);
          }
        }
      U[r + (size_t)m * c] = Ua;
      W[r + (size_t)m * c] = Wa;
    }
}

// ------------------------------------------------------------------- Wiener
// Orthonormal DCT-II matrix, T = dct(eye(d)): Tm[k + d*t] (row k, column t).
template <class T>
std::vector<T> dct_matrix(int d) {
  std::vector<T> Tm((size_t)d * d);
  for (int k = 0; k < d; ++k)
    for (int t = 0; t < d; ++t) {
      double ck = k == 0 ? std::sqrt(1.0 / d) : std::sqrt(2.0 / d);
      Tm[k + d * t] = (T)(ck * std::cos(M_PI * (2 * t + 1) * k / (2.0 * d)));
    }
  return Tm;
}

// myWavDec1D with Psi = [-c c], Phi = [c c] along the group dimension
// (waveTransforms3D/myWavDec1D.m:14-26): pairs (2t-1, 2t) circularly; output
// [level-1 details | level-2 details | ... | approximation].
template <class T>
void haar_dec(T* x, int R, int levels) {
  const T c = (T)(std::sqrt(2.0) / 2.0);
  std::vector<T> out(R), cur(x, x + R), lo;
  int cnt = 0, L = R;
  for (int l = 0; l < levels; ++l) {
    lo.assign(L / 2, T(0));
    for (int t = 0; t < L / 2; ++t) {
      T a = cur[2 * t], b = cur[pmod(2 * t - 1, L)];
      out[cnt + t] = -c * a + c * b;
      lo[t] = c * a + c * b;
    }
    cnt += L / 2;
    L /= 2;
    cur = lo;
  }
  for (int t = 0; t < L; ++t) out[cnt + t] = cur[t];
  std::copy(out.begin(), out.end(), x);
}

// myWavRec1D (waveTransforms3D/myWavRec1D.m:10-26): y[2t] = g[0] c[t],
// y[2t-1] = g[1] c[t] for each of the detail (g = Psi) and low (g = Phi) paths.
template <class T>
void haar_rec(T* x, int R, int levels) {
  const T c = (T)(std::sqrt(2.0) / 2.0);
  int Lc = R >> levels;
  int cnt = R - Lc;
  std::vector<T> cur(x + cnt, x + R), nxt;
  for (int l = levels - 1; l >= 0; --l) {
    int L = R >> l;  // output length of this level
    cnt -= L / 2;
    const T* det = x + cnt;
    nxt.assign(L, T(0));
    for (int t = 0; t < L / 2; ++t) {
      nxt[2 * t] += -c * det[t];
      nxt[pmod(2 * t - 1, L)] += c * det[t];
    }
    for (int t = 0; t < L / 2; ++t) {
      nxt[2 * t] += c * cur[t];
      nxt[pmod(2 * t - 1, L)] += c * cur[t];
    }
    cur = nxt;
  }
  std::copy(cur.begin(), cur.end(), x);
}

// denoiseAndAggregateWieFast.m:19-66 on G groups (V: d x d x K x G).
// sigWie (d x d, nullable): per-coefficient sigma (coloured noise; the d x d
// sigma of :49 expands over the slices and groups).
template <class T>
// clip01: denoiseAndAggregateWie.m (the dct() twin): V2 = min(max(V2,0),1)
// (:55) instead of max(V2,0); its dct/idct along dims 1 and 2 are the same
// orthonormal DCT-II pages (:34-35,44-45,53-54).
void wiener(const T* V1, const T* V2, int d, int K, long G, double sigma, T* V2f, T* gain2,
            const double* sigWie = nullptr, bool clip01 = false) {
  const int levels = K < 16 ? 2 : 3;  // :23-24
  std::vector<T> Tm = dct_matrix<T>(d);
  std::vector<T> s2tab((size_t)d * d);
  for (int e = 0; e < d * d; ++e) s2tab[e] = sigWie ? (T)(sigWie[e] * sigWie[e]) : (T)(sigma * sigma);
  const size_t gs = (size_t)d * d * K;
#pragma omp parallel
  {
    std::vector<T> A(gs), B(gs), tmp((size_t)d * d), col(K), gain(gs);
    // 2-D transform of each page: out = L * X * R^T  (L = R = Tm forward,
    // L = R = Tm^T inverse since inv(T) = T^T for the orthonormal DCT).
    auto page = [&](T* X, bool inverse) {
      for (int k = 0; k < K; ++k) {
        T* P = X + (size_t)d * d * k;
        for (int v = 0; v < d; ++v)
          for (int p = 0; p < d; ++p) {
            T acc = 0;
            for (int u = 0; u < d; ++u) acc += (inverse ? Tm[u + d * p] : Tm[p + d * u]) * P[u + d * v];
            tmp[p + d * v] = acc;
          }
        for (int q = 0; q < d; ++q)
          for (int p = 0; p < d; ++p) {
            T acc = 0;
            for (int v = 0; v < d; ++v) acc += tmp[p + d * v] * (inverse ? Tm[v + d * q] : Tm[q + d * v]);
            P[p + d * q] = acc;
          }
      }
    };
    auto along_k = [&](T* X, bool rec) {
      for (int e = 0; e < d * d; ++e) {
        for (int k = 0; k < K; ++k) col[k] = X[e + (size_t)d * d * k];
        if (rec)
          haar_rec<T>(col.data(), K, levels);
        else
          haar_dec<T>(col.data(), K, levels);
        for (int k = 0; k < K; ++k) X[e + (size_t)d * d * k] = col[k];
      }
    };
#pragma omp for schedule(static)
    for (long g = 0; g < G; ++g) {
      std::copy(V1 + gs * g, V1 + gs * (g + 1), A.begin());
      std::copy(V2 + gs * g, V2 + gs * (g + 1), B.begin());
      page(A.data(), false);
      along_k(A.data(), false);
      page(B.data(), false);
      along_k(B.data(), false);
      T n2 = 0;
      for (size_t e = 0; e < gs; ++e) {
        T a2 = A[e] * A[e];
        T f = a2 / (a2 + s2tab[e % (d * d)]);  // :49
        int k = (int)(e / (d * d));
        if (e % (d * d) == 0 && k >= 14) f = 1;  // filt(1,1,15:end) = 1 (:50)
        B[e] *= f;
        n2 += f * f;
      }
      along_k(B.data(), true);
      page(B.data(), true);
      for (size_t e = 0; e < gs; ++e) {
        T y = B[e] > 0 ? B[e] : T(0);  // :64
        if (clip01 && y > T(1)) y = T(1);
        V2f[gs * g + e] = y;
      }
      gain2[g] = n2;
    }
  }
}

// Aggregation of denoiseAndAggregateWieFast (:68-100): weight = w/|filt|_F^2,
// w = s'*s with s = sin(pi*linspace(.2,.8,d)); same placement and crop as step 1.
template <class T>
void aggregate_wie(const T* V2f, const T* gain2, const uint32_t* S, int m, int n, int d, int d2, int K,
                   T* U, T* W, bool gather) {
  const int Mt = m / d, Nt = n / d, s = d2 / 2;
  std::vector<double> wd = sine_mask(0.2, 0.8, d);
  std::vector<T> wm(wd.begin(), wd.end());
  auto blk = [&](size_t g, int k, int u, int v) { return V2f[u + (size_t)d * (v + (size_t)d * (k + (size_t)K * g))]; };
  if (!gather) {
    std::fill(U, U + (size_t)m * n, T(0));
    std::fill(W, W + (size_t)m * n, T(0));
    for (int y = 0; y < Mt; ++y)
      for (int x = 0; x < Nt; ++x) {
        size_t g = y + (size_t)Mt * x;
        for (int k = 0; k < K; ++k) {
          uint32_t idx = S[k + K * g] - 1u;
          int r0 = y * d + (int)(idx % d2) - s, c0 = x * d + (int)(idx / d2) - s;
          for (int v = 0; v < d; ++v)
            for (int u = 0; u < d; ++u) {
              int r = r0 + u, c = c0 + v;
              if (r < 0 || r >= m || c < 0 || c >= n) continue;
              T wt = wm[u + d * v] / gain2[g];
              U[r + (size_t)m * c] += wt * blk(g, k, u, v);
              W[r + (size_t)m * c] += wt;
            }
        }
      }
    return;
  }
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r) {
      T Ua = 0, Wa = 0;
      int ylo = first_tile(r, s, d, d2), yhi = std::min(Mt - 1, (r + s) / d);
      int xlo = first_tile(c, s, d, d2), xhi = std::min(Nt - 1, (c + s) / d);
      for (int y = ylo; y <= yhi; ++y)
        for (int x = xlo; x <= xhi; ++x) {
          size_t g = y + (size_t)Mt * x;
          for (int k = 0; k < K; ++k) {
            uint32_t idx = S[k + K * g] - 1u;
            int u = r - (y * d + (int)(idx % d2) - s), v = c - (x * d + (int)(idx / d2) - s);
            if (u < 0 || u >= d || v < 0 || v >= d) continue;
            T wt = wm[u + d * v] / gain2[g];
            Ua += wt * blk(g, k, u, v);
            Wa += wt;
          }
        }
      U[r + (size_t)m * c] = Ua;
      W[r + (size_t)m * c] = Wa;
    }
}

// --------------------------------------------------------------- the driver
// [U, out] = GBM3D(I, sigma, opts)  (source/GBM3D.m:21-83) on resolved opts.
// Returns 0, or a gbm3d_status for unsupported options.
// Vavg = medfilt3(Volume, [3 3 3]) (source/denoiseAndAggregateGPU.m:49-52):
// the median of the 27 values of the 3x3x3 neighbourhood, the volume extended
// by MATLAB's default 'symmetric' padding, which for a window of 3 repeats the
// edge value.  No clip at 0 on this branch.
template <class T>
void median3(const T* V, int m, int n, int K, T* out) {
  auto at = [&](int r, int c, int k) {
    r = std::min(std::max(r, 0), m - 1);
    c = std::min(std::max(c, 0), n - 1);
    k = std::min(std::max(k, 0), K - 1);
    return V[r + (size_t)m * (c + (size_t)n * k)];
  };
  std::vector<T> w(27);
  for (int k = 0; k < K; ++k)
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < m; ++r) {
        int e = 0;
        for (int dk = -1; dk <= 1; ++dk)
          for (int dc = -1; dc <= 1; ++dc)
            for (int dr = -1; dr <= 1; ++dr) w[e++] = at(r + dr, c + dc, k + dk);
        std::nth_element(w.begin(), w.begin() + 13, w.end());
        out[r + (size_t)m * (c + (size_t)n * k)] = w[13];
      }
}

// Coloured-noise thresholds (utilities/getColoredTau.m:9-57): 6 sqrt(sum(PSD
// |F|^2)/sum(|F|^2)) with F the 2-D DFT (d1 x d2) of the band's upsampled
// 2-D filter, evaluated directly as a 2-D DFT of the filter (no separation).
inline void colored_tau(int levels, const double* psd, int d1, int d2, double* tau) {
  const double lo[10] = {0.016572815184060, -0.016572815184060, -0.121533978016438, 0.121533978016438,
                         0.70710678118654752440, 0.70710678118654752440, 0.121533978016438, -0.121533978016438,
                         -0.016572815184060, 0.016572815184060};
  const double hi[10] = {0, 0, 0, 0, -0.70710678118654752440, 0.70710678118654752440, 0, 0, 0, 0};
  const double* rows[7] = {hi, lo, lo, hi, hi, lo, hi};  // Psi(:,:,j+1) = rows' * cols
  const double* cols[7] = {lo, hi, lo, hi, lo, hi, hi};
  for (int i = 0; i < levels; ++i) {
    const int s = 1 << i, L1 = std::min(10 * s, d1), L2 = std::min(10 * s, d2);
    for (int j = 0; j < 7; ++j) {
      double num = 0.0, den = 0.0;
      for (int k2 = 0; k2 < d2; ++k2)
        for (int k1 = 0; k1 < d1; ++k1) {
          double re = 0.0, im = 0.0;
          for (int t2 = 0; t2 < L2; ++t2)
            for (int t1 = 0; t1 < L1; ++t1) {
              const double v = rows[j][t1 / s] * cols[j][t2 / s];
              if (v == 0.0) continue;
              const double a = -2.0 * M_PI * ((double)k1 * t1 / d1 + (double)k2 * t2 / d2);
              re += v * std::cos(a);
              im += v * std::sin(a);
            }
          const double p = re * re + im * im;
          num += psd[k1 + (size_t)d1 * k2] * p;
          den += p;
        }
      tau[i + levels * j] = 6.0 * std::sqrt(num / den);
    }
  }
}

template <class T>
int gbm3d(const T* I0, int m, int n, double sigma, const gbm3d_opts& o, T* Uout, T* U1out, bool gather,
          const double* tau_in = nullptr, const double* sigWie = nullptr) {
  if (o.filtType != GBM3D_FILT_HT && o.filtType != GBM3D_FILT_MEDIAN) return GBM3D_ERR_UNSUPPORTED;
  if (o.numMin > o.numMax) return GBM3D_ERR_ARG;
  if (o.Wiener && o.nMatchSpinsWie == GBM3D_UNSET) return GBM3D_ERR_MISSING_FIELD;
  const int bs = o.blockSize, d2 = o.matchSize, K = o.numMax;
  const int M = (m + bs - 1) / bs * bs, N = (n + bs - 1) / bs * bs;  // :23-24
  std::vector<T> I = mirror_pad<T>(I0, m, n, M, N);
  std::vector<double> tau((size_t)o.levels * 7);
  if (tau_in)
    std::copy(tau_in, tau_in + tau.size(), tau.begin());
  else
    get_tau(o.levels, o.tauMode, sigma, tau.data());
  const int n1 = o.nMatchSpins, n2 = o.Wiener ? o.nMatchSpinsWie : 0;
  const int nsl = std::max(n1, n2);
  const size_t MN = (size_t)M * N;
  std::vector<double> Usl(MN * nsl, 0.0), Wsl(MN * nsl, 0.0), U1sl(MN * nsl, 0.0);
  {
    const int Mt = M / bs, Nt = N / bs;
    std::vector<uint32_t> S((size_t)K * Mt * Nt);
    std::vector<T> Vol(MN * K), Vavg(MN * K), U(MN), W(MN);
    for (int ii = 0; ii < n1; ++ii) {  // :50-55
      int sp = o.matchSpins[ii];
      std::vector<T> Is = circshift2<T>(I.data(), M, N, sp);
      const int st1 = match_step1<T>(Is.data(), M, N, bs, d2, K, S.data(), Vol.data(), nullptr, o.numMin, sigma);
      if (st1) return st1;
      if (o.filtType == GBM3D_FILT_MEDIAN)
        median3<T>(Vol.data(), M, N, K, Vavg.data());
      else
        ht_filter<T>(Vol.data(), M, N, K, o.levels, o.cycleSpin, tau.data(), Vavg.data(), nullptr);
      aggregate_ht<T>(Vavg.data(), S.data(), M, N, bs, d2, K, U.data(), W.data(), gather);
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < M; ++r) {
          size_t src = pmod(r + sp, M) + (size_t)M * pmod(c + sp, N);
          U1sl[MN * ii + r + (size_t)M * c] = (double)U[src];
          Wsl[MN * ii + r + (size_t)M * c] = (double)W[src];
        }
    }
  }
  std::vector<T> U1(MN);  // :57
  for (size_t e = 0; e < MN; ++e) {
    double a = 0, b = 0;
    for (int ii = 0; ii < nsl; ++ii) {
      a += U1sl[MN * ii + e];
      b += Wsl[MN * ii + e];
    }
    U1[e] = (T)(a / b);
  }
  auto crop = [&](const std::vector<T>& X, T* out) {
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < m; ++r) out[r + (size_t)m * c] = X[r + (size_t)M * c];
  };
  if (U1out) crop(U1, U1out);
  if (!o.Wiener) {  // :59-63
    crop(U1, Uout);
    return 0;
  }
  {
    const int d = o.blockSizeWie, Mt = M / d, Nt = N / d;
    const long G = (long)Mt * Nt;
    std::vector<uint32_t> S((size_t)K * G);
    std::vector<T> V1((size_t)d * d * K * G), V2(V1.size()), V2f(V1.size()), g2(G), U(MN), W(MN);
    for (int ii = 0; ii < n2; ++ii) {  // :71-79
      int sp = o.matchSpinsWie[ii];
      std::vector<T> Ps = circshift2<T>(U1.data(), M, N, sp);
      std::vector<T> Is = circshift2<T>(I.data(), M, N, sp);
      match_step2<T>(Ps.data(), Is.data(), M, N, d, d2, K, S.data(), V1.data(), V2.data(), nullptr);
      wiener<T>(V1.data(), V2.data(), d, K, G, sigma, V2f.data(), g2.data(), sigWie);
      aggregate_wie<T>(V2f.data(), g2.data(), S.data(), M, N, d, d2, K, U.data(), W.data(), gather);
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < M; ++r) {
          size_t src = pmod(r + sp, M) + (size_t)M * pmod(c + sp, N);
          Usl[MN * ii + r + (size_t)M * c] = (double)U[src];
          Wsl[MN * ii + r + (size_t)M * c] = (double)W[src];
        }
    }
  }
  // U = sum(U,3)./sum(W,3) (:82).  Slices ii >= numel(matchSpinsWie) keep
  // U = 0 and the step-1 W (U and W are sized by numel(matchSpins), :26-27).
  std::vector<T> U(MN);
  for (size_t e = 0; e < MN; ++e) {
    double a = 0, b = 0;
    for (int ii = 0; ii < nsl; ++ii) {
      a += Usl[MN * ii + e];
      b += Wsl[MN * ii + e];
    }
    U[e] = (T)(a / b);
  }
  crop(U, Uout);
  return 0;
}


// ------------------------------------------------- whole-image matching
// S = matchBlocksGPU(I, sigma, opts) (source/matchBlocksGPU.m:18-91), off the
// GBM3D() path: for every reference tile (stride d, floor(m/d) x floor(n/d))
// the normalised distance (|x|^2 + |y|^2 - 2<x,y>)/|x|^2 to the block at
// every one of the m*n positions of the image, circularly wrapped, and the
// numMax smallest by mink (ascending, NaN last, ties to the lower index; no
// centre override).  Indices are 1-based column-major into m x n (the block's
// top-left pixel).  When sigma > 30/255 the search runs on a prefiltered
// image (:34-50): the mean over 5 cycle shifts of a 3-level 2-D sym4 wavelet
// hard threshold at 2 sigma.
//
// Sums, in this order (the GPU forms the same chains): block norms and
// correlations column by column (v outer, u inner) with fused multiply-adds;
// the 2-D filters Psi_b(u, v) = h1[u] h2[v] rounded to T, analysis taps and
// synthesis taps likewise v outer, u inner; the synthesised bands added in
// the order of myWavRec2DGPUfast.m:15-31.

// sym4 decomposition filters, archive/waveletFilters/sym4Filters.mat (Lo_d, Hi_d)
constexpr double kSym4Lo[8] = {-0.07576571478927333, -0.02963552764599851, 0.49761866763201545, 0.8037387518059161,
                               0.29785779560527736,  -0.09921954357684722, -0.012603967262037833, 0.0322231006040427};
constexpr double kSym4Hi[8] = {-0.0322231006040427, -0.012603967262037833, 0.09921954357684722, 0.29785779560527736,
                               -0.8037387518059161, 0.49761866763201545,   0.02963552764599851,  -0.07576571478927333};

// Psi(:,:,b+1) of getWaveFilters2D.m:12-16: b = 0 Lo'Lo, 1 Hi'Lo (dim 1 high),
// 2 Lo'Hi (dim 2 high), 3 Hi'Hi; entry (u, v) = h1[u] h2[v], stored as T.
template <class T>
struct Sym4Bank {
  T psi[4][64];
  Sym4Bank() {
    for (int b = 0; b < 4; ++b)
      for (int v = 0; v < 8; ++v)
        for (int u = 0; u < 8; ++u) {
          const double h1 = (b & 1) ? kSym4Hi[u] : kSym4Lo[u];
          const double h2 = (b & 2) ? kSym4Hi[v] : kSym4Lo[v];
          psi[b][u + 8 * v] = (T)(h1 * h2);
        }
  }
};

// One level of myWavDec2DFFT.m:13-21: band b at (i, j) = sum_{u,v} Psi_b(u,v)
// X((2i - u) mod L1, (2j - v) mod L2) (circular convolution kept at the even
// samples; filters truncated to the level size like fftn(Psi, size(U))).
template <class T>
void wav2_analysis(const std::vector<T>& X, int L1, int L2, const Sym4Bank<T>& F, std::vector<T> out[4]) {
  const int h1 = L1 / 2, h2 = L2 / 2, n1 = std::min(8, L1), n2 = std::min(8, L2);
  for (int b = 0; b < 4; ++b) out[b].assign((size_t)h1 * h2, T(0));
#pragma omp parallel for schedule(static) if (h1 * h2 > 4096)
  for (int j = 0; j < h2; ++j)
    for (int i = 0; i < h1; ++i)
      for (int b = 0; b < 4; ++b) {
        T acc = 0;
        for (int v = 0; v < n2; ++v)
          for (int u = 0; u < n1; ++u)
            acc = madd<T>(F.psi[b][u + 8 * v], X[pmod(2 * i - u, L1) + (size_t)L1 * pmod(2 * j - v, L2)], acc);
        out[b][i + (size_t)h1 * j] = acc;
      }
}

// Circular correlation of the zero-upsampled band D (Lc1 x Lc2, values at the
// even positions of the 2Lc1 x 2Lc2 grid) with Psi_b (myWavRec2DGPUfast.m:16-18, 26-31).
template <class T>
T wav2_corr_up(const std::vector<T>& D, int Lc1, int Lc2, const T* psi, int t1, int t2) {
  const int L1 = 2 * Lc1, L2 = 2 * Lc2, n1 = std::min(8, L1), n2 = std::min(8, L2);
  T acc = 0;
  for (int v = 0; v < n2; ++v) {
    const int c = pmod(t2 + v, L2);
    if (c & 1) continue;
    for (int u = 0; u < n1; ++u) {
      const int r = pmod(t1 + u, L1);
      if (r & 1) continue;
      acc = madd<T>(psi[u + 8 * v], D[r / 2 + (size_t)Lc1 * (c / 2)], acc);
    }
  }
  return acc;
}

// I (m x n) -> the mean over shifts k = 0..4 of circshift(rec(thr(dec(circshift(I,[k k])))), [-k -k])
// (matchBlocksGPU.m:35-49), 3 levels, hard threshold at tau on the details.
template <class T>
std::vector<T> wav2_prefilter(const T* I, int m, int n, double tau) {
  const int levels = 3, shifts = 5;
  Sym4Bank<T> F;
  std::vector<T> Iavg((size_t)m * n, T(0));
  for (int k = 0; k < shifts; ++k) {
    std::vector<std::vector<T>> D(levels * 3);
    std::vector<int> L1s(levels + 1), L2s(levels + 1);
    std::vector<T> A((size_t)m * n);
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < m; ++r) A[r + (size_t)m * c] = I[pmod(r - k, m) + (size_t)m * pmod(c - k, n)];
    int L1 = m, L2 = n;
    for (int l = 0; l < levels; ++l) {
      L1s[l] = L1;
      L2s[l] = L2;
      std::vector<T> out[4];
      wav2_analysis<T>(A, L1, L2, F, out);
      for (int b = 1; b < 4; ++b) {
        for (auto& x : out[b])
          if (!(std::fabs((double)x) > tau)) x = 0;  // C.*(abs(C)>tau) (:44)
        D[l * 3 + b - 1] = std::move(out[b]);
      }
      A = std::move(out[0]);
      L1 /= 2;
      L2 /= 2;
    }
    // Cup{levels+1} = approximation; Cup{l} = sum_j corr(up(C{l,j})) + corr_lo(up(Cup{l+1}))
    std::vector<T> cup = std::move(A);
    for (int l = levels - 1; l >= 0; --l) {
      const int Lc1 = L1s[l] / 2, Lc2 = L2s[l] / 2, Lf1 = L1s[l], Lf2 = L2s[l];
      std::vector<T> nxt((size_t)Lf1 * Lf2);
#pragma omp parallel for schedule(static) if (Lf1 * Lf2 > 4096)
      for (int t2 = 0; t2 < Lf2; ++t2)
        for (int t1 = 0; t1 < Lf1; ++t1) {
          T a = wav2_corr_up<T>(D[l * 3 + 0], Lc1, Lc2, F.psi[1], t1, t2);
          a = a + wav2_corr_up<T>(D[l * 3 + 1], Lc1, Lc2, F.psi[2], t1, t2);
          a = a + wav2_corr_up<T>(D[l * 3 + 2], Lc1, Lc2, F.psi[3], t1, t2);
          a = a + wav2_corr_up<T>(cup, Lc1, Lc2, F.psi[0], t1, t2);
          nxt[t1 + (size_t)Lf1 * t2] = a;
        }
      cup = std::move(nxt);
    }
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < m; ++r) {
        T& o = Iavg[r + (size_t)m * c];
        o = o + cup[pmod(r + k, m) + (size_t)m * pmod(c + k, n)] / (T)shifts;
      }
  }
  return Iavg;
}

// Circular top-left-anchored d x d box sums of I^2 (blockNorms, matchBlocksGPU.m:58-60).
template <class T>
std::vector<T> circ_norm_map(const T* I, int m, int n, int d) {
  std::vector<T> bn((size_t)m * n);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < m; ++r) {
      T acc = 0;
      for (int v = 0; v < d; ++v)
        for (int u = 0; u < d; ++u) {
          const T y = I[pmod(r + u, m) + (size_t)m * pmod(c + v, n)];
          acc = madd<T>(y, y, acc);
        }
      bn[r + (size_t)m * c] = acc;
    }
  return bn;
}

// S (numMax x floor(m/d) x floor(n/d)), 1-based column-major indices into
// m x n; counts (floor(m/d) x floor(n/d)) entries of each tile's list, the
// slots after them 0.  numMax == numMin: the numMax smallest distances
// (matchBlocksGPU.m:69-70).  Otherwise the tolerance branch (:71-81): the
// positions with distance < tol = 0.005 sigma^2 d^2 in index order, or the
// numMax / numMin smallest when there are more / fewer of them.  Comparisons
// between the single distance or coefficient and the double tol / tau are
// exact (MATLAB compares the values, single(0.1) == 0.1 is false).
// filtered (nullable): receives the image the search ran on.
template <class T>
int match_global(const T* I0, int m, int n, double sigma, int d, int numMax, int numMin, uint32_t* S,
                 uint32_t* counts, T* filtered) {
  if (m < d || n < d || d < 1 || numMax < 1 || numMin < 1 || numMin > numMax || (size_t)numMax > (size_t)m * n)
    return GBM3D_ERR_ARG;
  std::vector<T> I(I0, I0 + (size_t)m * n);
  if (sigma > 30.0 / 255.0) {
    if (m % 8 || n % 8) return GBM3D_ERR_ARG;  // myWavRec2DGPUfast.m:27 size mismatch otherwise
    I = wav2_prefilter<T>(I0, m, n, 2.0 * sigma);
  }
  if (filtered) std::copy(I.begin(), I.end(), filtered);
  const double tol = 0.005 * sigma * sigma * d * d;
  const std::vector<T> bn = circ_norm_map<T>(I.data(), m, n, d);
  const int Mt = m / d, Nt = n / d;
  const size_t mn = (size_t)m * n;
#pragma omp parallel
  {
    std::vector<T> dist(mn);
    std::vector<int> idx(mn);
#pragma omp for collapse(2) schedule(dynamic)
    for (int tj = 0; tj < Nt; ++tj)
      for (int ti = 0; ti < Mt; ++ti) {
        const T nx = bn[(size_t)ti * d + (size_t)m * tj * d];
        size_t below = 0;
        for (int pc = 0; pc < n; ++pc)
          for (int pr = 0; pr < m; ++pr) {
            T acc = 0;
            for (int v = 0; v < d; ++v)
              for (int u = 0; u < d; ++u)
                acc = madd<T>(I[(ti * d + u) + (size_t)m * (tj * d + v)],
                              I[pmod(pr + u, m) + (size_t)m * pmod(pc + v, n)], acc);
            const size_t e = pr + (size_t)m * pc;
            dist[e] = ((nx + bn[e]) - T(2) * acc) / nx;
            below += (double)dist[e] < tol;
          }
        for (size_t e = 0; e < mn; ++e) idx[e] = (int)e;
        std::partial_sort(idx.begin(), idx.begin() + numMax, idx.end(),
                          [&](int a, int b) { return dist_less<T>(dist[a], a, dist[b], b); });
        int cnt = numMax;
        if (numMax != numMin) {
          if (below < (size_t)numMin)
            cnt = numMin;
          else if (below <= (size_t)numMax) {
            cnt = (int)below;  // find(): index order
            std::sort(idx.begin(), idx.begin() + cnt);
          }
        }
        uint32_t* out = S + (size_t)numMax * (ti + (size_t)Mt * tj);
        for (int k = 0; k < numMax; ++k) out[k] = k < cnt ? (uint32_t)idx[k] + 1u : 0u;
        if (counts) counts[ti + (size_t)Mt * tj] = (uint32_t)cnt;
      }
  }
  return 0;
}

// ------------------------------------------------------ distributed driver
// getBM3D_distribution (utilities/getBM3D_distribution.m:10-48): one case per
// (tile row, tile column, spin), tiles of d0 = 256 + k with stride d0 - k; a
// tile that reaches the last row/column is moved back so its size is a
// multiple of k (at least 2k).  Rows {y1, y2, x1, x2, spin}, 1-based inclusive.
struct TileCase {
  int y1, y2, x1, x2, spin;
};

inline std::vector<TileCase> distribution(int m, int n, const int32_t* spins, int nspins, int k) {
  const int d0 = 256 + k;
  const int Xcnt = (n - k + (d0 - k) - 1) / (d0 - k), Ycnt = (m - k + (d0 - k) - 1) / (d0 - k);
  std::vector<TileCase> R;
  for (int i = 1; i <= Ycnt; ++i) {
    const int indY = (i - 1) * (d0 - k);
    for (int j = 1; j <= Xcnt; ++j) {
      const int indX = (j - 1) * (d0 - k);
      for (int s = 0; s < nspins; ++s) {
        TileCase t{indY + 1, indY + d0, indX + 1, indX + d0, spins[s]};
        if (t.x2 >= n) t.x1 = std::min(n - (n - t.x1 + 1 + k - 1) / k * k + 1, n - 2 * k + 1);
        if (t.y2 >= m) t.y1 = std::min(m - (m - t.y1 + 1 + k - 1) / k * k + 1, m - 2 * k + 1);
        R.push_back(t);
      }
    }
  }
  for (auto& t : R) {
    t.y2 = std::min(t.y2, m);
    t.x2 = std::min(t.x2, n);
  }
  return R;
}

// makeInterpWindow (utilities/makeInterpWindow.m:8-13): a x b ones with
// linspace(.01, 1, k) ramps applied to the top rows, left columns, bottom rows
// (flipped) and right columns (flipped), in that order; column-major.
inline std::vector<double> interp_window(int a, int b, int k) {
  std::vector<double> w((size_t)a * b, 1.0);
  const std::vector<double> ramp = linspace(0.01, 1.0, k);
  for (int c = 0; c < b; ++c)
    for (int r = 0; r < k && r < a; ++r) w[r + (size_t)a * c] *= ramp[r];
  for (int c = 0; c < k && c < b; ++c)
    for (int r = 0; r < a; ++r) w[r + (size_t)a * c] *= ramp[c];
  for (int c = 0; c < b; ++c)
    for (int r = std::max(0, a - k); r < a; ++r) w[r + (size_t)a * c] *= ramp[a - 1 - r];
  for (int c = std::max(0, b - k); c < b; ++c)
    for (int r = 0; r < a; ++r) w[r + (size_t)a * c] *= ramp[b - 1 - c];
  return w;
}

// [U, out] = GBM3D_distributed(I, sigma, opts) (source/GBM3D_distributed.m:1-122):
// each case's tile (circshifted by its spin) runs matchBlocksCPU +
// denoiseAndAggregateGPU, the unnormalised U and W are shifted back and
// stitched with the tile's ramp window, U2 = sum(U w) / sum(w W); step 2 the
// same with matchBlocksWie on the U2 tile and denoiseAndAggregateWieFast.
// pilot != nullptr: skip step 1 and use it as U2.  Accumulation in double.
template <class T>
int gbm3d_distributed(const T* I, int m, int n, double sigma, const gbm3d_opts& o, T* Uout, T* U2out,
                      const T* pilot) {
  const int k = o.blockSize;
  if (m < 256 + k || n < 256 + k) {  // :22-25
    if (pilot) return GBM3D_ERR_ARG;
    return gbm3d<T>(I, m, n, sigma, o, Uout, U2out, true);
  }
  if (o.filtType != GBM3D_FILT_HT && o.filtType != GBM3D_FILT_MEDIAN) return GBM3D_ERR_UNSUPPORTED;
  if (o.numMax != o.numMin) return GBM3D_ERR_UNSUPPORTED;
  if (o.Wiener && o.nMatchSpinsWie == GBM3D_UNSET) return GBM3D_ERR_MISSING_FIELD;
  const int d2 = o.matchSize, K = o.numMax, bw = o.blockSizeWie;
  std::vector<double> tau((size_t)o.levels * 7);
  get_tau(o.levels, o.tauMode, sigma, tau.data());
  const size_t mn = (size_t)m * n;
  std::vector<double> num(mn), den(mn);
  std::vector<T> U2(mn);
  auto tile = [&](const T* src, const TileCase& t, int th, int tw, std::vector<T>& dst) {
    dst.resize((size_t)th * tw);
    for (int c = 0; c < tw; ++c)
      for (int r = 0; r < th; ++r)
        dst[r + (size_t)th * c] = src[(size_t)(t.y1 - 1 + pmod(r - t.spin, th)) + (size_t)m * (t.x1 - 1 + pmod(c - t.spin, tw))];
  };
  auto stitch = [&](const TileCase& t, int th, int tw, const std::vector<T>& U, const std::vector<T>& W) {
    const std::vector<double> w = interp_window(th, tw, k);
    for (int c = 0; c < tw; ++c)
      for (int r = 0; r < th; ++r) {
        const size_t src = pmod(r + t.spin, th) + (size_t)th * pmod(c + t.spin, tw);
        const size_t dst = (size_t)(t.y1 - 1 + r) + (size_t)m * (t.x1 - 1 + c);
        num[dst] += (double)U[src] * w[r + (size_t)th * c];
        den[dst] += w[r + (size_t)th * c] * (double)W[src];
      }
  };
  if (pilot) {
    std::copy(pilot, pilot + mn, U2.begin());
  } else {
    for (const TileCase& t : distribution(m, n, o.matchSpins, o.nMatchSpins, k)) {  // :53-61
      const int th = t.y2 - t.y1 + 1, tw = t.x2 - t.x1 + 1;
      if (th % (1 << o.levels) || tw % (1 << o.levels)) return GBM3D_ERR_UNSUPPORTED;
      std::vector<T> Il;
      tile(I, t, th, tw, Il);
      const int Mt = th / k, Nt = tw / k;
      std::vector<uint32_t> S((size_t)K * Mt * Nt);
      std::vector<T> Vol((size_t)th * tw * K), Vavg(Vol.size()), U((size_t)th * tw), W(U.size());
      match_step1<T>(Il.data(), th, tw, k, d2, K, S.data(), Vol.data(), nullptr);
      if (o.filtType == GBM3D_FILT_MEDIAN)
        median3<T>(Vol.data(), th, tw, K, Vavg.data());
      else
        ht_filter<T>(Vol.data(), th, tw, K, o.levels, o.cycleSpin, tau.data(), Vavg.data(), nullptr);
      aggregate_ht<T>(Vavg.data(), S.data(), th, tw, k, d2, K, U.data(), W.data(), true);
      stitch(t, th, tw, U, W);  // :65-79
    }
    for (size_t e = 0; e < mn; ++e) U2[e] = (T)(num[e] / den[e]);
  }
  if (U2out) std::copy(U2.begin(), U2.end(), U2out);
  if (!o.Wiener) {  // :81-84
    std::copy(U2.begin(), U2.end(), Uout);
    return 0;
  }
  std::fill(num.begin(), num.end(), 0.0);
  std::fill(den.begin(), den.end(), 0.0);
  for (const TileCase& t : distribution(m, n, o.matchSpinsWie, o.nMatchSpinsWie, k)) {  // :94-104
    const int th = t.y2 - t.y1 + 1, tw = t.x2 - t.x1 + 1;
    if (th % bw || tw % bw) return GBM3D_ERR_UNSUPPORTED;
    std::vector<T> Il, Pl;
    tile(I, t, th, tw, Il);
    tile(U2.data(), t, th, tw, Pl);
    const int Mt = th / bw, Nt = tw / bw;
    const long G = (long)Mt * Nt;
    std::vector<uint32_t> S((size_t)K * G);
    std::vector<T> V1((size_t)bw * bw * K * G), V2(V1.size()), V2f(V1.size()), g2(G), U((size_t)th * tw), W(U.size());
    match_step2<T>(Pl.data(), Il.data(), th, tw, bw, d2, K, S.data(), V1.data(), V2.data(), nullptr);
    wiener<T>(V1.data(), V2.data(), bw, K, G, sigma, V2f.data(), g2.data());
    aggregate_wie<T>(V2f.data(), g2.data(), S.data(), th, tw, bw, d2, K, U.data(), W.data(), true);
    stitch(t, th, tw, U, W);  // :107-120
  }
  for (size_t e = 0; e < mn; ++e) Uout[e] = (T)(num[e] / den[e]);
  return 0;
}

}  // namespace gbm3d_oracle
