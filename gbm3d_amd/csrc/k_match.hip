// k_match.hip -- block matching for gfx950: padded search images, block-norm
// maps, windowed FP32 distances with an exact top-K, and group gathers.
//
// Reference: source/matchBlocksCPU.m:56-111, source/matchBlocksWie.m:38-89.
// Distances are  dist = (|x|^2 + |y|^2 - 2<x,y>) / |x|^2  in FP32 with every
// sum formed as one fused multiply-add chain, column by column (v outer, u
// inner).  |x|^2 is read from the block-norm map at the window centre, which
// uses the same chain, so the centre candidate has distance exactly 0.
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

// ----------------------------------------------------------------- padding
// utilities/myMirrorPadImage.m:3-8 (padded row m+j = source row m-1-j).
__global__ void mirror_pad_kernel(const float* __restrict__ src, int m, int n, int M, int N,
                                  float* __restrict__ dst) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (r >= M) return;
  int sr = r < m ? r : 2 * m - 1 - r;
  int sc = c < n ? c : 2 * n - 1 - c;
  dst[r + (size_t)M * c] = src[sr + (size_t)m * sc];
}

int launch_mirror_pad(const float* src, int m, int n, int M, int N, float* dst, hipStream_t st) {
  if (M < m || N < n || M > 2 * m || N > 2 * n) return fail(GBM3D_ERR_ARG, "mirror pad: size");
  dim3 grid((M + 255) / 256, N);
  hipLaunchKernelGGL(mirror_pad_kernel, grid, dim3(256), 0, st, src, m, n, M, N, dst);
  return check_launch("mirror_pad");
}

// Search image of matchBlocksCPU.m:57-61 for the image circshifted by
// 'spin' (source/GBM3D.m:51): padded row r is image row
// (r < s ? r+s : (r-s) mod M) of the shifted image, s = d2/2.
__device__ __forceinline__ int padmap(int r, int m, int s) { return r < s ? r + s : (r - s) % m; }

__global__ void search_pad_kernel(const float* __restrict__ img, int M, int N, int spin, int d2,
                                  float* __restrict__ pad) {
  const int P = M + d2, s = d2 / 2;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (r >= P) return;
  int sr = pmod_i(padmap(r, M, s) - spin, M);
  int sc = pmod_i(padmap(c, N, s) - spin, N);
  pad[r + (size_t)P * c] = img[sr + (size_t)M * sc];
}

int launch_search_pad(const float* img, int M, int N, int spin, int d2, float* pad, hipStream_t st) {
  if (M < d2 || N < d2 || (d2 & 1)) return fail(GBM3D_ERR_ARG, "search pad: image smaller than matchSize");
  dim3 grid((M + d2 + 255) / 256, N + d2);
  hipLaunchKernelGGL(search_pad_kernel, grid, dim3(256), 0, st, img, M, N, spin, d2, pad);
  return check_launch("search_pad");
}

// ----------------------------------------------------------- block norms
// bn(p,q) = sum_{v<D} sum_{u<D} pad(p+u, q+v)^2 for p < P-D+1, q < Q-D+1
// (blockNormsPad, matchBlocksCPU.m:66).  A thread owns 4 consecutive rows.
template <int D>
__global__ __launch_bounds__(256) void norm_map_kernel(const float* __restrict__ pad, int P, int Q,
                                                       float* __restrict__ bn) {
  constexpr int TR = 64, TC = 16, RB = 4;
  constexpr int LR = TR + D - 1, LC = TC + D - 1, LDT = LR + 1;
  __shared__ float tile[LC * LDT];
  const int Pm = P - D + 1, Qm = Q - D + 1;
  const int r0 = blockIdx.x * TR, c0 = blockIdx.y * TC;
  for (int e = threadIdx.x; e < LR * LC; e += 256) {
    int rr = e % LR, cc = e / LR;
    int gr = min(r0 + rr, P - 1), gc = min(c0 + cc, Q - 1);
    tile[cc * LDT + rr] = pad[gr + (size_t)P * gc];
  }
  __syncthreads();
  const int tr = (threadIdx.x & 15) * RB, tc = threadIdx.x >> 4;
  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.f;
  for (int v = 0; v < D; ++v) {
    const float* col = &tile[(tc + v) * LDT + tr];
#pragma unroll
    for (int t = 0; t < RB + D - 1; ++t) {
      float y = col[t];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const int u = t - i;
        if (u >= 0 && u < D) acc[i] = fmaf(y, y, acc[i]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    int p = r0 + tr + i, q = c0 + tc;
    if (p < Pm && q < Qm) bn[p + (size_t)Pm * q] = acc[i];
  }
}

int launch_norm_map(const float* pad, int P, int Q, int d, float* bn, hipStream_t st) {
  const int Pm = P - d + 1, Qm = Q - d + 1;
  if (Pm < 1 || Qm < 1) return fail(GBM3D_ERR_ARG, "norm map: size");
  dim3 grid((Pm + 63) / 64, (Qm + 15) / 16);
  switch (d) {
    case 4: hipLaunchKernelGGL(norm_map_kernel<4>, grid, dim3(256), 0, st, pad, P, Q, bn); break;
    case 8: hipLaunchKernelGGL(norm_map_kernel<8>, grid, dim3(256), 0, st, pad, P, Q, bn); break;
    case 16: hipLaunchKernelGGL(norm_map_kernel<16>, grid, dim3(256), 0, st, pad, P, Q, bn); break;
    case 32: hipLaunchKernelGGL(norm_map_kernel<32>, grid, dim3(256), 0, st, pad, P, Q, bn); break;
    default: return fail(GBM3D_ERR_UNSUPPORTED, "block size must be 4, 8, 16 or 32");
  }
  return check_launch("norm_map");
}

// ---------------------------------------------------------------- matching
// Candidate keys: ordered float bits of the distance (NaN above +Inf, -0 as
// +0) in the high word, the 0-based column-major window index below, so one
// unsigned compare orders candidates as mink does (ascending, NaN last, ties
// to the lower index).
__device__ __forceinline__ uint64_t cand_key(float dist, int idx) {
  uint32_t b = __float_as_uint(dist);
  uint32_t o;
  if ((b & 0x7fffffffu) > 0x7f800000u) {
    o = 0xffffffffu;
  } else {
    if (b == 0x80000000u) b = 0u;
    o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  }
  return ((uint64_t)o << 32) | (uint32_t)idx;
}

// Exact K smallest of the wave's 64*NL keys; lane k (< K) returns the k-th.
// Fast path: theta = K-th smallest per-lane minimum distance bounds the
// answer from above, so only keys with distance <= theta (at least K of them)
// are compacted and bitonic-sorted across the wave.  More than 64 such keys
// (e.g. large runs of exact ties) take the exact K-round selection instead.
template <int NL>
__device__ uint64_t wave_topk(const uint64_t (&key)[NL], int K, uint64_t* scr, int lane) {
  uint32_t lmin = 0xffffffffu;
#pragma unroll
  for (int i = 0; i < NL; ++i) lmin = min(lmin, (uint32_t)(key[i] >> 32));
  const uint32_t sorted = wave_sort64<uint32_t>(lmin, lane);
  const uint32_t theta = (uint32_t)__shfl((int)sorted, K - 1, 64);
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) cnt += ((uint32_t)(key[i] >> 32) <= theta) ? 1 : 0;
  const int off = wave_excl_scan(cnt, lane);
  const int total = __shfl(off + cnt, 63, 64);
  if (total <= 64) {
    int pos = off;
#pragma unroll
    for (int i = 0; i < NL; ++i)
      if ((uint32_t)(key[i] >> 32) <= theta) scr[pos++] = key[i];
    wave_sync();
    uint64_t v = lane < total ? scr[lane] : ~0ull;
    wave_sync();
    return wave_sort64<uint64_t>(v, lane);
  }
  uint64_t res = ~0ull, prev = 0;
  for (int r = 0; r < K; ++r) {
    uint64_t lm = ~0ull;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const uint64_t k = key[i];
      if ((r == 0 || k > prev) && k < lm) lm = k;
    }
    lm = wave_min_u64(lm);
    if (lane == r) res = lm;
    prev = lm;
  }
  return res;
}

// Search geometry per (block size D, match size D2): each lane scores RA
// consecutive window rows of one window column per chunk, IT chunks.
template <int D, int D2>
struct MatchCfg {
  static constexpr int RA = (D2 == 48) ? 12 : 16;
  static constexpr int NCH = D2 * (D2 / RA);
  static constexpr int IT = NCH / 64;
  static constexpr int NL = IT * RA;
  static constexpr int WR = D2 + D;
  static constexpr int LDW = (WR % 2 == 0) ? WR + 1 : WR;  // odd: conflict-free columns
  static_assert(NCH % 64 == 0, "chunking");
};

// One wave per reference tile, 4 waves per workgroup.  The (D2+D)^2 search
// window is staged in LDS; window index a + D2*b <-> Ipad(i*D + a, j*D + b).
template <int D, int D2>
__global__ __launch_bounds__(256) void match_kernel(const float* __restrict__ pad, const float* __restrict__ bn,
                                                    int M, int N, int K, uint32_t* __restrict__ S,
                                                    float* __restrict__ vol) {
  using C = MatchCfg<D, D2>;
  constexpr int S2 = D2 / 2;
  __shared__ float win[4][C::WR * C::LDW];
  __shared__ uint64_t scr[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int Mt = M / D, Nt = N / D;
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
  if (g >= Mt * Nt) return;
  const int ti = g % Mt, tj = g / Mt;
  const int P = M + D2, Pm = P - D + 1;
  float* W = win[wave];
  const float* src = pad + (size_t)(ti * D) + (size_t)P * (tj * D);
  for (int e = lane; e < C::WR * C::WR; e += 64) {
    const int rr = e % C::WR, cc = e / C::WR;
    W[rr + C::LDW * cc] = src[rr + (size_t)P * cc];
  }
  wave_sync();
  const float nx = bn[(ti * D + S2) + (size_t)Pm * (tj * D + S2)];

  uint64_t key[C::NL];
#pragma unroll
  for (int it = 0; it < C::IT; ++it) {
    const int ch = lane + 64 * it;
    const int b = ch % D2, a0 = (ch / D2) * C::RA;
    float acc[C::RA];
#pragma unroll
    for (int i = 0; i < C::RA; ++i) acc[i] = 0.f;
    for (int v = 0; v < D; ++v) {
      float xs[D];
#pragma unroll
      for (int u = 0; u < D; ++u) xs[u] = W[(S2 + u) + C::LDW * (S2 + v)];
      const float* col = W + a0 + C::LDW * (b + v);
#pragma unroll
      for (int t = 0; t < C::RA + D - 1; ++t) {
        const float y = col[t];
#pragma unroll
        for (int i = 0; i < C::RA; ++i) {
          const int u = t - i;
          if (u >= 0 && u < D) acc[i] = fmaf(xs[u], y, acc[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < C::RA; ++i) {
      const int a = a0 + i;
      const float bnp = bn[(ti * D + a) + (size_t)Pm * (tj * D + b)];
      const float num = (nx + bnp) - 2.0f * acc[i];
      key[it * C::RA + i] = cand_key(num / nx, a + D2 * b);
    }
  }

  const uint64_t best = wave_topk<C::NL>(key, K, scr[wave], lane);
  int idx = (int)(uint32_t)best;
  if (lane == 0) idx = S2 + D2 * S2;  // slot 1 := window centre (matchBlocksCPU.m:103)
  if (lane < K) S[lane + (size_t)K * g] = (uint32_t)idx + 1u;
  if (vol == nullptr) return;
  // Volume((i-1)d+1:id, (j-1)d+1:jd, k) = Itmp(a:a+d-1, b:b+d-1)  (:107-111)
  const size_t MN = (size_t)M * N;
  for (int k = 0; k < K; ++k) {
    const int ik = __builtin_amdgcn_readlane(idx, k);
    const int a = ik % D2, b = ik / D2;
    float* dst = vol + MN * k + (size_t)(ti * D) + (size_t)M * (tj * D);
    for (int e = lane; e < D * D; e += 64) {
      const int u = e % D, v = e / D;
      dst[u + (size_t)M * v] = W[(a + u) + C::LDW * (b + v)];
    }
  }
}

template <int D, int D2>
static void launch_match_t(const float* pad, const float* bn, int M, int N, int K, uint32_t* S, float* vol,
                           hipStream_t st) {
  const int G = (M / D) * (N / D);
  hipLaunchKernelGGL((match_kernel<D, D2>), dim3((G + 3) / 4), dim3(256), 0, st, pad, bn, M, N, K, S, vol);
}

int launch_match(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                 hipStream_t st) {
  if (M % d || N % d || M < d2 || N < d2) return fail(GBM3D_ERR_ARG, "match: image size");
  if (K < 1 || K > 64 || K > d2 * d2) return fail(GBM3D_ERR_UNSUPPORTED, "match: numMax must be in [1, 64]");
#define GBM3D_MATCH_CASE(DD, DD2)                                     \
  if (d == DD && d2 == DD2) {                                         \
    launch_match_t<DD, DD2>(pad, bn, M, N, K, S, vol, st);            \
    return check_launch("match");                                     \
  }
  GBM3D_MATCH_CASE(4, 32)
  GBM3D_MATCH_CASE(8, 32)
  GBM3D_MATCH_CASE(16, 32)
  GBM3D_MATCH_CASE(32, 32)
  GBM3D_MATCH_CASE(8, 48)
  GBM3D_MATCH_CASE(16, 48)
#undef GBM3D_MATCH_CASE
  return fail(GBM3D_ERR_UNSUPPORTED, "match: (blockSize, matchSize) must be (4|8|16|32, 32) or (8|16, 48)");
}

// ------------------------------------------------------------ group gather
// V(:,:,k,i,j) = Ipad(i*d + a_k + (0:d-1), j*d + b_k + (0:d-1))  (matchBlocksWie.m:85-89)
__global__ void gather_groups_kernel(const float* __restrict__ pad, const uint32_t* __restrict__ S, int M, int N,
                                     int d, int d2, int K, float* __restrict__ V) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Mt = M / d, Nt = N / d;
  const size_t total = (size_t)d * d * K * Mt * Nt;
  if (e >= total) return;
  const int u = (int)(e % d), v = (int)((e / d) % d);
  const size_t gk = e / ((size_t)d * d);
  const size_t g = gk / K;
  const int ti = (int)(g % Mt), tj = (int)(g / Mt);
  const uint32_t idx = S[gk] - 1u;
  const int a = (int)(idx % d2), b = (int)(idx / d2);
  const int P = M + d2;
  V[e] = pad[(size_t)(ti * d + a + u) + (size_t)P * (tj * d + b + v)];
}

int launch_gather_groups(const float* pad, const uint32_t* S, int M, int N, int d, int d2, int K, float* V,
                         hipStream_t st) {
  const size_t total = (size_t)d * d * K * (M / d) * (N / d);
  if (total == 0) return GBM3D_OK;
  hipLaunchKernelGGL(gather_groups_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pad, S, M, N,
                     d, d2, K, V);
  return check_launch("gather_groups");
}

}  // namespace gbm3d
