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
template <int D, bool SQUARE = true>
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
        if (u >= 0 && u < D) acc[i] = SQUARE ? fmaf(y, y, acc[i]) : acc[i] + y;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    int p = r0 + tr + i, q = c0 + tc;
    if (p < Pm && q < Qm) bn[p + (size_t)Pm * q] = acc[i];
  }
}

int launch_sum_map(const float* pad, int P, int Q, int d, float* sm, hipStream_t st) {
  const int Pm = P - d + 1, Qm = Q - d + 1;
  if (Pm < 1 || Qm < 1) return fail(GBM3D_ERR_ARG, "sum map: size");
  dim3 grid((Pm + 63) / 64, (Qm + 15) / 16);
  switch (d) {
    case 4: hipLaunchKernelGGL((norm_map_kernel<4, false>), grid, dim3(256), 0, st, pad, P, Q, sm); break;
    case 8: hipLaunchKernelGGL((norm_map_kernel<8, false>), grid, dim3(256), 0, st, pad, P, Q, sm); break;
    case 16: hipLaunchKernelGGL((norm_map_kernel<16, false>), grid, dim3(256), 0, st, pad, P, Q, sm); break;
    case 32: hipLaunchKernelGGL((norm_map_kernel<32, false>), grid, dim3(256), 0, st, pad, P, Q, sm); break;
    default: return fail(GBM3D_ERR_UNSUPPORTED, "block size must be 4, 8, 16 or 32");
  }
  return check_launch("sum_map");
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
template <int D, int D2, int NW = 4>
__global__ __launch_bounds__(64 * NW) void match_kernel(const float* __restrict__ pad, const float* __restrict__ bn,
                                                        int M, int N, int K, uint32_t* __restrict__ S,
                                                        float* __restrict__ vol) {
  using C = MatchCfg<D, D2>;
  constexpr int S2 = D2 / 2;
  __shared__ float win[NW][C::WR * C::LDW];
  __shared__ uint64_t scr[NW][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int Mt = M / D, Nt = N / D;
  const int g = xcd_remap(blockIdx.x, gridDim.x) * NW + wave;
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
static void launch_match_t(const float* pad, const float* bn, int M, int N, int K, uint32_t* S, float* vol, int nw,
                           hipStream_t st) {
  const int G = (M / D) * (N / D);
  if (nw == 1)
    hipLaunchKernelGGL((match_kernel<D, D2, 1>), dim3(G), dim3(64), 0, st, pad, bn, M, N, K, S, vol);
  else if (nw == 2)
    hipLaunchKernelGGL((match_kernel<D, D2, 2>), dim3((G + 1) / 2), dim3(128), 0, st, pad, bn, M, N, K, S, vol);
  else
    hipLaunchKernelGGL((match_kernel<D, D2, 4>), dim3((G + 3) / 4), dim3(256), 0, st, pad, bn, M, N, K, S, vol);
}

// nw: reference tiles (waves) per workgroup, 1, 2 or 4 (a launch-shape choice
// of the batch autotuner; results do not depend on it).
int launch_match_w(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                   int nw, hipStream_t st) {
  if (M % d || N % d || M < d2 || N < d2) return fail(GBM3D_ERR_ARG, "match: image size");
  if (K < 1 || K > 64 || K > d2 * d2) return fail(GBM3D_ERR_UNSUPPORTED, "match: numMax must be in [1, 64]");
  if (nw != 1 && nw != 2 && nw != 4) return fail(GBM3D_ERR_ARG, "match: waves per workgroup must be 1, 2 or 4");
#define GBM3D_MATCH_CASE(DD, DD2)                                     \
  if (d == DD && d2 == DD2) {                                         \
    launch_match_t<DD, DD2>(pad, bn, M, N, K, S, vol, nw, st);        \
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

int launch_match(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                 hipStream_t st) {
  return launch_match_w(pad, bn, M, N, d, d2, K, S, vol, 4, st);
}

// ------------------------------------------------ BF16-MFMA matching
// Precision variant.  Cross terms on BF16 MFMA with FP32 accumulation, then
// the candidates whose approximate score is within the smallest R (R >= 2K)
// are rescored with exactly the FP32 chain of match_kernel, so their keys are
// bit-identical to the FP32 path's; the final top-K is taken over those keys.
//
// corr(a,b) = sum_v sum_r X_v(a,r) Y(r, b+v), X_v(a,r) = x(r-a, v) (a band
// Toeplitz matrix of block column v), Y the search window: per v two distinct
// 16x16 A tiles (r-a in [0,16) and [16,32)); B tiles are four consecutive
// window rows of one column, read as one 8-byte LDS load.  Both operands are
// centred on the reference block's mean c:
//   <x,y> = <x-c, y-c> + c*(sum y + sum x) - D^2 c^2   (sums in FP32).
typedef short bf4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float x) {  // round to nearest even
  uint32_t b = __float_as_uint(x);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((b >> 16) | 0x40);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}

__device__ __forceinline__ uint32_t ord_f32(float d) {
  uint32_t b = __float_as_uint(d);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

template <int D, int D2>
struct Bf16Cfg {
  static constexpr int WR = D2 + D;
  static constexpr int KS = (WR + 15) / 16;
  static constexpr int WRP = KS * 16;       // bf16 window rows, zero-padded
  static constexpr int LDB = WRP + 4;       // bf16 column stride (multiple of 4: 8-byte aligned B loads)
  static constexpr int LDW = (WR % 2 == 0) ? WR + 1 : WR;
  static constexpr int MT = D2 / 16, NT = D2 / 16;
  static constexpr int NC = MT * NT * 4;    // candidates per lane
  static_assert(D2 % 16 == 0, "match size must be a multiple of 16");
};

template <int D, int D2>
__global__ __launch_bounds__(256) void match_bf16_kernel(const float* __restrict__ pad,
                                                         const float* __restrict__ bn,
                                                         const float* __restrict__ smap, int M, int N, int K,
                                                         int R, uint32_t* __restrict__ S, float* __restrict__ vol) {
  using C = Bf16Cfg<D, D2>;
  constexpr int S2 = D2 / 2;
  __shared__ float win[4][C::WR * C::LDW];
  __shared__ __attribute__((aligned(16))) unsigned short wb[4][C::WRP * C::LDB];
  __shared__ uint64_t scr[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int Mt = M / D, Nt = N / D;
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
  if (g >= Mt * Nt) return;
  const int ti = g % Mt, tj = g / Mt;
  const int P = M + D2, Pm = P - D + 1;
  float* W = win[wave];
  unsigned short* Wb = wb[wave];
  const float* src = pad + (size_t)(ti * D) + (size_t)P * (tj * D);
  for (int e = lane; e < C::WR * C::WR; e += 64) {
    const int rr = e % C::WR, cc = e / C::WR;
    W[rr + C::LDW * cc] = src[rr + (size_t)P * cc];
  }
  const size_t cpos = (ti * D + S2) + (size_t)Pm * (tj * D + S2);
  const float nx = bn[cpos];
  const float sx = smap[cpos];
  const float cm = sx / (float)(D * D);
  wave_sync();
  for (int e = lane; e < C::WRP * C::WR; e += 64) {
    const int rr = e % C::WRP, cc = e / C::WRP;
    Wb[rr + C::LDB * cc] = rr < C::WR ? to_bf16(W[rr + C::LDW * cc] - cm) : (unsigned short)0;
  }
  // A tiles: X_v(a, r) for r - a = 16*delta + 4*(lane/16) + i - lane%16
  bf4_t A[D][2];
#pragma unroll
  for (int v = 0; v < D; ++v)
#pragma unroll
    for (int dl = 0; dl < 2; ++dl)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int t = 16 * dl + 4 * (lane >> 4) + i - (lane & 15);
        A[v][dl][i] = (t >= 0 && t < D) ? (short)to_bf16(W[(S2 + t) + C::LDW * (S2 + v)] - cm) : (short)0;
      }
  wave_sync();
  f4_t acc[C::MT][C::NT];
#pragma unroll
  for (int mt = 0; mt < C::MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < C::NT; ++nt) acc[mt][nt] = (f4_t){0.f, 0.f, 0.f, 0.f};
  for (int v = 0; v < D; ++v) {
#pragma unroll
    for (int nt = 0; nt < C::NT; ++nt) {
#pragma unroll
      for (int ks = 0; ks < C::KS; ++ks) {
        const bf4_t B = *reinterpret_cast<const bf4_t*>(&Wb[(16 * ks + 4 * (lane >> 4)) +
                                                            C::LDB * (16 * nt + (lane & 15) + v)]);
#pragma unroll
        for (int mt = 0; mt < C::MT; ++mt) {
          const int dl = ks - mt;
          if (dl == 0 || dl == 1)
            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[v][dl], B, acc[mt][nt], 0, 0, 0);
        }
      }
    }
  }
  // approximate scores num = |y|^2 - 2<x,y> (|x|^2 is common to the tile)
  const float ccorr = cm * sx - (float)(D * D) * cm * cm;
  uint32_t key[C::NC];
#pragma unroll
  for (int mt = 0; mt < C::MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < C::NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int a = 16 * mt + 4 * (lane >> 4) + i, b = 16 * nt + (lane & 15);
        const size_t o = (ti * D + a) + (size_t)Pm * (tj * D + b);
        const float corr = acc[mt][nt][i] + cm * smap[o] + ccorr;
        key[(mt * C::NT + nt) * 4 + i] = ord_f32(bn[o] - 2.0f * corr);
      }
  // candidates to rescore: every key <= theta, theta = R-th smallest lane minimum
  uint32_t lmin = 0xffffffffu;
#pragma unroll
  for (int j = 0; j < C::NC; ++j) lmin = min(lmin, key[j]);
  const uint32_t theta = (uint32_t)__shfl((int)wave_sort64<uint32_t>(lmin, lane), R - 1, 64);
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < C::NC; ++j) cnt += key[j] <= theta ? 1 : 0;
  const int off = wave_excl_scan(cnt, lane);
  const int total = __shfl(off + cnt, 63, 64);
  auto cand_idx = [&](int j) {
    const int mt = j / (4 * C::NT), nt = (j / 4) % C::NT, i = j % 4;
    return (16 * mt + 4 * (lane >> 4) + i) + D2 * (16 * nt + (lane & 15));
  };
  int mine = -1;  // window index this lane rescores
  if (total <= 64) {
    int pos = off;
#pragma unroll
    for (int j = 0; j < C::NC; ++j)
      if (key[j] <= theta) scr[wave][pos++] = (uint64_t)cand_idx(j);
    wave_sync();
    if (lane < total) mine = (int)scr[wave][lane];
    wave_sync();
  } else {
    // many equal scores: the R smallest (score, index) pairs, one round each
    uint64_t prev = 0;
    for (int r = 0; r < R; ++r) {
      uint64_t lm = ~0ull;
#pragma unroll
      for (int j = 0; j < C::NC; ++j) {
        const uint64_t kk = ((uint64_t)key[j] << 32) | (uint32_t)cand_idx(j);
        if ((r == 0 || kk > prev) && kk < lm) lm = kk;
      }
      lm = wave_min_u64(lm);
      if (lane == r) mine = (int)(uint32_t)lm;
      prev = lm;
    }
  }
  // exact FP32 rescoring (the FMA chain of match_kernel)
  uint64_t ek = ~0ull;
  if (mine >= 0) {
    const int a = mine % D2, b = mine / D2;
    float cr = 0.f;
    for (int v = 0; v < D; ++v)
#pragma unroll
      for (int u = 0; u < D; ++u)
        cr = fmaf(W[(S2 + u) + C::LDW * (S2 + v)], W[(a + u) + C::LDW * (b + v)], cr);
    const float bnp = bn[(ti * D + a) + (size_t)Pm * (tj * D + b)];
    const float num = (nx + bnp) - 2.0f * cr;
    ek = cand_key(num / nx, mine);
  }
  const uint64_t best = wave_sort64<uint64_t>(ek, lane);
  int idx = (int)(uint32_t)best;
  if (lane == 0) idx = S2 + D2 * S2;  // slot 1 := window centre (matchBlocksCPU.m:103)
  if (lane < K) S[lane + (size_t)K * g] = (uint32_t)idx + 1u;
  if (vol == nullptr) return;
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

int launch_match_bf16(const float* pad, const float* bn, const float* smap, int M, int N, int d, int d2, int K,
                      int R, uint32_t* S, float* vol, hipStream_t st) {
  if (M % d || N % d || M < d2 || N < d2) return fail(GBM3D_ERR_ARG, "match: image size");
  if (K < 1 || K > 64) return fail(GBM3D_ERR_UNSUPPORTED, "match: numMax must be in [1, 64]");
  if (R < K) R = K;
  if (R > 64) R = 64;
  const int G = (M / d) * (N / d);
  const dim3 grid((G + 3) / 4);
#define GBM3D_BF16_CASE(DD, DD2)                                                                              \
  if (d == DD && d2 == DD2) {                                                                                 \
    hipLaunchKernelGGL((match_bf16_kernel<DD, DD2>), grid, dim3(256), 0, st, pad, bn, smap, M, N, K, R, S, vol); \
    return check_launch("match_bf16");                                                                        \
  }
  GBM3D_BF16_CASE(8, 32)
  GBM3D_BF16_CASE(16, 32)
  GBM3D_BF16_CASE(8, 48)
  GBM3D_BF16_CASE(16, 48)
#undef GBM3D_BF16_CASE
  return fail(GBM3D_ERR_UNSUPPORTED, "BF16 matching: (blockSize, matchSize) must be (8|16, 32|48)");
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
