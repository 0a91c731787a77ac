// k_match.hip -- block matching for gfx950: padded search images, block-norm
// maps, windowed FP32 distances with an exact top-K, and group gathers.
//
// Reference: source/matchBlocksCPU.m:56-111, source/matchBlocksWie.m:38-89.
// Distances are  dist = (|x|^2 + |y|^2 - 2<x,y>) / |x|^2  in FP32 with every
// sum formed as one fused multiply-add chain, column by column (v outer, u
// inner).  |x|^2 is read from the block-norm map at the window centre, which
// uses the same chain, so the centre candidate has distance exactly 0.
#include "internal.hpp"
#include "select.hpp"
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
// (r < s ? r+s : (r-s) mod L) of the shifted image, s = d2/2, that is row
// (that - spin) mod L of the image; for r in [0, L + 2s), s <= L and spin in
// [0, L) one conditional subtraction or addition does each reduction.
__device__ __forceinline__ int padmap_spin(int r, int L, int s, int spin) {
  int x = r < s ? r + s : r - s;  // in [0, L + s)
  if (x >= L) x -= L;
  x -= spin;
  return x < 0 ? x + L : x;
}

// spinR = spin mod M, spinC = spin mod N.  A thread fills 4 consecutive rows
// of one padded column (one 16-byte store when the column allows it).
__global__ void search_pad_kernel(const float* __restrict__ img, int M, int N, int spinR, int spinC, int d2,
                                  float* __restrict__ pad) {
  const int P = M + d2, s = d2 / 2;
  const int r0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int c = blockIdx.y;
  if (r0 >= P) return;
  const float* src = img + (size_t)M * padmap_spin(c, N, s, spinC);
  float* dst = pad + (size_t)P * c + r0;
  if ((P & 3) == 0 && (reinterpret_cast<uintptr_t>(pad) & 15) == 0) {  // r0 + 3 < P, 16-byte aligned
    float4 v;
    v.x = src[padmap_spin(r0, M, s, spinR)];
    v.y = src[padmap_spin(r0 + 1, M, s, spinR)];
    v.z = src[padmap_spin(r0 + 2, M, s, spinR)];
    v.w = src[padmap_spin(r0 + 3, M, s, spinR)];
    *reinterpret_cast<float4*>(dst) = v;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (r0 + k < P) dst[k] = src[padmap_spin(r0 + k, M, s, spinR)];
  }
}

int launch_search_pad(const float* img, int M, int N, int spin, int d2, float* pad, hipStream_t st) {
  if (M < d2 || N < d2 || (d2 & 1)) return fail(GBM3D_ERR_ARG, "search pad: image smaller than matchSize");
  const int spinR = ((spin % M) + M) % M, spinC = ((spin % N) + N) % N;
  dim3 grid((M + d2 + 1023) / 1024, N + d2);
  hipLaunchKernelGGL(search_pad_kernel, grid, dim3(256), 0, st, img, M, N, spinR, spinC, d2, pad);
  return check_launch("search_pad");
}

// ----------------------------------------------------------- block norms
// bn(p,q) = sum_{v<D} sum_{u<D} pad(p+u, q+v)^2 for p < P-D+1, q < Q-D+1
// (blockNormsPad, matchBlocksCPU.m:66).  A thread owns 4 consecutive rows.
template <int D, bool SQUARE = true>
__global__ __launch_bounds__(256) void norm_map_kernel(const float* __restrict__ pad, int P, int Q,
                                                       float* __restrict__ bn) {
  constexpr int TR = 64, TC = 16, RB = 4;
  // LDT odd with LDT % 4 != 0: the four tile columns a wave reads at once
  // fall in distinct bank residues (lanes of a column are 4 rows apart)
  constexpr int LR = TR + D - 1, LC = TC + D - 1, LDT = LR + 2;
  static_assert(LDT % 2 == 1, "tile stride");
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
  // (block columns in a loop: the unrolled form holds every column's loads
  // in flight, 92 VGPRs and 5 waves per SIMD)
#pragma unroll 1
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

// The search image and its block norms in one pass: the norm tile's window
// comes from the image through the search-pad map (the values search_pad
// writes), each workgroup writes the pad elements of its tile's rows and
// columns (the last tile row and column also the D - 1 beyond them), and the
// norms are norm_map_kernel's arithmetic on the same window.
template <int D>
__global__ __launch_bounds__(256) void search_pad_norm_kernel(const float* __restrict__ img, int M, int N, int spinR,
                                                              int spinC, int d2, float* __restrict__ pad,
                                                              float* __restrict__ bn) {
  constexpr int TR = 64, TC = 16, RB = 4;
  constexpr int LR = TR + D - 1, LC = TC + D - 1, LDT = LR + 2;
  static_assert(LDT % 2 == 1, "tile stride");
  __shared__ float tile[LC * LDT];
  const int P = M + d2, Q = N + d2, s = d2 / 2;
  const int Pm = P - D + 1, Qm = Q - D + 1;
  const int r0 = blockIdx.x * TR, c0 = blockIdx.y * TC;
  for (int e = threadIdx.x; e < LR * LC; e += 256) {
    const int rr = e % LR, cc = e / LR;
    const int gr = min(r0 + rr, P - 1), gc = min(c0 + cc, Q - 1);
    tile[cc * LDT + rr] = img[padmap_spin(gr, M, s, spinR) + (size_t)M * padmap_spin(gc, N, s, spinC)];
  }
  __syncthreads();
  {
    const int nr = (r0 + TR >= Pm ? P : r0 + TR) - r0, nc = (c0 + TC >= Qm ? Q : c0 + TC) - c0;
    for (int e = threadIdx.x; e < nr * nc; e += 256) {
      const int rr = e % nr, cc = e / nr;
      pad[(r0 + rr) + (size_t)P * (c0 + cc)] = tile[cc * LDT + rr];
    }
  }
  const int tr = (threadIdx.x & 15) * RB, tc = threadIdx.x >> 4;
  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.f;
#pragma unroll 1
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

int launch_search_pad_norm(const float* img, int M, int N, int spin, int d2, int d, float* pad, float* bn,
                           hipStream_t st) {
  if (M < d2 || N < d2 || (d2 & 1)) return fail(GBM3D_ERR_ARG, "search pad: image smaller than matchSize");
  const int P = M + d2, Q = N + d2, Pm = P - d + 1, Qm = Q - d + 1;
  if (Pm < 1 || Qm < 1) return fail(GBM3D_ERR_ARG, "norm map: size");
  const int spinR = ((spin % M) + M) % M, spinC = ((spin % N) + N) % N;
  dim3 grid((Pm + 63) / 64, (Qm + 15) / 16);
  switch (d) {
    case 4: hipLaunchKernelGGL(search_pad_norm_kernel<4>, grid, dim3(256), 0, st, img, M, N, spinR, spinC, d2, pad, bn); break;
    case 8: hipLaunchKernelGGL(search_pad_norm_kernel<8>, grid, dim3(256), 0, st, img, M, N, spinR, spinC, d2, pad, bn); break;
    case 16: hipLaunchKernelGGL(search_pad_norm_kernel<16>, grid, dim3(256), 0, st, img, M, N, spinR, spinC, d2, pad, bn); break;
    case 32: hipLaunchKernelGGL(search_pad_norm_kernel<32>, grid, dim3(256), 0, st, img, M, N, spinR, spinC, d2, pad, bn); break;
    default: return fail(GBM3D_ERR_UNSUPPORTED, "block size must be 4, 8, 16 or 32");
  }
  return check_launch("search_pad_norm");
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

// Largest compacted candidate set the exact selection takes before falling
// back to K rounds over every candidate.
constexpr int kSelMax = 256;

// One wave per reference tile; a workgroup's NW waves take NW tiles adjacent
// along the tile rows (same tile column), whose (D2+D)^2 search windows are
// staged in LDS once as their union, (D2+D+(NW-1)D) x (D2+D); wave w's window
// starts w*D rows down.  Window index a + D2*b <-> Ipad(i*D + a, j*D + b).
// Occupancy hint: 8 waves per SIMD (64 VGPRs) where the LDS footprint admits
// it (160 KB per CU), otherwise none, leaving the register budget to the
// compiler for the large windows.
template <int D, int D2, int NW>
struct MatchOcc {
  static constexpr int RW = MatchCfg<D, D2>::WR + (NW - 1) * D;
  static constexpr int LDW = (RW % 2 == 0) ? RW + 1 : RW;
  static constexpr int kLds = MatchCfg<D, D2>::WR * LDW * 4 + NW * kSelMax * 8;
  static constexpr int kWaves = (163840 / kLds) * NW / 4 >= 8 ? 8 : 1;
};

// The exact top K of one reference tile over its D2 x D2 window candidates:
// W the tile's window in LDS (candidate (a, b) at W[a + LDW b], the reference
// block at (D2/2, D2/2)), bn the block-norm map (row stride Pm) with the
// tile's window origin at (ti D, tj D), nx = |x|^2, scr kSelMax keys of LDS
// for this wave.  Lane k (< K) returns the k-th smallest key.  Every
// candidate is scored with the FMA chain described at match_kernel.
template <int D, int D2, int LDW>
__device__ __forceinline__ uint64_t exact_search(const float* W, const float* __restrict__ bn, int Pm, int ti, int tj,
                                                 float nx, int K, uint64_t* scr, int lane, float tolu = 0.f,
                                                 uint32_t* below = nullptr) {
  using C = MatchCfg<D, D2>;
  constexpr int S2 = D2 / 2;
  float numv[C::NL];  // |x|^2 + |y|^2 - 2<x,y> of this lane's candidates
#pragma unroll
  for (int it = 0; it < C::IT; ++it) {
    const int ch = lane + 64 * it;
    const int b = ch % D2, a0 = (ch / D2) * C::RA;
    // acc[i] = sum over (v, u) in that order of fmaf(x(u, v), y, acc[i]) from
    // +0.  With the v loop unrolled (8x8 blocks) the first term is a product:
    // x*y differs from fmaf(x, y, +0) only in the sign of a zero, and the
    // numerator (nx + |y|^2) - 2 acc, nx + |y|^2 >= +0, is the same either way.
    float acc[C::RA];
    constexpr bool kFlat = D <= 8;
    constexpr int kUnrollV = kFlat ? D : 1;
    if (!kFlat) {
#pragma unroll
      for (int i = 0; i < C::RA; ++i) acc[i] = 0.f;
    }
#pragma unroll kUnrollV
    for (int v = 0; v < D; ++v) {
      float xs[D];
#pragma unroll
      for (int u = 0; u < D; ++u) xs[u] = W[(S2 + u) + LDW * (S2 + v)];
      const float* col = W + a0 + LDW * (b + v);
#pragma unroll
      for (int t = 0; t < C::RA + D - 1; ++t) {
        const float y = col[t];
#pragma unroll
        for (int i = 0; i < C::RA; ++i) {
          const int u = t - i;
          if (u >= 0 && u < D) acc[i] = (kFlat && v == 0 && u == 0) ? xs[0] * y : fmaf(xs[u], y, acc[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < C::RA; ++i) {
      const int a = a0 + i;
      const float bnp = bn[(ti * D + a) + (size_t)Pm * (tj * D + b)];
      numv[it * C::RA + i] = (nx + bnp) - 2.0f * acc[i];
    }
  }
  auto cand_idx = [&](int j) {
    const int ch = lane + 64 * (j / C::RA);
    return (ch / D2) * C::RA + j % C::RA + D2 * (ch % D2);
  };
  if (below) {
    // candidates with distance num / |x|^2 < tolu (the same for every lane):
    // num against tolu |x|^2 where that decides (2^-20 of room for the
    // roundings of the quotient and the product), the quotient otherwise
    const float tnx = tolu * nx;
    const float tz = fmaf(fabsf(tnx), 0x1p-20f, nx * 0x1p-126f);
    const bool fast = nx > 0.f && nx <= 3.4028235e38f && fabsf(tnx) <= 3.4028235e38f;
    uint32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < C::NL; ++j) {
      bool b;
      if (fast && numv[j] < tnx - tz)
        b = true;
      else if (fast && !(numv[j] <= tnx + tz))
        b = false;
      else
        b = numv[j] / nx < tolu;
      cnt += (uint32_t)__popcll(__ballot(b));
    }
    *below = cnt;
  }

  // The distance is num / |x|^2 with one |x|^2 per tile.  For a finite positive
  // |x|^2 the rounded quotient is non-decreasing in num, so with theta = the
  // K-th smallest lane minimum of num (at least K candidates have num <= theta),
  // every candidate whose quotient can reach the K smallest has
  // num <= theta + |theta| 2^-23 (the rounding of the quotient) plus a term for
  // quotients below the normal range; those (usually < 64) are divided, keyed
  // exactly like the FP32 restatement's candidates and sorted.
  uint64_t best = 0;
  bool done = false;
  if (nx > 0.f && nx <= 3.4028235e38f) {
    // ord_bits is monotone (NaN last, -0 as +0), so the lane minimum of the
    // ordered bits is the ordered bits of the NaN-ignoring float minimum
    float fmn = numv[0];
#pragma unroll
    for (int j = 1; j < C::NL; ++j) fmn = fminf(fmn, numv[j]);
    const uint32_t lmin = ord_bits(fmn);
    const uint32_t th = wave_kth_smallest(lmin, K);
    if (th != 0xffffffffu) {
      const float tnum = ord_to_f32(th);
      const float bound = fmaf(fabsf(tnum), 0x1p-20f, tnum) + nx * 0x1p-126f;
      // the survivors of each candidate slot j as a lane mask; they are
      // compacted slot by slot (the selection below depends on the set alone)
      uint64_t mk[C::NL];
      int total = 0;
#pragma unroll
      for (int j = 0; j < C::NL; ++j) {
        mk[j] = __ballot(numv[j] <= bound);
        total += __popcll(mk[j]);
      }
      if (total <= kSelMax) {
        uint32_t base = 0;
#pragma unroll
        for (int j = 0; j < C::NL; ++j) {
          if (numv[j] <= bound) {
            const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(mk[j] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk[j], 0u));
            scr[base + r] = ((uint64_t)__float_as_uint(numv[j]) << 32) | (uint32_t)cand_idx(j);
          }
          base += (uint32_t)__popcll(mk[j]);
        }
        wave_sync();
        auto key_at = [&](int q) {
          const uint64_t e = scr[q];
          return cand_key(__uint_as_float((uint32_t)(e >> 32)) / nx, (int)(uint32_t)e);
        };
        if (total <= 64) {
          const uint64_t ek = lane < total ? key_at(lane) : ~0ull;
          wave_sync();
          // at most 32 keys (most tiles): the lower half's sort is the answer
          best = (total <= 32 && K <= 32) ? wave_sort64<uint64_t, 32>(ek, lane) : wave_sort64<uint64_t>(ek, lane);
        } else {
          // up to kSelMax keys (best matches crowded into few lanes): each
          // lane sorts its four, then K rounds take the wave minimum of the
          // lanes' heads (keys are distinct: the owner alone pops)
          uint64_t h[kSelMax / 64];
#pragma unroll
          for (int r = 0; r < kSelMax / 64; ++r) h[r] = lane + 64 * r < total ? key_at(lane + 64 * r) : ~0ull;
          wave_sync();
#pragma unroll
          for (int a = 0; a < kSelMax / 64; ++a)
#pragma unroll
            for (int b = kSelMax / 64 - 1; b > a; --b)
              if (h[b] < h[b - 1]) {
                const uint64_t tmp = h[b];
                h[b] = h[b - 1];
                h[b - 1] = tmp;
              }
          uint64_t res = ~0ull;
          for (int r = 0; r < K; ++r) {
            const uint64_t m = wave_min_u64(h[0]);
            if (lane == r) res = m;
            if (h[0] == m) {
#pragma unroll
              for (int a = 0; a + 1 < kSelMax / 64; ++a) h[a] = h[a + 1];
              h[kSelMax / 64 - 1] = ~0ull;
            }
          }
          best = res;
        }
        done = true;
      }
    }
  }
  if (!done) {
    uint64_t key[C::NL];
#pragma unroll
    for (int j = 0; j < C::NL; ++j) key[j] = cand_key(numv[j] / nx, cand_idx(j));
    best = wave_topk<C::NL>(key, K, scr, lane);
  }
  return best;
}

template <int D, int D2, int NW = 4>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(MatchOcc<D, D2, NW>::kWaves))) void match_kernel(const float* __restrict__ pad, const float* __restrict__ bn,
                                                        int M, int N, int K, uint32_t* __restrict__ S,
                                                        float* __restrict__ vol, TolArgs ta) {
  using C = MatchCfg<D, D2>;
  constexpr int S2 = D2 / 2;
  constexpr int RW = C::WR + (NW - 1) * D;          // union rows
  constexpr int LDW = (RW % 2 == 0) ? RW + 1 : RW;  // odd: conflict-free columns
  __shared__ float win[C::WR * LDW];
  __shared__ uint64_t scr[NW][kSelMax];
  // the wave index in a scalar register: the tile's addresses are wave-uniform
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int Mt = M / D, Mb = (Mt + NW - 1) / NW;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ti0 = (blk % Mb) * NW, tj = blk / Mb;
  const int nval = min(NW, Mt - ti0);
  const int P = M + D2, Pm = P - D + 1;
  const int rows = C::WR + (nval - 1) * D;  // union rows inside the padded image
  const float* src = pad + (size_t)(ti0 * D) + (size_t)P * (tj * D);
  if constexpr ((64 * NW) % RW == 0) {
    // each thread keeps one union row and steps over the columns: 32-bit
    // offsets from the window's origin, LDS offsets immediate
    // (rows past the union, masked below, may lie past the padded image: the
    // resource ends with the image, so such loads return 0 and stay inside it)
    constexpr int CPR = 64 * NW / RW, NR = C::WR / CPR;
    static_assert(C::WR % CPR == 0, "window columns per round");
    const int rr = threadIdx.x % RW, c0 = threadIdx.x / RW;
    const size_t left = (size_t)P * (N + D2) - (size_t)(src - pad);
    const uint32_t span = 4u * (uint32_t)(RW + P * (C::WR - 1));
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(src, left < span / 4 ? 4u * (uint32_t)left : span);
    const uint32_t off = 4u * (uint32_t)(rr + P * c0);
    float x[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) x[i] = buf_load(rs, off, 4u * (uint32_t)(P * CPR * i));
#pragma unroll
    for (int i = 0; i < NR; ++i) win[rr + LDW * (c0 + CPR * i)] = rr < rows ? x[i] : 0.f;
  } else {
    for (int e = threadIdx.x; e < RW * C::WR; e += 64 * NW) {
      const int rr = e % RW, cc = e / RW;
      win[rr + LDW * cc] = rr < rows ? src[rr + (size_t)P * cc] : 0.f;
    }
  }
  if (NW > 1)
    __syncthreads();
  else
    wave_sync();
  if (wave >= nval) return;  // no barrier below
  const int ti = ti0 + wave, g = ti + Mt * tj;
  const float* W = win + D * wave;
  const float nx = bn[(ti * D + S2) + (size_t)Pm * (tj * D + S2)];

  uint32_t below = 0;
  const uint64_t best =
      exact_search<D, D2, LDW>(W, bn, Pm, ti, tj, nx, K, scr[wave], lane, ta.tolu, ta.Kmin ? &below : nullptr);
  int idx = (int)(uint32_t)best;
  if (ta.Kmin) {
    // tolerance branch (matchBlocksCPU.m:89-99): more than K below tol keeps
    // the K nearest; exactly K, those K in index order (find); fewer, the
    // reference fails (its list is shorter than numMax, :111)
    if (below < (uint32_t)K) {
      if (lane == 0) ta.flag[0] = 1;
    } else if (below == (uint32_t)K) {
      idx = (int)wave_sort64<uint32_t>(lane < K ? (uint32_t)idx : 0xffffffffu, lane);
    }
  }
  if (lane == 0) idx = S2 + D2 * S2;  // slot 1 := window centre (matchBlocksCPU.m:103)
  if (lane < K) S[lane + (size_t)K * g] = (uint32_t)idx + 1u;
  if (vol == nullptr) return;
  // Volume((i-1)d+1:id, (j-1)d+1:jd, k) = Itmp(a:a+d-1, b:b+d-1)  (:107-111)
  // (the block's (D-1)(M+1)+1 floats of slice k behind a buffer resource: one
  // 32-bit lane offset for all k)
  const size_t MN = (size_t)M * N;
  const uint32_t blk_bytes = 4u * (uint32_t)((D - 1) * (M + 1) + 1);
  for (int k = 0; k < K; ++k) {
    const int ik = __builtin_amdgcn_readlane(idx, k);
    const int a = ik % D2, b = ik / D2;
    const float* Wk = W + a + LDW * b;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(vol + MN * k + (size_t)(ti * D) + (size_t)M * (tj * D), blk_bytes);
    for (int e = lane; e < D * D; e += 64) {
      const int u = e % D, v = e / D;
      buf_store(rs, 4u * (uint32_t)(u + M * v), Wk[u + LDW * v]);
    }
  }
}

template <int D, int D2>
static void launch_match_t(const float* pad, const float* bn, int M, int N, int K, uint32_t* S, float* vol, int nw,
                           hipStream_t st, TolArgs ta) {
  const int Mt = M / D, Nt = N / D;
  const unsigned blocks = (unsigned)(((Mt + nw - 1) / nw) * Nt);
  if (nw == 1)
    hipLaunchKernelGGL((match_kernel<D, D2, 1>), dim3(blocks), dim3(64), 0, st, pad, bn, M, N, K, S, vol, ta);
  else if (nw == 2)
    hipLaunchKernelGGL((match_kernel<D, D2, 2>), dim3(blocks), dim3(128), 0, st, pad, bn, M, N, K, S, vol, ta);
  else
    hipLaunchKernelGGL((match_kernel<D, D2, 4>), dim3(blocks), dim3(256), 0, st, pad, bn, M, N, K, S, vol, ta);
}

// nw: reference tiles (waves) per workgroup, 1, 2 or 4 (a launch-shape choice
// of the batch autotuner; results do not depend on it).
int launch_match_w(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                   int nw, hipStream_t st, const TolArgs* tap) {
  if (M % d || N % d || M < d2 || N < d2) return fail(GBM3D_ERR_ARG, "match: image size");
  const TolArgs ta = tap ? *tap : TolArgs{0, 0.f, nullptr};
  if (ta.Kmin && (ta.Kmin > K || !ta.flag)) return fail(GBM3D_ERR_ARG, "match: tolerance arguments");
  if (K < 1 || K > 64 || K > d2 * d2) return fail(GBM3D_ERR_UNSUPPORTED, "match: numMax must be in [1, 64]");
  if (nw != 1 && nw != 2 && nw != 4) return fail(GBM3D_ERR_ARG, "match: waves per workgroup must be 1, 2 or 4");
#define GBM3D_MATCH_CASE(DD, DD2)                                     \
  if (d == DD && d2 == DD2) {                                         \
    launch_match_t<DD, DD2>(pad, bn, M, N, K, S, vol, nw, st, ta);    \
    return check_launch("match");                                     \
  }
  GBM3D_MATCH_CASE(4, 32)
  GBM3D_MATCH_CASE(8, 32)
  GBM3D_MATCH_CASE(16, 32)
  GBM3D_MATCH_CASE(32, 32)
  GBM3D_MATCH_CASE(4, 48)
  GBM3D_MATCH_CASE(8, 48)
  GBM3D_MATCH_CASE(16, 48)
  GBM3D_MATCH_CASE(32, 48)
#undef GBM3D_MATCH_CASE
  return fail(GBM3D_ERR_UNSUPPORTED, "match: (blockSize, matchSize) must be (4|8|16|32, 32|48)");
}

__global__ void set_i32_kernel(int* p, int v) {
  if (threadIdx.x == 0) p[0] = v;
}

int launch_set_i32(int* p, int v, hipStream_t st) {
  hipLaunchKernelGGL(set_i32_kernel, dim3(1), dim3(64), 0, st, p, v);
  return check_launch("set_i32");
}

bool match_supported(int d, int d2) { return (d == 4 || d == 8 || d == 16 || d == 32) && (d2 == 32 || d2 == 48); }
bool match_bf16_supported(int d, int d2) { return (d == 8 || d == 16) && (d2 == 32 || d2 == 48); }

int launch_match(const float* pad, const float* bn, int M, int N, int d, int d2, int K, uint32_t* S, float* vol,
                 hipStream_t st) {
  return launch_match_w(pad, bn, M, N, d, d2, K, S, vol, 4, st);
}

// ------------------------------------------------ BF16-MFMA matching
// Precision variant.  Cross terms on BF16 MFMA with FP32 accumulation give an
// approximate score and, per candidate, a lower bound of its exact FP32
// numerator up to a tile constant (error bounds below).  The R candidates
// with the smallest bounds (R >= K; every candidate at or below the R-th
// smallest lane minimum) are rescored with exactly the FP32 chain of
// match_kernel; their K-th exact key then excludes every candidate whose
// bound proves a strictly larger rounded distance, and any candidate not
// excluded is rescored as well (or, past 256 candidates, all are).  The
// final top K is taken over exact keys of a set that provably holds the FP32
// path's top K, so S equals the FP32 path's bit for bit, ties included.
//
// corr(a,b) = sum_v sum_r X_v(a,r) Y(r, b+v), X_v(a,r) = x(r-a, v): per block
// column v a band Toeplitz A operand (one fragment serves every 16-row M tile,
// since it depends on r - a only), and per (M tile, N tile) one
// v_mfma_f32_16x16x32_bf16 whose K = 32 window rows cover the band (r - a in
// [0, D), D <= 16).  B fragments are 8 consecutive window rows of one window
// column: one 16-byte LDS load.
//
// A workgroup holds NWT reference tiles adjacent along the tile columns; they
// share one staged search window (FP32 for the rescoring and the volume, BF16
// for the MFMA) and the norm / sum maps of their candidates.  The BF16 window
// is centred on the mean c of the workgroup's reference blocks, which keeps
// BF16's relative precision on the differences that rank the candidates:
//   <x,y> = <x-c, y-c> + c*(sum y) + (terms common to the tile),
// so the approximate score is |y|^2 - 2(<x-c,y-c> + c sum y).
// A reference block of zero (or non-finite) norm takes the exact FP32 search
// over all candidates.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float x) {  // round to nearest even
  uint32_t b = __float_as_uint(x);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((b >> 16) | 0x40);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}

// f = hi + lo + r with hi = bf16(f), lo = bf16(f - hi) (f - hi is exact in
// FP32), |r| <= 2^-16 |f|.
__device__ __forceinline__ void split_bf16(float f, unsigned short* hi, unsigned short* lo) {
  const unsigned short h = to_bf16(f);
  const float hf = __uint_as_float((uint32_t)h << 16);
  *hi = h;
  *lo = to_bf16(f - hf);
}

__device__ __forceinline__ uint32_t ord_f32(float d) {
  uint32_t b = __float_as_uint(d);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

constexpr int kRescoreMax = 128;  // rescored candidates per tile before the FP32 search over all

// ---- error bounds of the approximate scores
// Reals: x, y the reference and candidate blocks (FP32 values), c the FP32
// window centring constant, cx the FP32 centring constant of x, bn(y) and
// Sm(y) the FP32 norm and sum maps, nx = bn(x).
// The FP32 path ranks exnum(y) = fl(fl(nx + bn) - 2 cr), cr the FMA chain of
// <x, y> (D^2 roundings from 0); with N(y) = nx + bn(y) - 2<x,y> (nx, bn taken
// as exact), |exnum - N| <= E1 = (D^2 + 4) u (nx + bn), u = 2^-24.
// Identity (<x,y> = <x-cx, y-c> + c sum(x) + cx sum(y) - cx c D^2):
//   N(y) = T + Z(y),  Z(y) = bn(y) - 2<x-cx, y-c> - 2 cx sum(y),
// with the tile constant T = nx - 2c sum(x) + 2 cx c D^2.
// The centred values f = fl(v - c) (fl(v - cx) for x) are split
// f = hi + lo + r, hi = bf16(f), lo = bf16(f - hi) (f - hi is exact),
// |r| <= 2^-16 |f|, and the MFMA cross term is acc = sum (ha hb + ha lb + la hb)
// with FP32 accumulation (three MFMAs per block column; the dropped la lb and
// the r terms are at most 3.02 2^-16 |a||b| per product, the subtractions
// fl(v - c) add 2^-23):
//   |acc - <x-cx, y-c>| <= kappa_D |x-cx| |y-c|,
//   kappa_D = 3.1 2^-16 + 6.1 D^2 2^-24
// (the second term lets each of the 3 D^2 non-zero products' partial sums
// lose one ulp per addition, rounded or not; zero products add exactly),
// plus flushed subnormals (at most 2^-126 per product or addition: the
// slack term).  sc = fl(bn - 2 fl(acc + fl(cx Sm))) then satisfies
//   |sc - Z| <= E2 = 2 kappa |x-cx| |y-c| + 2|cx| |Sm - sum y| + u (bn + 5|acc| + 7|cx Sm|) + slack,
// |Sm - sum y| <= D^2 u sum|y| <= D^3 u sqrt(bn) (1.01).  B(y) >= E1 + E2
// (with the rounding of its own evaluation and of sc - B) gives
//   exnum(y) >= T + sc(y) - B(y) = T + lb(y).
constexpr float kU32 = 0x1p-24f;
template <int D>
__host__ __device__ constexpr float bf16x3_kappa() {
  return 3.1f * 0x1p-16f + 6.1f * (float)(D * D) * 0x1p-24f;
}

// Upper bound of |y - c| from b = bn(y), s = Sm(y) (sqb = sqrt(b)):
// |y - c|^2 = sum y^2 - 2c sum y + c^2 D^2, with sum y^2 <= b (1 + 2 D^2 u),
// |sum y - s| <= D^3 u sqrt(b) (1.01) and <= 4u (b + 2|c s| + c^2 D^2) for the
// FP32 evaluation (9u there also covers the rounding of this bound itself).
template <int D>
__device__ __forceinline__ float centred_norm_up(float b, float s, float c, float sqb) {
  constexpr float DD = (float)(D * D), D3 = (float)(D * D * D);
  const float cs = c * s, c2 = (c * c) * DD;
  const float qv = (b - 2.f * cs) + c2;
  const float tq = (b + 2.f * fabsf(cs)) + c2;
  const float qu = qv + kU32 * ((2.f * DD + 4.f) * b + 9.f * tq + 2.1f * fabsf(c) * D3 * sqb);
  return sqrtf(fmaxf(qu, 0.f)) * (1.f + 0x1p-20f);
}

// Approximate score sc and its bound B >= E1 + E2 (+ the rounding of sc - B)
// for one candidate: acc the MFMA cross term, bnp = bn(y), smp = Sm(y), nxc
// an upper bound of |x - cx|.
template <int D>
__device__ __forceinline__ void bf16_bound(float acc, float bnp, float smp, float c, float cx, float nx, float nxc,
                                           float* sc, float* bnd) {
  constexpr float DD = (float)(D * D), D3 = (float)(D * D * D);
  const float t1 = cx * smp;
  const float s = bnp - 2.0f * (acc + t1);
  const float sqb = sqrtf(fmaxf(bnp, 0.f));
  const float nyc = centred_norm_up<D>(bnp, smp, c, sqb);
  const float e = 2.f * bf16x3_kappa<D>() * nxc * nyc +
                  kU32 * ((DD + 5.f) * (nx + bnp) + 5.f * fabsf(acc) + 7.f * fabsf(t1) +
                          2.1f * fabsf(cx) * D3 * sqb + 4.f * fabsf(s)) +
                  DD * 0x1p-100f * (1.f + nxc + nyc);
  *sc = s;
  *bnd = e * (1.f + 0x1p-16f);
}

template <int D, int D2>
struct Bf16Cfg {
  static constexpr int NWT = (D == 16 && D2 == 48) ? 2 : 4;  // reference tiles (waves) per workgroup
  static constexpr int WR = D2 + D;                           // window rows
  static constexpr int WC = D2 + D + (NWT - 1) * D;           // union window columns
  static constexpr int LDW = (WR % 2 == 0) ? WR + 1 : WR;     // odd: conflict-free columns
  static constexpr int MT = D2 / 16, NT = D2 / 16;
  static constexpr int RB = D2 + 16;  // BF16 rows read by the B fragments (zero beyond WR)
  static constexpr int LB = RB + 8;   // BF16 column stride: 16-byte aligned, conflict-free 16-byte reads
  static constexpr int CC = D2 + (NWT - 1) * D;  // candidate columns of the workgroup
  static constexpr int LN = D2 + 1;
  static constexpr int NC = MT * NT * 4;  // candidates per lane
  static constexpr int XP = 48;           // zero-padded BF16 reference column, entry s <-> row s - 16
  static_assert(D2 % 16 == 0 && D <= 16, "match size a multiple of 16, block size <= 16");
  static_assert(RB >= WR, "B rows");
};

template <int D, int D2>
__global__ __launch_bounds__(256) void match_bf16_kernel(
    const float* __restrict__ pad, const float* __restrict__ bn, const float* __restrict__ smap, int M, int N, int K,
    int R, uint32_t* __restrict__ S, float* __restrict__ vol, uint32_t* __restrict__ stats) {
  using C = Bf16Cfg<D, D2>;
  constexpr int S2 = D2 / 2, NWT = C::NWT, NTH = 64 * NWT;
  constexpr float DD = (float)(D * D);
  __shared__ float Wf[C::WC * C::LDW];
  __shared__ __attribute__((aligned(16))) unsigned short Wb[C::WC * C::LB];  // bf16 hi part of the centred window
  __shared__ __attribute__((aligned(16))) unsigned short Wl[C::WC * C::LB];  // and its lo part
  __shared__ unsigned short xpad[NWT][2][D * C::XP];  // zero-padded centred reference columns, hi and lo
  // per wave: the exact search's keys, or the rescored set's (at most 64)
  // keys below its candidate list
  __shared__ uint64_t scr[NWT][kSelMax];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int Mt = M / D, Nt = N / D;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const int ti = blk % Mt, tj0 = (blk / Mt) * NWT;
  const int nval = min(NWT, Nt - tj0);
  const int P = M + D2, Pm = P - D + 1;
  // window centring constant: mean of the workgroup's reference blocks
  float csum = 0.f;
  for (int w = 0; w < nval; ++w) csum += smap[(ti * D + S2) + (size_t)Pm * ((tj0 + w) * D + S2)];
  const float cm = csum / (float)(nval * D * D);
  const float* src = pad + (size_t)(ti * D) + (size_t)P * (tj0 * D);
  const int wcv = D2 + D + (nval - 1) * D;  // window columns inside the padded image
  for (int e = tid; e < C::WR * C::WC; e += NTH) {
    const int rr = e % C::WR, cc = e / C::WR;
    Wf[rr + C::LDW * cc] = cc < wcv ? src[rr + (size_t)P * cc] : 0.f;
  }
  __syncthreads();
  for (int e = tid; e < C::RB * C::WC; e += NTH) {
    const int rr = e % C::RB, cc = e / C::RB;
    unsigned short h = 0, lo = 0;
    if (rr < C::WR) split_bf16(Wf[rr + C::LDW * cc] - cm, &h, &lo);
    Wb[rr + C::LB * cc] = h;
    Wl[rr + C::LB * cc] = lo;
  }
  const int tj = tj0 + wave;
  const size_t xo = (size_t)(ti * D + S2) + (size_t)Pm * (tj * D + S2);  // the reference block in the maps
  // the reference block is centred on its own mean cx: <x,y> = <x-cx, y-c> + cx sum(y) + (tile constants)
  const float nx = wave < nval ? bn[xo] : 1.f, smx = wave < nval ? smap[xo] : 0.f;
  const float cx = smx / DD;
  unsigned short* xp = xpad[wave][0];
  unsigned short* xl = xpad[wave][1];
  for (int e = lane; e < D * C::XP; e += 64) {
    const int v = e / C::XP, t = e % C::XP - 16;
    unsigned short h = 0, lo = 0;
    if (t >= 0 && t < D) split_bf16(Wf[(S2 + t) + C::LDW * (S2 + v + D * wave)] - cx, &h, &lo);
    xp[e] = h;
    xl[e] = lo;
  }
  __syncthreads();
  if (wave >= nval) return;  // no barrier below
  const int g = ti + Mt * tj;
  const int q = lane >> 4, l = lane & 15;
  const float* Wt = Wf + C::LDW * (D * wave);  // this tile's window
  const float* bnt = bn + (size_t)(ti * D) + (size_t)Pm * (tj * D);
  const float* smt = smap + (size_t)(ti * D) + (size_t)Pm * (tj * D);
  // exact numerator of window candidate (a, b): the FMA chain of match_kernel
  auto exact_num = [&](int a, int b) {
    float cr = 0.f;
#pragma unroll 1
    for (int v = 0; v < D; ++v)
#pragma unroll
      for (int u = 0; u < D; ++u) cr = fmaf(Wt[(S2 + u) + C::LDW * (S2 + v)], Wt[(a + u) + C::LDW * (b + v)], cr);
    return (nx + bnt[a + (size_t)Pm * b]) - 2.0f * cr;
  };
  auto exact_key = [&](int a, int b) { return cand_key(exact_num(a, b) / nx, a + D2 * b); };
  auto cand_ab = [&](int j, int& a, int& b) {
    const int mt = j / (4 * C::NT), nt = (j / 4) % C::NT, i = j % 4;
    a = 16 * mt + 4 * q + i;
    b = 16 * nt + l;
  };
  uint64_t* sk = scr[wave];
  unsigned short* cl = reinterpret_cast<unsigned short*>(sk + 64);  // the candidate list, above the keys
  // exact top K over the candidates j with in(j), `total` of them (<= kRescoreMax)
  auto rescore = [&](auto in, int total) {
    wave_sync();  // the list and keys may still be read by a previous pass
    uint32_t base = 0;
#pragma unroll
    for (int j = 0; j < C::NC; ++j) {
      const bool f = in(j);
      const uint64_t m = __ballot(f);
      if (f) {
        int a, b;
        cand_ab(j, a, b);
        cl[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
            (unsigned short)(a + D2 * b);
      }
      base += (uint32_t)__popcll(m);
    }
    wave_sync();
    uint64_t ek[kRescoreMax / 64];
#pragma unroll
    for (int rr = 0; rr < kRescoreMax / 64; ++rr) {
      const int c = lane + 64 * rr;
      ek[rr] = ~0ull;
      if (64 * rr < total && c < total) {
        const int w = cl[c];
        ek[rr] = exact_key(w % D2, w / D2);
      }
    }
    wave_sync();  // every lane has read the list before the keys are written
    return wave_topk<kRescoreMax / 64>(ek, K, sk, lane);
  };
  uint64_t best;
  int path = 2, nres = 0;  // 0: first set, 1: widened set, 2: every candidate
  int why = 4;             // for path 2: 4 first set too large, 5 no K-th key, 6 widened set too large
  bool done = false;
  if (nx > 0.f && nx <= 1e38f) {  // (2 nx finite: see the centre's numerator below)
    f4_t acc[C::MT][C::NT];
#pragma unroll
    for (int mt = 0; mt < C::MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < C::NT; ++nt) acc[mt][nt] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int v = 0; v < D; ++v) {
      unsigned short th[8], tl[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        th[j] = xp[v * C::XP + 8 * q - l + 16 + j];
        tl[j] = xl[v * C::XP + 8 * q - l + 16 + j];
      }
      const bf16x8_t Ah = __builtin_bit_cast(bf16x8_t, th), Al = __builtin_bit_cast(bf16x8_t, tl);
#pragma unroll
      for (int nt = 0; nt < C::NT; ++nt) {
        const int co = C::LB * (16 * nt + l + v + D * wave) + 8 * q;
#pragma unroll
        for (int mt = 0; mt < C::MT; ++mt) {
          const bf16x8_t Bh = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(Wb + co + 16 * mt));
          const bf16x8_t Bl = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(Wl + co + 16 * mt));
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc[mt][nt], 0, 0, 0);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc[mt][nt], 0, 0, 0);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc[mt][nt], 0, 0, 0);
        }
      }
    }
    // Approximate score sc(y) = |y|^2 - 2(<x-cx, y-c>_bf16x3 + cx sum(y)) and a
    // lower bound lb(y) = sc(y) - B(y) such that every candidate's exact FP32
    // numerator satisfies  exnum(y) >= T_low + lb(y),  T_low a lower bound of
    // the tile constant T (see bf16_bound).
    const float nxc = centred_norm_up<D>(nx, smx, cx, sqrtf(nx));
    float lb[C::NC];
    float scC = 0.f, bC = 0.f;
    constexpr int jc = ((S2 / 16) * C::NT + S2 / 16) * 4 + S2 % 4;  // the centre (a, b) = (S2, S2)
    constexpr int lc = 16 * ((S2 % 16) / 4) + S2 % 16;               // and its lane
#pragma unroll
    for (int mt = 0; mt < C::MT; ++mt)
#pragma unroll
      for (int nt = 0; nt < C::NT; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int a = 16 * mt + 4 * q + i, b = 16 * nt + l;
          const int j = (mt * C::NT + nt) * 4 + i;
          const size_t o = (size_t)a + (size_t)Pm * b;
          float sc, bnd;
          bf16_bound<D>(acc[mt][nt][i], bnt[o], smt[o], cm, cx, nx, nxc, &sc, &bnd);
          lb[j] = sc - bnd;
          if (j == jc) {
            scC = sc;
            bC = bnd;
          }
        }
    // first set: every candidate whose bound is at most the R-th smallest
    // lane minimum of the bounds (at least R of them)
    uint32_t lmin = 0xffffffffu;
#pragma unroll
    for (int j = 0; j < C::NC; ++j) lmin = min(lmin, ord_f32(lb[j]));
    const uint32_t theta = wave_kth_smallest(lmin, R);
    auto in0 = [&](int j) { return ord_f32(lb[j]) <= theta; };
    int total0 = 0;
#pragma unroll
    for (int j = 0; j < C::NC; ++j) total0 += __popcll(__ballot(in0(j)));
    if (theta != 0xffffffffu && total0 <= kRescoreMax) {
      const uint64_t bestA = rescore(in0, total0);
      nres = total0;
      // T_low from the centre candidate (x itself): T = N(x) - Z(x) with
      // N(x) >= exnum(x) - E1 and Z(x) <= sc(x) + E2
      scC = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(scC), lc));
      bC = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(bC), lc));
      // exnum(x) = fl(fl(2 nx) - 2 cr(x, x)) = 0 exactly: the norm map forms
      // nx with the very FMA chain of cr, so cr(x, x) = nx
      const float tlow = (-scC - bC) - 4.f * kU32 * (fabsf(scC) + bC);
      // the K-th exact key of the first set: a candidate whose numerator
      // exceeds thrN has a strictly larger rounded distance, so it is not in
      // the top K (the K-th of a subset is >= the K-th of all)
      const uint32_t kth = (uint32_t)(__builtin_amdgcn_readlane((uint32_t)(bestA >> 32), K - 1));
      why = kth != 0xffffffffu ? 6 : 5;
      if (kth != 0xffffffffu) {
        const float pk = ord_to_f32(kth) * nx;
        const float thrN = fmaf(fabsf(pk), 0x1p-19f, pk) + nx * 0x1p-123f;
        const float thr = (thrN - tlow) + 2.f * kU32 * (fabsf(thrN) + fabsf(tlow));
        // a candidate can enter the top K only if !(lb > thr) (NaN included)
        auto in1 = [&](int j) { return in0(j) || !(lb[j] > thr); };
        int extra = 0, total1 = 0;
#pragma unroll
        for (int j = 0; j < C::NC; ++j) {
          extra += __popcll(__ballot(!in0(j) && !(lb[j] > thr)));
          total1 += __popcll(__ballot(in1(j)));
        }
        if (extra == 0) {
          best = bestA;
          path = 0;
          done = true;
        } else if (total1 <= kRescoreMax) {
          best = rescore(in1, total1);
          nres += total1;
          path = 1;
          done = true;
        }
      }
    }
  } else {
    why = 7;
  }
  if (!done) {
    // the FP32 search itself (zero-norm reference blocks; tiles whose bounds
    // leave more than kRescoreMax candidates, e.g. along a straight edge)
    wave_sync();
    best = exact_search<D, D2, C::LDW>(Wt, bn, Pm, ti, tj, nx, K, sk, lane);
    nres += D2 * D2;
  }
  if (stats && lane == 0) {
    // counters spread over kBf16StatSlots cache lines (few tiles share one)
    uint32_t* sl = stats + (size_t)(blk % kBf16StatSlots) * kBf16StatStride;
    atomicAdd(&sl[path], 1u);
    atomicAdd(&sl[3], (uint32_t)nres);
    if (path == 2) atomicAdd(&sl[why], 1u);
  }
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
      dst[u + (size_t)M * v] = Wt[(a + u) + C::LDW * (b + v)];
    }
  }
}

int launch_match_bf16(const float* pad, const float* bn, const float* smap, int M, int N, int d, int d2, int K,
                      int R, uint32_t* S, float* vol, hipStream_t st, uint32_t* stats) {
  if (M % d || N % d || M < d2 || N < d2) return fail(GBM3D_ERR_ARG, "match: image size");
  if (K < 1 || K > 64) return fail(GBM3D_ERR_UNSUPPORTED, "match: numMax must be in [1, 64]");
  if (R < K) R = K;
  if (R > 64) R = 64;
  const int Mt = M / d, Nt = N / d;
#define GBM3D_BF16_CASE(DD, DD2)                                                                        \
  if (d == DD && d2 == DD2) {                                                                           \
    constexpr int nwt = Bf16Cfg<DD, DD2>::NWT;                                                          \
    const dim3 grid(Mt * ((Nt + nwt - 1) / nwt));                                                       \
    hipLaunchKernelGGL((match_bf16_kernel<DD, DD2>), grid, dim3(64 * nwt), 0, st, pad, bn, smap, M, N, K, R, \
                       S, vol, stats);                                                                  \
    return check_launch("match_bf16");                                                                  \
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

// Volume(ti d + u, tj d + v, k) = Ipad(ti d + a_k + u, tj d + b_k + v): the
// mosaic volume the matcher writes (matchBlocksCPU.m:105-111), rebuilt from
// S and the search image; only when *run != 0 (a gated launch).
__global__ void mosaic_gather_kernel(const float* __restrict__ pad, const uint32_t* __restrict__ S, int M, int N,
                                     int d, int d2, int K, float* __restrict__ V, const int* __restrict__ run) {
  if (*run == 0) return;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t MN = (size_t)M * N;
  if (e >= MN * K) return;
  const int r = (int)(e % M), c = (int)((e / M) % N), k = (int)(e / MN);
  const int ti = r / d, tj = c / d;
  const uint32_t idx = S[k + (size_t)K * (ti + (size_t)(M / d) * tj)] - 1u;
  const int a = (int)(idx % d2), b = (int)(idx / d2);
  V[e] = pad[(size_t)(ti * d + a + r % d) + (size_t)(M + d2) * (tj * d + b + c % d)];
}

int launch_mosaic_gather(const float* pad, const uint32_t* S, int M, int N, int d, int d2, int K, float* V,
                         const int* run, hipStream_t st) {
  if (!run || M % d || N % d) return fail(GBM3D_ERR_ARG, "mosaic gather: geometry");
  const size_t total = (size_t)M * N * K;
  hipLaunchKernelGGL(mosaic_gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pad, S, M, N, d,
                     d2, K, V, run);
  return check_launch("mosaic_gather");
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
