// k_gmatch.hip -- matchBlocksGPU for gfx950: the whole-image circular block
// search (source/matchBlocksGPU.m:56-91) and its 2-D sym4 wavelet prefilter
// (:34-50).
//
// Every sum is one FP32 fused multiply-add chain from +0, column by column
// (v outer, u inner): block norms and correlations as in the windowed
// matcher, the wavelet bands as sum_v sum_u Psi(u, v) X(., .).  The FFT
// products of the reference compute the same circular convolutions; the
// direct chains make the result independent of FFT sizes and round-off.
#include "internal.hpp"
#include "select.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {

// sym4 decomposition taps Lo_d, Hi_d (archive/waveletFilters/sym4Filters.mat,
// loaded by getWaveFilters2D('sym', 4))
__constant__ double kSym4Lo[8] = {-0.07576571478927333, -0.02963552764599851, 0.49761866763201545, 0.8037387518059161,
                                  0.29785779560527736,  -0.09921954357684722, -0.012603967262037833, 0.0322231006040427};
__constant__ double kSym4Hi[8] = {-0.0322231006040427, -0.012603967262037833, 0.09921954357684722, 0.29785779560527736,
                                  -0.8037387518059161, 0.49761866763201545,   0.02963552764599851,  -0.07576571478927333};

// The four 8 x 8 filters of getWaveFilters2D.m:12-16 in LDS: P[64 b + u + 8 v]
// = h1[u] h2[v] rounded to float, b = 0 Lo'Lo, 1 Hi'Lo, 2 Lo'Hi, 3 Hi'Hi
// (the first factor runs along dimension 1).
__device__ __forceinline__ void load_psi(float* P) {
  for (int e = threadIdx.x; e < 256; e += blockDim.x) {
    const int b = e >> 6, u = e & 7, v = (e >> 3) & 7;
    const double h1 = (b & 1) ? kSym4Hi[u] : kSym4Lo[u];
    const double h2 = (b & 2) ? kSym4Hi[v] : kSym4Lo[v];
    P[e] = (float)(h1 * h2);
  }
}

// One level of myWavDec2DFFT.m:13-21 at output (i, j) of the L1/2 x L2/2
// bands: band b = sum_v sum_u Psi_b(u, v) X((2i - u) mod L1, (2j - v) mod L2)
// (the circular convolution kept at the odd 1-based samples; filters cut to
// the level size like fftn(Psi, size(U))).  The details are hard-thresholded:
// C .* (abs(C) > tau) (matchBlocksGPU.m:44), |c| > tau <=> |c| > taul for the
// largest float taul <= tau.  Level 0 reads the image circshifted by [k k]
// (X(r, c) = I((r - k) mod L1, (c - k) mod L2)); deeper levels pass k = 0.
__global__ __launch_bounds__(256) void wav2_dec_kernel(const float* __restrict__ X, int L1, int L2, int k,
                                                       float taul, float* __restrict__ A, float* __restrict__ D1,
                                                       float* __restrict__ D2, float* __restrict__ D3) {
  __shared__ float P[256];
  load_psi(P);
  __syncthreads();
  const int h1 = L1 / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= h1) return;
  const int n1 = min(8, L1), n2 = min(8, L2);
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int v = 0; v < n2; ++v) {
    const float* col = X + (size_t)L1 * pmod_i(2 * j - v - k, L2);
    for (int u = 0; u < n1; ++u) {
      const float x = col[pmod_i(2 * i - u - k, L1)];
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[b] = fmaf(P[64 * b + u + 8 * v], x, acc[b]);
    }
  }
  const size_t o = i + (size_t)h1 * j;
  A[o] = acc[0];
  D1[o] = fabsf(acc[1]) > taul ? acc[1] : 0.f;
  D2[o] = fabsf(acc[2]) > taul ? acc[2] : 0.f;
  D3[o] = fabsf(acc[3]) > taul ? acc[3] : 0.f;
}

// sum_v sum_u Psi(u, v) up((t1 + u) mod L1, (t2 + v) mod L2) for the
// zero-upsampled band (Lc1 x Lc2 values at the even positions of L1 x L2):
// real(ifftn(fftn(tmp) .* conj(fftn(Psi, size(tmp))))) at (t1, t2).
__device__ __forceinline__ float corr_up(const float* __restrict__ Dc, int Lc1, int L1, int L2, int n1, int n2,
                                         const float* p, int t1, int t2) {
  float acc = 0.f;
  for (int v = 0; v < n2; ++v) {
    const int c = pmod_i(t2 + v, L2);
    if (c & 1) continue;
    const float* col = Dc + (size_t)Lc1 * (c >> 1);
    for (int u = 0; u < n1; ++u) {
      const int r = pmod_i(t1 + u, L1);
      if (r & 1) continue;
      acc = fmaf(p[u + 8 * v], col[r >> 1], acc);
    }
  }
  return acc;
}

// One level of myWavRec2DGPUfast.m:12-31 at output (t1, t2) of the
// 2 Lc1 x 2 Lc2 level: ((detail 1 + detail 2) + detail 3) + the coarser
// level, each its correlation with the band's filter.  avg != nullptr (the
// finest level of shift k): avg((t1 - k) mod L1, (t2 - k) mod L2) += a / 5
// instead (circshift(., [1-k 1-k]) / shiftCnt, matchBlocksGPU.m:48), from +0
// for the first shift.
__global__ __launch_bounds__(256) void wav2_rec_kernel(const float* __restrict__ D1, const float* __restrict__ D2,
                                                       const float* __restrict__ D3, const float* __restrict__ Cc,
                                                       int Lc1, int Lc2, int k, float* __restrict__ out,
                                                       float* __restrict__ avg) {
  __shared__ float P[256];
  load_psi(P);
  __syncthreads();
  const int L1 = 2 * Lc1, L2 = 2 * Lc2;
  const int t1 = blockIdx.x * blockDim.x + threadIdx.x, t2 = blockIdx.y;
  if (t1 >= L1) return;
  const int n1 = min(8, L1), n2 = min(8, L2);
  float a = corr_up(D1, Lc1, L1, L2, n1, n2, P + 64, t1, t2);
  a = a + corr_up(D2, Lc1, L1, L2, n1, n2, P + 128, t1, t2);
  a = a + corr_up(D3, Lc1, L1, L2, n1, n2, P + 192, t1, t2);
  a = a + corr_up(Cc, Lc1, L1, L2, n1, n2, P, t1, t2);
  if (avg == nullptr) {
    out[t1 + (size_t)L1 * t2] = a;
  } else {
    float* o = avg + pmod_i(t1 - k, L1) + (size_t)L1 * pmod_i(t2 - k, L2);
    *o = (k == 0 ? 0.f : *o) + a / 5.0f;
  }
}

// cpad(r, c) = I(r mod m, c mod n), (m + d - 1) x (n + d - 1): the norm map
// of this image is blockNorms (matchBlocksGPU.m:58-60), the circular
// top-left-anchored d x d sums of I^2.
__global__ void circ_pad_kernel(const float* __restrict__ I, int m, int n, int P, int Q, float* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (r >= P) return;
  out[r + (size_t)P * c] = I[(r < m ? r : r - m) + (size_t)m * (c < n ? c : c - n)];
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Candidate windows of 32 x 32 positions; a lane scores RA consecutive rows
// of one window column.
constexpr int kGW = 32, kGRA = 16;

template <int D>
struct GCfg {
  static constexpr int WR = kGW + D;                      // staged rows / columns
  static constexpr int LDW = (WR % 2 == 0) ? WR + 1 : WR;  // odd: conflict-free columns
};

// One wave per reference tile, NW tiles per workgroup sharing each staged
// window of the image.  The wave keeps its K best keys so far (lane k the
// k-th, ~0 past K).  A window's candidates are keyed and selected only when
// one can still enter: the distance num / |x|^2 rounds to at most the K-th
// best distance dK only if num <= |x|^2 dK (1 + 2^-23) (+ a subnormal term),
// which the bound below covers with room for its own rounding; the window's
// exact top K (wave_topk) is then merged with the running list (the K
// smallest of two sorted lists: min(a_k, b_(63-k)), sorted).  Without a
// finite positive |x|^2 or a finite dK every candidate is selected.
//
// Tolerance branch (Kmin < K): also counts the positions with distance < tol
// (dist < tolu exactly, tolu the smallest float >= tol; num compared with
// tolu |x|^2 where that decides, divided otherwise), then keeps the best K
// (more than K below tol), the best Kmin (fewer than Kmin) or the ones below
// tol in index order (matchBlocksGPU.m:71-81).
template <int D, int NW>
__global__ __launch_bounds__(64 * NW) void gmatch_kernel(const float* __restrict__ I, const float* __restrict__ bn,
                                                         int m, int n, int K, int Kmin, float tolu,
                                                         uint32_t* __restrict__ S, uint32_t* __restrict__ counts) {
  using C = GCfg<D>;
  __shared__ float W[C::WR * C::LDW];
  __shared__ float X[NW][D * D];
  __shared__ uint64_t scr[NW][64];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int Mt = m / D, Nt = n / D, G = Mt * Nt;
  const int g = blockIdx.x * NW + wave;
  const bool live = g < G;
  const int ti = live ? g % Mt : 0, tj = live ? g / Mt : 0;
  for (int e = lane; e < D * D; e += 64) X[wave][e] = I[(ti * D + e % D) + (size_t)m * (tj * D + e / D)];
  const float nx = bn[ti * D + (size_t)m * (tj * D)];
  const bool fin = nx > 0.f && nx <= 3.4028235e38f;
  const bool tolmode = Kmin < K;
  const float tnx = tolu * nx;
  const float tzone = fmaf(fabsf(tnx), 0x1p-20f, nx * 0x1p-126f);
  const bool tfast = fin && tnx <= 3.4028235e38f;
  uint32_t below = 0;
  uint64_t best = ~0ull;
  bool all = true;
  float bound = 0.f;
  const int b = lane % kGW, a0 = (lane / kGW) * kGRA;
  for (int wc = 0; wc < n; wc += kGW)
    for (int wr = 0; wr < m; wr += kGW) {
      __syncthreads();
      for (int e = threadIdx.x; e < C::WR * C::WR; e += 64 * NW) {
        const int rr = e % C::WR, cc = e / C::WR;  // constant divisor
        int r = wr + rr, c = wc + cc;               // < m + WR, < n + WR: a few wraps at most
        while (r >= m) r -= m;
        while (c >= n) c -= n;
        W[rr + C::LDW * cc] = I[r + (size_t)m * c];
      }
      __syncthreads();
      if (!live) continue;
      float acc[kGRA];
#pragma unroll
      for (int i = 0; i < kGRA; ++i) acc[i] = 0.f;
      constexpr int kUnrollV = D <= 4 ? D : 1;
#pragma unroll kUnrollV
      for (int v = 0; v < D; ++v) {
        float xs[D];
#pragma unroll
        for (int u = 0; u < D; ++u) xs[u] = X[wave][u + D * v];
        const float* col = W + a0 + C::LDW * (b + v);
#pragma unroll
        for (int t = 0; t < kGRA + D - 1; ++t) {
          const float y = col[t];
#pragma unroll
          for (int i = 0; i < kGRA; ++i) {
            const int u = t - i;
            if (u >= 0 && u < D) acc[i] = fmaf(xs[u], y, acc[i]);
          }
        }
      }
      const int pc = wc + b;
      float num[kGRA];
      bool in[kGRA];
      bool any = false;
#pragma unroll
      for (int i = 0; i < kGRA; ++i) {
        const int pr = wr + a0 + i;
        in[i] = pr < m && pc < n;
        num[i] = in[i] ? (nx + bn[pr + (size_t)m * pc]) - 2.0f * acc[i] : 0.f;
        if (tolmode && in[i]) {
          if (tfast && num[i] < tnx - tzone)
            ++below;
          else if (!tfast || num[i] <= tnx + tzone)
            below += (num[i] / nx < tolu) ? 1u : 0u;
        }
        in[i] = in[i] && (all || num[i] <= bound);
        any = any || in[i];
      }
      if (__ballot(any) == 0) continue;
      uint64_t key[kGRA];
#pragma unroll
      for (int i = 0; i < kGRA; ++i) key[i] = in[i] ? cand_key(num[i] / nx, (wr + a0 + i) + m * pc) : ~0ull;
      uint64_t w = wave_topk<kGRA>(key, K, scr[wave], lane);
      if (lane >= K) w = ~0ull;
      const uint64_t wrev = shfl_xor_u64(w, 63);  // lane 63 - l
      best = wave_sort64<uint64_t>(best < wrev ? best : wrev, lane);
      if (lane >= K) best = ~0ull;
      const uint32_t ok = (uint32_t)(readlane_u64(best, K - 1) >> 32);
      if (fin && ok < ord_bits(__builtin_inff())) {
        const float t = nx * ord_to_f32(ok);
        bound = fmaf(fabsf(t), 0x1p-20f, t) + nx * 0x1p-126f;
        all = false;
      }
    }
  if (!live) return;
  int cnt = K;
  uint32_t idx = (uint32_t)best;
  if (tolmode) {
    uint32_t tot = below;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) tot += shfl_xor_u32(tot, o);
    if (tot < (uint32_t)Kmin) {
      cnt = Kmin;
    } else if (tot <= (uint32_t)K) {
      cnt = (int)tot;  // find(dist < tol): index order
      idx = wave_sort64<uint32_t>(lane < cnt ? idx : 0xffffffffu, lane);
    }
  }
  if (lane < K) S[lane + (size_t)K * g] = lane < cnt ? idx + 1u : 0u;
  if (lane == 0 && counts) counts[g] = (uint32_t)cnt;
}

template <int D>
void launch_gmatch_t(const float* I, const float* bn, int m, int n, int K, int Kmin, float tolu, uint32_t* S,
                     uint32_t* counts, hipStream_t st) {
  constexpr int NW = 4;
  const int G = (m / D) * (n / D);
  hipLaunchKernelGGL((gmatch_kernel<D, NW>), dim3((G + NW - 1) / NW), dim3(64 * NW), 0, st, I, bn, m, n, K, Kmin,
                     tolu, S, counts);
}

}  // namespace

bool match_global_supported(int d) { return d == 4 || d == 8 || d == 16 || d == 32; }

int launch_wav2_prefilter(const float* I, int m, int n, double tau, float* Iavg, float* work, hipStream_t st) {
  if (m % 8 || n % 8 || m < 8 || n < 8)
    return fail(GBM3D_ERR_ARG, "matchBlocksGPU prefilter: the 3-level transform needs m and n multiples of 8");
  float taul = (float)tau;
  if ((double)taul > tau) taul = nextafterf(taul, -__builtin_inff());
  // per level l (sizes L1 x L2 halving): A, D1, D2, D3 of (L1/2)(L2/2);
  // reconstruction scratch of the two coarser levels
  const size_t q0 = (size_t)(m / 2) * (n / 2), q1 = q0 / 4, q2 = q1 / 4;
  float* A[3];
  float* Dl[3][3];
  float* p = work;
  const size_t qs[3] = {q0, q1, q2};
  for (int l = 0; l < 3; ++l) {
    A[l] = p;
    p += qs[l];
    for (int j = 0; j < 3; ++j) {
      Dl[l][j] = p;
      p += qs[l];
    }
  }
  float* rec1 = p;  // level 1 reconstruction, (m/2) x (n/2)
  p += q0;
  float* rec2 = p;  // level 2, (m/4) x (n/4)
  for (int k = 0; k < 5; ++k) {
    const float* src = I;
    int L1 = m, L2 = n;
    for (int l = 0; l < 3; ++l) {
      dim3 grid((L1 / 2 + 255) / 256, L2 / 2);
      hipLaunchKernelGGL(wav2_dec_kernel, grid, dim3(256), 0, st, src, L1, L2, l == 0 ? k : 0, taul, A[l], Dl[l][0],
                         Dl[l][1], Dl[l][2]);
      src = A[l];
      L1 /= 2;
      L2 /= 2;
    }
    // level 3 -> 2 (coarse = A[2]), 2 -> 1 (coarse = rec2), 1 -> 0 into Iavg
    hipLaunchKernelGGL(wav2_rec_kernel, dim3((m / 4 + 255) / 256, n / 4), dim3(256), 0, st, Dl[2][0], Dl[2][1],
                       Dl[2][2], A[2], m / 8, n / 8, 0, rec2, nullptr);
    hipLaunchKernelGGL(wav2_rec_kernel, dim3((m / 2 + 255) / 256, n / 2), dim3(256), 0, st, Dl[1][0], Dl[1][1],
                       Dl[1][2], rec2, m / 4, n / 4, 0, rec1, nullptr);
    hipLaunchKernelGGL(wav2_rec_kernel, dim3((m + 255) / 256, n), dim3(256), 0, st, Dl[0][0], Dl[0][1], Dl[0][2],
                       rec1, m / 2, n / 2, k, nullptr, Iavg);
  }
  return check_launch("wav2_prefilter");
}

int launch_match_global(const float* I, int m, int n, int d, int K, int Kmin, double tol, uint32_t* S,
                        uint32_t* counts, float* work, hipStream_t st) {
  if (!match_global_supported(d)) return fail(GBM3D_ERR_UNSUPPORTED, "matchBlocksGPU: block size must be 4, 8, 16 or 32");
  if (m < d || n < d) return fail(GBM3D_ERR_ARG, "matchBlocksGPU: image smaller than one block");
  if ((size_t)m * n >= ((size_t)1 << 31)) return fail(GBM3D_ERR_ARG, "matchBlocksGPU: image of 2^31 pixels or more");
  if (K < 1 || K > 64 || Kmin < 1 || Kmin > K || (size_t)K > (size_t)m * n)
    return fail(GBM3D_ERR_ARG, "matchBlocksGPU: need 1 <= numMin <= numMax <= min(64, m n)");
  float tolu = (float)tol;
  if ((double)tolu < tol) tolu = nextafterf(tolu, __builtin_inff());
  const int P = m + d - 1, Q = n + d - 1;
  float* cpad = work;
  float* bn = cpad + (size_t)P * Q;
  hipLaunchKernelGGL(circ_pad_kernel, dim3((P + 255) / 256, Q), dim3(256), 0, st, I, m, n, P, Q, cpad);
  GBM3D_TRY(launch_norm_map(cpad, P, Q, d, bn, st));
  switch (d) {
    case 4: launch_gmatch_t<4>(I, bn, m, n, K, Kmin, tolu, S, counts, st); break;
    case 8: launch_gmatch_t<8>(I, bn, m, n, K, Kmin, tolu, S, counts, st); break;
    case 16: launch_gmatch_t<16>(I, bn, m, n, K, Kmin, tolu, S, counts, st); break;
    default: launch_gmatch_t<32>(I, bn, m, n, K, Kmin, tolu, S, counts, st); break;
  }
  return check_launch("match_global");
}

}  // namespace gbm3d
