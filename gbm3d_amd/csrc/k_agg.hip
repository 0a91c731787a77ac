// k_agg.hip -- step-1 aggregation weights and the deterministic gather
// aggregation shared by both steps, plus the spin-sum normalisation.
//
// Reference: source/denoiseAndAggregateGPU.m:58-92 (weights 1/(TV+eps) and
// a sine mask), source/denoiseAndAggregateWieFast.m:68-100 (weight
// mask/|filt|^2), utilities/getBM3DAggWeights.m:15-19, source/GBM3D.m:57,82.
//
// The reference scatters every (tile y, tile x, slot k) block onto a padded
// canvas in (y outer, x inner, k) order and crops it.  Here one wave owns a
// 16x16 output tile: it enumerates the candidate (y, x, k) triples in that
// same order, keeps those whose block meets its tile, and adds each block's
// pixels into an LDS accumulator one block after another.  Every pixel thus
// receives its contributions in the reference's order, in FP32, with no
// atomics: results are bitwise reproducible run to run.
#include <cstdlib>

#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

// -------------------------------------------------------- step-1 prep
// One 256-thread workgroup per group (tile) g: Vavg = max(mean_s vrec_s, 0)
// over the tile's d x d x K stack, then w_g = 1/(D + eps) with
// D = sum|diff3_v|/4 + sum|diff3_u|/4 + sum|diff1_k| (getBM3DAggWeights(3,1)).

// 8 x 8 blocks, 16 slices: the group's 4 KB stack is read with 16-byte loads
// (four per lane), staged in LDS with one exchange, and each lane then takes
// its pixel's terms slice by slice in the order of the general form below.
constexpr int kPrepStack = 16;

template <int D>
__device__ __forceinline__ void agg1_prep_stack(const float* __restrict__ vg, int G, float* __restrict__ wgt) {
  constexpr int DD = D * D, NF = kPrepStack * DD / 4 / 64;  // float4 per lane
  static_assert(DD == 64 && NF == 4, "8x8 blocks, 16 slices");
  __shared__ __attribute__((aligned(16))) float st[4][kPrepStack * DD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
  if (g >= G) return;
  const float4* blk = reinterpret_cast<const float4*>(vg + (size_t)DD * kPrepStack * g);
  float* sw = st[wave];
  float4 q[NF];
#pragma unroll
  for (int r = 0; r < NF; ++r) q[r] = blk[lane + 64 * r];
#pragma unroll
  for (int r = 0; r < NF; ++r) reinterpret_cast<float4*>(sw)[lane + 64 * r] = q[r];
  wave_sync();
  const int u = lane % D, v = lane / D;
  float s1 = 0.f, s2 = 0.f, s3 = 0.f, prev = 0.f;
#pragma unroll
  for (int k = 0; k < kPrepStack; ++k) {
    const float* c = sw + DD * k + lane;
    const float x = c[0];
    if (k > 0) s3 += fabsf(x - prev);  // |diff(U,1,3)|
    prev = x;
    if (v + 3 < D) s2 += fabsf(third_diff(c[0], c[D], c[2 * D], c[3 * D]));
    if (u + 3 < D) s1 += fabsf(third_diff(c[0], c[1], c[2], c[3]));
  }
  s1 = wave_sum_f32(s1);
  s2 = wave_sum_f32(s2);
  s3 = wave_sum_f32(s3);
  if (lane == 0) {
    float Dv = s2 * 0.25f;
    Dv = Dv + s1 * 0.25f;
    Dv = Dv + s3;
    wgt[g] = 1.0f / (Dv + 2.220446049250313e-16f);  // 1/(D(blocks)+eps)
  }
}

// One wave per group; lane e < d*d owns pixel (u, v) = (e % d, e / d) of
// every slice, slices staged one at a time in LDS for the in-plane differences.
// The slices are fetched in batches of KB whose loads are all issued before
// the batch is processed, so one global latency covers KB slices.
template <int D>
__global__ __launch_bounds__(256) void agg1_prep_wave_kernel(const float* __restrict__ vg, int G, int K,
                                                             float* __restrict__ wgt) {
  constexpr int DD = D * D, PER = (DD + 63) / 64, KB = PER >= 16 ? 2 : 16 / PER;
  if constexpr (DD == 64) {
    if (K == kPrepStack) {  // 8x8 blocks, 16 slices: the whole 4 KB stack staged at once
      agg1_prep_stack<D>(vg, G, wgt);
      return;
    }
  }
  __shared__ float sl[4][DD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
  if (g >= G) return;
  const float* blk = vg + (size_t)DD * K * g;
  float* cur = sl[wave];
  float prev[PER];
  float s1 = 0.f, s2 = 0.f, s3 = 0.f;
  for (int k0 = 0; k0 < K; k0 += KB) {
    float xb[KB][PER];
#pragma unroll
    for (int j = 0; j < KB; ++j)
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int e = lane + 64 * i;
        xb[j][i] = (e < DD && k0 + j < K) ? blk[(size_t)DD * (k0 + j) + e] : 0.f;
      }
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      const int k = k0 + j;
      if (k < K) {  // wave-uniform
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          const int e = lane + 64 * i;
          if (e < DD) {
            const float x = xb[j][i];
            if (k > 0) s3 += fabsf(x - prev[i]);  // |diff(U,1,3)|
            prev[i] = x;
            cur[e] = x;
          }
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          const int e = lane + 64 * i;
          if (e < DD) {
            const int u = e % D, v = e / D;
            if (v + 3 < D) s2 += fabsf(third_diff(cur[e], cur[e + D], cur[e + 2 * D], cur[e + 3 * D]));
            if (u + 3 < D) s1 += fabsf(third_diff(cur[e], cur[e + 1], cur[e + 2], cur[e + 3]));
          }
        }
        wave_sync();
      }
    }
  }
  s1 = wave_sum_f32(s1);
  s2 = wave_sum_f32(s2);
  s3 = wave_sum_f32(s3);
  if (lane == 0) {
    float Dv = s2 * 0.25f;
    Dv = Dv + s1 * 0.25f;
    Dv = Dv + s3;
    wgt[g] = 1.0f / (Dv + 2.220446049250313e-16f);  // 1/(D(blocks)+eps)
  }
}

int launch_agg1_prep(const float* vg, int M, int N, int K, int d, float* wgt, hipStream_t st) {
  if (M % d || N % d) return fail(GBM3D_ERR_ARG, "agg1 prep: block size");
  const int G = (M / d) * (N / d);
  if (G == 0) return GBM3D_OK;
  const dim3 grid((G + 3) / 4);
  switch (d) {
    case 4: hipLaunchKernelGGL(agg1_prep_wave_kernel<4>, grid, dim3(256), 0, st, vg, G, K, wgt); break;
    case 8: hipLaunchKernelGGL(agg1_prep_wave_kernel<8>, grid, dim3(256), 0, st, vg, G, K, wgt); break;
    case 16: hipLaunchKernelGGL(agg1_prep_wave_kernel<16>, grid, dim3(256), 0, st, vg, G, K, wgt); break;
    case 32: hipLaunchKernelGGL(agg1_prep_wave_kernel<32>, grid, dim3(256), 0, st, vg, G, K, wgt); break;
    default: return fail(GBM3D_ERR_UNSUPPORTED, "block size must be 4, 8, 16 or 32");
  }
  return check_launch("agg1_prep");
}

// dst[u + d*(v + d*(k + K*g))] = src(i*d + u, j*d + v, k), g = i + Mt*j.
__global__ void mosaic_to_groups_kernel(const float* __restrict__ src, int M, int N, int K, int d,
                                        float* __restrict__ dst) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)M * N * K;
  if (e >= total) return;
  const int dd = d * d, Mt = M / d;
  const int u = (int)(e % d), v = (int)((e / d) % d);
  const size_t gk = e / dd;
  const int k = (int)(gk % K);
  const size_t g = gk / K;
  const int ti = (int)(g % Mt), tj = (int)(g / Mt);
  dst[e] = src[(size_t)(ti * d + u) + (size_t)M * ((size_t)(tj * d + v) + (size_t)N * k)];
}

int launch_mosaic_to_groups(const float* src, int M, int N, int K, int d, float* dst, hipStream_t st) {
  if (M % d || N % d) return fail(GBM3D_ERR_ARG, "mosaic_to_groups: size");
  const size_t total = (size_t)M * N * K;
  if (total == 0) return GBM3D_OK;
  hipLaunchKernelGGL(mosaic_to_groups_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, src, M, N, K,
                     d, dst);
  return check_launch("mosaic_to_groups");
}


// ---------------------------------------------------------- aggregation
constexpr int AT = 16;  // output tile edge per wave

// Lane e of the wave owns pixel (u, v) = (e % D, e / D) (+64i) of every block.
// Candidates come 64 per round in (y, x, k) order; the round's hits (blocks
// that meet the tile) are listed in LDS in that order, so that lane j holds
// hit j's block, packed origin, weight and accumulator offset.  Hits are taken
// eight at a time: their loads are issued together (whole blocks, no bounds
// tests), then each block is added into the LDS tile accumulator in order,
// predicated on the pixel lying in the tile (one packed 16-bit add and a
// mask).  U and W are interleaved (one 64-bit read and write per pixel and
// block); the accumulator's column stride AS makes a half-wave's 64-bit
// accesses to one block hit distinct banks.  The next round's index and
// weight loads are issued before the current round's hits are processed.

// two 16-bit lanes of one register (v_pk_add_u16)
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
static_assert((AT & (AT - 1)) == 0, "AT a power of two");
constexpr uint32_t kOutsideAT = ~((uint32_t)(AT - 1) | ((uint32_t)(AT - 1) << 16));

template <int D>
struct AggCfg {
  static constexpr int AS = D <= 8 ? 24 : 16;
};

// BUF: the blocks array spans fewer than 2^31 bytes and the image fewer than
// 2^26 pixels, so a block is read through one buffer resource with its byte
// offset in a scalar register, the slot lists and weights through buffer
// resources with 32-bit offsets, and every index product has 24-bit factors.
template <int STEP, int D, bool BUF>
__global__ __launch_bounds__(256) void aggregate_batched_kernel(const float* __restrict__ blocks,
                                                                const float* __restrict__ wgt,
                                                                const uint32_t* __restrict__ S,
                                                                const float* __restrict__ masks, int M, int N,
                                                                int d2, int K, int spin, float* __restrict__ U,
                                                                float* __restrict__ W) {
  constexpr int DD = D * D, PER = (DD + 63) / 64, NB = 8, AS = AggCfg<D>::AS;
  __shared__ float2 acc[4][AT * AS];
  // per wave: the round's hits (block, packed origin, weight, LDS offset),
  // NB entries of slack for a batch's reads past the last hit
  __shared__ int4 hl[4][64 + NB];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tilesR = (M + AT - 1) / AT, tilesC = (N + AT - 1) / AT;
  const int tile = xcd_remap(blockIdx.x, gridDim.x) * 4 + wave;
  if (tile >= tilesR * tilesC) return;
  const int R0 = (tile % tilesR) * AT, C0 = (tile / tilesR) * AT;
  float2* ac = acc[wave];
  for (int e = lane; e < AT * AS; e += 64) ac[e] = make_float2(0.f, 0.f);
  const __amdgpu_buffer_rsrc_t rB = buf_rsrc(blocks, BUF ? 4u * (uint32_t)DD * (uint32_t)(K * (M / D) * (N / D)) : 0u);
  const __amdgpu_buffer_rsrc_t rS = buf_rsrc(reinterpret_cast<const float*>(S), BUF ? 4u * (uint32_t)(K * (M / D) * (N / D)) : 0u);
  const __amdgpu_buffer_rsrc_t rW = buf_rsrc(wgt, BUF ? 4u * (uint32_t)((M / D) * (N / D)) : 0u);
  // a * b with 24-bit factors where the sizes allow it (full-rate multiply-add)
  auto imul = [](int a, int b) { return BUF ? __mul24(a, b) : a * b; };
  float mk[PER];
  uint32_t puv[PER];  // this lane's pixel (u, v) of a block as two 16-bit halves
  float2* pa[PER];    // the same pixel of a block at the tile origin
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = lane + 64 * i;
    puv[i] = (uint32_t)(e % D) | ((uint32_t)(e / D) << 16);
    pa[i] = ac + (e % D + AS * (e / D));
    mk[i] = e < DD ? masks[(STEP == 1 ? 0 : DD) + e] : 0.f;
  }
  wave_sync();
  const int s = d2 / 2, Mt = M / D, Nt = N / D;
  auto first = [&](int r) {
    const int x = r + s - d2 - D + 2;
    return x > 0 ? (x + D - 1) / D : 0;
  };
  const int ylo = first(R0), yhi = min(Mt - 1, (R0 + AT - 1 + s) / D);
  const int xlo = first(C0), xhi = min(Nt - 1, (C0 + AT - 1 + s) / D);
  const int ny = yhi - ylo + 1, nx = xhi - xlo + 1;
  const int ncand = (ny > 0 && nx > 0) ? ny * nx * K : 0;
  // candidate t = ((y - ylo) * nx + (x - xlo)) * K + k.  Quotients without
  // integer division: K a power of two takes shifts; yx / nx (yx < 2^16, small
  // nx) is floor((yx + 1/2) / nx) from a float reciprocal, exact at these sizes.
  const bool kpow2 = (K & (K - 1)) == 0;
  const int lk = 31 - __clz(K);
  const float rnx = 1.0f / (float)max(nx, 1);
  // window index idx < d2^2 (< 2^22): idx / d2 = floor((idx + 1/2) / d2) from
  // the float reciprocal, exact (the quotient is at least 1/(2 d2) from an
  // integer, far above the rounding error)
  const float rd2 = 1.0f / (float)d2;
  auto split = [&](int t, int& yx, int& k) {
    if (kpow2) {
      yx = t >> lk;
      k = t & (K - 1);
    } else {
      yx = t / K;
      k = t - yx * K;
    }
  };
  auto tile_of = [&](int yx, int& y, int& x) {
    const int qy = (int)(((float)yx + 0.5f) * rnx);
    y = ylo + qy;
    x = xlo + (yx - imul(qy, nx));
  };
  // its reference tile (y, x), slot k, slot index and weight
  struct Cand {
    int y, x, k;
    uint32_t sidx;
    float w;
  };
  auto fetch = [&](int t, Cand& c) {
    c.y = c.x = c.k = 0;
    c.sidx = 0;
    c.w = 0.f;
    if (t < ncand) {
      int yx;
      split(t, yx, c.k);
      tile_of(yx, c.y, c.x);
      const int g = c.y + imul(Mt, c.x);
      float wg;
      if constexpr (BUF) {
        c.sidx = __float_as_uint(buf_load(rS, 4u * (uint32_t)(imul(K, g) + c.k)));
        wg = buf_load(rW, 4u * (uint32_t)g);
      } else {
        c.sidx = S[(size_t)K * g + c.k];
        wg = wgt[g];
      }
      // step 2 uses 1/weight (w / norm(filt)^2): divided here, once per
      // candidate lane, rather than once per hit
      c.w = STEP == 1 ? wg : 1.0f / wg;
    }
  };
  Cand cn;
  fetch(lane, cn);
  for (int base = 0; base < ncand; base += 64) {
    const int t = base + lane;
    const Cand cc0 = cn;
    const float wv = cc0.w;
    if (base + 64 < ncand) fetch(t + 64, cn);
    int blk = 0, bob = 0;
    uint32_t brc = 0;
    bool hit = false;
    if (t < ncand) {
      const uint32_t idx = cc0.sidx - 1u;
      const int bq = (int)(((float)idx + 0.5f) * rd2);
      const int br = cc0.y * D + ((int)idx - imul(bq, d2)) - s - R0;  // block origin relative to the tile
      const int bc = cc0.x * D + bq - s - C0;
      hit = br < AT && br + D > 0 && bc < AT && bc + D > 0;
      blk = cc0.k + imul(K, cc0.y + imul(Mt, cc0.x));
      brc = ((uint32_t)br & 0xffffu) | ((uint32_t)bc << 16);  // 16-bit two's complement halves
      bob = br + AS * bc;                                      // its LDS offset (float2 units)
    }
    // this round's hits, in candidate order, listed in LDS so that lane j
    // holds the j-th hit's block, packed origin, weight and accumulator offset
    const uint64_t mask = __ballot(hit);
    const int nh = __popcll(mask);
    if (nh == 0) continue;
    {
      // BUF: the block's byte offset in the blocks array (below 2^31).  The
      // other lanes fill entries nh..63 with the first hit's block, which a
      // batch's loads past the last hit then read again from the cache.
      const int bx = BUF ? (int)(4u * (uint32_t)DD * (uint32_t)blk) : blk;
      const int b0 = __builtin_amdgcn_readlane(bx, __builtin_ctzll(mask));
      const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
      hl[wave][hit ? rank : nh + (lane - rank)] = make_int4(hit ? bx : b0, (int)brc, __float_as_int(wv), bob);
    }
    wave_sync();
    const int hblk = hl[wave][lane].x;
    for (int j0 = 0; j0 < nh; j0 += NB) {
      // all NB loads issued (past the last hit: some block, not used), so
      // each hit waits only for its own block
      float val[NB][PER];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if constexpr (BUF) {
          // past the last hit: the first hit's block (entries nh..63) or,
          // wrapped, a hit of this round
          const uint32_t bo = (uint32_t)__builtin_amdgcn_readlane(hblk, (j0 + j) & 63);
#pragma unroll
          for (int i = 0; i < PER; ++i)
            val[j][i] = (lane + 64 * i < DD) ? buf_load(rB, 4u * (uint32_t)(lane + 64 * i), bo) : 0.f;
        } else {
          const int b = __builtin_amdgcn_readlane(hblk, (j0 + j) & 63);
          const float* bp = blocks + (size_t)DD * b;
#pragma unroll
          for (int i = 0; i < PER; ++i) val[j][i] = (lane + 64 * i < DD) ? bp[lane + 64 * i] : 0.f;
        }
      }
      // the batch's hit fields by broadcast LDS reads (one address in every
      // lane), all issued before the first is used; entries past the last
      // hit are read and not used
      int4 hh[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) hh[j] = hl[wave][j0 + j];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j0 + j >= nh) break;
        const uint32_t rc = (uint32_t)hh[j].y;
        const float wg = __int_as_float(hh[j].z);  // step 2: 1/weight
        const int ob = hh[j].w;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          // pixel (r + u, c + v) in the tile: both 16-bit sums (wrapping) below AT
          const us2 t = __builtin_bit_cast(us2, puv[i]) + __builtin_bit_cast(us2, rc);
          if (lane + 64 * i < DD && (__builtin_bit_cast(uint32_t, t) & kOutsideAT) == 0u) {
            float2 a = pa[i][ob];
            if (STEP == 1) {
              const float c = (wg * val[j][i]) * mk[i];  // weight*blocks(:,:,i).*mask in single
              a.x = a.x + c;
              a.y = a.y + wg * mk[i];
            } else {
              const float wt = mk[i] * wg;  // w / norm(filt)^2
              a.x = a.x + wt * val[j][i];
              a.y = a.y + wt;
            }
            pa[i][ob] = a;
          }
        }
        // no barrier between blocks: LDS operations of one wave execute in
        // program order, and the compiler keeps these possibly-aliasing
        // read-modify-writes in order, so each pixel sums in (y, x, k) order
      }
    }
    wave_sync();  // the hit list read before the next round rewrites it
  }
  wave_sync();
  for (int e = lane; e < AT * AT; e += 64) {
    const int r = e % AT, c = e / AT;
    const int pr = R0 + r, pc = C0 + c;
    if (pr >= M || pc >= N) continue;
    const size_t o = (size_t)pmod_i(pr - spin, M) + (size_t)M * pmod_i(pc - spin, N);
    const float2 a = ac[r + AS * c];
    U[o] = a.x;
    W[o] = a.y;
  }
}

int launch_aggregate_m(int step, const float* blocks, const float* wgt, const uint32_t* S, const float* masks, int M,
                       int N, int d, int d2, int K, int spin, float* U, float* W, hipStream_t st) {
  if (M % d || N % d) return fail(GBM3D_ERR_ARG, "aggregate: size");
  const int tiles = ((M + AT - 1) / AT) * ((N + AT - 1) / AT);
  dim3 grid((tiles + 3) / 4);
  // the blocks array: K slices of M x N floats.  GBM3D_AGG_POINTER=1 keeps
  // the 64-bit pointer loads below 2 GB too, for cross-checking the two.
  const char* ptr_env = getenv("GBM3D_AGG_POINTER");
  const bool buf = (size_t)K * M * N * sizeof(float) < ((size_t)1 << 31) && (size_t)M * N < ((size_t)1 << 26) &&
                   !(ptr_env && ptr_env[0] == '1');
#define GBM3D_AGG_LAUNCH(S_, DD, B_)                                                                         \
  hipLaunchKernelGGL((aggregate_batched_kernel<S_, DD, B_>), grid, dim3(256), 0, st, blocks, wgt, S, masks, M, N, \
                     d2, K, spin, U, W)
#define GBM3D_AGG_CASE(DD)                                                                                   \
  if (d == DD) {                                                                                             \
    if (step == 1) {                                                                                         \
      if (buf)                                                                                               \
        GBM3D_AGG_LAUNCH(1, DD, true);                                                                       \
      else                                                                                                   \
        GBM3D_AGG_LAUNCH(1, DD, false);                                                                      \
    } else {                                                                                                 \
      if (buf)                                                                                               \
        GBM3D_AGG_LAUNCH(2, DD, true);                                                                       \
      else                                                                                                   \
        GBM3D_AGG_LAUNCH(2, DD, false);                                                                      \
    }                                                                                                        \
    return check_launch("aggregate");                                                                        \
  }
  GBM3D_AGG_CASE(4)
  GBM3D_AGG_CASE(8)
  GBM3D_AGG_CASE(16)
  GBM3D_AGG_CASE(32)
#undef GBM3D_AGG_CASE
#undef GBM3D_AGG_LAUNCH
  return fail(GBM3D_ERR_UNSUPPORTED, "aggregate: block size must be 4, 8, 16 or 32");
}

// --------------------------------------------------------- normalisation
// out(r,c) = sum_s U_s(r,c) / (sum_s W_s(r,c) + sum_t Wx_t(r,c)), double
// arithmetic (sum(U,3)./sum(W,3) of source/GBM3D.m:57,82), cropped to m x n.
// A thread takes 4 consecutive rows of one column; vec: 16-byte loads and
// stores (M, m, m2 multiples of 4 and 16-byte aligned arrays, checked on the
// host).  The per-element sums are those of the scalar form.
__global__ void normalize_kernel(const float* __restrict__ U, int nU, const float* __restrict__ W, int nW,
                                 const float* __restrict__ Wx, int nWx, int M, int N, int m, int n,
                                 float* __restrict__ out, float* __restrict__ out2, int m2, int n2, int vec) {
  const int r0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  const int c = blockIdx.y;
  if (r0 >= m) return;
  const bool v4 = vec && r0 + 3 < m;
  const size_t MN = (size_t)M * N, o = (size_t)r0 + (size_t)M * c;
  double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
  if (v4) {
    auto add4 = [&](double (&acc)[4], const float* p) {
      const float4 x = *reinterpret_cast<const float4*>(p + o);
      acc[0] += (double)x.x;
      acc[1] += (double)x.y;
      acc[2] += (double)x.z;
      acc[3] += (double)x.w;
    };
    for (int s = 0; s < nU; ++s) add4(a, U + MN * s);
    for (int s = 0; s < nW; ++s) add4(b, W + MN * s);
    for (int s = 0; s < nWx; ++s) add4(b, Wx + MN * s);
  } else {
    auto add1 = [&](double (&acc)[4], const float* p) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (r0 + k < m) acc[k] += (double)p[o + k];
    };
    for (int s = 0; s < nU; ++s) add1(a, U + MN * s);
    for (int s = 0; s < nW; ++s) add1(b, W + MN * s);
    for (int s = 0; s < nWx; ++s) add1(b, Wx + MN * s);
  }
  float v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = (float)(a[k] / b[k]);
  float* d1 = out + (size_t)r0 + (size_t)m * c;
  if (v4) {
    *reinterpret_cast<float4*>(d1) = make_float4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (r0 + k < m) d1[k] = v[k];
  }
  if (out2 != nullptr && c < n2) {
    float* d2 = out2 + (size_t)r0 + (size_t)m2 * c;
    if (vec && r0 + 3 < m2) {
      *reinterpret_cast<float4*>(d2) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (r0 + k < m2) d2[k] = v[k];
    }
  }
}

int launch_normalize(const float* U, int nU, const float* W, int nW, const float* Wextra, int nWextra, int M, int N,
                     int m, int n, float* out, hipStream_t st, float* out2, int m2, int n2) {
  if (m > M || n > N) return fail(GBM3D_ERR_ARG, "normalize: crop");
  if (out2 != nullptr && (m2 > m || n2 > n)) return fail(GBM3D_ERR_ARG, "normalize: second crop");
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const int vec = M % 4 == 0 && m % 4 == 0 && al(U) && al(W) && (nWextra == 0 || al(Wextra)) && al(out) &&
                  (out2 == nullptr || (m2 % 4 == 0 && al(out2)));
  dim3 grid((m + 1023) / 1024, n);
  hipLaunchKernelGGL(normalize_kernel, grid, dim3(256), 0, st, U, nU, W, nW, Wextra, nWextra, M, N, m, n, out, out2,
                     m2, n2, vec);
  return check_launch("normalize");
}

}  // namespace gbm3d
