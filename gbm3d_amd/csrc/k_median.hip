// k_median.hip -- the 'median' step-1 filter: Vavg = medfilt3(Volume, [3 3 3])
// (source/denoiseAndAggregateGPU.m:49-52).
//
// Each output is the median of the 27 values of its 3x3x3 neighbourhood in the
// mosaic volume (M x N x K); outside the volume the nearest edge value is
// repeated (MATLAB's default 'symmetric' padding, which for a window of 3 is
// the edge value).  One thread per (row, column) walks the K slices, keeping
// the previous and current 3x3 planes in registers, so each slice is read
// once.  The median is the 14th smallest of 27, selected with the forgetful
// scheme: start from 15 values, repeatedly move the minimum and the maximum
// to the ends with compare-exchange passes, drop both and take in the next
// value; the median of the last three is the result.  Exact (a selection of
// inputs, no arithmetic).
#include "internal.hpp"
#include "wave.hpp"

namespace gbm3d {

namespace {

__device__ __forceinline__ void cswap(float& a, float& b) {
  const float lo = fminf(a, b), hi = fmaxf(a, b);
  a = lo;
  b = hi;
}

__device__ __forceinline__ float median27(const float (&v)[27]) {
  float s[15];
#pragma unroll
  for (int i = 0; i < 15; ++i) s[i] = v[i];
#pragma unroll
  for (int it = 0; it < 12; ++it) {
    const int n = 15 - it;
#pragma unroll
    for (int i = 0; i < n - 1; ++i) cswap(s[i], s[i + 1]);  // maximum to s[n-1]
#pragma unroll
    for (int i = n - 2; i > 0; --i) cswap(s[i - 1], s[i]);  // minimum of the rest to s[0]
#pragma unroll
    for (int i = 0; i < n - 2; ++i) s[i] = s[i + 1];
    s[n - 2] = v[15 + it];
  }
  return fmaxf(fminf(s[0], s[1]), fminf(fmaxf(s[0], s[1]), s[2]));
}

// Offset of voxel (r, c, k): mosaic (gd == 0) or group-contiguous gd x gd blocks.
__device__ __forceinline__ size_t out_off(int r, int c, int k, int M, int N, int K, int gd) {
  if (gd == 0) return (size_t)r + (size_t)M * ((size_t)c + (size_t)N * k);
  const size_t g = (size_t)(r / gd) + (size_t)(M / gd) * (c / gd);
  return (size_t)(r % gd) + (size_t)gd * ((size_t)(c % gd) + (size_t)gd * ((size_t)k + (size_t)K * g));
}

}  // namespace

__global__ __launch_bounds__(256) void median3_kernel(const float* __restrict__ V, int M, int N, int K, int gd,
                                                      float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)M * N) return;
  const int r = (int)(e % M), c = (int)(e / M);
  int rr[3], cc[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    rr[i] = min(max(r + i - 1, 0), M - 1);
    cc[i] = min(max(c + i - 1, 0), N - 1);
  }
  const size_t MN = (size_t)M * N;
  auto plane = [&](int k, float (&p)[9]) {
    const float* Vk = V + MN * (size_t)k;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) p[i + 3 * j] = Vk[(size_t)rr[i] + (size_t)M * cc[j]];
  };
  float prev[9], cur[9], nxt[9];
  plane(0, cur);
#pragma unroll
  for (int i = 0; i < 9; ++i) prev[i] = cur[i];  // slice -1 repeats slice 0
  for (int k = 0; k < K; ++k) {
    if (k + 1 < K)
      plane(k + 1, nxt);
    else
#pragma unroll
      for (int i = 0; i < 9; ++i) nxt[i] = cur[i];  // slice K repeats slice K-1
    float v[27];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      v[i] = prev[i];
      v[9 + i] = cur[i];
      v[18 + i] = nxt[i];
    }
    out[out_off(r, c, k, M, N, K, gd)] = median27(v);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      prev[i] = cur[i];
      cur[i] = nxt[i];
    }
  }
}

int launch_median3(const float* V, int M, int N, int K, int gd, float* out, hipStream_t st) {
  if (M < 1 || N < 1 || K < 1) return fail(GBM3D_ERR_ARG, "median3: empty volume");
  if (gd && (M % gd || N % gd)) return fail(GBM3D_ERR_ARG, "median3: group layout block size");
  const size_t n = (size_t)M * N;
  hipLaunchKernelGGL(median3_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, V, M, N, K, gd, out);
  return check_launch("median3");
}

}  // namespace gbm3d
