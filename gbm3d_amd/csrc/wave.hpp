// wave.hpp -- wavefront (64-lane) primitives for gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gbm3d {

__device__ __forceinline__ int pmod_i(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map dispatch index b (of nb) to a logical index so that XCD x works on
// one contiguous range of logical indices; neighbouring tiles, whose halos
// overlap, then share an L2.  A bijection for any nb.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Tile (bx, by) of an nbx x nby tile grid for logical index lb, in 8 x 8
// super-tiles (bx fastest inside one, super-tiles down the bx axis first) when
// both counts are multiples of 8, else bx fastest.  The 64 workgroups an XCD
// runs at once then cover a square-ish region whose halos they share in its
// L2, instead of two full tile columns.  A bijection on [0, nbx nby).
__device__ __forceinline__ void tile_order8(int lb, int nbx, int nby, int& bx, int& by) {
  if ((nbx & 7) == 0 && (nby & 7) == 0) {
    const int st = lb >> 6, w = lb & 63, sx = nbx >> 3;
    bx = (st % sx) * 8 + (w & 7);
    by = (st / sx) * 8 + (w >> 3);
  } else {
    bx = lb % nbx;
    by = lb / nbx;
  }
}

// Make LDS writes of this wave visible to its other lanes (a wave executes in
// lockstep; the fences stop the compiler from moving LDS accesses across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// This lane's lane index (0..63).
__device__ __forceinline__ int lane_id() {
  return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// The value of lane (this lane ^ m), the whole wave taking part: one
// ds_bpermute on an address formed once per m (no clamp to a sub-width).
__device__ __forceinline__ uint32_t shfl_xor_u32(uint32_t v, int m) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((lane_id() ^ m) << 2, (int)v);
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = shfl_xor_u32((uint32_t)v, m);
  uint32_t hi = shfl_xor_u32((uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_xor_any(uint32_t v, int m) { return shfl_xor_u32(v, m); }
__device__ __forceinline__ uint64_t shfl_xor_any(uint64_t v, int m) { return shfl_xor_u64(v, m); }

// Raw buffer resource over `bytes` bytes from `base` (word 3: 32-bit data
// format); a load through it takes the base in scalar registers, a 32-bit
// byte offset in a vector register and, optionally, a uniform byte offset in
// a scalar register (no per-load 64-bit address arithmetic).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const float* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes, uint32_t soff_bytes = 0) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off_bytes, soff_bytes, 0));
}
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off_bytes, float v, uint32_t soff_bytes = 0) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off_bytes, soff_bytes, 0);
}

// diff(x, 3) of four consecutive samples in the order diff() takes it
// (getBM3DAggWeights.m:15-19), shared by the step-1 weights wherever they
// are computed so that every place gives the same bits
__device__ __forceinline__ float third_diff(float x0, float x1, float x2, float x3) {
  const float a = x1 - x0, b = x2 - x1, c = x3 - x2;
  const float e = b - a, f = c - b;
  return f - e;
}

// a + b in a scalar register, opaque to the optimiser (which would otherwise
// hoist every multiple of a loop-invariant stride and keep them all live)
__device__ __forceinline__ uint32_t sadd_u32(uint32_t a, uint32_t b) {
  uint32_t r;
  asm volatile("s_add_u32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
  return r;
}

// Ascending bitonic sort of one value per lane across the 64 lanes.
// N = 32: each half of the wave sorted on its own, the lower ascending.
template <class T, int N = 64>
__device__ __forceinline__ T wave_sort64(T v, int lane) {
  static_assert(N == 32 || N == 64, "sort width");
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const T p = shfl_xor_any(v, j);
      // this lane keeps the smaller of the pair where its ascending/lower
      // flags agree, else the larger: one compare and a select (equal values
      // are interchangeable)
      const bool want_min = ((lane & k) == 0) == ((lane & j) == 0);
      v = ((v < p) == want_min) ? v : p;
    }
  }
  return v;
}

// The K-th smallest (1 <= K <= 64) of the 64 lanes' values, the same in every
// lane: the largest t with fewer than K values below t, built from the top
// bit down with one ballot per bit.
__device__ __forceinline__ uint32_t wave_kth_smallest(uint32_t v, int K) {
  uint32_t t = 0;
#pragma unroll
  for (int b = 31; b >= 0; --b) {
    const uint32_t c = t | (1u << b);
    if (__popcll(__ballot(v < c)) < K) t = c;
  }
  return t;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    uint64_t p = shfl_xor_u64(v, o);
    v = p < v ? p : v;
  }
  return v;
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __uint_as_float(shfl_xor_u32(__float_as_uint(v), o));
  return v;
}

}  // namespace gbm3d
