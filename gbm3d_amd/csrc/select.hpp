// select.hpp -- candidate keys and the exact wave-wide top-K shared by the
// windowed matchers (k_match.hip) and the whole-image matcher (k_gmatch.hip).
#pragma once

#include "wave.hpp"

namespace gbm3d {

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

// The ordered bits of cand_key's high word, and back.
__device__ __forceinline__ uint32_t ord_bits(float dist) { return (uint32_t)(cand_key(dist, 0) >> 32); }
__device__ __forceinline__ float ord_to_f32(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
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
  const uint32_t theta = wave_kth_smallest(lmin, K);
  int total = 0;
#pragma unroll
  for (int i = 0; i < NL; ++i) total += __popcll(__ballot((uint32_t)(key[i] >> 32) <= theta));
  if (total <= 64) {
    // compacted slot by slot (the sort below depends on the set alone)
    uint32_t base = 0;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const bool in = (uint32_t)(key[i] >> 32) <= theta;
      const uint64_t m = __ballot(in);
      if (in) scr[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = key[i];
      base += (uint32_t)__popcll(m);
    }
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

}  // namespace gbm3d
