// Device-side primitives shared by the gfx950 kernels: wave64 ballot/mbcnt
// helpers, canonical ordinal keys, block combines and the LDS-resident stable
// LSD radix sort of (key, column index) pairs used by every ranking kernel.
//
// Every kernel that includes this header runs 512 threads = 8 wave64 waves.
// A segment of n items is split into 8 contiguous per-wave spans whose length
// is a multiple of 64, so "step" s of a span covers the 64 consecutive items
// [64*s, 64*s+64) and a 64-bit ballot of a step is exactly one bitmap word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fcd {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kBins = 256;           // 8-bit radix digits
constexpr uint32_t kQuietNaN = 0x7fc00000u;

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63u); }
__device__ __forceinline__ int wave_id() { return static_cast<int>(threadIdx.x >> 6); }

__device__ __forceinline__ uint64_t ballot(bool p) {
  return static_cast<uint64_t>(__ballot(p ? 1 : 0));
}

// Number of bits of m set in lanes strictly below the calling lane.
__device__ __forceinline__ uint32_t below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Bits at positions <= b (b in [0,63]); (2<<63) wraps to 0 so b=63 -> all ones.
__device__ __forceinline__ uint64_t mask_le(int b) { return (2ull << b) - 1ull; }
__device__ __forceinline__ uint64_t mask_lt(int b) { return (1ull << b) - 1ull; }

// Canonical ordinal key: ascending unsigned key order == ascending numeric
// order, +0.0 and -0.0 fold to one key (reference semantics:
// src/cross_sectional_rank.cu:35-39, src/rolling_ic.cu:34-38). Callers only
// form keys of finite values.
__device__ __forceinline__ uint32_t ord_key(float v) {
  uint32_t b = (v == 0.0f) ? 0u : __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint64_t ord_key(double v) {
  uint64_t b = (v == 0.0) ? 0ull : static_cast<uint64_t>(__double_as_longlong(v));
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

template <typename V> struct KeyOf;
template <> struct KeyOf<float> { using type = uint32_t; };
template <> struct KeyOf<double> { using type = uint64_t; };

// Per-wave contiguous span [lo, hi) of a segment of n items (64-aligned lo).
__device__ __forceinline__ void wave_span(int n, int& lo, int& hi) {
  const int C = ((n + kThreads - 1) / kThreads) * 64;
  lo = min(wave_id() * C, n);
  hi = min(lo + C, n);
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int wave_incl_max(int v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (l >= o) v = max(v, t);
  }
  return v;
}

__device__ __forceinline__ int wave_incl_min(int v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (l >= o) v = min(v, t);
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), o, 64);
    uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), o, 64);
    v += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  return v;
}

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Exclusive prefix over waves of a wave-uniform value; also returns the block
// total. scratch needs kWaves words. Contains two block barriers.
__device__ __forceinline__ uint32_t block_wave_prefix(uint32_t v_wave, uint32_t* scratch,
                                                      uint32_t& total) {
  if (lane_id() == 0) scratch[wave_id()] = v_wave;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    uint32_t x = scratch[w];
    pre += (w < wave_id()) ? x : 0u;
    tot += x;
  }
  total = tot;
  __syncthreads();
  return pre;
}

// Stable LSD radix sort of n (key, idx) pairs held in (ka, ia), 8-bit digits,
// entirely in LDS. Stability: each wave owns a contiguous span and walks it in
// order; inside a 64-item step the rank among equal digits comes from an
// 8-ballot match mask counted in lane order, and wave w's per-digit base is
// the digit's global offset plus the counts of waves < w. A pass whose digit
// is the same for all n items is the identity and is skipped.
// hist: kWaves*kBins words; flag: 1 word. Returns 0 if the sorted result is in
// (ka, ia), 1 if in (kb, ib). Entry and exit are block-synchronous.
template <typename K>
__device__ __forceinline__ int lds_radix_sort(K* ka, uint16_t* ia, K* kb, uint16_t* ib,
                                              uint32_t* hist, uint32_t* scan_tmp,
                                              uint32_t* flag, int n) {
  const int lane = lane_id();
  const int w = wave_id();
  int lo, hi;
  wave_span(n, lo, hi);
  uint32_t* hw = hist + w * kBins;
  int cur = 0;
  if (n <= 1) return 0;
  for (int shift = 0; shift < static_cast<int>(sizeof(K) * 8); shift += 8) {
    K* ks = cur ? kb : ka;
    uint16_t* is = cur ? ib : ia;
    K* kd = cur ? ka : kb;
    uint16_t* id = cur ? ia : ib;
#pragma unroll
    for (int d = lane; d < kBins; d += 64) hw[d] = 0u;
    if (threadIdx.x == 0) *flag = 0u;
    for (int s = lo + lane; s < hi; s += 64) {
      const uint32_t d = static_cast<uint32_t>(ks[s] >> shift) & 255u;
      atomicAdd(&hw[d], 1u);
    }
    __syncthreads();
    uint32_t tot = 0, incl = 0;
    if (threadIdx.x < kBins) {
      const int d = threadIdx.x;
      uint32_t run = 0;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) {
        uint32_t h = hist[ww * kBins + d];
        hist[ww * kBins + d] = run;
        run += h;
      }
      tot = run;
      incl = wave_incl_sum(tot);
      if (lane == 63) scan_tmp[w] = incl;
    }
    __syncthreads();
    if (threadIdx.x < kBins) {
      uint32_t pre = 0;
      for (int ww = 0; ww < w; ++ww) pre += scan_tmp[ww];
      const uint32_t base = pre + incl - tot;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) hist[ww * kBins + threadIdx.x] += base;
      if (tot == static_cast<uint32_t>(n)) *flag = 1u;
    }
    __syncthreads();
    const bool skip = (*flag != 0u);
    if (!skip) {
      for (int s0 = lo; s0 < hi; s0 += 64) {
        const int s = s0 + lane;
        const bool act = s < hi;
        const K key = act ? ks[s] : K(0);
        const uint16_t ix = act ? is[s] : uint16_t(0);
        const uint32_t d = static_cast<uint32_t>(key >> shift) & 255u;
        uint64_t M = ballot(act);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t bb = ballot(bit);
          M &= bit ? bb : ~bb;
        }
        const uint32_t rk = below(M);
        const uint32_t base = hw[d];
        if (act) {
          kd[base + rk] = key;
          id[base + rk] = ix;
          if (rk == 0) hw[d] = base + static_cast<uint32_t>(__popcll(M));
        }
      }
      cur ^= 1;
    }
    __syncthreads();
  }
  return cur;
}

}  // namespace fcd
