// Device-side primitives shared by the gfx950 kernels: wave64 ballot/mbcnt
// helpers, canonical ordinal keys, block combines and the register-resident
// stable LSD radix sort of (key, column index) pairs used by every ranking
// kernel.
//
// Every ranking kernel runs 512 threads = 8 wave64 waves. A segment of n
// items is split into 8 contiguous per-wave spans whose length is a multiple
// of 64; lane l of a wave holds items lo + 64*s + l (s = 0..E-1) in
// registers, so a 64-bit ballot of step s is exactly one bitmap word.
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

// Per-wave contiguous span [lo, hi) of a segment of n items (64-aligned lo)
// and its number of 64-item steps.
struct Span {
  int lo, hi, steps;
};
__device__ __forceinline__ Span span_of(int n) {
  const int C = ((n + kThreads - 1) / kThreads) * 64;
  Span s;
  s.lo = min(wave_id() * C, n);
  s.hi = min(s.lo + C, n);
  s.steps = (s.hi - s.lo + 63) >> 6;
  return s;
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

// Lanes of the calling wave whose 8-bit digit equals the caller's, among the
// lanes where `act` holds.
__device__ __forceinline__ uint64_t match8(uint32_t d, bool act) {
  uint64_t M = ballot(act);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = ballot(bit);
    M &= bit ? bb : ~bb;
  }
  return M;
}

// Stable LSD radix sort (8-bit digits) of the n (key, column) pairs held in
// (kbuf, ibuf)[0, n). On return the sorted pairs are in (kbuf, ibuf) and in
// the caller's registers key[s]/idx[s] at sorted position lo + 64*s + lane.
//
// Per pass, per wave, per 64-item step: every active lane ORs its lane bit
// into its digit's slot of the wave's 256-entry match table (one 64-bit LDS
// atomic), reads the slot back (the lanes sharing its digit) and the lowest
// of those lanes (the group leader) clears the slot and advances the wave's
// running count of that digit with one returning LDS add; the group reads
// the returned base through a lane permute. LDS executes one wave's
// instructions in order, so successive steps chain without the wave waiting
// on any of them; ranks are stable inside a step (lane order), across steps
// (instruction order) and across waves (exclusive prefix of the per-wave
// digit counts). A pass whose digit is equal on all n items is the identity
// and is skipped.
//
// hist: kWaves*kBins words; tab: kWaves*kBins 64-bit slots, all zero on
// entry and left zero on exit; flags: one word per pass (zeroed by the caller
// before a barrier that precedes this call). Entry and exit are
// block-synchronous. Requires ceil(n/512) <= E.
template <typename K, int E>
__device__ __forceinline__ void radix_sort_regs(K (&key)[E], uint32_t (&idx)[E], int n, K* kbuf,
                                                uint16_t* ibuf, uint32_t* hist, uint64_t* tab,
                                                uint32_t* scan_tmp, uint32_t* flags) {
  const int lane = lane_id();
  const int w = wave_id();
  const Span sp = span_of(n);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int p = sp.lo + s * 64 + lane;
    const bool a = (s < sp.steps) && (p < sp.hi);
    key[s] = a ? kbuf[p] : static_cast<K>(~static_cast<K>(0));
    idx[s] = a ? static_cast<uint32_t>(ibuf[p]) : 0u;
  }
  if (n <= 1) return;
  uint32_t* hw = hist + w * kBins;
  uint64_t* tw = tab + w * kBins;
  const uint64_t lanebit = 1ull << lane;
  int pass = 0;
  for (int shift = 0; shift < static_cast<int>(sizeof(K) * 8); shift += 8, ++pass) {
#pragma unroll
    for (int d = lane; d < kBins; d += 64) hw[d] = 0u;
    uint32_t loc[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      loc[s] = 0u;
      if (s < sp.steps) {
        const int p = sp.lo + s * 64 + lane;
        const bool a = p < sp.hi;
        const uint32_t d = static_cast<uint32_t>(key[s] >> shift) & 255u;
        uint64_t M = 0ull;
        if (a) {
          atomicOr(reinterpret_cast<unsigned long long*>(&tw[d]),
                   static_cast<unsigned long long>(lanebit));
          M = tw[d];
        }
        const uint32_t r = below(M);
        uint32_t base = 0u;
        if (a && r == 0u) {
          base = atomicAdd(&hw[d], static_cast<uint32_t>(__popcll(M)));
          tw[d] = 0ull;
        }
        const int leader = a ? (__ffsll(static_cast<unsigned long long>(M)) - 1) : lane;
        base = __shfl(base, leader, 64);
        loc[s] = base + r;
      }
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
      if (tot == static_cast<uint32_t>(n)) flags[pass] = 1u;
    }
    __syncthreads();
    if (flags[pass] != 0u) continue;
#pragma unroll
    for (int s = 0; s < E; ++s) {
      if (s < sp.steps) {
        const int p = sp.lo + s * 64 + lane;
        if (p < sp.hi) {
          const uint32_t d = static_cast<uint32_t>(key[s] >> shift) & 255u;
          const uint32_t pos = hw[d] + loc[s];
          kbuf[pos] = key[s];
          ibuf[pos] = static_cast<uint16_t>(idx[s]);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int p = sp.lo + s * 64 + lane;
      if ((s < sp.steps) && (p < sp.hi)) {
        key[s] = kbuf[p];
        idx[s] = static_cast<uint32_t>(ibuf[p]);
      }
    }
  }
  __syncthreads();
}

}  // namespace fcd
