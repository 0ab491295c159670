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

// Word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle): loads past the resource's byte count return 0 and stores past
// it are dropped. The range check covers the lane offset (the scalar offset
// is kept 0).
constexpr int kBufWord3 = 0x00020000;

// A raw buffer over `bytes` bytes from p (bytes < 2^31).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_buffer(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, static_cast<int>(bytes), kBufWord3);
}

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & 63u); }
// The wave's index in the workgroup, as a scalar (uniform) value, so that
// what derives from it (spans, step counts) stays in SGPRs.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
}

// The predicate's lane mask as is (no 0/1 round trip through a register).
__device__ __forceinline__ uint64_t ballot(bool p) {
  return static_cast<uint64_t>(__builtin_amdgcn_ballot_w64(p));
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

// ---- cross-lane primitives; every lane of the wave must be active --------

// v of lane (lane ^ O), O a power of two: ds_swizzle in xor mode inside each
// 32-lane half (no address arithmetic), ds_bpermute across the halves.
template <int O>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
  static_assert(O >= 1 && O <= 32 && (O & (O - 1)) == 0, "O must be a power of two <= 32");
  if constexpr (O < 32) {
    return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x1F | (O << 10)));
  } else {
    return static_cast<uint32_t>(
        __builtin_amdgcn_ds_bpermute((lane_id() ^ 32) << 2, static_cast<int>(v)));
  }
}
template <int O>
__device__ __forceinline__ uint64_t lane_xor(uint64_t v) {
  const uint32_t lo = lane_xor<O>(static_cast<uint32_t>(v));
  const uint32_t hi = lane_xor<O>(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Whole-wave reduction with the xor butterfly 32, 16, ..., 1; every lane
// gets the result. T is uint32_t or uint64_t.
template <typename T, typename Op>
__device__ __forceinline__ T wave_allreduce(T v, Op op) {
  v = op(v, lane_xor<32>(v));
  v = op(v, lane_xor<16>(v));
  v = op(v, lane_xor<8>(v));
  v = op(v, lane_xor<4>(v));
  v = op(v, lane_xor<2>(v));
  v = op(v, lane_xor<1>(v));
  return v;
}

// DPP move of v (row_shr / row_bcast control Ctrl, rows RowMask); lanes
// whose source is outside the row, or whose row is masked off, get ident.
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_or(uint32_t v, uint32_t ident) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(ident), static_cast<int>(v), Ctrl, RowMask, 0xF, false));
}

// Inclusive scan over the wave's 64 lanes in lane order: Kogge-Stone inside
// each 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row 1 and
// row 2's into row 3 (row_bcast:15), then lane 31 into rows 2 and 3
// (row_bcast:31).
template <typename Op>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t ident, Op op) {
  v = op(v, dpp_or<0x111, 0xF>(v, ident));
  v = op(v, dpp_or<0x112, 0xF>(v, ident));
  v = op(v, dpp_or<0x114, 0xF>(v, ident));
  v = op(v, dpp_or<0x118, 0xF>(v, ident));
  v = op(v, dpp_or<0x142, 0xA>(v, ident));
  v = op(v, dpp_or<0x143, 0xC>(v, ident));
  return v;
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
  return wave_incl_scan(v, 0u, [](uint32_t x, uint32_t y) { return x + y; });
}
__device__ __forceinline__ int wave_incl_max(int v) {
  return static_cast<int>(wave_incl_scan(static_cast<uint32_t>(v), 0x80000000u,
                                         [](uint32_t x, uint32_t y) {
                                           return static_cast<uint32_t>(
                                               max(static_cast<int>(x), static_cast<int>(y)));
                                         }));
}
__device__ __forceinline__ int wave_incl_min(int v) {
  return static_cast<int>(wave_incl_scan(static_cast<uint32_t>(v), 0x7FFFFFFFu,
                                         [](uint32_t x, uint32_t y) {
                                           return static_cast<uint32_t>(
                                               min(static_cast<int>(x), static_cast<int>(y)));
                                         }));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  return wave_allreduce(v, [](uint64_t x, uint64_t y) { return x + y; });
}
__device__ __forceinline__ int wave_min_i(int v) {
  return static_cast<int>(wave_allreduce(static_cast<uint32_t>(v), [](uint32_t x, uint32_t y) {
    return static_cast<uint32_t>(min(static_cast<int>(x), static_cast<int>(y)));
  }));
}
__device__ __forceinline__ int wave_max_i(int v) {
  return static_cast<int>(wave_allreduce(static_cast<uint32_t>(v), [](uint32_t x, uint32_t y) {
    return static_cast<uint32_t>(max(static_cast<int>(x), static_cast<int>(y)));
  }));
}
// lane 63's value, wave-uniform
__device__ __forceinline__ uint32_t last_lane(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
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

// match8 over the full wave. Bit b of the digit replicated to a word (t = 0
// or ~0) XOR the ballot of that bit marks the lanes that disagree on it; the
// OR over the 8 bits is accumulated with one v_bitop3 per half and bit
// (truth table 0xF6 = S0 | (S1 ^ S2)), and its complement is the match.
__device__ __forceinline__ uint64_t match8_all(uint32_t d) {
  uint32_t xlo = 0u, xhi = 0u;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(d), b, 1));
    // ballot of t != 0 taken on t itself (one compare per bit)
    const uint64_t bb = __builtin_amdgcn_uicmp(t, 0u, 33 /* ICMP_NE */);
    xlo = __builtin_amdgcn_bitop3_b32(xlo, static_cast<uint32_t>(bb), t, 0xF6);
    xhi = __builtin_amdgcn_bitop3_b32(xhi, static_cast<uint32_t>(bb >> 32), t, 0xF6);
  }
  // last bit: ~(S0 | (S1 ^ S2)) = 0x09 gives the match itself
  const uint32_t t7 = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(d), 7, 1));
  const uint64_t b7 = __builtin_amdgcn_uicmp(t7, 0u, 33 /* ICMP_NE */);
  const uint32_t mlo = __builtin_amdgcn_bitop3_b32(xlo, static_cast<uint32_t>(b7), t7, 0x09);
  const uint32_t mhi = __builtin_amdgcn_bitop3_b32(xhi, static_cast<uint32_t>(b7 >> 32), t7, 0x09);
  return (static_cast<uint64_t>(mhi) << 32) | mlo;
}

// Stable LSD radix sort (8-bit digits) of the n (key, column) pairs held in
// (kbuf, ibuf)[0, n). On return the sorted pairs are in (kbuf, ibuf)[0, n) and
// in the caller's registers key[s]/idx[s] at sorted position
// p = w*C + 64*s + lane (C = 64*ceil(n/512), the span_of layout); registers of
// positions p >= n hold sentinels.
//
// The segment is padded to 512*ceil(n/512) items with all-ones sentinel keys.
// Every canonical key of a finite value is below all-ones, so sentinels sort
// after every real key and leave the real keys' positions unchanged, and every
// lane of every step is active (no exec masking in the passes).
//
// Per pass, per wave, per 64-item step: eight ballots give each lane the set
// M of lanes sharing its digit (registers only). Then every lane reads the
// wave's running count of its digit and the lowest lane of M (the group
// leader) adds |M| to it. The reads and adds of all steps of the pass are
// issued back to back, with no wait in between: LDS executes one wave's
// instructions in issue order, so step s reads the count after exactly the
// steps before it (one LDS round trip per pass instead of one per step). A
// lane's offset is that count plus mbcnt(M). Rank inside a step is lane
// order, across steps issue order, across waves the exclusive prefix of the
// per-wave counts, so the sort is stable. A pass whose digit is equal on all
// real keys (block AND == OR) is the identity and is skipped.
//
// table: kWaves*kBins words, one 256-word row per wave, any contents on entry
// and exit. Each pass zeroes a wave's row, counts the wave's digits into it,
// and the scan overwrites it with the wave's digit offsets for the scatter
// (the thread of digit d reads column d of every row before the barrier that
// precedes its writes of the same column). scan_tmp: 8 words. Entry and exit
// are block-synchronous. Requires ceil(n/512) <= E.
template <typename K, int E>
__device__ __forceinline__ void radix_sort_regs(K (&key)[E], uint32_t (&idx)[E], int n, K* kbuf,
                                                uint16_t* ibuf, uint32_t* table,
                                                uint32_t* scan_tmp) {
  constexpr K kSent = static_cast<K>(~static_cast<K>(0));
  const int lane = lane_id();
  const int w = wave_id();
  const int nsteps = (n + kThreads - 1) / kThreads;
  const int base = w * 64 * nsteps;
  K kand = kSent, kor = 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int p = base + s * 64 + lane;
    const bool a = (s < nsteps) && (p < n);
    key[s] = a ? kbuf[p] : kSent;
    idx[s] = a ? static_cast<uint32_t>(ibuf[p]) : 0u;
    if (a) {
      kand &= key[s];
      kor |= key[s];
    }
  }
  if (n <= 1) return;
  // bits that differ between real keys: block AND / OR
  {
    const K wa = wave_allreduce(kand, [](K x, K y) { return x & y; });
    const K wo = wave_allreduce(kor, [](K x, K y) { return x | y; });
    K* red = reinterpret_cast<K*>(table);  // free until the first pass
    if (lane == 0) {
      red[2 * w] = wa;
      red[2 * w + 1] = wo;
    }
    __syncthreads();
    K ba = kSent, bo = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      ba &= red[2 * ww];
      bo |= red[2 * ww + 1];
    }
    kand = ba;
    kor = bo;
    __syncthreads();
  }
  const K diff = kand ^ kor;
  uint32_t* hw = table + w * kBins;  // this wave's row: counts, then offsets
  for (int shift = 0; shift < static_cast<int>(sizeof(K) * 8); shift += 8) {
    if (((diff >> shift) & static_cast<K>(255)) == 0) continue;  // uniform: identity pass
#pragma unroll
    for (int d = lane; d < kBins; d += 64) hw[d] = 0u;
    // (a) lanes sharing each lane's digit: rank in the step and group size
    uint32_t loc[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      loc[s] = 0u;
      if (s < nsteps) {
        const uint32_t d = static_cast<uint32_t>(key[s] >> shift) & 255u;
        const uint64_t M = match8_all(d);
        loc[s] = below(M) | (static_cast<uint32_t>(__popcll(M)) << 8);
      }
    }
    // (b) per step, every lane reads the wave's running count of its digit,
    // then the group's leader adds the group size; all issued back to back
    // (LDS executes one wave's instructions in issue order)
#pragma unroll
    for (int s = 0; s < E; ++s) {
      if (s < nsteps) {
        const uint32_t d = static_cast<uint32_t>(key[s] >> shift) & 255u;
        const uint32_t pre = __hip_atomic_load(&hw[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if ((loc[s] & 255u) == 0u)
          __hip_atomic_fetch_add(&hw[d], loc[s] >> 8, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        loc[s] = pre + (loc[s] & 255u);
      }
    }
    __syncthreads();
    // digit offsets, exclusive over (digit, wave) in digit-major order,
    // replace the counts in each wave's row
    uint32_t tot = 0, incl = 0, cw[kWaves];
    if (threadIdx.x < kBins) {
      const int d = threadIdx.x;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) {
        cw[ww] = table[ww * kBins + d];
        tot += cw[ww];
      }
      incl = wave_incl_sum(tot);
      if (lane == 63) scan_tmp[w] = incl;
    }
    __syncthreads();
    if (threadIdx.x < kBins) {
      uint32_t run = 0;
      for (int ww = 0; ww < w; ++ww) run += scan_tmp[ww];
      run += incl - tot;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) {
        table[ww * kBins + threadIdx.x] = run;
        run += cw[ww];
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      if (s < nsteps) {
        const uint32_t d = static_cast<uint32_t>(key[s] >> shift) & 255u;
        const uint32_t pos = hw[d] + loc[s];
        kbuf[pos] = key[s];
        ibuf[pos] = static_cast<uint16_t>(idx[s]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      if (s < nsteps) {
        const int p = base + s * 64 + lane;
        key[s] = kbuf[p];
        idx[s] = static_cast<uint32_t>(ibuf[p]);
      }
    }
  }
  __syncthreads();
}

}  // namespace fcd
