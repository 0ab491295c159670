// Segmented cross-sectional ranking on gfx950: one 512-thread workgroup per
// (factor, date) segment, the segment's items held in registers (E per lane)
// with one LDS exchange buffer, ranked by the stable radix sort of
// fc_device.h. Three kernels:
//
//   rank_scan_kernel  cross_sectional_rank + the four parameter_scan variants
//                     (asc/desc x masked/unmasked) derived from ONE sort.
//   fwd_info_kernel   per date: ordinal position and tie group of every valid
//                     forward return (shared by all factors of that date).
//   rank_ic_kernel    per (factor, date): factor sort, optional masked
//                     ascending cs_rank output, and the Spearman IC on the
//                     pairwise-valid set against fwd_info, with an exact
//                     integer Sum(rank_f*rank_r) and an FP64 finalize.
//
// Rank semantics follow the reference contract: stable ordinal ranks (ties
// broken by ascending column index), descending = ascending rank of the
// negated values, invalid cells = quiet NaN 0x7fc00000
// (src/cross_sectional_rank.cu:1-15, :35-72; fc/_cpu_core.py:57-78).
#include "fc_device.h"
#include "fc_internal.h"

#include <algorithm>
#include <climits>
#include <type_traits>

namespace fcr {

using namespace fcd;

// ---- LDS layout (identical on host and device) -----------------------------
struct Lds {
  size_t kbuf, ibuf, aux, tab, misc, bmask, stb, mbs, cnt, fst, lst, cp, ls, fs, total;
};

__host__ __device__ inline Lds lds_layout(int N, int ksz) {
  Lds L;
  const size_t S = (static_cast<size_t>(N) + 63) / 64;
  // the sort pads a segment to a multiple of 512 items (sentinel keys)
  const size_t Np = (static_cast<size_t>(N) + 511) & ~static_cast<size_t>(511);
  size_t o = 0;
  L.kbuf = o; o += Np * ksz;
  // output staging / return-order rows reuse the key buffer once the sorted
  // keys have been read (Np * ksz >= 4 * N)
  L.aux = L.kbuf;
  L.ibuf = o; o += (Np * 2 + 7) & ~static_cast<size_t>(7);
  // the sort's per-wave digit count / offset rows (8 KB)
  L.tab = o; o += static_cast<size_t>(kWaves) * kBins * 4;
  L.misc = o; o += 128 * 4;
  L.bmask = o; o += S * 8;   // written before the sort, read after it
  // sorted-order structure: built after the sort, in the (then idle) sort
  // table (40 * S bytes <= 8 KB for N <= 12288)
  size_t t = L.tab;
  L.stb = t; t += S * 8;
  L.mbs = t; t += S * 8;
  L.cnt = t; t += S * 4;
  L.fst = t; t += S * 4;
  L.lst = t; t += S * 4;
  L.cp = t; t += S * 4;
  L.ls = t; t += S * 4;
  L.fs = t; t += S * 4;
  L.total = (o + 15) & ~static_cast<size_t>(15);
  return L;
}

// the sorted-order structure (40 bytes per 64 items) fits the sort table up
// to the largest register-resident segment, 512 * 24 items
static_assert(40 * ((512 * 24 + 63) / 64) <= kWaves * kBins * 4, "sort table too small");

size_t rank_lds_bytes(int N, int key_bytes) { return lds_layout(N, key_bytes).total; }

// Register budgets (minimum waves per SIMD) follow the segments that the LDS
// holds per CU (160 KB): with 4-byte keys, three up to E = 10 (40 KB at
// N = 5000; 6 waves per SIMD, 80 VGPRs) and two up to E = 16; with 8-byte
// keys two up to E = 10. Every kernel fits its budget without spills.
template <typename V>
constexpr int rank_waves_per_simd(int E) {
  return sizeof(V) == 4 ? (E <= 10 ? 6 : (E <= 16 ? 4 : 2)) : (E <= 10 ? 4 : 2);
}
// The scan and return-order kernels fit 64 VGPRs at E <= 10 with 4-byte
// keys, so four of their 40 KB segments share a CU.
template <typename V>
constexpr int light_waves_per_simd(int E) {
  return (sizeof(V) == 4 && E <= 10) ? 8 : rank_waves_per_simd<V>(E);
}

// misc word slots
constexpr int kMiscScan = 0;    // 8 words: block_wave_prefix scratch
constexpr int kMiscSort = 8;    // 8 words: radix scan scratch
constexpr int kMiscRed = 24;    // 4 x 8 words: per-wave reductions (24..55)
constexpr int kMiscKm = 56;     // masked count after analyze
constexpr int kMiscOut = 57;    // small broadcast results
constexpr int kMiscKey = 64;    // 32 words: per-wave key AND / OR (up to 8-byte keys)

struct Bufs {
  unsigned char* base;
  Lds L;
  template <typename T> __device__ T* at(size_t off) const { return reinterpret_cast<T*>(base + off); }
};

// Load one row (E items per lane, all loads in flight together), mark the
// masked-valid set (bitmap by column index) and compact the sort set's
// (key, column) pairs into (kbuf, ibuf) in column order. Also zeroes the
// radix match table. Returns the sort-set size; block-synchronous on exit.
template <typename V, typename K, int E>
__device__ int load_compact(const V* __restrict__ xrow, long long es,
                            const uint8_t* __restrict__ mrow, bool masked_only, int N,
                            const Bufs& B) {
  K* kbuf = B.at<K>(B.L.kbuf);
  uint16_t* ibuf = B.at<uint16_t>(B.L.ibuf);
  uint64_t* bmask = B.at<uint64_t>(B.L.bmask);
  uint32_t* misc = B.at<uint32_t>(B.L.misc);
  const Span sp = span_of(N);
  const int lane = lane_id();
  // Every load of the row is issued before any is consumed: indexes past the
  // span are clamped to column 0 (always valid) and masked off afterwards, so
  // no load sits behind a branch or a wait of its own.
  V v[E];
  uint32_t mraw[E];
  bool act[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int i = sp.lo + s * 64 + lane;
    act[s] = (s < sp.steps) && (i < sp.hi);
    v[s] = xrow[static_cast<long long>(act[s] ? i : 0) * es];
  }
  if (mrow != nullptr) {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int i = sp.lo + s * 64 + lane;
      mraw[s] = mrow[act[s] ? i : 0];
    }
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) mraw[s] = 1u;
  }
  bool mk[E];
#pragma unroll
  for (int s = 0; s < E; ++s) mk[s] = act[s] && mraw[s] != 0u;
  uint64_t sb[E];
  uint32_t cnt = 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    sb[s] = 0ull;
    if (s < sp.steps) {
      const bool fin = act[s] && __builtin_isfinite(v[s]);
      const bool inset = masked_only ? (fin && mk[s]) : fin;
      const uint64_t mb = ballot(fin && mk[s]);
      sb[s] = ballot(inset);
      if (lane == 0) bmask[(sp.lo >> 6) + s] = mb;
      cnt += static_cast<uint32_t>(__popcll(sb[s]));
    }
  }
  uint32_t total;
  uint32_t pos = block_wave_prefix(cnt, misc + kMiscScan, total);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    if (s < sp.steps) {
      if ((sb[s] >> lane) & 1ull) {
        const uint32_t q = pos + below(sb[s]);
        kbuf[q] = ord_key(v[s]);
        ibuf[q] = static_cast<uint16_t>(sp.lo + s * 64 + lane);
      }
      pos += static_cast<uint32_t>(__popcll(sb[s]));
    }
  }
  __syncthreads();
  return static_cast<int>(total);
}

// Prefix sum / prefix max / suffix min over per-step words (S steps), wave 0
// and wave 1 working in parallel. Block-synchronous on exit.
__device__ void step_scans(int S, const uint32_t* CNT, uint32_t* CP, const int* LST, int* LS,
                           const int* FST, int* FS, uint32_t* total_out) {
  if (wave_id() == 0) {
    uint32_t carry = 0;
    int cmax = -1;
    for (int j0 = 0; j0 < S; j0 += 64) {
      const int j = j0 + lane_id();
      const bool a = j < S;
      const uint32_t c = a ? CNT[j] : 0u;
      const uint32_t inc = wave_incl_sum(c);
      int lm = -1;
      if (LST != nullptr) lm = wave_incl_max(a ? LST[j] : -1);
      if (a) {
        CP[j] = carry + inc - c;
        if (LS != nullptr) LS[j] = max(cmax, lm);
      }
      carry += last_lane(inc);
      cmax = max(cmax, static_cast<int>(last_lane(static_cast<uint32_t>(lm))));
    }
    if (lane_id() == 0) *total_out = carry;
  } else if (wave_id() == 1 && FST != nullptr) {
    int cmin = INT_MAX;
    for (int j0 = 0; j0 < S; j0 += 64) {
      const int j = S - 1 - (j0 + lane_id());
      const bool a = j >= 0;
      const int lm = wave_incl_min(a ? FST[j] : INT_MAX);
      if (a) FS[j] = min(cmin, lm);
      cmin = min(cmin, static_cast<int>(last_lane(static_cast<uint32_t>(lm))));
    }
  }
  __syncthreads();
}

// Tie-group and masked-count structure of the sorted order: start-of-group
// bits (STB), masked bits by sorted position (MBS), exclusive masked counts
// per step (CP), last group start at or before each step (LS) and first group
// start at or after each step (FS). Km -> misc[kMiscKm].
template <typename K, int E>
__device__ void analyze_sorted(const K (&key)[E], const uint32_t (&idx)[E], int Kn,
                               bool all_masked, const Bufs& B) {
  const K* kbuf = B.at<K>(B.L.kbuf);
  const uint64_t* bmask = B.at<uint64_t>(B.L.bmask);
  uint64_t* STB = B.at<uint64_t>(B.L.stb);
  uint64_t* MBS = B.at<uint64_t>(B.L.mbs);
  uint32_t* CNT = B.at<uint32_t>(B.L.cnt);
  int* FST = B.at<int>(B.L.fst);
  int* LST = B.at<int>(B.L.lst);
  const Span sp = span_of(Kn);
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < E; ++s) {
    if (s < sp.steps) {
      const int s0 = sp.lo + s * 64;
      const int p = s0 + lane;
      const bool act = p < sp.hi;
      bool st = false, mv = false;
      if (act) {
        st = (p == 0) || (key[s] != kbuf[p - 1]);
        const uint32_t i = idx[s];
        mv = all_masked || ((bmask[i >> 6] >> (i & 63)) & 1ull);
      }
      const uint64_t sb = ballot(st);
      const uint64_t mb = ballot(mv);
      if (lane == 0) {
        const int j = s0 >> 6;
        STB[j] = sb;
        MBS[j] = mb;
        CNT[j] = static_cast<uint32_t>(__popcll(mb));
        FST[j] = sb ? s0 + __ffsll(static_cast<unsigned long long>(sb)) - 1 : INT_MAX;
        LST[j] = sb ? s0 + 63 - __clzll(static_cast<long long>(sb)) : -1;
      }
    }
  }
  __syncthreads();
  const int S = (Kn + 63) >> 6;
  step_scans(S, CNT, B.at<uint32_t>(B.L.cp), LST, B.at<int>(B.L.ls), FST, B.at<int>(B.L.fs),
             B.at<uint32_t>(B.L.misc) + kMiscKm);
}

struct Groups {
  const uint64_t* STB;
  const uint64_t* MBS;
  const uint32_t* CP;
  const int* LS;
  const int* FS;
  int Kn;
  uint32_t Km;
  __device__ bool masked(int p) const { return (MBS[p >> 6] >> (p & 63)) & 1ull; }
  // masked items at sorted positions < p (p may be Kn)
  __device__ uint32_t cbefore(int p) const {
    if (p >= Kn) return Km;
    return CP[p >> 6] + static_cast<uint32_t>(__popcll(MBS[p >> 6] & mask_lt(p & 63)));
  }
  __device__ int gstart(int p) const {
    const uint64_t w = STB[p >> 6] & mask_le(p & 63);
    return w ? ((p & ~63) + 63 - __clzll(static_cast<long long>(w))) : LS[(p >> 6) - 1];
  }
  __device__ int gend(int p) const {
    const uint64_t w = STB[p >> 6] & ~mask_le(p & 63);
    if (w) return (p & ~63) + __ffsll(static_cast<unsigned long long>(w)) - 1;
    const int j = (p >> 6) + 1;
    return (j < ((Kn + 63) >> 6)) ? min(FS[j], Kn) : Kn;
  }
};

__device__ __forceinline__ Groups make_groups(const Bufs& B, int Kn) {
  Groups G;
  G.STB = B.at<uint64_t>(B.L.stb);
  G.MBS = B.at<uint64_t>(B.L.mbs);
  G.CP = B.at<uint32_t>(B.L.cp);
  G.LS = B.at<int>(B.L.ls);
  G.FS = B.at<int>(B.L.fs);
  G.Kn = Kn;
  G.Km = B.at<uint32_t>(B.L.misc)[kMiscKm];
  return G;
}

// Row staging in LDS: fill, then a coalesced 16-byte-per-lane store (scalar
// tail when the row is not 16-byte aligned).
__device__ __forceinline__ void fill_row(uint32_t* stage, int N, uint32_t v) {
  for (int i = threadIdx.x; i < N; i += kThreads) stage[i] = v;
}
__device__ __forceinline__ void store_row(const uint32_t* stage, uint32_t* __restrict__ out, int N) {
  if ((reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
    const int n4 = N >> 2;
    const uint4* s4 = reinterpret_cast<const uint4*>(stage);
    uint4* o4 = reinterpret_cast<uint4*>(out);
#pragma unroll 4
    for (int i = threadIdx.x; i < n4; i += kThreads) o4[i] = s4[i];
    for (int i = (n4 << 2) + threadIdx.x; i < N; i += kThreads) out[i] = stage[i];
  } else {
#pragma unroll 4
    for (int i = threadIdx.x; i < N; i += kThreads) out[i] = stage[i];
  }
}

// ---- kernel 1: cs_rank / parameter_scan ------------------------------------
// XCD-aware segment map (workgroups are dispatched round-robin over the 8
// XCDs, b % 8): the F planes of one date run back to back on one XCD, so a
// (T,N,F) panel read plane by plane (element stride F) pulls each row's cache
// lines from HBM once and the other planes hit that XCD's L2. False for the
// padding slots (t >= T).
__device__ __forceinline__ bool scan_segment_of(const ScanArgs& a, int& t, int& f) {
  const int b = blockIdx.x;
  const int j = b >> 3;
  t = (b & 7) + 8 * (j / a.F);
  f = j - (j / a.F) * a.F;
  return t < a.T;
}

// One (t, f) segment of cs_rank / parameter_scan through the register radix
// sort. smem holds lds_layout(N, sizeof key) bytes.
template <typename V, int E>
__device__ __forceinline__ void rank_scan_radix(const ScanArgs& a, int t, int f, unsigned char* smem) {
  using K = typename KeyOf<V>::type;
  Bufs B{smem, lds_layout(a.N, sizeof(K))};
  const int N = a.N;
  uint32_t* misc = B.at<uint32_t>(B.L.misc);
  const V* xrow = static_cast<const V*>(a.x) + f * a.x_ps + t * a.x_rs;
  const uint8_t* mrow = a.mask ? a.mask + f * a.m_ps + t * a.m_rs : nullptr;
  const int Kn = load_compact<V, K, E>(xrow, a.x_es, mrow, a.masked_only != 0, N, B);
  K key[E];
  uint32_t idx[E];
  radix_sort_regs<K, E>(key, idx, Kn, B.at<K>(B.L.kbuf), B.at<uint16_t>(B.L.ibuf),
                        B.at<uint32_t>(B.L.tab), misc + kMiscSort);
  analyze_sorted<K, E>(key, idx, Kn, a.masked_only != 0, B);
  const Groups G = make_groups(B, Kn);
  uint32_t* stage = B.at<uint32_t>(B.L.aux);
  const Span sp = span_of(Kn);
  const size_t orow = (static_cast<size_t>(f) * a.T + t) * static_cast<size_t>(N);
  // one specialised pass per requested group (a runtime group switch inside
  // the unrolled item loop doubles the kernel's registers)
  auto emit = [&](auto gc) {
    constexpr int g = decltype(gc)::value;
    float* out = a.out[g];
    if (out == nullptr) return;
    fill_row(stage, N, kQuietNaN);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int p = sp.lo + s * 64 + lane_id();
      if ((s < sp.steps) && (p < sp.hi)) {
        int r = -1;
        if constexpr (g == 0) {
          if (G.masked(p)) r = static_cast<int>(G.cbefore(p)) + 1;
        } else if constexpr (g == 1) {
          r = p + 1;
        } else if constexpr (g == 2) {
          if (G.masked(p)) {
            const int gs = G.gstart(p), ge = G.gend(p);
            r = static_cast<int>(G.Km - G.cbefore(ge)) +
                static_cast<int>(G.cbefore(p) - G.cbefore(gs)) + 1;
          }
        } else {
          const int gs = G.gstart(p), ge = G.gend(p);
          r = (Kn - ge) + (p - gs) + 1;
        }
        if (r > 0) stage[idx[s]] = __float_as_uint(static_cast<float>(r));
      }
    }
    __syncthreads();
    store_row(stage, reinterpret_cast<uint32_t*>(out + orow), N);
    __syncthreads();
  };
  emit(std::integral_constant<int, 0>{});
  emit(std::integral_constant<int, 1>{});
  emit(std::integral_constant<int, 2>{});
  emit(std::integral_constant<int, 3>{});
}

template <typename V, int E>
__global__ __launch_bounds__(kThreads, light_waves_per_simd<V>(E)) void rank_scan_kernel(ScanArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  int t, f;
  if (!scan_segment_of(a, t, f)) return;
  rank_scan_radix<V, E>(a, t, f, smem);
}

// ---- kernel 2: per-date forward-return order info --------------------------
// info[t][i] = (position among valid returns, 1-based) | (1 + start position
// of i's tie group) << 16; 0 for invalid cells. kr[t] = valid count.
template <typename V, int E>
__device__ __forceinline__ void fwd_info_radix(const FwdArgs& a, int t, unsigned char* smem) {
  using K = typename KeyOf<V>::type;
  Bufs B{smem, lds_layout(a.N, sizeof(K))};
  const int N = a.N;
  uint32_t* misc = B.at<uint32_t>(B.L.misc);
  const V* rrow = static_cast<const V*>(a.r) + t * a.r_rs;
  const uint8_t* mrow = a.rmask ? a.rmask + t * a.m_rs : nullptr;
  const int Kn = load_compact<V, K, E>(rrow, 1, mrow, true, N, B);
  K key[E];
  uint32_t idx[E];
  radix_sort_regs<K, E>(key, idx, Kn, B.at<K>(B.L.kbuf), B.at<uint16_t>(B.L.ibuf),
                        B.at<uint32_t>(B.L.tab), misc + kMiscSort);
  analyze_sorted<K, E>(key, idx, Kn, true, B);
  const Groups G = make_groups(B, Kn);
  uint32_t* stage = B.at<uint32_t>(B.L.aux);
  fill_row(stage, N, 0u);
  __syncthreads();
  const Span sp = span_of(Kn);
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int p = sp.lo + s * 64 + lane_id();
    if ((s < sp.steps) && (p < sp.hi))
      stage[idx[s]] = static_cast<uint32_t>(p + 1) | (static_cast<uint32_t>(G.gstart(p) + 1) << 16);
  }
  __syncthreads();
  store_row(stage, a.info + static_cast<size_t>(t) * N, N);
  if (threadIdx.x == 0) a.kr[t] = Kn;
}

template <typename V, int E>
__global__ __launch_bounds__(kThreads, light_waves_per_simd<V>(E)) void fwd_info_kernel(FwdArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  fwd_info_radix<V, E>(a, static_cast<int>(blockIdx.x), smem);
}

// ---- kernel 3: fused rank + Spearman IC ------------------------------------
// Block b -> (t, f) so that every factor of one date lands on the same XCD
// (blocks are dispatched round-robin over the 8 XCDs, b % 8), keeping the
// date's fwd_info row and mask row in that XCD's L2.
// Segment schedule (workgroup -> (t, f)); segments are independent, so the
// schedule changes speed only, never results.
//   kSchedXcd   consecutive factors of one date on one XCD (workgroups are
//               dispatched round-robin over the 8 XCDs, b % 8), sharing the
//               date's return-order row in that XCD's L2
//   kSchedFMaj  factor-major linear order
//   kSchedTMaj  date-major linear order
__device__ __forceinline__ void ic_segment_of(const IcArgs& a, int& t, int& f) {
  const int b = blockIdx.x;
  if (a.sched == kSchedXcd) {
    const int j = b >> 3;
    t = (b & 7) + 8 * (j / a.F);
    f = j - (j / a.F) * a.F;
  } else if (a.sched == kSchedFMaj) {
    f = b / a.T;
    t = b - f * a.T;
  } else {
    t = b / a.F;
    f = b - t * a.F;
  }
}

// Spearman IC of m ordinal rank pairs from S = Sum rank_f * rank_r; NaN when
// degenerate. Ordinal ranks are permutations of 1..m: Sum r = m(m+1)/2, so the
// centred sums are exact multiples of 1/4 below 2^53:
//   sxy = S - m(m+1)^2/4,  sxx = syy = m(m^2-1)/12.
__device__ __forceinline__ double spearman_from_rank_products(uint64_t Ssum, uint32_t m,
                                                              bool degenerate) {
  double ic = __longlong_as_double(0x7ff8000000000000ll);
  if (!degenerate) {
    const long long mm = static_cast<long long>(m);
    const long long A = 4ll * static_cast<long long>(Ssum) - mm * (mm + 1) * (mm + 1);
    const long long Bq = mm * (mm * mm - 1);
    const double sxy = static_cast<double>(A) * 0.25;
    const double sxx = static_cast<double>(Bq) / 12.0;
    // reference finalize order: (sxy / sqrt(sxx)) / sqrt(syy)
    ic = (sxx > 0.0) ? (sxy / sqrt(sxx)) / sqrt(sxx) : ic;
  }
  return ic;
}

// One (t, f) segment of rank + IC through the register radix sort. smem holds
// lds_layout(N, sizeof key) bytes.
template <typename V, int E>
__device__ __forceinline__ void rank_ic_radix(const IcArgs& a, int t, int f, unsigned char* smem) {
  using K = typename KeyOf<V>::type;
  Bufs B{smem, lds_layout(a.N, sizeof(K))};
  const int N = a.N;
  const int lane = lane_id();
  uint32_t* misc = B.at<uint32_t>(B.L.misc);
  const V* xrow = static_cast<const V*>(a.x) + f * a.x_ps + t * a.x_rs;
  const uint8_t* mrow = a.fmask ? a.fmask + f * a.m_ps + t * a.m_rs : nullptr;

  // forward-return info row, read into registers ahead of its use. With
  // 8-byte keys at E <= 10 (two segments per CU, registers to spare) the
  // loads are issued first and overlap the sort; otherwise they are issued
  // right after the sort (whose offset registers are then free) and overlap
  // the rank output.
  constexpr bool kPrefetchRI = sizeof(V) == 8 && E <= 10;
  const uint32_t* irow = a.info + static_cast<size_t>(t) * N;
  const Span rsp = span_of(N);
  uint32_t ri_reg[E];
  auto load_ri = [&]() {
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int i = rsp.lo + s * 64 + lane;
      const bool ok = (s < rsp.steps) && (i < rsp.hi);
      ri_reg[s] = irow[ok ? i : 0];
    }
  };
  if constexpr (kPrefetchRI) load_ri();
  const int Kr = a.kr[t];

  const int Kn = load_compact<V, K, E>(xrow, a.x_es, mrow, true, N, B);
  K key[E];
  uint32_t idx[E];
  radix_sort_regs<K, E>(key, idx, Kn, B.at<K>(B.L.kbuf), B.at<uint16_t>(B.L.ibuf),
                        B.at<uint32_t>(B.L.tab), misc + kMiscSort);
  if constexpr (!kPrefetchRI) load_ri();
  const Span sp = span_of(Kn);
  const size_t frow = (static_cast<size_t>(f) * a.T + t) * static_cast<size_t>(N);
  uint32_t* aux = B.at<uint32_t>(B.L.aux);

  // (1) optional ascending masked cs_rank of the factor
  if (a.rank_out != nullptr) {
    fill_row(aux, N, kQuietNaN);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const int p = sp.lo + s * 64 + lane;
      if ((s < sp.steps) && (p < sp.hi)) aux[idx[s]] = __float_as_uint(static_cast<float>(p + 1));
    }
    __syncthreads();
    store_row(aux, reinterpret_cast<uint32_t*>(a.rank_out + frow), N);
    __syncthreads();
  }

  // (2) return info row into LDS
  uint32_t* RI = aux;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int i = rsp.lo + s * 64 + lane;
    if ((s < rsp.steps) && (i < rsp.hi)) RI[i] = ri_reg[s];
  }
  __syncthreads();

  // (3) pass A over the factor order: pairwise-valid bits P, AND / OR of
  // the factor keys over P (the factor is constant on P iff they agree),
  // min/max return tie group over P
  uint64_t* PB = B.at<uint64_t>(B.L.stb);
  uint32_t* PCNT = B.at<uint32_t>(B.L.cnt);
  uint32_t* PC = B.at<uint32_t>(B.L.cp);
  uint32_t rinf[E];
  int gmin = INT_MAX, gmax = -1;
  K kand = static_cast<K>(~static_cast<K>(0)), kor = 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    rinf[s] = 0u;
    if (s < sp.steps) {
      const int p = sp.lo + s * 64 + lane;
      const bool act = p < sp.hi;
      rinf[s] = act ? RI[idx[s]] : 0u;
      const bool P = rinf[s] != 0u;
      const uint64_t pb = ballot(P);
      if (lane == 0) {
        PB[(sp.lo >> 6) + s] = pb;
        PCNT[(sp.lo >> 6) + s] = static_cast<uint32_t>(__popcll(pb));
      }
      if (P) {
        const int g = static_cast<int>(rinf[s] >> 16);
        gmin = min(gmin, g);
        gmax = max(gmax, g);
        kand &= key[s];
        kor |= key[s];
      }
    }
  }
  gmin = wave_min_i(gmin);
  gmax = wave_max_i(gmax);
  kand = wave_allreduce(kand, [](K x, K y) { return x & y; });
  kor = wave_allreduce(kor, [](K x, K y) { return x | y; });
  int* red = reinterpret_cast<int*>(misc + kMiscRed);
  K* kred = reinterpret_cast<K*>(misc + kMiscKey);
  if (lane == 0) {
    red[16 + wave_id()] = gmin;
    red[24 + wave_id()] = gmax;
    kred[2 * wave_id()] = kand;
    kred[2 * wave_id() + 1] = kor;
  }
  __syncthreads();
  const int S = (Kn + 63) >> 6;
  step_scans(S, PCNT, PC, nullptr, nullptr, nullptr, nullptr, misc + kMiscOut);
  const uint32_t m = misc[kMiscOut];
  bool degenerate = static_cast<int>(m) < a.min_valid;
  K* kbuf = B.at<K>(B.L.kbuf);
  if (m > 0) {
    int G0 = INT_MAX, G1 = -1;
    K ka = static_cast<K>(~static_cast<K>(0)), ko = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      G0 = min(G0, red[16 + w]);
      G1 = max(G1, red[24 + w]);
      ka &= kred[2 * w];
      ko |= kred[2 * w + 1];
    }
    if (ka == ko || G0 == G1) degenerate = true;
  }
  const bool need_ranks = (a.rank_f != nullptr);
  if (degenerate && !need_ranks) {
    if (threadIdx.x == 0)
      a.ic[static_cast<size_t>(f) * a.T + t] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }
  __syncthreads();  // every read of kbuf done before it is reused as FW

  // (4) pass B: factor ranks on P, scattered into the return order
  uint32_t* FW = reinterpret_cast<uint32_t*>(kbuf);
  for (int q = threadIdx.x; q < Kr; q += kThreads) FW[q] = 0u;
  __syncthreads();
#pragma unroll
  for (int s = 0; s < E; ++s) {
    if (s < sp.steps) {
      const uint64_t pb = PB[(sp.lo >> 6) + s];
      if ((pb >> lane) & 1ull) {
        const uint32_t rf = PC[(sp.lo >> 6) + s] + below(pb) + 1u;
        const uint32_t i = idx[s];
        FW[(rinf[s] & 0xFFFFu) - 1u] = rf | (i << 16);
        if (need_ranks) a.rank_f[static_cast<size_t>(t) * N + i] = static_cast<double>(rf);
      }
    }
  }
  __syncthreads();

  // (5) pass C over the return order: return ranks on P and Sum(rf*rr)
  const Span qsp = span_of(Kr);
  uint32_t fw[E];
  uint64_t qb[E];
  uint32_t cnt = 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int q = qsp.lo + s * 64 + lane;
    fw[s] = ((s < qsp.steps) && (q < qsp.hi)) ? FW[q] : 0u;
    qb[s] = 0ull;
    if (s < qsp.steps) {
      qb[s] = ballot(fw[s] != 0u);
      cnt += static_cast<uint32_t>(__popcll(qb[s]));
    }
  }
  uint32_t tot;
  uint32_t base = block_wave_prefix(cnt, misc + kMiscScan, tot);
  uint64_t acc = 0;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    if (s < qsp.steps) {
      if (fw[s] != 0u) {
        const uint32_t rr = base + below(qb[s]) + 1u;
        acc += static_cast<uint64_t>(fw[s] & 0xFFFFu) * rr;
        if (need_ranks) a.rank_r[static_cast<size_t>(t) * N + (fw[s] >> 16)] = static_cast<double>(rr);
      }
      base += static_cast<uint32_t>(__popcll(qb[s]));
    }
  }
  acc = wave_sum_u64(acc);
  uint64_t* red64 = reinterpret_cast<uint64_t*>(misc + kMiscRed);
  if (lane == 0) red64[wave_id()] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t Ssum = 0;
    for (int w = 0; w < kWaves; ++w) Ssum += red64[w];
    a.ic[static_cast<size_t>(f) * a.T + t] = spearman_from_rank_products(Ssum, m, degenerate);
  }
}

template <typename V, int E>
__global__ __launch_bounds__(kThreads, rank_waves_per_simd<V>(E)) void rank_ic_kernel(IcArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  int t, f;
  ic_segment_of(a, t, f);
  if (t >= a.T || f >= a.F) return;
  rank_ic_radix<V, E>(a, t, f, smem);
}

// The calling kernel's argument struct, read in place from its kernarg
// segment (it is the kernel's only parameter, at offset 0). The out-of-line
// fallbacks below take this pointer: a struct passed by value is copied to
// scratch at kernel entry on every launch, whether or not the call happens.
template <typename Args>
__device__ __forceinline__ const Args* kernel_args() {
  return reinterpret_cast<const Args*>(
      reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr()));
}

// The radix path as an out-of-line call, so that the bucketed kernel's
// registers are allocated for its own phases only.
template <int E>
__device__ __noinline__ void rank_ic_radix_call(const IcArgs* a, int t, int f, unsigned char* smem) {
  rank_ic_radix<float, E>(*a, t, f, smem);
}

// ---- kernel 3b: fused rank + Spearman IC by bucketed counting (float rows) --
//
// Ranks a row without sorting it. Every valid value x gets the bucket
//   b(x) = trunc(med3(fma(x, 1/2, -lo/2) * s, 0, NB - 1)),  s = NB / (hi/2 - lo/2)
// over the row's valid range [lo, hi]. Each step is an IEEE round-to-nearest
// operation with a non-negative coefficient, hence non-decreasing in x, so b
// is non-decreasing in x: every key of a lower bucket is below every key of a
// higher one. A histogram of the buckets (LDS atomics whose returned counts
// give each item a slot in its bucket), an exclusive scan and one scatter put
// the items in bucket order as 64-bit words
//   key << 32 | column << 18 | pairwise-valid << 16 | return position.
// An item's 0-based rank among the valid items is its bucket's start plus the
// number of words of its bucket below its own word: smaller key, or equal key
// and smaller column. That is the stable ordinal rank of the reference's
// segmented sort (src/cross_sectional_rank.cu:152-164) without a sort. Words
// past the bucket's end belong to higher buckets or are all-ones sentinels,
// so the compare loop may read past it. Buckets hold about N/NB items on
// spread-out rows; a row with a bucket of more than kMaxBucket items
// (constant or heavily clustered values) is ranked by the radix sort instead,
// with identical results.
//
// The IC side needs no sort either: the pairwise-valid items' return
// positions (from fwd_info) are set in a bitmap over return order, their
// factor ranks in a bitmap over factor order; with the exclusive popcount
// prefix of each bitmap's words, an item's rank on the pairwise-valid set is a
// prefix lookup plus one popcount on either side.
namespace bucket {
constexpr int kPad = 8;          // all-ones sentinel words after the bucket-ordered array
constexpr int kJunk = 64;        // per-lane words that take inactive lanes' writes
constexpr int kMaxBucket = 48;   // a larger bucket sends the row to the radix sort
constexpr int kHead = 5;         // words of every bucket compared with no range check
constexpr int kScanHead = 6;     // the same for the parameter scan's four-rank compares
constexpr int kMinBuckets = 1024;
constexpr int kMaxBuckets = 4096;

struct Layout {
  size_t A, H, RB, NP, misc, total;
  int NB;
};

// Buckets: the smallest power of two >= 1.5 N in [1024, 4096]. LDS: the
// bucket-ordered words with their sentinels and junk words (later the rank
// staging row), the 16-bit counts / prefixes (NB + 1 entries) and junk words,
// two bitmaps of (bits, exclusive prefix) word pairs (return positions of
// P; ranks of the valid items outside P), scratch. At N = 5000 that is 52 KB,
// so three segments share a CU.
__host__ __device__ inline Layout layout(int N) {
  Layout L;
  int nb = kMinBuckets;
  while (nb < kMaxBuckets && 2 * nb < 3 * N) nb <<= 1;
  L.NB = nb;
  const size_t W = (static_cast<size_t>(N) + 31) / 32;
  size_t o = 0;
  L.A = o; o += (static_cast<size_t>(N) + kPad + kJunk) * 8;
  // counts (NB / 2 + 1 words, 16-byte aligned for the scan's uint4 reads),
  // then kJunk junk words (one per lane)
  o = (o + 15) & ~static_cast<size_t>(15);
  L.H = o; o += (static_cast<size_t>(nb / 2 + 1 + kJunk) * 4 + 15) & ~static_cast<size_t>(15);
  L.RB = o; o += W * 8;
  L.NP = o; o += W * 8;
  L.misc = o; o += 128 * 4;
  L.total = (o + 15) & ~static_cast<size_t>(15);
  return L;
}

// misc word slots
constexpr int kRed = 0;      // 8 x 4: per-wave value range and counts
constexpr int kHist = 64;    // 8: per-wave count totals
constexpr int kRbit = 72;    // 8: per-wave return-bitmap popcounts
constexpr int kBig = 80;     // 8: per-wave largest bucket
constexpr int kSum = 96;     // 8 x u64: per-wave Sum rank_f * rank_r
constexpr int kDeg = 112;    // 4: keys of P's first / last factor rank, return tie
                             // groups of P's first / last return rank

// ord_key of a finite float without a zero test: v + 0 turns -0 into +0
// (round to nearest), then negative values invert all bits and the others
// set the sign bit.
// (v_bitop3 with operands (a, b, c) and truth-table bit (a << 2 | b << 1 | c):
// 0x1E is a ^ (b | c), 0x4B is a ^ (~b | c))
__device__ __forceinline__ uint32_t canonical_key(float v) {
  const uint32_t b = __float_as_uint(v + 0.0f);
  const uint32_t t = static_cast<uint32_t>(static_cast<int32_t>(b) >> 31);
  return __builtin_amdgcn_bitop3_b32(b, t, 0x80000000u, 0x1E);  // b ^ (t | sign)
}
__device__ __forceinline__ float key_value(uint32_t k) {
  const uint32_t t = static_cast<uint32_t>(static_cast<int32_t>(k) >> 31);
  return __uint_as_float(__builtin_amdgcn_bitop3_b32(k, t, 0x80000000u, 0x4B));  // k ^ (~t | sign)
}

struct Map {
  float nlo, scale, top;
};
__device__ __forceinline__ Map make_map(float lo, float hi, int NB) {
  Map m;
  m.nlo = -0.5f * lo;
  const float dh = fmaf(hi, 0.5f, m.nlo);  // hi/2 - lo/2 never overflows
  const float sc = dh > 0.0f ? static_cast<float>(NB) / dh : 0.0f;
  // a range of a few subnormal steps gives an infinite scale; a finite one
  // keeps d * scale free of 0 * inf (it saturates to +inf instead)
  m.scale = sc <= 3.0e38f ? sc : 3.0e38f;
  m.top = static_cast<float>(NB - 1);
  return m;
}
// Non-decreasing in x (see above); in [0, NB) for finite x. +0 and -0 share
// a bucket: d is +-0 or the same nonzero value for both.
__device__ __forceinline__ uint32_t bucket_of_value(float x, const Map& m) {
  const float d = fmaf(x, 0.5f, m.nlo);
  return static_cast<uint32_t>(__builtin_amdgcn_fmed3f(d * m.scale, 0.0f, m.top));
}
// The same bucket from a word's canonical key.
__device__ __forceinline__ uint32_t bucket_of(uint32_t k, const Map& m) {
  return bucket_of_value(key_value(k), m);
}

__device__ __forceinline__ void lds_or(uint32_t* p, uint32_t bits) {
  __hip_atomic_fetch_or(p, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_add(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// op over lanes 0..7 (one value per wave, lane l holding wave l's; lanes
// 8.. ignored), as a uniform value: three row_shr steps inside the first
// 16-lane row, then lane 7. One LDS read per lane instead of eight
// dependent ones.
__device__ __forceinline__ uint32_t add_u(uint32_t x, uint32_t y) { return x + y; }
__device__ __forceinline__ uint32_t fmin_u(uint32_t x, uint32_t y) {
  return __float_as_uint(fminf(__uint_as_float(x), __uint_as_float(y)));
}
__device__ __forceinline__ uint32_t fmax_u(uint32_t x, uint32_t y) {
  return __float_as_uint(fmaxf(__uint_as_float(x), __uint_as_float(y)));
}
template <typename Op>
__device__ __forceinline__ uint32_t comb8(uint32_t v, uint32_t ident, Op op) {
  v = op(v, dpp_or<0x111, 0xF>(v, ident));
  v = op(v, dpp_or<0x112, 0xF>(v, ident));
  v = op(v, dpp_or<0x114, 0xF>(v, ident));
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 7));
}
// Sum of the per-wave values slot8[0..7] over the waves below w.
__device__ __forceinline__ uint32_t below_sum8(const uint32_t* slot8, int w, int lane) {
  const uint32_t v = slot8[lane & 7];
  return comb8(lane < w ? v : 0u, 0u, [](uint32_t x, uint32_t y) { return x + y; });
}
// Largest of the per-wave values slot8[0..7].
__device__ __forceinline__ uint32_t max8(const uint32_t* slot8, int lane) {
  return comb8(slot8[lane & 7], 0u, [](uint32_t x, uint32_t y) { return max(x, y); });
}

// In-place exclusive scan of the 16-bit counts (two per word, NB / 2
// words), first half: wave w scans its contiguous eighth of the words, each
// lane 1, 2 or 4 consecutive words (NB = 1024, 2048, 4096; one 16-byte read
// at 4096). The lane keeps its words as packed prefixes from its own start
// and its offset in the wave; the wave's total goes to misc[kHist + w], and
// misc[kBig + w] exceeds kMaxBucket iff a bucket of the wave does.
struct CountScan {
  uint32_t lp[4];  // per word: prefix before its low count | before its high count << 16
  uint32_t pre;    // counts of the wave's lanes below this one
};
__device__ __forceinline__ CountScan count_scan_wave(const uint32_t* H, int NB, int w, int lane,
                                                     uint32_t* misc) {
  const int per_wave = (NB >> 1) / kWaves;  // 64, 128 or 256 words
  const int wpl = per_wave >> 6;
  const int base = w * per_wave + lane * wpl;
  uint32_t c[4] = {0u, 0u, 0u, 0u};
  if (wpl == 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(H + base);
    c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
  } else if (wpl == 2) {
    const uint2 v = *reinterpret_cast<const uint2*>(H + base);
    c[0] = v.x; c[1] = v.y;
  } else {
    c[0] = H[base];
  }
  CountScan cs;
  uint32_t run = 0u, big = 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t lo = c[k] & 0xFFFFu, hi = c[k] >> 16;
    cs.lp[k] = run | ((run + lo) << 16);
    run += lo + hi;
    big = max(big, max(lo, hi));
  }
  const uint32_t inc = wave_incl_sum(run);
  cs.pre = inc - run;
  const bool over = ballot(big > static_cast<uint32_t>(kMaxBucket)) != 0ull;
  if (lane == 63) misc[kHist + w] = inc;
  if (lane == 0) misc[kBig + w] = over ? static_cast<uint32_t>(kMaxBucket) + 1u : 0u;
  return cs;
}
// Second half, after a barrier: off = the totals of the waves below.
__device__ __forceinline__ void count_scan_store(uint32_t* H, int NB, int w, int lane,
                                                 const CountScan& cs, uint32_t off) {
  const int per_wave = (NB >> 1) / kWaves;
  const int wpl = per_wave >> 6;
  const int base = w * per_wave + lane * wpl;
  const uint32_t add = __umul24(off + cs.pre, 0x10001u);  // both halves (positions < 2^16)
  if (wpl == 4) {
    *reinterpret_cast<uint4*>(H + base) =
        make_uint4(cs.lp[0] + add, cs.lp[1] + add, cs.lp[2] + add, cs.lp[3] + add);
  } else if (wpl == 2) {
    *reinterpret_cast<uint2*>(H + base) = make_uint2(cs.lp[0] + add, cs.lp[1] + add);
  } else {
    H[base] = cs.lp[0] + add;
  }
}

// Exclusive popcount prefix of bitmap word pair d (bits in .x), from the
// pair's stored prefix (.y).
__device__ __forceinline__ uint32_t rank_below(uint2 d, uint32_t bit) {
  return d.y + static_cast<uint32_t>(__popc(d.x & ((1u << bit) - 1u)));
}
}  // namespace bucket

// Register budget: three 51.8 KB segments share a CU up to E = 10 (6 waves
// per SIMD), two up to E = 16, one beyond.
constexpr int bucket_waves_per_simd(int E) { return E <= 10 ? 6 : (E <= 16 ? 4 : 2); }

// Items of a float row per lane in column order (c = s * 512 + tid); E as for
// the radix kernels. Invalid items carry the all-ones key (no finite value's
// key). Each phase issues its LDS accesses for all of a lane's items before it
// uses any result, and inactive lanes write to their junk words instead of
// branching, so a lane waits about once per phase. The row is read and its
// ranks written through raw buffers with 32-bit byte offsets (4 * es * N <
// 2^31 is checked on the host).
template <int E, bool kRankOut>
__global__ __launch_bounds__(kThreads, bucket_waves_per_simd(E)) void rank_ic_bucket_kernel(IcArgs a) {
  // positions handled together in the rank and IC phases (register budget)
  constexpr int kChunk = E <= 5 ? E : (E % 5 == 0 ? 5 : 8);
  static_assert(E % kChunk == 0, "E must be a multiple of the chunk");
  constexpr uint32_t kInvalid = ~0u;
  // rows of at most 5120 carry the slot in the bucket-ordered words
  constexpr bool kSlotWord = E <= 10;
  static_assert(!kSlotWord || (E * kThreads <= 8192 && bucket::kMaxBucket < 64), "slot word fields");
  extern __shared__ __align__(16) unsigned char smem[];
  int t, f;
  ic_segment_of(a, t, f);
  if (t >= a.T || f >= a.F) return;
  const int N = a.N;
  const bucket::Layout L = bucket::layout(N);
  const int NB = L.NB;
  uint64_t* A = reinterpret_cast<uint64_t*>(smem + L.A);
  uint32_t* H = reinterpret_cast<uint32_t*>(smem + L.H);
  const uint16_t* H16 = reinterpret_cast<const uint16_t*>(smem + L.H);
  uint2* RB = reinterpret_cast<uint2*>(smem + L.RB);
  uint2* NP = reinterpret_cast<uint2*>(smem + L.NP);
  uint32_t* misc = reinterpret_cast<uint32_t*>(smem + L.misc);
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = lane_id();
  const int w = wave_id();
  const int nw = (N + 31) >> 5;
  const float* xrow = static_cast<const float*>(a.x) + f * a.x_ps + t * a.x_rs;
  const uint8_t* mrow = a.fmask ? a.fmask + f * a.m_ps + t * a.m_rs : nullptr;
  const uint32_t* irow = a.info + static_cast<size_t>(t) * N;
  const uint32_t es = static_cast<uint32_t>(a.x_es);
  // the row's values, mask bytes and return info as raw buffers: a lane's
  // item s sits at a 32-bit byte offset, and lanes past the row read 0
  const __amdgpu_buffer_rsrc_t xb = raw_buffer(xrow, (static_cast<uint32_t>(N - 1) * es + 1u) * 4u);
  const __amdgpu_buffer_rsrc_t ib = raw_buffer(irow, static_cast<uint32_t>(N) * 4u);
  const uint32_t xo = static_cast<uint32_t>(tid) * es * 4u, xstep = es * (4u * kThreads);

  // (0) every load of the row in flight; counts and bitmaps cleared meanwhile
  float v[E];
  uint32_t mr[E], ri[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    v[s] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xb, xo + s * xstep, 0, 0));
    ri[s] = __builtin_amdgcn_raw_buffer_load_b32(ib, static_cast<uint32_t>(s * kThreads + tid) * 4u, 0, 0);
  }
  // one uniform branch around all mask loads (a per-item branch would wait
  // for every load issued before it)
  if (mrow != nullptr) {
    const __amdgpu_buffer_rsrc_t mb = raw_buffer(mrow, static_cast<uint32_t>(N));
#pragma unroll
    for (int s = 0; s < E; ++s)
      mr[s] = __builtin_amdgcn_raw_buffer_load_b8(mb, static_cast<uint32_t>(s * kThreads + tid), 0, 0);
  } else {
    // no mask: the row bound stands in for it (the buffer reads 0 past the
    // row, so a mask byte there is 0 already)
#pragma unroll
    for (int s = 0; s < E; ++s) mr[s] = s * kThreads + tid < N ? 1u : 0u;
  }
  for (int i = tid; i <= NB / 2; i += kThreads) H[i] = 0u;
  for (int i = tid; i < nw; i += kThreads) {
    RB[i] = make_uint2(0u, 0u);
    NP[i] = make_uint2(0u, 0u);
  }
  __syncthreads();

  // (1) keys (all-ones when invalid: masked, non-finite or past the row);
  // pairwise-valid items (return valid too) set their return position in RB;
  // the valid value range (invalid items enter min / max as NaN, which fminf /
  // fmaxf skip); the counts
  uint32_t key[E];
  float vlo = __uint_as_float(kQuietNaN), vhi = vlo;
  uint32_t nF = 0u, nP = 0u;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    // bitwise, not short-circuit: no branch per item
    const bool fv = (mr[s] != 0u) & __builtin_isfinite(v[s]);  // (0 past the row)
    const bool P = fv & (ri[s] != 0u);
    key[s] = fv ? bucket::canonical_key(v[s]) : kInvalid;
    nF += static_cast<uint32_t>(__popcll(ballot(fv)));
    nP += static_cast<uint32_t>(__popcll(ballot(P)));
    const float vv = fv ? v[s] : __uint_as_float(kQuietNaN);
    vlo = fminf(vlo, vv);
    vhi = fmaxf(vhi, vv);
    if (P) {
      const uint32_t rp = ri[s] - 1u;  // low half: 0-based return position
      bucket::lds_or(&RB[__builtin_amdgcn_ubfe(rp, 5, 11)].x, 1u << (rp & 31u));
    }
  }
  {
    auto fmn = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fminf(__uint_as_float(x), __uint_as_float(y)));
    };
    auto fmx = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fmaxf(__uint_as_float(x), __uint_as_float(y)));
    };
    const uint32_t r0 = last_lane(wave_incl_scan(__float_as_uint(vlo), kQuietNaN, fmn));
    const uint32_t r1 = last_lane(wave_incl_scan(__float_as_uint(vhi), kQuietNaN, fmx));
    if (lane == 0) {
      uint32_t* r = misc + bucket::kRed + 4 * w;
      r[0] = r0;
      r[1] = r1;
      r[2] = nF;
      r[3] = nP;
    }
  }
  __syncthreads();
  // every wave combines the 8 x 4 partials: lane l < 32 holds value l & 3 of
  // wave l >> 2
  uint32_t Kn, m;
  {
    const uint32_t pv = misc[bucket::kRed + (lane & 31)];
    const int k = lane & 3;
    float lo = k == 0 ? __uint_as_float(pv) : __uint_as_float(kQuietNaN);
    float hi = k == 1 ? __uint_as_float(pv) : __uint_as_float(kQuietNaN);
    uint32_t y = k >= 2 ? pv : 0u;
    lo = fminf(lo, __uint_as_float(lane_xor<4>(__float_as_uint(lo))));
    hi = fmaxf(hi, __uint_as_float(lane_xor<4>(__float_as_uint(hi))));
    y += lane_xor<4>(y);
    lo = fminf(lo, __uint_as_float(lane_xor<8>(__float_as_uint(lo))));
    hi = fmaxf(hi, __uint_as_float(lane_xor<8>(__float_as_uint(hi))));
    y += lane_xor<8>(y);
    lo = fminf(lo, __uint_as_float(lane_xor<16>(__float_as_uint(lo))));
    hi = fmaxf(hi, __uint_as_float(lane_xor<16>(__float_as_uint(hi))));
    y += lane_xor<16>(y);
    auto rl = [](uint32_t v2, int l) {
      return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v2), l));
    };
    vlo = __uint_as_float(rl(__float_as_uint(lo), 0));
    vhi = __uint_as_float(rl(__float_as_uint(hi), 1));
    Kn = rl(y, 2);
    m = rl(y, 3);
  }
  if (!kRankOut && static_cast<int>(m) < a.min_valid) {
    if (tid == 0) a.ic[static_cast<size_t>(f) * a.T + t] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }

  // (2) bucket histogram: 16-bit counts, two per word; the count an item's
  // atomic returns is its slot in the bucket. Invalid items add to their
  // lane's junk word after the counts.
  const bucket::Map mp = bucket::make_map(vlo, vhi, NB);
  uint32_t bs[E];  // bucket | slot << 16
  {
    const uint32_t junk_word = static_cast<uint32_t>(NB / 2 + 1 + lane);
    uint32_t old[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t b = bucket::bucket_of_value(v[s], mp);
      bs[s] = b;
      old[s] = bucket::lds_add(&H[key[s] != kInvalid ? (b >> 1) : junk_word],
                               __umul24(b & 1u, 0xFFFFu) + 1u);
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t sh = (bs[s] & 1u) << 4;
      bs[s] = __builtin_amdgcn_ubfe(old[s], sh, 16) << 16 | bs[s];
    }
  }
  __syncthreads();

  // (3) in-place exclusive scan of the counts (H16[b] = items in buckets
  // below b, H16[NB] = Kn), each wave over a contiguous eighth of the
  // words, 64 consecutive words per step; the largest bucket; the
  // exclusive popcount prefix of the return bitmap's words
  uint32_t maxc = 0u;
  {
    const bucket::CountScan cs = bucket::count_scan_wave(H, NB, w, lane, misc);
    const uint32_t rc = tid < nw ? static_cast<uint32_t>(__popc(RB[tid].x)) : 0u;
    const uint32_t inc2 = wave_incl_sum(rc);
    if (lane == 63) misc[bucket::kRbit + w] = inc2;
    __syncthreads();
    const uint32_t off = bucket::below_sum8(misc + bucket::kHist, w, lane);
    const uint32_t off2 = inc2 - rc + bucket::below_sum8(misc + bucket::kRbit, w, lane);
    maxc = bucket::max8(misc + bucket::kBig, lane);
    bucket::count_scan_store(H, NB, w, lane, cs, off);
    if (tid == 0) H[NB >> 1] = Kn;
    if (tid < nw) RB[tid].y = off2;
  }
  __syncthreads();
  if (maxc > static_cast<uint32_t>(bucket::kMaxBucket)) {
    // clustered row: the radix sort (same results), from a fresh LDS layout
    rank_ic_radix_call<E>(kernel_args<IcArgs>(), t, f, smem);
    return;
  }

  // (4) scatter into bucket order (invalid items to the junk words), then
  // the sentinels. Word: key << 32 | column << 18 | P << 16 | return
  // position; for rows of at most 5120 (E <= 10), key << 32 | column << 19 |
  // slot << 13 | return position + 1 (0: not P), whose slot gives the
  // bucket's start from the word's position.
  {
    const uint32_t junk_pos = static_cast<uint32_t>(N + bucket::kPad + lane);
    uint32_t pre[E];
#pragma unroll
    for (int s = 0; s < E; ++s) pre[s] = H16[bs[s] & 0xFFFFu];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t col = static_cast<uint32_t>(s * kThreads + tid);
      uint32_t low;
      if constexpr (kSlotWord)
        low = (col << 19) | ((bs[s] >> 16) << 13) | (ri[s] & 0x1FFFu);
      else
        low = (col << 18) | (ri[s] != 0u ? 0x10000u | ((ri[s] - 1u) & 0xFFFFu) : 0u);
      const uint32_t pos = key[s] != kInvalid ? pre[s] + (bs[s] >> 16) : junk_pos;
      A[pos] = (static_cast<uint64_t>(key[s]) << 32) | low;
    }
  }
  if (tid < bucket::kPad) A[Kn + tid] = ~0ull;
  __syncthreads();

  // (5) in bucket order: each word counts the words of its bucket below it,
  // which gives its rank r (0-based). Pairwise-valid items (P) take their
  // return rank on P from RB; the other valid items mark r in NP, so that
  // the factor rank on P is r + 1 - (non-P items ranked below). The first
  // kHead words of a bucket are compared with no range check (words past its
  // end are larger); longer buckets finish in a loop.
  // Rows of at most 8 * 512 (9 * 512) valid items (E = 10) leave the last
  // two (one) bucket-order steps of every wave empty: the rank and IC phases
  // are compiled for EA steps (8, 9 or E) and the row takes the shortest copy
  // that holds it.
  auto rank_and_ic = [&](auto ea) {
    constexpr int EA = decltype(ea)::value;
    const Span sp = span_of(static_cast<int>(Kn));
    uint32_t outp[EA];  // column | (rank + 1) << 16; 0 for no item
    uint32_t rrs[EA];   // return rank on P; 0 for non-P
    // word fields: column, return position + 1 (0: not P)
    auto w_col = [](uint32_t lo32) { return kSlotWord ? lo32 >> 19 : lo32 >> 18; };
    auto w_rp1 = [](uint32_t lo32) {
      return kSlotWord ? (lo32 & 0x1FFFu) : ((lo32 & 0x10000u) ? (lo32 & 0xFFFFu) + 1u : 0u);
    };
#pragma unroll
    for (int s0 = 0; s0 < EA; s0 += kChunk) {
      uint64_t x[kChunk];
      uint32_t se[kChunk], nb[kChunk];
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        if (s0 + g >= EA) continue;
        const int p = sp.lo + (s0 + g) * 64 + lane;
        const bool act = s0 + g < sp.steps && p < sp.hi;
        x[g] = A[act ? p : static_cast<int>(Kn)];
      }
      // se: bucket start | end << 16 (slot words: start | 1 << 31); 0 for the
      // sentinel of an inactive position (key all-ones)
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        if (s0 + g >= EA) continue;
        const bool valid = static_cast<uint32_t>(x[g] >> 32) != kInvalid;
        if constexpr (kSlotWord) {
          const uint32_t p = static_cast<uint32_t>(sp.lo + (s0 + g) * 64 + lane);
          const uint32_t st = p - ((static_cast<uint32_t>(x[g]) >> 13) & 63u);
          se[g] = valid ? (st | 0x80000000u) : 0u;
        } else {
          const uint32_t b = bucket::bucket_of(static_cast<uint32_t>(x[g] >> 32), mp);
          const uint32_t st = H16[b], en = H16[b + 1];
          se[g] = valid ? (st | (en << 16)) : 0u;
        }
      }
      bool more[kChunk];  // (slot words) the bucket goes on past the head
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        if (s0 + g >= EA) continue;
        const uint64_t* h = A + (se[g] & 0xFFFFu);
        uint32_t c = 0u;
#pragma unroll
        for (int i = 0; i < bucket::kHead; ++i) c += (h[i] < x[g]) ? 1u : 0u;
        nb[g] = c;
        if constexpr (kSlotWord)
          more[g] = se[g] != 0u &&
                    ((static_cast<uint32_t>(h[bucket::kHead - 1]) >> 13) & 63u) == bucket::kHead - 1;
      }
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        if (s0 + g >= EA) continue;
        const uint32_t st = se[g] & 0xFFFFu;
        if constexpr (kSlotWord) {
          // pairs while the pair's second word is still in the bucket (its
          // slot is its distance from the start); words past the bucket are
          // larger, so the last pair's extra word adds nothing
          for (uint32_t q = st + bucket::kHead; more[g]; q += 2) {
            const uint64_t w0 = A[q], w1 = A[q + 1];
            nb[g] += (w0 < x[g]) ? 1u : 0u;
            nb[g] += (w1 < x[g]) ? 1u : 0u;
            more[g] = ((static_cast<uint32_t>(w1) >> 13) & 63u) == q + 1u - st;
          }
        } else {
          const uint32_t en = se[g] >> 16;
          for (uint32_t q = st + bucket::kHead; q < en; q += 2) {
            nb[g] += (A[q] < x[g]) ? 1u : 0u;
            nb[g] += (A[q + 1] < x[g]) ? 1u : 0u;
          }
        }
      }
      uint2 rw[kChunk];
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        if (s0 + g >= EA) continue;
        const uint32_t rp1 = w_rp1(static_cast<uint32_t>(x[g]));
        const bool P = se[g] != 0u && rp1 != 0u;
        rw[g] = RB[P ? ((rp1 - 1u) >> 5) : 0u];
      }
      // the bitmap reads land before the branches below (otherwise each
      // branch's join waits for whatever is still outstanding)
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
#pragma unroll
      for (int g = 0; g < kChunk; ++g) {
        const int s = s0 + g;
        if (s >= EA) continue;
        const uint32_t lo32 = static_cast<uint32_t>(x[g]);
        const uint32_t rp1 = w_rp1(lo32);
        const bool act = se[g] != 0u;
        const bool P = act && rp1 != 0u;
        const uint32_t r0 = (se[g] & 0xFFFFu) + nb[g];
        const uint32_t rr = bucket::rank_below(rw[g], (rp1 - 1u) & 31u) + 1u;
        rrs[s] = P ? rr : 0u;
        if (act && !P) bucket::lds_or(&NP[r0 >> 5].x, 1u << (r0 & 31u));
        if (P && (rr == 1u || rr == m)) {
          // ends of P in return order, for the constant check
          const uint32_t gs = irow[w_col(lo32)] >> 16;
          if (rr == 1u) misc[bucket::kDeg + 2] = gs;
          if (rr == m) misc[bucket::kDeg + 3] = gs;
        }
        outp[s] = act ? (w_col(lo32) | ((r0 + 1u) << 16)) : 0u;
      }
    }
    __syncthreads();

    // (6) exclusive popcount prefix of the non-P marks, by wave 0 alone (at
    // most 384 words: no cross-wave exchange or second barrier)
    if (w == 0) {
      uint32_t carry = 0u;
      for (int base = 0; base < nw; base += 64) {
        const int i = base + lane;
        const uint32_t rc = i < nw ? static_cast<uint32_t>(__popc(NP[i].x)) : 0u;
        const uint32_t inc = wave_incl_sum(rc);
        if (i < nw) NP[i].y = carry + inc - rc;
        carry += last_lane(inc);
      }
    }
    __syncthreads();

    // (7) factor rank on P and Sum rank_f * rank_r
    // a lane's sum stays below 2^32: at most E items of rank <= E * 512 each
    static_assert(static_cast<unsigned long long>(E) * (E * kThreads) * (E * kThreads) < (1ull << 32),
                  "per-lane rank-product sum in 32 bits");
    uint32_t acc32 = 0u;
    {
      uint2 nw2[EA];
#pragma unroll
      for (int s = 0; s < EA; ++s) nw2[s] = NP[rrs[s] != 0u ? (((outp[s] >> 16) - 1u) >> 5) : 0u];
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): all of them before the branches below
#pragma unroll
      for (int s = 0; s < EA; ++s) {
        const uint32_t r0 = (outp[s] >> 16) - 1u;
        const uint32_t rf = r0 + 1u - bucket::rank_below(nw2[s], r0 & 31u);
        acc32 += __umul24(rf, rrs[s]);
        if (rrs[s] != 0u && (rf == 1u || rf == m)) {
          // ends of P in factor order, for the constant check
          const uint32_t k = static_cast<uint32_t>(A[sp.lo + s * 64 + lane] >> 32);
          if (rf == 1u) misc[bucket::kDeg] = k;
          if (rf == m) misc[bucket::kDeg + 1] = k;
        }
      }
    }
    // the wave's sum from two DPP scans of the 16-bit halves (each total
    // below 2^22), with no LDS round trips
    const uint32_t s_lo = last_lane(wave_incl_sum(acc32 & 0xFFFFu));
    const uint32_t s_hi = last_lane(wave_incl_sum(acc32 >> 16));
    const uint64_t acc = static_cast<uint64_t>(s_lo) + (static_cast<uint64_t>(s_hi) << 16);
    uint64_t* sums = reinterpret_cast<uint64_t*>(misc + bucket::kSum);
    if (lane == 0) sums[w] = acc;
    __syncthreads();  // also ends every read of the bucket-ordered words

    // (8) the ascending masked rank row, staged by column over the words,
    // stored in column order
    if constexpr (kRankOut) {
      uint32_t* stage = reinterpret_cast<uint32_t*>(A);
#pragma unroll
      for (int s = 0; s < EA; ++s)
        if (outp[s] != 0u) stage[outp[s] & 0xFFFFu] = __float_as_uint(static_cast<float>(outp[s] >> 16));
        __syncthreads();
      // stores past the row are dropped by the buffer's range check
      const __amdgpu_buffer_rsrc_t ob = raw_buffer(
          a.rank_out + (static_cast<size_t>(f) * a.T + t) * static_cast<size_t>(N), static_cast<uint32_t>(N) * 4u);
      uint32_t r[E];
#pragma unroll
      for (int s = 0; s < E; ++s) r[s] = stage[min(s * kThreads + tid, N - 1)];  // (in the row)
      int tid_o = tid;  // (offsets recomputed, not held from the loads)
      asm volatile("" : "+v"(tid_o));
#pragma unroll
      for (int s = 0; s < E; ++s)
        __builtin_amdgcn_raw_buffer_store_b32(key[s] != kInvalid ? r[s] : kQuietNaN, ob,
                                              static_cast<uint32_t>(s * kThreads + tid_o) * 4u, 0, 0);
    }
    if (tid == 0) {
      uint64_t Ssum = 0;
      for (int ww = 0; ww < kWaves; ++ww) Ssum += sums[ww];
      // NaN below min_valid, or when the factor or the return is constant on
      // P: its first and last items in that order share a key / tie group
      const bool degenerate =
          static_cast<int>(m) < a.min_valid ||
          (m > 0u && (misc[bucket::kDeg] == misc[bucket::kDeg + 1] ||
                      misc[bucket::kDeg + 2] == misc[bucket::kDeg + 3]));
      a.ic[static_cast<size_t>(f) * a.T + t] = spearman_from_rank_products(Ssum, m, degenerate);
    }
  };
  if constexpr (E == 10) {
    if (Kn <= 8u * kThreads)
      rank_and_ic(std::integral_constant<int, 8>{});
    else if (Kn <= 9u * kThreads)
      rank_and_ic(std::integral_constant<int, 9>{});
    else
      rank_and_ic(std::integral_constant<int, E>{});
  } else {
    rank_and_ic(std::integral_constant<int, E>{});
  }
}

// ---- kernel 2b: per-date forward-return order info by bucketed counting ----
// The info of fwd_info_kernel for float rows, ranked like rank_ic_bucket_kernel:
// a word's position among the valid returns is its bucket's start plus the
// words of its bucket below it, and its tie group's start is the bucket's
// start plus the words of the bucket with a smaller key (equal keys share a
// bucket). Rows with a bucket of more than kMaxBucket items take the radix
// sort.
template <int E>
__device__ __noinline__ void fwd_info_radix_call(const FwdArgs* a, int t, unsigned char* smem) {
  fwd_info_radix<float, E>(*a, t, smem);
}

template <int E>
__global__ __launch_bounds__(kThreads, bucket_waves_per_simd(E)) void fwd_info_bucket_kernel(FwdArgs a) {
  constexpr uint32_t kInvalid = ~0u;
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = static_cast<int>(blockIdx.x);
  const int N = a.N;
  const bucket::Layout L = bucket::layout(N);
  const int NB = L.NB;
  uint64_t* A = reinterpret_cast<uint64_t*>(smem + L.A);
  uint32_t* H = reinterpret_cast<uint32_t*>(smem + L.H);
  const uint16_t* H16 = reinterpret_cast<const uint16_t*>(smem + L.H);
  uint32_t* misc = reinterpret_cast<uint32_t*>(smem + L.misc);
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = lane_id();
  const int w = wave_id();
  const float* rrow = static_cast<const float*>(a.r) + t * a.r_rs;
  const uint8_t* mrow = a.rmask ? a.rmask + t * a.m_rs : nullptr;

  // (0) the row's loads (raw buffers: 32-bit byte offsets, lanes past the row
  // read 0); counts cleared meanwhile
  float v[E];
  uint32_t mr[E];
  const __amdgpu_buffer_rsrc_t rb = raw_buffer(rrow, static_cast<uint32_t>(N) * 4u);
#pragma unroll
  for (int s = 0; s < E; ++s)
    v[s] = __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(rb, static_cast<uint32_t>(s * kThreads + tid) * 4u, 0, 0));
  if (mrow != nullptr) {
    const __amdgpu_buffer_rsrc_t mb = raw_buffer(mrow, static_cast<uint32_t>(N));
#pragma unroll
    for (int s = 0; s < E; ++s)
      mr[s] = __builtin_amdgcn_raw_buffer_load_b8(mb, static_cast<uint32_t>(s * kThreads + tid), 0, 0);
  } else {
    // no mask: the row bound stands in for it (the buffer reads 0 past the
    // row, so a mask byte there is 0 already)
#pragma unroll
    for (int s = 0; s < E; ++s) mr[s] = s * kThreads + tid < N ? 1u : 0u;
  }
  for (int i = tid; i <= NB / 2; i += kThreads) H[i] = 0u;
  __syncthreads();

  // (1) keys (all-ones when invalid), the valid value range, the count
  uint32_t key[E];
  float vlo = __uint_as_float(kQuietNaN), vhi = vlo;
  uint32_t nF = 0u;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const bool fv = (mr[s] != 0u) & __builtin_isfinite(v[s]);  // (0 past the row)
    key[s] = fv ? bucket::canonical_key(v[s]) : kInvalid;
    nF += static_cast<uint32_t>(__popcll(ballot(fv)));
    const float vv = fv ? v[s] : __uint_as_float(kQuietNaN);
    vlo = fminf(vlo, vv);
    vhi = fmaxf(vhi, vv);
  }
  {
    auto fmn = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fminf(__uint_as_float(x), __uint_as_float(y)));
    };
    auto fmx = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fmaxf(__uint_as_float(x), __uint_as_float(y)));
    };
    const uint32_t r0 = last_lane(wave_incl_scan(__float_as_uint(vlo), kQuietNaN, fmn));
    const uint32_t r1 = last_lane(wave_incl_scan(__float_as_uint(vhi), kQuietNaN, fmx));
    if (lane == 0) {
      uint32_t* r = misc + bucket::kRed + 4 * w;
      r[0] = r0;
      r[1] = r1;
      r[2] = nF;
    }
  }
  __syncthreads();
  uint32_t Kn = 0u;
  {
    const uint32_t* r = misc + bucket::kRed + 4 * (lane & 7);
    vlo = fminf(vlo, __uint_as_float(bucket::comb8(r[0], kQuietNaN, bucket::fmin_u)));
    vhi = fmaxf(vhi, __uint_as_float(bucket::comb8(r[1], kQuietNaN, bucket::fmax_u)));
    Kn += bucket::comb8(lane < kWaves ? r[2] : 0u, 0u, bucket::add_u);
  }

  // (2) bucket histogram (slots from the atomics' returns)
  const bucket::Map mp = bucket::make_map(vlo, vhi, NB);
  uint32_t bs[E];  // bucket | slot << 16
  {
    const uint32_t junk_word = static_cast<uint32_t>(NB / 2 + 1 + lane);
    uint32_t old[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t b = bucket::bucket_of_value(v[s], mp);
      bs[s] = b;
      old[s] = bucket::lds_add(&H[key[s] != kInvalid ? (b >> 1) : junk_word],
                               __umul24(b & 1u, 0xFFFFu) + 1u);
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t sh = (bs[s] & 1u) << 4;
      bs[s] = __builtin_amdgcn_ubfe(old[s], sh, 16) << 16 | bs[s];
    }
  }
  __syncthreads();

  // (3) exclusive scan of the counts and the largest bucket (as in rank_ic)
  uint32_t maxc = 0u;
  {
    const bucket::CountScan cs = bucket::count_scan_wave(H, NB, w, lane, misc);
    __syncthreads();
    const uint32_t off = bucket::below_sum8(misc + bucket::kHist, w, lane);
    maxc = bucket::max8(misc + bucket::kBig, lane);
    bucket::count_scan_store(H, NB, w, lane, cs, off);
    if (tid == 0) H[NB >> 1] = Kn;
  }
  __syncthreads();
  if (maxc > static_cast<uint32_t>(bucket::kMaxBucket)) {
    fwd_info_radix_call<E>(kernel_args<FwdArgs>(), t, smem);
    return;
  }

  // (4) scatter into bucket order: key << 32 | column << 18
  {
    const uint32_t junk_pos = static_cast<uint32_t>(N + bucket::kPad + lane);
    uint32_t pre[E];
#pragma unroll
    for (int s = 0; s < E; ++s) pre[s] = H16[bs[s] & 0xFFFFu];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t col = static_cast<uint32_t>(s * kThreads + tid);
      const uint32_t pos = key[s] != kInvalid ? pre[s] + (bs[s] >> 16) : junk_pos;
      A[pos] = (static_cast<uint64_t>(key[s]) << 32) | (col << 18);
    }
  }
  if (tid < bucket::kPad) A[Kn + tid] = ~0ull;
  __syncthreads();

  // (5) position = start + words below; tie-group start = start + words of
  // a smaller key
  const Span sp = span_of(static_cast<int>(Kn));
  uint32_t outp[E];  // column | info: (pos + 1) | (tie start + 1) << 16 kept in two halves
  uint32_t inf[E];
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int p = sp.lo + s * 64 + lane;
    const bool act = s < sp.steps && p < sp.hi;
    const uint64_t x = A[act ? p : static_cast<int>(Kn)];
    const uint32_t xk = static_cast<uint32_t>(x >> 32);
    const uint32_t b = bucket::bucket_of(xk, mp);
    const uint32_t st = act ? H16[b] : 0u, en = act ? H16[b + 1] : 0u;
    uint32_t nb = 0u, nk = 0u;
    for (uint32_t q = st; q < en; q += 2) {
      const uint64_t y0 = A[q], y1 = A[q + 1];
      nb += (y0 < x ? 1u : 0u) + (y1 < x ? 1u : 0u);
      nk += (static_cast<uint32_t>(y0 >> 32) < xk ? 1u : 0u) +
            (static_cast<uint32_t>(y1 >> 32) < xk ? 1u : 0u);
    }
    outp[s] = act ? (static_cast<uint32_t>(x) >> 18) + 1u : 0u;  // column + 1
    inf[s] = (st + nb + 1u) | ((st + nk + 1u) << 16);
  }
  __syncthreads();  // every read of the words done
  uint32_t* stage = reinterpret_cast<uint32_t*>(A);
#pragma unroll
  for (int s = 0; s < E; ++s)
    if (outp[s] != 0u) stage[outp[s] - 1u] = inf[s];
  __syncthreads();
  // (stores past the row are dropped by the buffer's range check)
  const __amdgpu_buffer_rsrc_t ob = raw_buffer(a.info + static_cast<size_t>(t) * N, static_cast<uint32_t>(N) * 4u);
  int tid_o = tid;  // (offsets recomputed, not held from the loads)
  asm volatile("" : "+v"(tid_o));
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const uint32_t c = static_cast<uint32_t>(s * kThreads + tid_o);
    const uint32_t r = stage[min(c, static_cast<uint32_t>(N - 1))];  // (in the row)
    __builtin_amdgcn_raw_buffer_store_b32(key[s] != kInvalid ? r : 0u, ob, c * 4u, 0, 0);
  }
  if (tid == 0) a.kr[t] = static_cast<int>(Kn);
}

// ---- kernel 1b: cs_rank / parameter_scan by bucketed counting (float rows) --
// The ranked set S is the finite cells (masked_only: the finite masked
// cells), ranked like rank_ic_bucket_kernel. Per word of S, over the words of
// its bucket: words below (ascending position p = start + that), words with a
// smaller key (its tie group starts at gs = start + that) and words with a
// key at most its own (the group ends at ge = start + that); equal keys share
// a bucket. Unmasked words of S mark p in a bitmap U, so that the masked
// words before sorted position x are c(x) = x - (U bits below x). The four
// groups are then those of rank_scan_kernel (src/cross_sectional_rank.cu:
// 152-164 evaluated four ways):
//   asc-masked    c(p) + 1                      (masked cells)
//   asc-unmasked  p + 1
//   desc-masked   (Km - c(ge)) + (c(p) - c(gs)) + 1   (masked cells)
//   desc-unmasked (Kn - ge) + (p - gs) + 1
// Rows with a bucket of more than kMaxBucket items take the radix sort.
template <int E>
__device__ __noinline__ void rank_scan_radix_call(const ScanArgs* a, int t, int f, unsigned char* smem) {
  rank_scan_radix<float, E>(*a, t, f, smem);
}

template <int E>
__global__ __launch_bounds__(kThreads, bucket_waves_per_simd(E)) void rank_scan_bucket_kernel(ScanArgs a) {
  constexpr int kChunk = E <= 5 ? E : (E % 5 == 0 ? 5 : (E % 4 == 0 ? 4 : 8));
  static_assert(E % kChunk == 0, "E must be a multiple of the chunk");
  constexpr uint32_t kInvalid = ~0u;
  extern __shared__ __align__(16) unsigned char smem[];
  int t, f;
  if (!scan_segment_of(a, t, f)) return;
  const int N = a.N;
  const bucket::Layout L = bucket::layout(N);
  const int NB = L.NB;
  uint64_t* A = reinterpret_cast<uint64_t*>(smem + L.A);
  uint32_t* H = reinterpret_cast<uint32_t*>(smem + L.H);
  const uint16_t* H16 = reinterpret_cast<const uint16_t*>(smem + L.H);
  uint2* U = reinterpret_cast<uint2*>(smem + L.RB);  // W + 1 words: RB and NP back to back
  uint32_t* misc = reinterpret_cast<uint32_t*>(smem + L.misc);
  const int tid = static_cast<int>(threadIdx.x);
  const int lane = lane_id();
  const int w = wave_id();
  const int nw = (N + 31) / 32 + 1;  // bits up to position N inclusive
  const bool masked_only = a.masked_only != 0;
  const float* xrow = static_cast<const float*>(a.x) + f * a.x_ps + t * a.x_rs;
  const uint8_t* mrow = a.mask ? a.mask + f * a.m_ps + t * a.m_rs : nullptr;
  const uint32_t es = static_cast<uint32_t>(a.x_es);
  // the row's values as a raw buffer: 32-bit byte offsets, lanes past the
  // row read 0
  const __amdgpu_buffer_rsrc_t xb = raw_buffer(xrow, (static_cast<uint32_t>(N - 1) * es + 1u) * 4u);
  const uint32_t xo = static_cast<uint32_t>(tid) * es * 4u, xstep = es * (4u * kThreads);

  // (0) the row's loads; counts and bitmap cleared meanwhile
  float v[E];
  uint32_t mr[E];
#pragma unroll
  for (int s = 0; s < E; ++s)
    v[s] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(xb, xo + s * xstep, 0, 0));
  if (mrow != nullptr) {
    const __amdgpu_buffer_rsrc_t mb = raw_buffer(mrow, static_cast<uint32_t>(N));
#pragma unroll
    for (int s = 0; s < E; ++s)
      mr[s] = __builtin_amdgcn_raw_buffer_load_b8(mb, static_cast<uint32_t>(s * kThreads + tid), 0, 0);
  } else {
#pragma unroll
    for (int s = 0; s < E; ++s) mr[s] = 1u;
  }
  for (int i = tid; i <= NB / 2; i += kThreads) H[i] = 0u;
  for (int i = tid; i < nw; i += kThreads) U[i] = make_uint2(0u, 0u);
  __syncthreads();

  // (1) keys of S (all-ones outside), masked bits, the range of S, counts
  uint32_t key[E];
  uint32_t mbits = 0u;
  float vlo = __uint_as_float(kQuietNaN), vhi = vlo;
  uint32_t nS = 0u, nM = 0u;
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const bool mk = mr[s] != 0u;
    const bool fin = (s * kThreads + tid < N) && __builtin_isfinite(v[s]);
    const bool inS = fin && (mk || !masked_only);
    key[s] = inS ? bucket::canonical_key(v[s]) : kInvalid;
    mbits |= mk ? (1u << s) : 0u;
    nS += static_cast<uint32_t>(__popcll(ballot(inS)));
    nM += static_cast<uint32_t>(__popcll(ballot(inS && mk)));
    const float vv = inS ? v[s] : __uint_as_float(kQuietNaN);
    vlo = fminf(vlo, vv);
    vhi = fmaxf(vhi, vv);
  }
  {
    auto fmn = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fminf(__uint_as_float(x), __uint_as_float(y)));
    };
    auto fmx = [](uint32_t x, uint32_t y) {
      return __float_as_uint(fmaxf(__uint_as_float(x), __uint_as_float(y)));
    };
    const uint32_t r0 = last_lane(wave_incl_scan(__float_as_uint(vlo), kQuietNaN, fmn));
    const uint32_t r1 = last_lane(wave_incl_scan(__float_as_uint(vhi), kQuietNaN, fmx));
    if (lane == 0) {
      uint32_t* r = misc + bucket::kRed + 4 * w;
      r[0] = r0;
      r[1] = r1;
      r[2] = nS;
      r[3] = nM;
    }
  }
  __syncthreads();
  uint32_t Kn = 0u, Km = 0u;
  {
    const uint32_t* r = misc + bucket::kRed + 4 * (lane & 7);
    vlo = fminf(vlo, __uint_as_float(bucket::comb8(r[0], kQuietNaN, bucket::fmin_u)));
    vhi = fmaxf(vhi, __uint_as_float(bucket::comb8(r[1], kQuietNaN, bucket::fmax_u)));
    Kn += bucket::comb8(r[2], 0u, bucket::add_u);
    Km += bucket::comb8(r[3], 0u, bucket::add_u);
  }
  // summed here rather than at the group ranks, which would hold the
  // per-wave counts through the sort
  asm volatile("" : "+v"(Km));

  // (2) bucket histogram (slots from the atomics' returns)
  const bucket::Map mp = bucket::make_map(vlo, vhi, NB);
  uint32_t bs[E];  // bucket | slot << 16
  {
    const uint32_t junk_word = static_cast<uint32_t>(NB / 2 + 1 + lane);
    uint32_t old[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t b = bucket::bucket_of_value(v[s], mp);
      bs[s] = b;
      old[s] = bucket::lds_add(&H[key[s] != kInvalid ? (b >> 1) : junk_word],
                               __umul24(b & 1u, 0xFFFFu) + 1u);
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t sh = (bs[s] & 1u) << 4;
      bs[s] = __builtin_amdgcn_ubfe(old[s], sh, 16) << 16 | bs[s];
    }
  }
  __syncthreads();

  // (3) exclusive scan of the counts and the largest bucket (as in rank_ic)
  uint32_t maxc = 0u;
  {
    const bucket::CountScan cs = bucket::count_scan_wave(H, NB, w, lane, misc);
    __syncthreads();
    const uint32_t off = bucket::below_sum8(misc + bucket::kHist, w, lane);
    maxc = bucket::max8(misc + bucket::kBig, lane);
    bucket::count_scan_store(H, NB, w, lane, cs, off);
    if (tid == 0) H[NB >> 1] = Kn;
  }
  __syncthreads();
  if (maxc > static_cast<uint32_t>(bucket::kMaxBucket)) {
    rank_scan_radix_call<E>(kernel_args<ScanArgs>(), t, f, smem);
    return;
  }

  // (4) scatter into bucket order: key << 32 | column << 18 | masked << 16
  {
    const uint32_t junk_pos = static_cast<uint32_t>(N + bucket::kPad + lane);
    uint32_t pre[E];
#pragma unroll
    for (int s = 0; s < E; ++s) pre[s] = H16[bs[s] & 0xFFFFu];
#pragma unroll
    for (int s = 0; s < E; ++s) {
      const uint32_t col = static_cast<uint32_t>(s * kThreads + tid);
      const uint32_t low = (col << 18) | (((mbits >> s) & 1u) << 16);
      const uint32_t pos = key[s] != kInvalid ? pre[s] + (bs[s] >> 16) : junk_pos;
      A[pos] = (static_cast<uint64_t>(key[s]) << 32) | low;
    }
  }
  if (tid < bucket::kPad) A[Kn + tid] = ~0ull;
  __syncthreads();

  // (5) in bucket order: position p, tie group [gs, ge); unmasked words of S
  // mark p in U
  const Span sp = span_of(static_cast<int>(Kn));
  uint32_t pg[E];  // p | gs << 16
  uint32_t ec[E];  // ge | column << 16 | masked << 31; 0 for no item
#pragma unroll
  for (int s0 = 0; s0 < E; s0 += kChunk) {
    uint64_t x[kChunk];
    uint32_t se[kChunk], nb[kChunk], nk[kChunk];
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const int p = sp.lo + (s0 + g) * 64 + lane;
      const bool act = s0 + g < sp.steps && p < sp.hi;
      x[g] = A[act ? p : static_cast<int>(Kn)];
    }
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const uint32_t b = bucket::bucket_of(static_cast<uint32_t>(x[g] >> 32), mp);
      const uint32_t st = H16[b], en = H16[b + 1];
      se[g] = (static_cast<uint32_t>(x[g] >> 32) != kInvalid) ? (st | (en << 16)) : 0u;
    }
    // word below: low half of nb; key below / key at most: nk halves
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const uint64_t* h = A + (se[g] & 0xFFFFu);
      const uint32_t xk = static_cast<uint32_t>(x[g] >> 32);
      uint32_t c = 0u, k = 0u;
#pragma unroll
      for (int i = 0; i < bucket::kScanHead; ++i) {
        const uint64_t y = h[i];
        const uint32_t yk = static_cast<uint32_t>(y >> 32);
        c += (y < x[g]) ? 1u : 0u;
        k += (yk < xk ? 1u : 0u) + (yk <= xk ? 0x10000u : 0u);
      }
      nb[g] = c;
      nk[g] = k;
    }
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const uint32_t st = se[g] & 0xFFFFu, en = se[g] >> 16;
      const uint32_t xk = static_cast<uint32_t>(x[g] >> 32);
      for (uint32_t q = st + bucket::kScanHead; q < en; ++q) {
        const uint64_t y = A[q];
        const uint32_t yk = static_cast<uint32_t>(y >> 32);
        nb[g] += (y < x[g]) ? 1u : 0u;
        nk[g] += (yk < xk ? 1u : 0u) + (yk <= xk ? 0x10000u : 0u);
      }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): reads done before the mark branches
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const int s = s0 + g;
      const uint32_t st = se[g] & 0xFFFFu;
      const bool act = se[g] != 0u;
      const uint32_t lo32 = static_cast<uint32_t>(x[g]);
      const bool mk = lo32 & 0x10000u;
      const uint32_t p = st + nb[g];
      // words past the bucket (head reads) are larger in both key and word
      const uint32_t gs = st + (nk[g] & 0xFFFFu), ge = st + (nk[g] >> 16);
      pg[s] = p | (gs << 16);
      ec[s] = act ? (ge | ((lo32 >> 18) << 16) | (mk ? 0x80000000u : 0u)) : 0u;
      if (act && !mk) bucket::lds_or(&U[p >> 5].x, 1u << (p & 31u));
    }
  }
  __syncthreads();

  // (6) exclusive popcount prefix of U, by wave 0 alone (at most 385 words)
  if (w == 0) {
    uint32_t carry = 0u;
    for (int base = 0; base < nw; base += 64) {
      const int i = base + lane;
      const uint32_t rc = i < nw ? static_cast<uint32_t>(__popc(U[i].x)) : 0u;
      const uint32_t inc = wave_incl_sum(rc);
      if (i < nw) U[i].y = carry + inc - rc;
      carry += last_lane(inc);
    }
  }
  __syncthreads();

  // (7) the four groups, staged two rows at a time over the words, stored in
  // column order
  uint32_t g01[E], g23[E];  // ranks (0: none) of groups 0 | 1 << 16 and 2 | 3 << 16
#pragma unroll
  for (int s0 = 0; s0 < E; s0 += kChunk) {
    uint2 up[kChunk], ugs[kChunk], uge[kChunk];
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      // no select on activity: an inactive position has p, gs <= kScanHead and
      // ec = 0, so its reads stay in word 0 (a select let the compiler split
      // the three reads into a branch per item)
      const int s = s0 + g;
      up[g] = U[(pg[s] & 0xFFFFu) >> 5];
      ugs[g] = U[pg[s] >> 21];
      uge[g] = U[(ec[s] & 0xFFFFu) >> 5];
    }
#pragma unroll
    for (int g = 0; g < kChunk; ++g) {
      const int s = s0 + g;
      const uint32_t p = pg[s] & 0xFFFFu, gs = pg[s] >> 16, ge = ec[s] & 0xFFFFu;
      const bool act = ec[s] != 0u;
      const bool mk = ec[s] >> 31;
      const uint32_t cp = p - bucket::rank_below(up[g], p & 31u);
      const uint32_t cgs = gs - bucket::rank_below(ugs[g], gs & 31u);
      const uint32_t cge = ge - bucket::rank_below(uge[g], ge & 31u);
      const uint32_t r0 = mk ? cp + 1u : 0u;
      const uint32_t r1 = p + 1u;
      const uint32_t r2 = mk ? (Km - cge) + (cp - cgs) + 1u : 0u;
      const uint32_t r3 = (Kn - ge) + (p - gs) + 1u;
      g01[s] = act ? (r0 | (r1 << 16)) : 0u;
      g23[s] = act ? (r2 | (r3 << 16)) : 0u;
      // packed here, so the word counts are not held into the stores
      asm volatile("" : "+v"(g01[s]), "+v"(g23[s]));
    }
  }
  uint32_t* stage = reinterpret_cast<uint32_t*>(A);  // 2 rows of N words
  const size_t orow = (static_cast<size_t>(f) * a.T + t) * static_cast<size_t>(N);
  int tid_o = tid;
  asm volatile("" : "+v"(tid_o));
  // the staged rows start at 0 (no rank: NaN), so the output needs no
  // validity of its own
  auto emit2 = [&](const uint32_t (&gg)[E], float* out_lo, float* out_hi) {
    if (out_lo == nullptr && out_hi == nullptr) return;
    __syncthreads();  // earlier reads of the stage / words done
    uint4* st4 = reinterpret_cast<uint4*>(stage);
    for (int i = tid; i < (2 * N + 3) / 4; i += kThreads) st4[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < E; ++s) {
      if (ec[s] != 0u) {
        const uint32_t col = (ec[s] >> 16) & 0x3FFFu;
        stage[col] = gg[s] & 0xFFFFu;
        stage[N + col] = gg[s] >> 16;
      }
    }
    __syncthreads();
    // the rows as raw buffers (stores past the row are dropped)
    const uint32_t rowb = static_cast<uint32_t>(N) * 4u;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float* const o = h == 0 ? out_lo : out_hi;
      if (o == nullptr) continue;
      const __amdgpu_buffer_rsrc_t ob = raw_buffer(o + orow, rowb);
#pragma unroll
      for (int s = 0; s < E; ++s) {
        const uint32_t c = static_cast<uint32_t>(s * kThreads + tid_o);
        const uint32_t r = stage[h * N + min(c, static_cast<uint32_t>(N - 1))];  // (in the row)
        __builtin_amdgcn_raw_buffer_store_b32(r != 0u ? __float_as_uint(static_cast<float>(r)) : kQuietNaN, ob,
                                              c * 4u, 0, 0);
      }
    }
  };
  emit2(g01, a.out[0], a.out[1]);
  emit2(g23, a.out[2], a.out[3]);
}

// ---- launchers --------------------------------------------------------------
template <typename Kern>
static hipError_t prep_lds(Kern k, size_t bytes) {
  if (bytes > 65536) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(bytes));
  }
  return hipSuccess;
}

// Items per lane: smallest instantiated E with 512*E >= N.
int rank_items_per_lane(int N) {
  const int need = (N + kThreads - 1) / kThreads;
  if (need <= 2) return 2;
  if (need <= 4) return 4;
  if (need <= 10) return 10;
  if (need <= 16) return 16;
  if (need <= 24) return 24;
  return 0;
}

#define FC_DISPATCH_E(N, KERNEL, V, ...)                                  \
  switch (rank_items_per_lane(N)) {                                      \
    case 2: return launch_one(KERNEL<V, 2>, __VA_ARGS__);                \
    case 4: return launch_one(KERNEL<V, 4>, __VA_ARGS__);                \
    case 10: return launch_one(KERNEL<V, 10>, __VA_ARGS__);              \
    case 16: return launch_one(KERNEL<V, 16>, __VA_ARGS__);              \
    case 24: return launch_one(KERNEL<V, 24>, __VA_ARGS__);              \
    default: return hipErrorLaunchOutOfResources;                        \
  }

template <typename Kern, typename Args>
static hipError_t launch_one(Kern k, dim3 grid, size_t lds, hipStream_t s, const Args& a) {
  hipError_t e = prep_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, grid, dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_rank_scan(const ScanArgs& a, int value_bytes, hipStream_t s) {
  const size_t lds = lds_layout(a.N, value_bytes == 8 ? 8 : 4).total;
  const dim3 grid(8u * static_cast<unsigned>((a.T + 7) / 8) * static_cast<unsigned>(a.F));
  if (value_bytes == 8) {
    FC_DISPATCH_E(a.N, rank_scan_kernel, double, grid, lds, s, a)
  }
  // float rows: the bucketed kernel (32-bit byte offsets, its LDS layout
  // and its fallback's radix layout within the device's per-workgroup LDS)
  const size_t blds = std::max(bucket::layout(a.N).total, lds);
  int dev = 0, cap = 0;
  const bool short_row = a.x_es > 0 && 4ll * a.x_es * a.N < (1ll << 31);
  if (short_row && hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      blds <= static_cast<size_t>(cap)) {
    switch (rank_items_per_lane(a.N)) {
      case 2: return launch_one(rank_scan_bucket_kernel<2>, grid, blds, s, a);
      case 4: return launch_one(rank_scan_bucket_kernel<4>, grid, blds, s, a);
      case 10: return launch_one(rank_scan_bucket_kernel<10>, grid, blds, s, a);
      case 16: return launch_one(rank_scan_bucket_kernel<16>, grid, blds, s, a);
      case 24: return launch_one(rank_scan_bucket_kernel<24>, grid, blds, s, a);
      default: return hipErrorLaunchOutOfResources;
    }
  }
  FC_DISPATCH_E(a.N, rank_scan_kernel, float, grid, lds, s, a)
}

hipError_t launch_fwd_info(const FwdArgs& a, int value_bytes, hipStream_t s) {
  const size_t lds = lds_layout(a.N, value_bytes == 8 ? 8 : 4).total;
  const dim3 grid(static_cast<unsigned>(a.T));
  if (value_bytes == 8) {
    FC_DISPATCH_E(a.N, fwd_info_kernel, double, grid, lds, s, a)
  }
  // float rows: the bucketed kernel when the device's per-workgroup LDS holds
  // its layout (and the radix layout of its fallback)
  const size_t blds = std::max(bucket::layout(a.N).total, lds);
  int dev = 0, cap = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      blds <= static_cast<size_t>(cap)) {
    switch (rank_items_per_lane(a.N)) {
      case 2: return launch_one(fwd_info_bucket_kernel<2>, grid, blds, s, a);
      case 4: return launch_one(fwd_info_bucket_kernel<4>, grid, blds, s, a);
      case 10: return launch_one(fwd_info_bucket_kernel<10>, grid, blds, s, a);
      case 16: return launch_one(fwd_info_bucket_kernel<16>, grid, blds, s, a);
      case 24: return launch_one(fwd_info_bucket_kernel<24>, grid, blds, s, a);
      default: return hipErrorLaunchOutOfResources;
    }
  }
  FC_DISPATCH_E(a.N, fwd_info_kernel, float, grid, lds, s, a)
}

hipError_t launch_rank_ic(const IcArgs& a, int value_bytes, hipStream_t s) {
  const size_t lds = lds_layout(a.N, value_bytes == 8 ? 8 : 4).total;
  const unsigned tgroups = static_cast<unsigned>((a.T + 7) / 8);
  const dim3 grid(a.sched == kSchedXcd ? 8u * tgroups * static_cast<unsigned>(a.F)
                                       : static_cast<unsigned>(a.T) * static_cast<unsigned>(a.F));
  if (value_bytes == 8) {
    FC_DISPATCH_E(a.N, rank_ic_kernel, double, grid, lds, s, a)
  }
  // float rows: the bucketed kernel, whose LDS also holds the radix layout
  // of its fallback, when the device's per-workgroup LDS allows
  const size_t blds = std::max(bucket::layout(a.N).total, lds);
  int dev = 0, cap = 0;
  // (the bucketed kernel addresses a row with 32-bit byte offsets; the
  // pairwise-set rank diagnostic, F = 1 only, stays on the radix kernel)
  const bool short_row = a.x_es > 0 && 4ll * a.x_es * a.N < (1ll << 31);
  if (short_row && a.rank_f == nullptr && a.rank_r == nullptr && hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cap, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      blds <= static_cast<size_t>(cap)) {
    const bool ro = a.rank_out != nullptr;
    switch (rank_items_per_lane(a.N)) {
      case 2: return ro ? launch_one(rank_ic_bucket_kernel<2, true>, grid, blds, s, a)
                        : launch_one(rank_ic_bucket_kernel<2, false>, grid, blds, s, a);
      case 4: return ro ? launch_one(rank_ic_bucket_kernel<4, true>, grid, blds, s, a)
                        : launch_one(rank_ic_bucket_kernel<4, false>, grid, blds, s, a);
      case 10: return ro ? launch_one(rank_ic_bucket_kernel<10, true>, grid, blds, s, a)
                         : launch_one(rank_ic_bucket_kernel<10, false>, grid, blds, s, a);
      case 16: return ro ? launch_one(rank_ic_bucket_kernel<16, true>, grid, blds, s, a)
                         : launch_one(rank_ic_bucket_kernel<16, false>, grid, blds, s, a);
      case 24: return ro ? launch_one(rank_ic_bucket_kernel<24, true>, grid, blds, s, a)
                         : launch_one(rank_ic_bucket_kernel<24, false>, grid, blds, s, a);
      default: return hipErrorLaunchOutOfResources;
    }
  }
  FC_DISPATCH_E(a.N, rank_ic_kernel, float, grid, lds, s, a)
}

}  // namespace fcr
