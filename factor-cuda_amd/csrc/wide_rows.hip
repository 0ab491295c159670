// Rows longer than the register-resident rank kernels hold (more than
// 512 * 24 items, up to the 2^24 cap): every row's (canonical key, column)
// pairs are sorted by rocPRIM's segmented radix sort, which is stable, so
// equal keys keep column order -- the reference's stable ordinal ties
// (src/cross_sectional_rank.cu:42-372 and src/rolling_ic.cu:183-540 sort the
// same pairs with CUB) -- and the ranks are scattered back by column.
// Invalid cells carry the all-ones key, above the canonical key of every
// finite value, so they sort after the valid ones and take no rank.
//
// Float rows store (row << 32 | 32-bit key) and go through ONE device-wide
// stable radix sort over 32 + log2(rows) bits: a segmented sort runs one
// workgroup per row, which leaves the chip idle when a few long rows are all
// the work. Double rows (64-bit keys, no room for the row) keep rocPRIM's
// segmented sort. Either way row t's pairs end up at [t*N, t*N + N).
//
// Work goes plane by plane, so one sort never holds more than T*N pairs
// (T*N <= INT32_MAX is an argument contract). These rows are the rare case;
// the register-resident kernels serve everything up to 12288 columns.
#include "fc_device.h"
#include "fc_internal.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace fcw {

using fcd::ord_key;
using fcr::ScanArgs;

constexpr int kBlock = 256;
constexpr uint32_t kNaNBits = 0x7fc00000u;

struct RowStart {
  int n;
  __host__ __device__ int operator()(int s) const { return s * n; }
};

static unsigned row_bits(long long rows) {
  unsigned b = 0;
  while ((1ll << b) < rows) ++b;
  return b;
}

// Sort nseg rows of n stored keys each (row-major) with their columns.
// COMP: keys are (row << 32 | key32), one device-wide sort; otherwise the
// 64-bit keys of double rows, segmented sort.
template <bool COMP>
static hipError_t sort_rows(void* temp, size_t& temp_bytes, const uint64_t* keys,
                            uint64_t* keys_out, const uint32_t* cols, uint32_t* cols_out, int nseg,
                            int n, hipStream_t s) {
  const unsigned items = static_cast<unsigned>(static_cast<size_t>(nseg) * static_cast<size_t>(n));
  if (COMP)
    return rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, cols, cols_out, items, 0u,
                                     32u + row_bits(nseg), s);
  auto off = rocprim::make_transform_iterator(rocprim::make_counting_iterator<int>(0), RowStart{n});
  return rocprim::segmented_radix_sort_pairs(temp, temp_bytes, keys, keys_out, cols, cols_out,
                                             items, static_cast<unsigned>(nseg), off, off + 1, 0u,
                                             64u, s);
}

size_t wide_temp_bytes(int T, int N, int key_bytes, bool ic) {
  size_t b = 0;
  const hipError_t e =
      key_bytes == 8
          ? sort_rows<false>(nullptr, b, nullptr, nullptr, nullptr, nullptr, T, N, nullptr)
          : sort_rows<true>(nullptr, b, nullptr, nullptr, nullptr, nullptr, ic ? 2 * T : T, N,
                            nullptr);
  return e == hipSuccess ? b : 0;
}

// Stored sort key of a cell of row `row`: float rows (row << 32 | key32),
// double rows the 64-bit key; invalid cells all-ones in the key bits.
__device__ __forceinline__ uint64_t stored_key(float v, bool ok, long long row) {
  const uint32_t k = ok ? ord_key(v) : ~0u;
  return (static_cast<uint64_t>(row) << 32) | k;
}
__device__ __forceinline__ uint64_t stored_key(double v, bool ok, long long) {
  return ok ? ord_key(v) : ~0ull;
}
template <bool COMP>
__device__ __forceinline__ bool stored_none(uint64_t k) {
  return COMP ? static_cast<uint32_t>(k) == ~0u : k == ~0ull;
}

static unsigned grid_for(long long total) {
  const long long b = (total + kBlock - 1) / kBlock;
  return static_cast<unsigned>(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

// ---- rank scan (cs_rank / parameter_scan groups) ---------------------------
// Plane f, group g (0 asc-masked, 1 asc-unmasked, 2 desc-masked,
// 3 desc-unmasked): keys of row t at [t*N, t*N + N), the output plane
// prefilled with the invalid-cell NaN. Descending ranks are ascending ranks
// of the negated values.
template <typename V>
__global__ __launch_bounds__(kBlock) void wide_scan_keys(ScanArgs a, int f, int g, uint64_t* keys,
                                                         uint32_t* cols) {
  const bool masked = (g & 1) == 0;
  const bool desc = g >= 2;
  const long long total = static_cast<long long>(a.T) * a.N;
  const V* x = static_cast<const V*>(a.x) + f * a.x_ps;
  const uint8_t* m = (masked && a.mask != nullptr) ? a.mask + f * a.m_ps : nullptr;
  uint32_t* out = reinterpret_cast<uint32_t*>(a.out[g]) + static_cast<size_t>(f) * total;
  for (long long c = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; c < total;
       c += static_cast<long long>(gridDim.x) * kBlock) {
    const int t = static_cast<int>(c / a.N);
    const int i = static_cast<int>(c - static_cast<long long>(t) * a.N);
    const V v = x[t * a.x_rs + i * a.x_es];
    const bool ok = __builtin_isfinite(v) && (m == nullptr || m[t * a.m_rs + i] != 0);
    keys[c] = stored_key(desc ? -v : v, ok, t);
    cols[c] = static_cast<uint32_t>(i);
    out[c] = kNaNBits;
  }
}

// Sorted position p of row seg -> rank p + 1 at the pair's column.
template <bool COMP>
__global__ __launch_bounds__(kBlock) void wide_scatter_f32(const uint64_t* keys_out,
                                                           const uint32_t* cols_out,
                                                           long long total, int N, float* out) {
  for (long long c = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; c < total;
       c += static_cast<long long>(gridDim.x) * kBlock) {
    if (stored_none<COMP>(keys_out[c])) continue;
    const long long row = c / N;
    const int p = static_cast<int>(c - row * N);
    out[row * N + cols_out[c]] = static_cast<float>(p + 1);
  }
}

template <typename V>
static hipError_t wide_rank_scan_t(const ScanArgs& a, const WideBufs& b, hipStream_t s) {
  constexpr bool COMP = sizeof(V) == 4;
  const long long total = static_cast<long long>(a.T) * a.N;
  uint64_t* keys = static_cast<uint64_t*>(b.keys);
  uint64_t* keys_out = static_cast<uint64_t*>(b.keys_out);
  for (int f = 0; f < a.F; ++f) {
    for (int g = 0; g < 4; ++g) {
      if (a.out[g] == nullptr) continue;
      hipLaunchKernelGGL((wide_scan_keys<V>), dim3(grid_for(total)), dim3(kBlock), 0, s, a, f, g,
                         keys, b.cols);
      size_t tb = b.temp_bytes;
      hipError_t e = sort_rows<COMP>(b.temp, tb, keys, keys_out, b.cols, b.cols_out, a.T, a.N, s);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((wide_scatter_f32<COMP>), dim3(grid_for(total)), dim3(kBlock), 0, s,
                         keys_out, b.cols_out, total, a.N,
                         a.out[g] + static_cast<size_t>(f) * static_cast<size_t>(total));
    }
  }
  return hipGetLastError();
}

hipError_t wide_rank_scan(const ScanArgs& a, int value_bytes, const WideBufs& b, hipStream_t s) {
  return value_bytes == 8 ? wide_rank_scan_t<double>(a, b, s) : wide_rank_scan_t<float>(a, b, s);
}

// ---- Spearman IC ---------------------------------------------------------
// Plane f: T factor rows, then T return rows, both keyed on
// the pairwise-valid set P (factor finite and unmasked, return finite and
// unmasked); everything outside P takes the all-ones key. Ranks of both
// sides are cleared here and written by the scatter for P only.
template <typename V>
__global__ __launch_bounds__(kBlock) void wide_ic_keys(WideIcArgs a, int f, uint64_t* keys,
                                                       uint32_t* cols, uint32_t* rf, uint32_t* rr) {
  const long long total = static_cast<long long>(a.T) * a.N;
  const V* x = static_cast<const V*>(a.x) + f * a.x_ps;
  const V* r = static_cast<const V*>(a.r);
  const uint8_t* fm = a.fmask ? a.fmask + f * a.m_ps : nullptr;
  for (long long c = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; c < total;
       c += static_cast<long long>(gridDim.x) * kBlock) {
    const int t = static_cast<int>(c / a.N);
    const int i = static_cast<int>(c - static_cast<long long>(t) * a.N);
    const V xv = x[t * a.x_rs + i * a.x_es];
    const V rv = r[c];
    const bool p = __builtin_isfinite(xv) && (fm == nullptr || fm[t * a.m_rs + i] != 0) &&
                   __builtin_isfinite(rv) && (a.rmask == nullptr || a.rmask[c] != 0);
    // factor rows 0..T-1, return rows T..2T-1
    keys[c] = stored_key(xv, p, t);
    keys[total + c] = stored_key(rv, p, static_cast<long long>(a.T) + t);
    cols[c] = static_cast<uint32_t>(i);
    cols[total + c] = static_cast<uint32_t>(i);
    rf[c] = 0u;
    rr[c] = 0u;
  }
}

template <bool COMP>
__global__ __launch_bounds__(kBlock) void wide_ic_scatter(const uint64_t* keys_out,
                                                          const uint32_t* cols_out, long long total,
                                                          int N, uint32_t* rf, uint32_t* rr) {
  for (long long c = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; c < 2 * total;
       c += static_cast<long long>(gridDim.x) * kBlock) {
    if (stored_none<COMP>(keys_out[c])) continue;
    const bool ret = c >= total;
    const long long cc = ret ? c - total : c;
    const long long row = cc / N;
    const int p = static_cast<int>(cc - row * N);
    (ret ? rr : rf)[row * N + cols_out[c]] = static_cast<uint32_t>(p + 1);
  }
}

// One workgroup per date: m = |P|, S = Sum rf*rr exactly (128-bit), constant
// sides from the first and last sorted keys of P, and the closed-form
// Spearman of ordinal ranks, finalized as rank_ic_kernel does.
// Equal stored keys of one row are equal canonical keys (float rows share
// the row bits), so first == last detects a constant side.
__global__ __launch_bounds__(1024) void wide_ic_rows(WideIcArgs a, int f, const uint64_t* keys_out,
                                                     const uint32_t* rf, const uint32_t* rr) {
  __shared__ unsigned long long part_s[1024];
  __shared__ unsigned long long part_m[1024];
  const int t = blockIdx.x;
  const size_t row = static_cast<size_t>(t) * a.N;
  // per thread at most N/1024 <= 2^14 products below 2^48 each: 64 bits hold it
  unsigned long long m = 0, s = 0;
  for (int i = threadIdx.x; i < a.N; i += 1024) {
    const uint32_t x = rf[row + i];
    const uint32_t y = rr[row + i];
    if (x != 0u) {
      ++m;
      s += static_cast<unsigned long long>(x) * y;
    }
    if (a.rank_f) a.rank_f[row + i] = static_cast<double>(x);
    if (a.rank_r) a.rank_r[row + i] = static_cast<double>(y);
  }
  part_s[threadIdx.x] = s;
  part_m[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  unsigned long long mt = 0;
  unsigned __int128 St = 0;
  for (int k = 0; k < 1024; ++k) {  // fixed order
    mt += part_m[k];
    St += part_s[k];
  }
  const uint64_t* kf = keys_out + row;
  const uint64_t* kr = keys_out + static_cast<size_t>(a.T) * a.N + row;
  const bool degenerate = mt == 0 || static_cast<long long>(mt) < a.min_valid ||
                          kf[0] == kf[mt - 1] || kr[0] == kr[mt - 1];
  double ic = __longlong_as_double(0x7ff8000000000000ll);
  if (!degenerate) {
    const __int128 mm = static_cast<__int128>(mt);
    const __int128 A = 4 * static_cast<__int128>(St) - mm * (mm + 1) * (mm + 1);
    const __int128 Bq = mm * (mm * mm - 1);
    const double sxy = static_cast<double>(A) * 0.25;
    const double sxx = static_cast<double>(Bq) / 12.0;
    ic = (sxx > 0.0) ? (sxy / sqrt(sxx)) / sqrt(sxx) : ic;
  }
  a.ic[static_cast<size_t>(f) * a.T + t] = ic;
}

template <typename V>
static hipError_t wide_ic_t(const WideIcArgs& a, const WideBufs& b, hipStream_t s) {
  constexpr bool COMP = sizeof(V) == 4;
  const long long total = static_cast<long long>(a.T) * a.N;
  uint64_t* keys = static_cast<uint64_t*>(b.keys);
  uint64_t* keys_out = static_cast<uint64_t*>(b.keys_out);
  for (int f = 0; f < a.F; ++f) {
    hipLaunchKernelGGL((wide_ic_keys<V>), dim3(grid_for(total)), dim3(kBlock), 0, s, a, f, keys,
                       b.cols, b.rf, b.rr);
    if (COMP) {  // factor rows 0..T-1 and return rows T..2T-1 in one sort
      size_t tb = b.temp_bytes;
      hipError_t e =
          sort_rows<true>(b.temp, tb, keys, keys_out, b.cols, b.cols_out, 2 * a.T, a.N, s);
      if (e != hipSuccess) return e;
    } else {  // factor rows, then return rows (two sorts of T rows: offsets stay in int)
      for (int side = 0; side < 2; ++side) {
        const size_t o = side == 0 ? 0 : static_cast<size_t>(total);
        size_t tb = b.temp_bytes;
        hipError_t e = sort_rows<false>(b.temp, tb, keys + o, keys_out + o, b.cols + o,
                                        b.cols_out + o, a.T, a.N, s);
        if (e != hipSuccess) return e;
      }
    }
    hipLaunchKernelGGL((wide_ic_scatter<COMP>), dim3(grid_for(2 * total)), dim3(kBlock), 0, s,
                       keys_out, b.cols_out, total, a.N, b.rf, b.rr);
    hipLaunchKernelGGL(wide_ic_rows, dim3(a.T), dim3(1024), 0, s, a, f, keys_out, b.rf, b.rr);
  }
  return hipGetLastError();
}

hipError_t wide_rolling_ic(const WideIcArgs& a, int value_bytes, const WideBufs& b, hipStream_t s) {
  return value_bytes == 8 ? wide_ic_t<double>(a, b, s) : wide_ic_t<float>(a, b, s);
}

}  // namespace fcw
