// Internal launch interfaces of the correlation kernels (stock_corr and
// factor_corr). Nothing here is exported.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fcc {

constexpr int kTierContract = 0;  // FP64 MFMA sums, shifted/2^k-scaled columns
constexpr int kTierFp32 = 1;      // FP32 MFMA sums of standardised columns
constexpr int kTierBf16x3 = 2;    // bf16 hi/lo split MFMA + INT8 MFMA counts

// Per-column statistics over the column's valid cells (finite && mask).
struct ColStat {
  double n;       // valid count
  double mean;    // shift applied before the contraction
  double scale;   // multiplier applied after the shift (power of two in the contract tier)
  double sxx;     // sum of squared deviations from mean
  double minv, maxv;
  double amax;    // max |x|
  double aminnz;  // min nonzero |x| (1e308 if none)
  double diag;    // 1.0 iff the self-correlation is finite, else NaN
  double pad[7];
};
static_assert(sizeof(ColStat) == 128, "ColStat is 128 bytes");

// Operand planes of the stock_corr contraction, layout [Np][Tp] (row = stock).
struct StockPlanes {
  void* z;      // contract: double, fp32: float, bf16x3: bf16 hi
  void* lo;     // bf16x3: bf16 lo
  void* qh;     // bf16x3: bf16 hi of (hi + lo)^2
  void* ql;     // bf16x3: bf16 lo of (hi + lo)^2
  void* m;      // contract: double 0/1, fp32: float 0/1, bf16x3: bf16 0/1
  int8_t* m8;   // bf16x3: int8 0/1
  void* xt;     // original values, NaN where invalid (exact fallback pass),
                // in the input's precision
  int xt_bytes; // 4 (float input) or 8 (double input)
  int Np, Tp;   // rows padded to a multiple of 128 (tile height), T to 64
};
inline int stock_rows_padded(int N) { return (N + 127) / 128 * 128; }
inline int stock_t_padded(int T) { return (T + 63) / 64 * 64; }
struct ColPart;  // per-chunk partial column statistics

size_t stock_plane_bytes(int tier, int N, int T, int value_bytes);
int stock_tile(int tier);

size_t stock_colpart_bytes(int T, int N);
// fb_count (may be null): the tile kernels' fallback counter, cleared by the
// column-statistics kernel
hipError_t launch_stock_colprep(const void* X, int value_bytes, const uint8_t* mask, int T, int N,
                                int tier, ColStat* cs, ColPart* parts, const StockPlanes& P,
                                uint32_t* fb_count, hipStream_t s);
hipError_t launch_stock_fullsums(const StockPlanes& P, int N, ColStat* cs, hipStream_t s);
// full: 1 when every cell is valid (the fully-valid path: only Sum(xy) is contracted)
hipError_t launch_stock_tiles(int tier, int full, const StockPlanes& P, const ColStat* cs, int T,
                              int N, double* out, uint32_t* fb_list, uint32_t* fb_count,
                              uint32_t fb_cap, hipStream_t s);
hipError_t launch_stock_fallback(const StockPlanes& P, int T, int N, const uint32_t* fb_list,
                                 const uint32_t* fb_count, uint32_t fb_cap, double* out,
                                 hipStream_t s);

// factor_corr: rows r = t*N + i of a (T*N, F) panel, mask per row.
struct FactorArgs {
  const void* x;
  int value_bytes;
  const uint8_t* mask;
  long long R;
  int F;
};
size_t factor_scratch_bytes(int tier, long long R, int F, int* n_groups);
hipError_t launch_factor_corr(int tier, const FactorArgs& a, void* scratch, size_t scratch_bytes,
                              double* out, hipStream_t s);

}  // namespace fcc
