// Internal (library-private) launch interfaces between the C-ABI host drivers
// and the gfx950 kernels. Nothing here is exported.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fcr {

// Largest row length the LDS ranking kernels accept (column indices are
// stored as 16-bit values in LDS).
constexpr int kMaxLdsN = 65535;

// cs_rank / parameter_scan over F planes of (T, N). Element (f, t, i) of x is
// at x[f*x_ps + t*x_rs + i*x_es]; mask (f, t, i) at mask[f*m_ps + t*m_rs + i]
// (m_ps = 0 shares one (T, N) mask across planes). out[g] (dense [F][T][N]) is
// written for each non-null group g: 0 asc-masked, 1 asc-unmasked,
// 2 desc-masked, 3 desc-unmasked. masked_only = 1 sorts only mask-valid cells
// (allowed when only groups 0/2 are requested).
struct ScanArgs {
  const void* x;
  long long x_ps, x_rs, x_es;
  const uint8_t* mask;
  long long m_ps, m_rs;
  float* out[4];
  int F, T, N;
  int masked_only;
};

// Per-date order information of the forward returns r (T, N), row stride r_rs.
struct FwdArgs {
  const void* r;
  long long r_rs;
  const uint8_t* rmask;
  long long m_rs;
  uint32_t* info;  // [T][N]
  int* kr;         // [T]
  int T, N;
};

// Rank + Spearman IC of F factor planes against the forward-return info.
struct IcArgs {
  const void* x;
  long long x_ps, x_rs, x_es;
  const uint8_t* fmask;
  long long m_ps, m_rs;
  const uint32_t* info;
  const int* kr;
  double* ic;        // [F][T]
  float* rank_out;   // optional [F][T][N] ascending masked ranks
  double* rank_f;    // optional [T][N] (F == 1 only), 0 outside the pairwise set
  double* rank_r;    // optional [T][N]
  int F, T, N;
  int min_valid;
  int sched;         // kSched*: workgroup -> segment schedule
};
constexpr int kSchedXcd = 0;
constexpr int kSchedFMaj = 1;
constexpr int kSchedTMaj = 2;

size_t rank_lds_bytes(int N, int key_bytes);
// Register items per lane chosen for a row of N (0 = N too large for the
// register-resident kernels, whose limit is 512*24 = 12288).
int rank_items_per_lane(int N);
hipError_t launch_rank_scan(const ScanArgs& a, int value_bytes, hipStream_t s);
hipError_t launch_fwd_info(const FwdArgs& a, int value_bytes, hipStream_t s);
hipError_t launch_rank_ic(const IcArgs& a, int value_bytes, hipStream_t s);

}  // namespace fcr

// Rows beyond the register-resident kernels (wide_rows.hip): stable
// segmented radix sort of each row's (key, column) pairs.
namespace fcw {

// Device buffers: keys and keys_out hold 2*T*N 64-bit stored keys (T*N for
// the rank scan), cols and cols_out 2*T*N column indices, rf and rr T*N
// ranks each (IC only), temp temp_bytes of sort storage (wide_temp_bytes).
struct WideBufs {
  void* keys;
  void* keys_out;
  uint32_t* cols;
  uint32_t* cols_out;
  uint32_t* rf;
  uint32_t* rr;
  void* temp;
  size_t temp_bytes;
};

// Spearman IC of F factor planes against one (T, N) return panel (dense,
// row stride N; its mask likewise). Factor element (f,t,i) at
// x[f*x_ps + t*x_rs + i*x_es], factor mask at fmask[f*m_ps + t*m_rs + i].
// ic is [F][T]; rank_f / rank_r (F == 1 only, optional) get the ordinal
// ranks on the pairwise-valid set, 0 elsewhere.
struct WideIcArgs {
  const void* x;
  long long x_ps, x_rs, x_es;
  const uint8_t* fmask;
  long long m_ps, m_rs;
  const void* r;
  const uint8_t* rmask;
  double* ic;
  double* rank_f;
  double* rank_r;
  int F, T, N;
  int min_valid;
};

// rocPRIM temporary storage for sorting nseg rows of n keys of key_bytes
// (0 if the query fails).
size_t wide_temp_bytes(int T, int N, int key_bytes, bool ic);
// The scan groups of ScanArgs over values of value_bytes (4 or 8).
hipError_t wide_rank_scan(const fcr::ScanArgs& a, int value_bytes, const WideBufs& b,
                          hipStream_t s);
hipError_t wide_rolling_ic(const WideIcArgs& a, int value_bytes, const WideBufs& b, hipStream_t s);

}  // namespace fcw
