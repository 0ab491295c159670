// Device-memory plans of the host drivers: the workspace slots each driver
// fills and their sizes. One plan serves both the exact per-op byte formula
// (fc_op_device_bytes) and the free-HBM check every host driver runs before
// it allocates. Library-private.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

struct fc_workspace;

namespace fcm {

struct MemPlan {
  static constexpr int kMax = 16;  // the largest plan (rolling_ic, wide rows) has 14 entries
  int slot[kMax];
  size_t bytes[kMax];
  int n = 0;
  bool overflow = false;  // an add beyond kMax: the plan is unusable
  void add(int s, size_t b) {
    if (n >= kMax) {
      overflow = true;
      return;
    }
    slot[n] = s;
    bytes[n] = b;
    ++n;
  }
};

// Bytes the tracker charges for an allocation of b bytes (256-byte granules,
// at least one).
inline size_t granule(size_t b) {
  const size_t g = (b + 255) & ~static_cast<size_t>(255);
  return g == 0 ? 256 : g;
}

inline size_t plan_total(const MemPlan& p) {
  size_t t = 0;
  for (int i = 0; i < p.n; ++i) t += granule(p.bytes[i]);
  return t;
}

// Plans (0 on success, a negative contract code on invalid arguments).
int plan_cs_rank(int T, int N, bool mask, MemPlan& p);
int plan_parameter_scan(int T, int N, unsigned groups, MemPlan& p);
int plan_rolling_ic(int T, int N, int value_bytes, bool fmask, bool rmask, bool ranks,
                    MemPlan& p);
int plan_stock_corr(int T, int N, int value_bytes, bool mask, int precision, MemPlan& p);
int plan_factor_corr(int T, int N, int F, int value_bytes, bool mask, int precision,
                     MemPlan& p);

// hipSuccess when the bytes the plan adds to ws (slots it must grow; all of
// it for a new workspace) fit the device's free HBM, hipErrorOutOfMemory
// otherwise. Nothing is allocated.
hipError_t plan_fits(const MemPlan& p, const fc_workspace* ws);

}  // namespace fcm
