// C-ABI host drivers of stock_corr and factor_corr (include/fc_hip.h):
// argument contracts, scratch from the workspace, H2D/D2H for the
// host-pointer entry points, domain check, fast/general dispatch and the
// exact fallback pass.
#include "fc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "corr_internal.h"
#include "mem_plan.h"

hipError_t fc_ws_get(fc_workspace* ws, int s, size_t bytes, void** out);
fc_workspace* fc_ws_scoped_create();
void fc_ws_scoped_release(fc_workspace* ws);

namespace {

// workspace slots used here (capi.hip owns slots 0..12)
enum { SC_X = 13, SC_MASK, SC_STATS, SC_PLANES, SC_OUT, SC_FBL, SC_FBC, SC_FSCR, SC_PARTS };

#define FC_TRY(expr)                                      \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return static_cast<int>(_e);    \
  } while (0)

struct ScopedWs {
  fc_workspace* ws;
  bool own;
  explicit ScopedWs(fc_workspace* w) : ws(w ? w : fc_ws_scoped_create()), own(w == nullptr) {}
  ~ScopedWs() {
    if (own) fc_ws_scoped_release(ws);
  }
};

bool valid_tier(int p) { return p == FC_PREC_CONTRACT || p == FC_PREC_FP32 || p == FC_PREC_BF16X3; }

// Stock correlation on device-resident inputs. When stats_host != nullptr the
// column statistics are copied back (synchronising the stream) to apply the
// float64 domain check and select the fully-valid path.
int stock_corr_core(const void* d_X, int vb, const uint8_t* d_mask, int T, int N, int tier,
                    double* d_out, fc_workspace* ws, hipStream_t s, bool host_checks,
                    fc_stock_corr_stats* stats) {
  const int Np = fcc::stock_rows_padded(N);
  const int Tp = fcc::stock_t_padded(T);
  void *dstats = nullptr, *dplanes = nullptr, *dfbl = nullptr, *dfbc = nullptr, *dparts = nullptr;
  FC_TRY(fc_ws_get(ws, SC_PARTS, fcc::stock_colpart_bytes(T, N), &dparts));
  FC_TRY(fc_ws_get(ws, SC_STATS, static_cast<size_t>(Np) * sizeof(fcc::ColStat), &dstats));
  FC_TRY(fc_ws_get(ws, SC_PLANES, fcc::stock_plane_bytes(tier, N, T, vb), &dplanes));
  const uint64_t pairs = static_cast<uint64_t>(N) * (N - 1) / 2;
  const uint32_t cap = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(pairs, 0xFFFFFFF0ull)));
  FC_TRY(fc_ws_get(ws, SC_FBL, static_cast<size_t>(cap) * 8, &dfbl));
  FC_TRY(fc_ws_get(ws, SC_FBC, 256, &dfbc));
  fcc::StockPlanes P{};
  const size_t plane = static_cast<size_t>(Np) * Tp;
  char* base = static_cast<char*>(dplanes);
  P.xt = base;
  P.xt_bytes = vb == 8 ? 8 : 4;
  base += plane * P.xt_bytes;
  if (tier == fcc::kTierContract) {
    P.z = base;
    P.m = base + plane * 8;
  } else if (tier == fcc::kTierFp32) {
    P.z = base;
    P.m = base + plane * 4;
  } else {
    P.z = base;
    P.lo = base + plane * 2;
    P.m = base + plane * 4;
    P.qh = base + plane * 6;
    P.ql = base + plane * 8;
    P.m8 = reinterpret_cast<int8_t*>(base + plane * 10);
  }
  P.Np = Np;
  P.Tp = Tp;
  fcc::ColStat* cs = static_cast<fcc::ColStat*>(dstats);
  FC_TRY(fcc::launch_stock_colprep(d_X, vb, d_mask, T, N, tier, cs,
                                   static_cast<fcc::ColPart*>(dparts), P,
                                   static_cast<uint32_t*>(dfbc), s));
  int full = 0;
  if (host_checks) {
    std::vector<fcc::ColStat> h(static_cast<size_t>(N));
    FC_TRY(hipMemcpyAsync(h.data(), cs, static_cast<size_t>(N) * sizeof(fcc::ColStat),
                          hipMemcpyDeviceToHost, s));
    FC_TRY(hipStreamSynchronize(s));
    bool all_valid = true;
    for (int i = 0; i < N; ++i) {
      if (vb == 8 && h[i].n > 0.0 && (h[i].amax > 1e150 || h[i].aminnz < 1e-150)) return -4;
      if (h[i].n != static_cast<double>(T)) all_valid = false;
    }
    full = (all_valid && tier == fcc::kTierBf16x3) ? 1 : 0;
    if (stats) stats->selected_path = all_valid ? 1 : 0;
  }
  if (full) FC_TRY(fcc::launch_stock_fullsums(P, N, cs, s));
  FC_TRY(fcc::launch_stock_tiles(tier, full, P, cs, T, N, d_out, static_cast<uint32_t*>(dfbl),
                                 static_cast<uint32_t*>(dfbc), cap, s));
  FC_TRY(fcc::launch_stock_fallback(P, T, N, static_cast<const uint32_t*>(dfbl),
                                    static_cast<const uint32_t*>(dfbc), cap, d_out, s));
  if (stats) {
    uint32_t c = 0;
    FC_TRY(hipMemcpyAsync(&c, dfbc, 4, hipMemcpyDeviceToHost, s));
    FC_TRY(hipStreamSynchronize(s));
    stats->fallback_count = static_cast<int>(std::min<uint32_t>(c, cap));
  }
  return 0;
}

}  // namespace

// ---- device-memory plans of the correlation host drivers -------------------
namespace fcm {

int plan_stock_corr(int T, int N, int value_bytes, bool mask, int precision, MemPlan& p) {
  if (T < 1 || N < 1 || (value_bytes != 4 && value_bytes != 8) || !valid_tier(precision))
    return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (static_cast<int64_t>(N) * N > INT32_MAX) return -3;
  const size_t total = static_cast<size_t>(T) * N;
  p.add(SC_X, total * value_bytes);
  p.add(SC_OUT, static_cast<size_t>(N) * N * 8);
  if (mask) p.add(SC_MASK, total);
  p.add(SC_PARTS, fcc::stock_colpart_bytes(T, N));
  p.add(SC_STATS, static_cast<size_t>(fcc::stock_rows_padded(N)) * sizeof(fcc::ColStat));
  p.add(SC_PLANES, fcc::stock_plane_bytes(precision, N, T, value_bytes));
  const uint64_t pairs = static_cast<uint64_t>(N) * (N - 1) / 2;
  const uint64_t cap = std::max<uint64_t>(1, std::min<uint64_t>(pairs, 0xFFFFFFF0ull));
  p.add(SC_FBL, static_cast<size_t>(cap) * 8);
  p.add(SC_FBC, 256);
  return 0;
}

int plan_factor_corr(int T, int N, int F, int value_bytes, bool mask, int precision,
                     MemPlan& p) {
  if (T < 1 || N < 1 || F < 1 || (value_bytes != 4 && value_bytes != 8) || !valid_tier(precision))
    return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (F > 128) return -3;
  const long long R = static_cast<long long>(T) * N;
  p.add(SC_X, static_cast<size_t>(R) * F * value_bytes);
  p.add(SC_OUT, static_cast<size_t>(F) * F * 8);
  if (mask) p.add(SC_MASK, static_cast<size_t>(R));
  int groups = 0;
  p.add(SC_FSCR, fcc::factor_scratch_bytes(precision, R, F, &groups));
  return 0;
}

}  // namespace fcm

extern "C" int fc_stock_corr(const void* h_X, int value_bytes, const uint8_t* h_mask, int T, int N,
                             int precision, double* h_out, fc_stock_corr_stats* stats) {
  if (h_X == nullptr || h_out == nullptr || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (!valid_tier(precision)) return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (static_cast<int64_t>(N) * N > INT32_MAX) return -3;
  ScopedWs sw(nullptr);
  fc_workspace* ws = sw.ws;
  if (ws == nullptr) return static_cast<int>(hipErrorOutOfMemory);
  {
    fcm::MemPlan mp;
    (void)fcm::plan_stock_corr(T, N, value_bytes, h_mask != nullptr, precision, mp);
    FC_TRY(fcm::plan_fits(mp, ws));
  }
  const size_t total = static_cast<size_t>(T) * N;
  void *dx = nullptr, *dm = nullptr, *dout = nullptr;
  FC_TRY(fc_ws_get(ws, SC_X, total * value_bytes, &dx));
  FC_TRY(fc_ws_get(ws, SC_OUT, static_cast<size_t>(N) * N * 8, &dout));
  hipStream_t s = nullptr;
  FC_TRY(hipMemcpyAsync(dx, h_X, total * value_bytes, hipMemcpyHostToDevice, s));
  if (h_mask) {
    FC_TRY(fc_ws_get(ws, SC_MASK, total, &dm));
    FC_TRY(hipMemcpyAsync(dm, h_mask, total, hipMemcpyHostToDevice, s));
  }
  fc_stock_corr_stats st{0, 0};
  int rc = stock_corr_core(dx, value_bytes, static_cast<const uint8_t*>(dm), T, N, precision,
                           static_cast<double*>(dout), ws, s, true, &st);
  if (rc != 0) return rc;
  FC_TRY(hipMemcpyAsync(h_out, dout, static_cast<size_t>(N) * N * 8, hipMemcpyDeviceToHost, s));
  FC_TRY(hipStreamSynchronize(s));
  if (stats) *stats = st;
  return 0;
}

extern "C" int fc_stock_corr_dev(const float* d_X, const uint8_t* d_mask, int T, int N,
                                 int precision, double* d_out, fc_workspace* ws, void* stream) {
  if (d_X == nullptr || d_out == nullptr || ws == nullptr || T < 1 || N < 1) return -1;
  if (!valid_tier(precision)) return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (static_cast<int64_t>(N) * N > INT32_MAX) return -3;
  return stock_corr_core(d_X, 4, d_mask, T, N, precision, d_out, ws,
                         reinterpret_cast<hipStream_t>(stream), false, nullptr);
}

extern "C" int fc_stock_corr_dev_ex(const void* d_X, int value_bytes, const uint8_t* d_mask, int T,
                                    int N, int precision, double* d_out, fc_stock_corr_stats* stats,
                                    fc_workspace* ws, void* stream) {
  if (d_X == nullptr || d_out == nullptr || ws == nullptr || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (!valid_tier(precision)) return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (static_cast<int64_t>(N) * N > INT32_MAX) return -3;
  fc_stock_corr_stats st{0, 0};
  const int rc = stock_corr_core(d_X, value_bytes, d_mask, T, N, precision, d_out, ws,
                                 reinterpret_cast<hipStream_t>(stream), true, &st);
  if (rc == 0 && stats != nullptr) *stats = st;
  return rc;
}

extern "C" int fc_factor_corr(const void* h_F3, int value_bytes, const uint8_t* h_mask, int T,
                              int N, int F, int precision, double* h_out) {
  if (h_F3 == nullptr || h_out == nullptr || T < 1 || N < 1 || F < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (!valid_tier(precision)) return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (F > 128) return -3;
  ScopedWs sw(nullptr);
  fc_workspace* ws = sw.ws;
  if (ws == nullptr) return static_cast<int>(hipErrorOutOfMemory);
  {
    fcm::MemPlan mp;
    (void)fcm::plan_factor_corr(T, N, F, value_bytes, h_mask != nullptr, precision, mp);
    FC_TRY(fcm::plan_fits(mp, ws));
  }
  const long long R = static_cast<long long>(T) * N;
  void *dx = nullptr, *dm = nullptr, *dout = nullptr, *scr = nullptr;
  FC_TRY(fc_ws_get(ws, SC_X, static_cast<size_t>(R) * F * value_bytes, &dx));
  FC_TRY(fc_ws_get(ws, SC_OUT, static_cast<size_t>(F) * F * 8, &dout));
  hipStream_t s = nullptr;
  FC_TRY(hipMemcpyAsync(dx, h_F3, static_cast<size_t>(R) * F * value_bytes, hipMemcpyHostToDevice, s));
  if (h_mask) {
    FC_TRY(fc_ws_get(ws, SC_MASK, static_cast<size_t>(R), &dm));
    FC_TRY(hipMemcpyAsync(dm, h_mask, static_cast<size_t>(R), hipMemcpyHostToDevice, s));
  }
  int groups = 0;
  const size_t sb = fcc::factor_scratch_bytes(precision, R, F, &groups);
  FC_TRY(fc_ws_get(ws, SC_FSCR, sb, &scr));
  fcc::FactorArgs a{dx, value_bytes, static_cast<const uint8_t*>(dm), R, F};
  FC_TRY(fcc::launch_factor_corr(precision, a, scr, sb, static_cast<double*>(dout), s));
  FC_TRY(hipMemcpyAsync(h_out, dout, static_cast<size_t>(F) * F * 8, hipMemcpyDeviceToHost, s));
  FC_TRY(hipStreamSynchronize(s));
  return 0;
}

extern "C" int fc_factor_corr_dev(const float* d_F3, const uint8_t* d_mask, int T, int N, int F,
                                  int precision, double* d_out, fc_workspace* ws, void* stream) {
  return fc_factor_corr_dev_ex(d_F3, 4, d_mask, T, N, F, precision, d_out, ws, stream);
}

extern "C" int fc_factor_corr_dev_ex(const void* d_F3, int value_bytes, const uint8_t* d_mask, int T,
                                     int N, int F, int precision, double* d_out, fc_workspace* ws,
                                     void* stream) {
  if (d_F3 == nullptr || d_out == nullptr || ws == nullptr || T < 1 || N < 1 || F < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (!valid_tier(precision)) return -1;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -2;
  if (F > 128) return -3;
  const long long R = static_cast<long long>(T) * N;
  int groups = 0;
  const size_t sb = fcc::factor_scratch_bytes(precision, R, F, &groups);
  void* scr = nullptr;
  FC_TRY(fc_ws_get(ws, SC_FSCR, sb, &scr));
  fcc::FactorArgs a{d_F3, value_bytes, d_mask, R, F};
  FC_TRY(fcc::launch_factor_corr(precision, a, scr, sb, d_out, reinterpret_cast<hipStream_t>(stream)));
  return 0;
}
