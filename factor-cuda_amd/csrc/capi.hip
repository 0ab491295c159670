// C-ABI host drivers of libfc_hip.so: argument contracts, workspace and
// allocation accounting, H2D/D2H for the host-pointer entry points, and the
// launches of the gfx950 kernels. See include/fc_hip.h for the interface.
#include "fc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>
#include <new>

#include "fc_internal.h"
#include "corr_internal.h"
#include "mem_plan.h"

// ---- allocation accounting (MemTracker equivalent) -------------------------
namespace {

std::atomic<int64_t> g_live{0}, g_peak{0}, g_nalloc{0}, g_nfree{0};

hipError_t tracked_malloc(void** p, size_t bytes) {
  // 256-byte granularity like the reference's accounting (src/mem_tracker.cu:15-68);
  // a zero-byte request takes (and is charged) one granule, as tracked_free
  // and the memory plans count it
  const size_t b = fcm::granule(bytes);
  hipError_t e = hipMalloc(p, b);
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  const int64_t live = g_live.fetch_add(static_cast<int64_t>(b)) + static_cast<int64_t>(b);
  int64_t pk = g_peak.load();
  while (live > pk && !g_peak.compare_exchange_weak(pk, live)) {
  }
  g_nalloc.fetch_add(1);
  return hipSuccess;
}

void tracked_free(void* p, size_t bytes) {
  if (p == nullptr) return;
  (void)hipFree(p);
  g_live.fetch_sub(static_cast<int64_t>(fcm::granule(bytes)));
  g_nfree.fetch_add(1);
}

}  // namespace

// ---- workspace --------------------------------------------------------------
struct fc_workspace {
  static constexpr int kSlots = 32;
  void* p[kSlots] = {};
  size_t cap[kSlots] = {};
  int device = -1;
  // shape of the forward-return order currently held in the INFO/KR slots
  int order_T = -1, order_N = -1, order_vb = 0;
};

namespace fcws {
enum Slot {
  X = 0, MASK, OUT0, OUT1, OUT2, OUT3, R, RMASK, INFO, KR, IC, RANKF, RANKR,
  C0, C1, C2, C3, C4, C5, C6, C7, C8, C9, C10,
  // sort buffers of rows beyond the register-resident kernels
  WK, WK2, WC, WC2, WRF, WRR, WT, kSlotCount
};
static_assert(kSlotCount <= fc_workspace::kSlots, "workspace slot table too small");
}

static void ws_release(fc_workspace* ws) {
  if (ws == nullptr) return;
  int cur = -1;
  (void)hipGetDevice(&cur);
  if (ws->device >= 0 && ws->device != cur) (void)hipSetDevice(ws->device);
  for (int i = 0; i < fc_workspace::kSlots; ++i) {
    tracked_free(ws->p[i], ws->cap[i]);
    ws->p[i] = nullptr;
    ws->cap[i] = 0;
  }
  if (ws->device >= 0 && ws->device != cur && cur >= 0) (void)hipSetDevice(cur);
  ws->device = -1;
  ws->order_T = ws->order_N = -1;
}

// Device buffer of at least `bytes` in slot s, reused while large enough and
// on the same device. Contents are not preserved across growth.
hipError_t fc_ws_get(fc_workspace* ws, int s, size_t bytes, void** out) {
  if (s < 0 || s >= fc_workspace::kSlots) return hipErrorInvalidValue;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (ws->device != dev) {
    if (ws->device >= 0) ws_release(ws);
    ws->device = dev;
  }
  if (ws->cap[s] < bytes || ws->p[s] == nullptr) {
    tracked_free(ws->p[s], ws->cap[s]);
    ws->p[s] = nullptr;
    ws->cap[s] = 0;
    e = tracked_malloc(&ws->p[s], bytes);
    if (e != hipSuccess) return e;
    ws->cap[s] = bytes;
  }
  *out = ws->p[s];
  return hipSuccess;
}

// A workspace that lives for one call when the caller passes none.
struct ScopedWs {
  fc_workspace own;
  fc_workspace* ws;
  explicit ScopedWs(fc_workspace* w) : ws(w ? w : &own) {}
  ~ScopedWs() {
    if (ws == &own) ws_release(&own);
  }
};

static int rank_max_n_for(int value_bytes);

// ---- device-memory plans of the rank-family host drivers --------------------
namespace fcm {

// Sort buffers of rows beyond the register-resident kernels: keys and
// column indices (in and out; both sides of the IC), the IC's two rank
// planes, and rocPRIM's temporary storage (queried; 0 without a device).
static void plan_wide(int T, int N, int key_bytes, bool ic, MemPlan& p) {
  const size_t pairs = static_cast<size_t>(T) * N * (ic ? 2 : 1);
  p.add(fcws::WK, pairs * 8);  // stored keys are 64-bit: (row, key32) or the double key
  p.add(fcws::WK2, pairs * 8);
  p.add(fcws::WC, pairs * 4);
  p.add(fcws::WC2, pairs * 4);
  if (ic) {
    p.add(fcws::WRF, static_cast<size_t>(T) * N * 4);
    p.add(fcws::WRR, static_cast<size_t>(T) * N * 4);
  }
  const size_t temp = fcw::wide_temp_bytes(T, N, key_bytes, ic);
  if (temp == 0) p.overflow = true;  // size query failed: the plan is unusable
  p.add(fcws::WT, temp);
}

static int rank_shape_ok(int T, int N) {
  if (T < 1 || N < 1) return -1;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  return 0;
}

int plan_cs_rank(int T, int N, bool mask, MemPlan& p) {
  const int rc = rank_shape_ok(T, N);
  if (rc != 0) return rc;
  const size_t total = static_cast<size_t>(T) * N;
  p.add(fcws::X, total * 4);
  p.add(fcws::OUT0, total * 4);
  if (mask) p.add(fcws::MASK, total);
  if (N > rank_max_n_for(4)) plan_wide(T, N, 4, false, p);
  return 0;
}

int plan_parameter_scan(int T, int N, unsigned groups, MemPlan& p) {
  const int rc = rank_shape_ok(T, N);
  if (rc != 0) return rc;
  const size_t total = static_cast<size_t>(T) * N;
  p.add(fcws::X, total * 4);
  for (int g = 0; g < 4; ++g)
    if ((groups >> g) & 1u) p.add(fcws::OUT0 + g, total * 4);
  if (groups & 5u) p.add(fcws::MASK, total);  // the masked groups 0 and 2
  if (N > rank_max_n_for(4)) plan_wide(T, N, 4, false, p);
  return 0;
}

int plan_rolling_ic(int T, int N, int value_bytes, bool fmask, bool rmask, bool ranks,
                    MemPlan& p) {
  if (T < 1 || N < 1 || (value_bytes != 4 && value_bytes != 8)) return -1;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  const size_t total = static_cast<size_t>(T) * N;
  p.add(fcws::X, total * value_bytes);
  p.add(fcws::R, total * value_bytes);
  p.add(fcws::IC, static_cast<size_t>(T) * 8);
  if (fmask) p.add(fcws::MASK, total);
  if (rmask) p.add(fcws::RMASK, total);
  if (ranks) {
    p.add(fcws::RANKF, total * 8);
    p.add(fcws::RANKR, total * 8);
  }
  if (N > rank_max_n_for(value_bytes)) {
    plan_wide(T, N, value_bytes, true, p);
  } else {
    p.add(fcws::INFO, total * 4);
    p.add(fcws::KR, static_cast<size_t>(T) * 4);
  }
  return 0;
}

hipError_t plan_fits(const MemPlan& p, const fc_workspace* ws) {
  if (p.overflow) return hipErrorInvalidValue;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const bool same = ws != nullptr && ws->device == dev;
  size_t add = 0;
  for (int i = 0; i < p.n; ++i) {
    const int sl = p.slot[i];
    const bool held = same && ws->p[sl] != nullptr && ws->cap[sl] >= p.bytes[i];
    if (!held) add += granule(p.bytes[i]);
  }
  if (add == 0) return hipSuccess;
  size_t free_b = 0, total_b = 0;
  e = hipMemGetInfo(&free_b, &total_b);
  if (e != hipSuccess) return e;
  return add <= free_b ? hipSuccess : hipErrorOutOfMemory;
}

}  // namespace fcm

extern "C" long long fc_op_device_bytes(int op, int T, int N, int F, int value_bytes,
                                        int precision, int flags) {
  fcm::MemPlan p;
  int rc = -1;
  const bool mask = (flags & FC_PLAN_MASK) != 0;
  switch (op) {
    case FC_OP_CS_RANK:
      rc = fcm::plan_cs_rank(T, N, mask, p);
      break;
    case FC_OP_PARAMETER_SCAN: {
      unsigned groups = static_cast<unsigned>(flags >> 8) & 15u;
      if (groups == 0) groups = 15u;
      rc = fcm::plan_parameter_scan(T, N, groups, p);
      break;
    }
    case FC_OP_ROLLING_IC:
      rc = fcm::plan_rolling_ic(T, N, value_bytes, mask, (flags & FC_PLAN_RMASK) != 0,
                                (flags & FC_PLAN_RANKS) != 0, p);
      break;
    case FC_OP_STOCK_CORR:
      rc = fcm::plan_stock_corr(T, N, value_bytes, mask, precision, p);
      break;
    case FC_OP_FACTOR_CORR:
      rc = fcm::plan_factor_corr(T, N, F, value_bytes, mask, precision, p);
      break;
    default:
      break;
  }
  if (rc != 0) return rc;
  if (p.overflow) return -1;
  return static_cast<long long>(fcm::plan_total(p));
}

extern "C" int fc_mem_info(size_t* free_bytes, size_t* total_bytes) {
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  if (e != hipSuccess) return static_cast<int>(e);
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return 0;
}

fc_workspace* fc_ws_scoped_create() { return new (std::nothrow) fc_workspace(); }
void fc_ws_scoped_release(fc_workspace* ws) {
  ws_release(ws);
  delete ws;
}

extern "C" int fc_workspace_create(fc_workspace** out) {
  if (out == nullptr) return -1;
  *out = new (std::nothrow) fc_workspace();
  return *out ? 0 : static_cast<int>(hipErrorOutOfMemory);
}
extern "C" void fc_workspace_clear(fc_workspace* ws) { ws_release(ws); }
extern "C" void fc_workspace_destroy(fc_workspace* ws) {
  ws_release(ws);
  delete ws;
}

extern "C" void fc_mem_get_stats(fc_mem_stats* out) {
  if (out == nullptr) return;
  out->live_bytes = g_live.load();
  out->peak_bytes = g_peak.load();
  out->n_alloc = g_nalloc.load();
  out->n_free = g_nfree.load();
}
extern "C" void fc_mem_reset_peak(void) { g_peak.store(g_live.load()); }
extern "C" const char* fc_version(void) { return "fc_hip 1.1.1-mi355x"; }

// ---- device capabilities ---------------------------------------------------
static int max_lds_bytes() {
  static int cached[64];
  static std::once_flag once[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  std::call_once(once[dev], [dev]() {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 65536;
    cached[dev] = v;
  });
  return cached[dev];
}

static int rank_max_n_for(int value_bytes) {
  const int kb = value_bytes == 8 ? 8 : 4;
  const size_t cap = static_cast<size_t>(max_lds_bytes());
  int lo = 0, hi = fcr::kMaxLdsN;
  while (hi > 0 && fcr::rank_items_per_lane(hi) == 0) hi = std::min(hi - 1, 512 * 24);
  while (lo < hi) {
    const int mid = lo + (hi - lo + 1) / 2;
    if (fcr::rank_lds_bytes(mid, kb) <= cap) lo = mid; else hi = mid - 1;
  }
  return lo;
}

extern "C" int fc_rank_max_n(int value_bytes) { return rank_max_n_for(value_bytes); }

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define FC_TRY(expr)                             \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return static_cast<int>(_e); \
  } while (0)

// Sort buffers of rows beyond the register-resident kernels, from ws.
static int wide_bufs(fc_workspace* ws, int T, int N, int key_bytes, bool ic, fcw::WideBufs& b) {
  const size_t pairs = static_cast<size_t>(T) * N * (ic ? 2 : 1);
  void *k = nullptr, *k2 = nullptr, *c = nullptr, *c2 = nullptr, *t = nullptr;
  FC_TRY(fc_ws_get(ws, fcws::WK, pairs * 8, &k));
  FC_TRY(fc_ws_get(ws, fcws::WK2, pairs * 8, &k2));
  FC_TRY(fc_ws_get(ws, fcws::WC, pairs * 4, &c));
  FC_TRY(fc_ws_get(ws, fcws::WC2, pairs * 4, &c2));
  b.keys = k;
  b.keys_out = k2;
  b.cols = static_cast<uint32_t*>(c);
  b.cols_out = static_cast<uint32_t*>(c2);
  b.rf = b.rr = nullptr;
  if (ic) {
    void *rf = nullptr, *rr = nullptr;
    FC_TRY(fc_ws_get(ws, fcws::WRF, static_cast<size_t>(T) * N * 4, &rf));
    FC_TRY(fc_ws_get(ws, fcws::WRR, static_cast<size_t>(T) * N * 4, &rr));
    b.rf = static_cast<uint32_t*>(rf);
    b.rr = static_cast<uint32_t*>(rr);
  }
  b.temp_bytes = fcw::wide_temp_bytes(T, N, key_bytes, ic);
  // 0 = rocPRIM's size query failed: no sort may run on a guessed buffer
  if (b.temp_bytes == 0) return static_cast<int>(hipErrorInvalidValue);
  FC_TRY(fc_ws_get(ws, fcws::WT, b.temp_bytes, &t));
  b.temp = t;
  return 0;
}

// ---- rank family: device-pointer entry points ------------------------------
// The scan of F planes; rows beyond the register-resident kernels go through
// the segmented sort with buffers from ws (a scoped workspace, freed after
// the stream drains, when ws is NULL).
static int rank_scan_core(const float* d_X, int F, int T, int N, long long plane_stride,
                          long long row_stride, long long elem_stride, const uint8_t* d_mask,
                          long long mask_plane_stride, float* d_out0, float* d_out1,
                          float* d_out2, float* d_out3, hipStream_t s, fc_workspace* ws) {
  if (d_X == nullptr || F < 1 || T < 1 || N < 1) return -1;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  const bool masked_groups = (d_out0 != nullptr) || (d_out2 != nullptr);
  const bool unmasked_groups = (d_out1 != nullptr) || (d_out3 != nullptr);
  if (masked_groups && d_mask == nullptr) return -1;
  if (!masked_groups && !unmasked_groups) return 0;
  fcr::ScanArgs a{};
  a.x = d_X;
  a.x_ps = plane_stride;
  a.x_rs = row_stride;
  a.x_es = elem_stride;
  a.mask = masked_groups ? d_mask : nullptr;
  a.m_ps = mask_plane_stride;
  a.m_rs = N;
  a.out[0] = d_out0;
  a.out[1] = d_out1;
  a.out[2] = d_out2;
  a.out[3] = d_out3;
  a.F = F;
  a.T = T;
  a.N = N;
  a.masked_only = (masked_groups && !unmasked_groups) ? 1 : 0;
  if (N > rank_max_n_for(4)) {
    ScopedWs sw(ws);
    fcw::WideBufs b{};
    const int rc = wide_bufs(sw.ws, T, N, 4, false, b);
    if (rc != 0) return rc;
    FC_TRY(fcw::wide_rank_scan(a, 4, b, s));
    if (ws == nullptr) FC_TRY(hipStreamSynchronize(s));  // before the scoped buffers go
    return 0;
  }
  FC_TRY(fcr::launch_rank_scan(a, 4, s));
  return 0;
}

extern "C" int fc_rank_scan_dev(const float* d_X, int F, int T, int N, long long plane_stride,
                                long long row_stride, long long elem_stride,
                                const uint8_t* d_mask, long long mask_plane_stride, float* d_out0,
                                float* d_out1, float* d_out2, float* d_out3, void* stream) {
  return rank_scan_core(d_X, F, T, N, plane_stride, row_stride, elem_stride, d_mask,
                        mask_plane_stride, d_out0, d_out1, d_out2, d_out3, as_stream(stream),
                        nullptr);
}

static int fwd_order_core(const void* d_R, int vb, const uint8_t* d_rmask, int T, int N,
                          fc_workspace* ws, hipStream_t s) {
  if (N > rank_max_n_for(vb)) return static_cast<int>(hipErrorLaunchOutOfResources);
  void* info = nullptr;
  void* kr = nullptr;
  ws->order_T = ws->order_N = -1;
  FC_TRY(fc_ws_get(ws, fcws::INFO, static_cast<size_t>(T) * N * 4, &info));
  FC_TRY(fc_ws_get(ws, fcws::KR, static_cast<size_t>(T) * 4, &kr));
  fcr::FwdArgs fa{};
  fa.r = d_R;
  fa.r_rs = N;
  fa.rmask = d_rmask;
  fa.m_rs = N;
  fa.info = static_cast<uint32_t*>(info);
  fa.kr = static_cast<int*>(kr);
  fa.T = T;
  fa.N = N;
  FC_TRY(fcr::launch_fwd_info(fa, vb, s));
  ws->order_T = T;
  ws->order_N = N;
  ws->order_vb = vb;
  return 0;
}

static int rank_ic_ordered_core(const void* d_X, int vb, int F, int T, int N, long long ps,
                                long long rs, long long es, const uint8_t* d_fmask,
                                long long fmask_ps, int min_valid, double* d_ic,
                                float* d_rank_out, double* d_rank_f, double* d_rank_r,
                                fc_workspace* ws, hipStream_t s, int sched = fcr::kSchedXcd) {
  if (ws->order_T != T || ws->order_N != N) return -1;
  if (N > rank_max_n_for(vb)) return static_cast<int>(hipErrorLaunchOutOfResources);
  fcr::IcArgs ia{};
  ia.x = d_X;
  ia.x_ps = ps;
  ia.x_rs = rs;
  ia.x_es = es;
  ia.fmask = d_fmask;
  ia.m_ps = fmask_ps;
  ia.m_rs = N;
  ia.info = static_cast<const uint32_t*>(ws->p[fcws::INFO]);
  ia.kr = static_cast<const int*>(ws->p[fcws::KR]);
  ia.ic = d_ic;
  ia.rank_out = d_rank_out;
  ia.rank_f = d_rank_f;
  ia.rank_r = d_rank_r;
  ia.F = F;
  ia.T = T;
  ia.N = N;
  ia.min_valid = min_valid;
  ia.sched = sched;
  FC_TRY(fcr::launch_rank_ic(ia, vb, s));
  return 0;
}

// Internal entry points of the batched schedule (batch.hip): the checked
// device-pointer paths with the segment schedule exposed.
int fc_internal_fwd_order(const void* d_R, int vb, const uint8_t* d_rmask, int T, int N,
                          fc_workspace* ws, hipStream_t s) {
  return fwd_order_core(d_R, vb, d_rmask, T, N, ws, s);
}
int fc_internal_rank_ic_ordered(const void* d_X, int vb, int F, int T, int N, long long ps,
                                long long rs, long long es, const uint8_t* d_fmask,
                                long long fmask_ps, int min_valid, double* d_ic,
                                float* d_rank_out, fc_workspace* ws, hipStream_t s, int sched) {
  return rank_ic_ordered_core(d_X, vb, F, T, N, ps, rs, es, d_fmask, fmask_ps, min_valid, d_ic,
                              d_rank_out, nullptr, nullptr, ws, s, sched);
}

extern "C" int fc_fwd_order_dev(const void* d_R, int value_bytes, const uint8_t* d_rmask, int T,
                                int N, fc_workspace* ws, void* stream) {
  if (d_R == nullptr || ws == nullptr || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  return fwd_order_core(d_R, value_bytes, d_rmask, T, N, ws, as_stream(stream));
}

extern "C" int fc_rank_ic_ordered_dev(const void* d_X, int value_bytes, int F, int T, int N,
                                      long long plane_stride, long long row_stride,
                                      long long elem_stride, const uint8_t* d_fmask,
                                      long long fmask_plane_stride, int min_valid, double* d_ic,
                                      float* d_rank_out, fc_workspace* ws, void* stream) {
  if (d_X == nullptr || d_ic == nullptr || ws == nullptr || F < 1 || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (min_valid < 2) return -4;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  return rank_ic_ordered_core(d_X, value_bytes, F, T, N, plane_stride, row_stride, elem_stride,
                              d_fmask, fmask_plane_stride, min_valid, d_ic, d_rank_out, nullptr,
                              nullptr, ws, as_stream(stream));
}

extern "C" int fc_rank_ic_dev(const void* d_X, int value_bytes, int F, int T, int N,
                              long long plane_stride, long long row_stride, long long elem_stride,
                              const uint8_t* d_fmask, long long fmask_plane_stride,
                              const void* d_R, const uint8_t* d_rmask, int min_valid,
                              double* d_ic, float* d_rank_out, fc_workspace* ws, void* stream) {
  if (d_X == nullptr || d_R == nullptr || d_ic == nullptr || ws == nullptr || F < 1 || T < 1 ||
      N < 1)
    return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (min_valid < 2) return -4;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  if (N > rank_max_n_for(value_bytes)) {
    // rows beyond the register-resident kernels: segmented sorts
    hipStream_t s = as_stream(stream);
    fcw::WideBufs b{};
    int rc = wide_bufs(ws, T, N, value_bytes, true, b);
    if (rc != 0) return rc;
    fcw::WideIcArgs wa{};
    wa.x = d_X;
    wa.x_ps = plane_stride;
    wa.x_rs = row_stride;
    wa.x_es = elem_stride;
    wa.fmask = d_fmask;
    wa.m_ps = fmask_plane_stride;
    wa.m_rs = N;
    wa.r = d_R;
    wa.rmask = d_rmask;
    wa.ic = d_ic;
    wa.F = F;
    wa.T = T;
    wa.N = N;
    wa.min_valid = min_valid;
    FC_TRY(fcw::wide_rolling_ic(wa, value_bytes, b, s));
    if (d_rank_out != nullptr) {
      // ascending masked cs_rank of each factor plane (finite cells without
      // a mask), from the same buffers
      fcr::ScanArgs sa{};
      sa.x = d_X;
      sa.x_ps = plane_stride;
      sa.x_rs = row_stride;
      sa.x_es = elem_stride;
      sa.mask = d_fmask;
      sa.m_ps = fmask_plane_stride;
      sa.m_rs = N;
      sa.out[0] = d_rank_out;
      sa.F = F;
      sa.T = T;
      sa.N = N;
      FC_TRY(fcw::wide_rank_scan(sa, value_bytes, b, s));
    }
    return 0;
  }
  int rc = fwd_order_core(d_R, value_bytes, d_rmask, T, N, ws, as_stream(stream));
  if (rc != 0) return rc;
  return rank_ic_ordered_core(d_X, value_bytes, F, T, N, plane_stride, row_stride, elem_stride,
                              d_fmask, fmask_plane_stride, min_valid, d_ic, d_rank_out, nullptr,
                              nullptr, ws, as_stream(stream));
}

// ---- allocation-free rank + IC -----------------------------------------------
// fc_rank_ic_dev without a workspace: the return order (info, 4*T*N bytes,
// and the per-date counts, 4*T bytes) lives in one file-scope device array
// that every call on a GPU shares. fwd_info_kernel writes every info cell
// and count of the call's shape before rank_ic_kernel reads them, so no
// clearing is needed and nothing of an earlier call is read. Overlapping
// calls on one GPU race on the array.
namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(32) << 20;  // 32 MiB per GPU
constexpr int kMaxScratchDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

void* g_scratch_ptr[kMaxScratchDevices];
hipError_t g_scratch_err[kMaxScratchDevices];
std::once_flag g_scratch_once[kMaxScratchDevices];

// This device's address of `scratch`, looked up once per device with that
// device current (each GPU holds its own copy of the array).
hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? p : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess) return g_scratch_err[dev];
  if (g_scratch_ptr[dev] == nullptr) return hipErrorInvalidSymbol;
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// info [T][N] u32 at offset 0, counts [T] at the next 256-byte boundary
size_t scratch_counts_offset(int T, int N) {
  return (static_cast<size_t>(T) * N * 4 + 255) & ~static_cast<size_t>(255);
}

}  // namespace

extern "C" long long fc_rank_ic_noalloc_scratch_bytes(void) {
  return static_cast<long long>(kScratchBytes);
}

extern "C" int fc_rank_ic_dev_noalloc(const void* d_X, int value_bytes, int F, int T, int N,
                                      long long plane_stride, long long row_stride,
                                      long long elem_stride, const uint8_t* d_fmask,
                                      long long fmask_plane_stride, const void* d_R,
                                      const uint8_t* d_rmask, int min_valid, double* d_ic,
                                      float* d_rank_out, void* stream) {
  if (d_X == nullptr || d_R == nullptr || d_ic == nullptr || F < 1 || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (min_valid < 2) return -4;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  // beyond the array, or rows that need the long-row sort buffers
  if (scratch_counts_offset(T, N) + static_cast<size_t>(T) * 4 > kScratchBytes) return -3;
  if (N > rank_max_n_for(value_bytes)) return -3;
  void* base = nullptr;
  FC_TRY(scratch_address(&base));
  hipStream_t s = as_stream(stream);
  uint32_t* info = static_cast<uint32_t*>(base);
  int* kr = reinterpret_cast<int*>(static_cast<unsigned char*>(base) +
                                   scratch_counts_offset(T, N));
  fcr::FwdArgs fa{};
  fa.r = d_R;
  fa.r_rs = N;
  fa.rmask = d_rmask;
  fa.m_rs = N;
  fa.info = info;
  fa.kr = kr;
  fa.T = T;
  fa.N = N;
  FC_TRY(fcr::launch_fwd_info(fa, value_bytes, s));
  fcr::IcArgs ia{};
  ia.x = d_X;
  ia.x_ps = plane_stride;
  ia.x_rs = row_stride;
  ia.x_es = elem_stride;
  ia.fmask = d_fmask;
  ia.m_ps = fmask_plane_stride;
  ia.m_rs = N;
  ia.info = info;
  ia.kr = kr;
  ia.ic = d_ic;
  ia.rank_out = d_rank_out;
  ia.F = F;
  ia.T = T;
  ia.N = N;
  ia.min_valid = min_valid;
  ia.sched = fcr::kSchedXcd;
  FC_TRY(fcr::launch_rank_ic(ia, value_bytes, s));
  return 0;
}

// ---- rank family: host-pointer drop-in drivers -----------------------------
extern "C" int fc_cs_rank_f32(const float* h_X, const uint8_t* h_mask, int T, int N,
                              int descending, float* h_out, fc_workspace* wsp) {
  if (h_X == nullptr || h_out == nullptr || T < 1 || N < 1) return -1;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  ScopedWs sw(wsp);
  fc_workspace* ws = sw.ws;
  {
    fcm::MemPlan mp;
    (void)fcm::plan_cs_rank(T, N, h_mask != nullptr, mp);
    FC_TRY(fcm::plan_fits(mp, ws));
  }
  const size_t total = static_cast<size_t>(T) * N;
  void *dx = nullptr, *dm = nullptr, *dout = nullptr;
  FC_TRY(fc_ws_get(ws, fcws::X, total * 4, &dx));
  FC_TRY(fc_ws_get(ws, fcws::OUT0, total * 4, &dout));
  hipStream_t s = nullptr;
  FC_TRY(hipMemcpyAsync(dx, h_X, total * 4, hipMemcpyHostToDevice, s));
  if (h_mask != nullptr) {
    FC_TRY(fc_ws_get(ws, fcws::MASK, total, &dm));
    FC_TRY(hipMemcpyAsync(dm, h_mask, total, hipMemcpyHostToDevice, s));
  }
  float* outs[4] = {nullptr, nullptr, nullptr, nullptr};
  const int g = h_mask ? (descending ? 2 : 0) : (descending ? 3 : 1);
  outs[g] = static_cast<float*>(dout);
  int rc = rank_scan_core(static_cast<const float*>(dx), 1, T, N, 0, N, 1,
                          static_cast<const uint8_t*>(dm), 0, outs[0], outs[1], outs[2], outs[3], s,
                          ws);
  if (rc != 0) return rc;
  FC_TRY(hipMemcpyAsync(h_out, dout, total * 4, hipMemcpyDeviceToHost, s));
  FC_TRY(hipStreamSynchronize(s));
  return 0;
}

static bool whitelisted(int e) {
  return e == static_cast<int>(hipErrorInvalidConfiguration) ||
         e == static_cast<int>(hipErrorLaunchOutOfResources);
}

extern "C" int fc_parameter_scan_f32(const float* h_X, const uint8_t* h_mask, int T, int N,
                                     float* const h_out[4], int group_status[4],
                                     double* h_time_ms, float* h_time_gpu_ms,
                                     const int* h_active, fc_workspace* wsp) {
  if (group_status != nullptr)
    for (int g = 0; g < 4; ++g) group_status[g] = FC_GROUP_NOT_EXECUTED;
  if (h_time_ms)
    for (int g = 0; g < 4; ++g) h_time_ms[g] = 0.0;
  if (h_time_gpu_ms)
    for (int g = 0; g < 4; ++g) h_time_gpu_ms[g] = 0.0f;
  if (h_X == nullptr || h_out == nullptr || group_status == nullptr || T < 1 || N < 1) return -1;
  bool act[4];
  int nact = 0;
  for (int g = 0; g < 4; ++g) {
    act[g] = (h_active == nullptr) || (h_active[g] != 0);
    if (act[g] && h_out[g] == nullptr) return -1;
    nact += act[g] ? 1 : 0;
  }
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  if ((act[0] || act[2]) && h_mask == nullptr) return -1;
  if (nact == 0) return 0;
  ScopedWs sw(wsp);
  fc_workspace* ws = sw.ws;
  {
    fcm::MemPlan mp;
    unsigned groups = 0;
    for (int g = 0; g < 4; ++g) groups |= act[g] ? (1u << g) : 0u;
    (void)fcm::plan_parameter_scan(T, N, groups, mp);
    FC_TRY(fcm::plan_fits(mp, ws));
  }
  const size_t total = static_cast<size_t>(T) * N;
  void *dx = nullptr, *dm = nullptr;
  void* dout[4] = {nullptr, nullptr, nullptr, nullptr};
  FC_TRY(fc_ws_get(ws, fcws::X, total * 4, &dx));
  for (int g = 0; g < 4; ++g)
    if (act[g]) FC_TRY(fc_ws_get(ws, fcws::OUT0 + g, total * 4, &dout[g]));
  hipStream_t s = nullptr;
  FC_TRY(hipMemcpyAsync(dx, h_X, total * 4, hipMemcpyHostToDevice, s));
  if (act[0] || act[2]) {
    FC_TRY(fc_ws_get(ws, fcws::MASK, total, &dm));
    FC_TRY(hipMemcpyAsync(dm, h_mask, total, hipMemcpyHostToDevice, s));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const bool timing = (h_time_ms != nullptr) || (h_time_gpu_ms != nullptr);
  if (timing) {
    FC_TRY(hipEventCreate(&e0));
    FC_TRY(hipEventCreate(&e1));
    FC_TRY(hipStreamSynchronize(s));
  }
  const auto w0 = std::chrono::steady_clock::now();
  if (timing) FC_TRY(hipEventRecord(e0, s));
  int rc = rank_scan_core(static_cast<const float*>(dx), 1, T, N, 0, N, 1,
                          static_cast<const uint8_t*>(dm), 0, static_cast<float*>(dout[0]),
                          static_cast<float*>(dout[1]), static_cast<float*>(dout[2]),
                          static_cast<float*>(dout[3]), s, ws);
  if (rc == 0 && timing) rc = static_cast<int>(hipEventRecord(e1, s));
  if (rc == 0) rc = static_cast<int>(hipStreamSynchronize(s));
  const auto w1 = std::chrono::steady_clock::now();
  float gpu_ms = 0.0f;
  if (rc == 0 && timing) (void)hipEventElapsedTime(&gpu_ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc != 0) {
    // One launch serves every active group: a whitelisted launch failure
    // downgrades all of them, anything else fails the scan.
    if (rc > 0 && whitelisted(rc)) {
      for (int g = 0; g < 4; ++g)
        if (act[g]) group_status[g] = rc;
      (void)hipGetLastError();
      return 0;
    }
    return rc;
  }
  for (int g = 0; g < 4; ++g) {
    if (!act[g]) continue;
    FC_TRY(hipMemcpyAsync(h_out[g], dout[g], total * 4, hipMemcpyDeviceToHost, s));
  }
  FC_TRY(hipStreamSynchronize(s));
  const double wall = std::chrono::duration<double, std::milli>(w1 - w0).count();
  for (int g = 0; g < 4; ++g) {
    if (!act[g]) continue;
    group_status[g] = 0;
    if (h_time_ms) h_time_ms[g] = wall / nact;
    if (h_time_gpu_ms) h_time_gpu_ms[g] = gpu_ms / static_cast<float>(nact);
  }
  return 0;
}

extern "C" int fc_rolling_ic(const void* h_F, const void* h_R, int value_bytes,
                             const uint8_t* h_fmask, const uint8_t* h_rmask, int T, int N,
                             int min_valid, double* h_ic, double* h_rank_f, double* h_rank_r,
                             fc_workspace* wsp) {
  if (h_F == nullptr || h_R == nullptr || h_ic == nullptr || T < 1 || N < 1) return -1;
  if (value_bytes != 4 && value_bytes != 8) return -1;
  if (min_valid < 2) return -4;
  if (N > (1 << 24)) return -2;
  if (static_cast<int64_t>(T) * N > INT32_MAX) return -3;
  ScopedWs sw(wsp);
  fc_workspace* ws = sw.ws;
  {
    fcm::MemPlan mp;
    (void)fcm::plan_rolling_ic(T, N, value_bytes, h_fmask != nullptr, h_rmask != nullptr,
                               h_rank_f != nullptr || h_rank_r != nullptr, mp);
    FC_TRY(fcm::plan_fits(mp, ws));
  }
  const size_t total = static_cast<size_t>(T) * N;
  const size_t vbytes = total * static_cast<size_t>(value_bytes);
  void *df = nullptr, *dr = nullptr, *dfm = nullptr, *drm = nullptr, *dic = nullptr;
  void *drf = nullptr, *drr = nullptr;
  FC_TRY(fc_ws_get(ws, fcws::X, vbytes, &df));
  FC_TRY(fc_ws_get(ws, fcws::R, vbytes, &dr));
  FC_TRY(fc_ws_get(ws, fcws::IC, static_cast<size_t>(T) * 8, &dic));
  hipStream_t s = nullptr;
  FC_TRY(hipMemcpyAsync(df, h_F, vbytes, hipMemcpyHostToDevice, s));
  FC_TRY(hipMemcpyAsync(dr, h_R, vbytes, hipMemcpyHostToDevice, s));
  if (h_fmask) {
    FC_TRY(fc_ws_get(ws, fcws::MASK, total, &dfm));
    FC_TRY(hipMemcpyAsync(dfm, h_fmask, total, hipMemcpyHostToDevice, s));
  }
  if (h_rmask) {
    FC_TRY(fc_ws_get(ws, fcws::RMASK, total, &drm));
    FC_TRY(hipMemcpyAsync(drm, h_rmask, total, hipMemcpyHostToDevice, s));
  }
  const bool ranks = (h_rank_f != nullptr) || (h_rank_r != nullptr);
  if (ranks) {
    FC_TRY(fc_ws_get(ws, fcws::RANKF, total * 8, &drf));
    FC_TRY(fc_ws_get(ws, fcws::RANKR, total * 8, &drr));
    FC_TRY(hipMemsetAsync(drf, 0, total * 8, s));
    FC_TRY(hipMemsetAsync(drr, 0, total * 8, s));
  }
  if (N > rank_max_n_for(value_bytes)) {
    fcw::WideBufs b{};
    int rc = wide_bufs(ws, T, N, value_bytes, true, b);
    if (rc != 0) return rc;
    fcw::WideIcArgs wa{};
    wa.x = df;
    wa.x_rs = N;
    wa.x_es = 1;
    wa.fmask = static_cast<const uint8_t*>(dfm);
    wa.m_rs = N;
    wa.r = dr;
    wa.rmask = static_cast<const uint8_t*>(drm);
    wa.ic = static_cast<double*>(dic);
    wa.rank_f = static_cast<double*>(drf);
    wa.rank_r = static_cast<double*>(drr);
    wa.F = 1;
    wa.T = T;
    wa.N = N;
    wa.min_valid = min_valid;
    FC_TRY(fcw::wide_rolling_ic(wa, value_bytes, b, s));
  } else {
    int rc = fwd_order_core(dr, value_bytes, static_cast<const uint8_t*>(drm), T, N, ws, s);
    if (rc != 0) return rc;
    rc = rank_ic_ordered_core(df, value_bytes, 1, T, N, 0, N, 1, static_cast<const uint8_t*>(dfm),
                              0, min_valid, static_cast<double*>(dic), nullptr,
                              static_cast<double*>(drf), static_cast<double*>(drr), ws, s);
    if (rc != 0) return rc;
  }
  FC_TRY(hipMemcpyAsync(h_ic, dic, static_cast<size_t>(T) * 8, hipMemcpyDeviceToHost, s));
  if (h_rank_f) FC_TRY(hipMemcpyAsync(h_rank_f, drf, total * 8, hipMemcpyDeviceToHost, s));
  if (h_rank_r) FC_TRY(hipMemcpyAsync(h_rank_r, drr, total * 8, hipMemcpyDeviceToHost, s));
  FC_TRY(hipStreamSynchronize(s));
  return 0;
}
