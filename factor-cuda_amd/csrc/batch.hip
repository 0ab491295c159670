// Batched multi-panel schedule (include/fc_hip.h, fc_batch_*): rank +
// rolling Spearman IC + factor_corr of many device-resident panels of
// different T and N, enqueued on a few branch streams and captured once into
// a hipGraph that later calls replay.
//
// Inside a panel all F factors x all dates are one launch (rank_ic_kernel);
// across panels the graph's branches run independent panels concurrently.
// Panels are dealt to branches largest-first (T*N*F), so the branches finish
// close together. Two result-preserving knobs (fc_batch_set_tuning): the
// workgroup -> segment schedule of the rank/IC launch and the number of
// graph branches. (factor_corr's row partition fixes its summation order,
// so its tiling is not a knob: every setting gives bitwise-identical
// results.) The timing autotuner over them is switched off (fc_batch_tune).
#include "fc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <new>
#include <numeric>
#include <vector>

#include "fc_internal.h"

int fc_internal_fwd_order(const void* d_R, int vb, const uint8_t* d_rmask, int T, int N,
                          fc_workspace* ws, hipStream_t s);
int fc_internal_rank_ic_ordered(const void* d_X, int vb, int F, int T, int N, long long ps,
                                long long rs, long long es, const uint8_t* d_fmask,
                                long long fmask_ps, int min_valid, double* d_ic,
                                float* d_rank_out, fc_workspace* ws, hipStream_t s, int sched);

namespace {

constexpr int kMaxBranches = 4;
constexpr int kBatchEnabled = 0;  // see fc_batch_create

#define FCB_TRY(expr)                                     \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return static_cast<int>(_e);    \
  } while (0)
#define FCB_RC(expr)             \
  do {                           \
    const int _r = (expr);       \
    if (_r != 0) return _r;      \
  } while (0)

}  // namespace

struct fc_batch {
  std::vector<fc_panel> panels;
  std::vector<fc_workspace*> ws;   // one per panel: return order + correlation scratch
  std::vector<int> order;          // panels by decreasing T*N*F
  int precision = FC_PREC_CONTRACT;
  int sched = fcr::kSchedXcd;
  int branches = kMaxBranches;
  hipStream_t cap = nullptr;                  // capture stream
  hipStream_t br[kMaxBranches] = {};          // branch streams, default priority
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxBranches] = {};
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;      // recorded after each graph launch
  bool launched = false;
  int device = -1;
};

namespace {

int enqueue_panel(fc_batch* b, int p, hipStream_t s) {
  const fc_panel& P = b->panels[static_cast<size_t>(p)];
  fc_workspace* ws = b->ws[static_cast<size_t>(p)];
  FCB_RC(fc_internal_fwd_order(P.R, 4, P.rmask, P.T, P.N, ws, s));
  // (T,N,F) panel, factor fastest: plane f starts at X + f, row stride N*F,
  // element stride F; one (T,N) mask shared by every factor
  FCB_RC(fc_internal_rank_ic_ordered(P.X, 4, P.F, P.T, P.N, 1, static_cast<long long>(P.N) * P.F,
                                     P.F, P.mask, 0, P.min_valid, P.ic, P.rank_out, ws, s,
                                     b->sched));
  if (P.corr != nullptr)
    FCB_RC(fc_factor_corr_dev(P.X, P.mask, P.T, P.N, P.F, b->precision, P.corr, ws, s));
  return 0;
}

// Every panel, forked over the branch streams and joined back into s. The
// same code runs eagerly and under stream capture.
int enqueue_all(fc_batch* b, hipStream_t s) {
  const int B = b->branches;
  if (B <= 1) {
    for (int p : b->order) FCB_RC(enqueue_panel(b, p, s));
    return 0;
  }
  FCB_TRY(hipEventRecord(b->fork, s));
  for (int k = 0; k < B; ++k) FCB_TRY(hipStreamWaitEvent(b->br[k], b->fork, 0));
  for (size_t i = 0; i < b->order.size(); ++i)
    FCB_RC(enqueue_panel(b, b->order[i], b->br[i % static_cast<size_t>(B)]));
  for (int k = 0; k < B; ++k) {
    FCB_TRY(hipEventRecord(b->join[k], b->br[k]));
    FCB_TRY(hipStreamWaitEvent(s, b->join[k], 0));
  }
  return 0;
}

// The exec is destroyed only after its last launch (on whatever stream the
// caller gave fc_batch_run) has completed.
void drop_graph(fc_batch* b) {
  if (b->exec != nullptr) {
    if (b->launched && b->done != nullptr) (void)hipEventSynchronize(b->done);
    (void)hipGraphExecDestroy(b->exec);
  }
  b->exec = nullptr;
  b->launched = false;
}

int build_graph(fc_batch* b) {
  drop_graph(b);
  FCB_TRY(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_all(b, b->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->cap, &g);
  if (rc != 0) {
    if (g != nullptr) (void)hipGraphDestroy(g);
    return rc;
  }
  FCB_TRY(ec);
  const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    b->exec = nullptr;
    return static_cast<int>(ei);
  }
  return 0;
}

void release(fc_batch* b) {
  drop_graph(b);
  for (fc_workspace* w : b->ws) fc_workspace_destroy(w);
  b->ws.clear();
  for (int k = 0; k < kMaxBranches; ++k) {
    if (b->br[k] != nullptr) (void)hipStreamDestroy(b->br[k]);
    if (b->join[k] != nullptr) (void)hipEventDestroy(b->join[k]);
  }
  if (b->fork != nullptr) (void)hipEventDestroy(b->fork);
  if (b->done != nullptr) (void)hipEventDestroy(b->done);
  if (b->cap != nullptr) (void)hipStreamDestroy(b->cap);
}

int validate(const fc_panel& P) {
  if (P.X == nullptr || P.R == nullptr || P.ic == nullptr) return -1;
  if (P.T < 1 || P.N < 1 || P.F < 1) return -1;
  if (P.min_valid < 2) return -4;
  if (P.N > (1 << 24)) return -2;
  if (static_cast<int64_t>(P.T) * P.N > INT32_MAX) return -3;
  if (static_cast<int64_t>(P.T) * P.N * P.F > INT32_MAX) return -3;
  if (P.corr != nullptr && P.F > 128) return -3;
  return 0;
}

}  // namespace

extern "C" int fc_batch_create(const fc_panel* panels, int n_panels, int corr_precision,
                               fc_batch** out) {
  if (out == nullptr || panels == nullptr || n_panels < 1) return -1;
  *out = nullptr;
  if (corr_precision != FC_PREC_CONTRACT && corr_precision != FC_PREC_FP32 &&
      corr_precision != FC_PREC_BF16X3)
    return -1;
  for (int p = 0; p < n_panels; ++p) FCB_RC(validate(panels[p]));
  // Switched off: runs of this schedule ended two GPU processes with host
  // crashes (a segmentation fault in the timing autotuner at 7 panels, a
  // glibc heap-consistency abort at 64 panels). An AddressSanitizer replay of
  // the same call sequence through this library's host code and bindings,
  // on a recording HIP stand-in, found no memory error, so no cause is
  // known; valid arguments are refused until one is.
  if (kBatchEnabled == 0) return static_cast<int>(hipErrorNotSupported);
  fc_batch* b = new (std::nothrow) fc_batch();
  if (b == nullptr) return static_cast<int>(hipErrorOutOfMemory);
  b->panels.assign(panels, panels + n_panels);
  b->precision = corr_precision;
  b->order.resize(static_cast<size_t>(n_panels));
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
    const auto w = [&](int p) {
      const fc_panel& P = b->panels[static_cast<size_t>(p)];
      return static_cast<long long>(P.T) * P.N * P.F;
    };
    return w(x) > w(y);
  });
  int rc = 0;
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
  for (int k = 0; k < kMaxBranches && e == hipSuccess; ++k) {
    e = hipStreamCreateWithFlags(&b->br[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  for (int p = 0; p < n_panels && e == hipSuccess && rc == 0; ++p) {
    fc_workspace* w = nullptr;
    rc = fc_workspace_create(&w);
    if (rc == 0) b->ws.push_back(w);
  }
  // one eager pass allocates every workspace buffer, so capture never allocates
  if (e == hipSuccess && rc == 0) rc = enqueue_all(b, b->cap);
  if (e == hipSuccess && rc == 0) e = hipStreamSynchronize(b->cap);
  if (e != hipSuccess || rc != 0) {
    release(b);
    delete b;
    return rc != 0 ? rc : static_cast<int>(e);
  }
  *out = b;
  return 0;
}

extern "C" int fc_batch_run(fc_batch* b, void* stream) {
  if (b == nullptr) return -1;
  if (b->exec == nullptr) FCB_RC(build_graph(b));
  const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  FCB_TRY(hipGraphLaunch(b->exec, s));
  b->launched = true;
  FCB_TRY(hipEventRecord(b->done, s));
  return 0;
}

extern "C" int fc_batch_run_panel(fc_batch* b, int panel, void* stream) {
  if (b == nullptr || panel < 0 || panel >= static_cast<int>(b->panels.size())) return -1;
  return enqueue_panel(b, panel, reinterpret_cast<hipStream_t>(stream));
}

// The autotuner (time every schedule x branch-count candidate as a captured
// graph, keep the fastest) is switched off: a run of it at 7 panels ended
// the process with a host segmentation fault, and a 64-panel run (create,
// then tune) with a glibc heap-consistency abort, with no GPU fault and no
// cause found in this file's host code. Until one is found it returns
// hipErrorNotSupported; fc_batch_set_tuning still selects the knobs.
extern "C" int fc_batch_tune(fc_batch* b, int reps, void* stream) {
  (void)stream;
  if (b == nullptr || reps < 1) return -1;
  return static_cast<int>(hipErrorNotSupported);
}

extern "C" int fc_batch_get_tuning(const fc_batch* b, int* sched, int* branches) {
  if (b == nullptr) return -1;
  if (sched) *sched = b->sched;
  if (branches) *branches = b->branches;
  return 0;
}

extern "C" int fc_batch_set_tuning(fc_batch* b, int sched, int branches) {
  if (b == nullptr) return -1;
  if (sched != fcr::kSchedXcd && sched != fcr::kSchedFMaj && sched != fcr::kSchedTMaj) return -1;
  if (branches < 1 || branches > kMaxBranches) return -1;
  b->sched = sched;
  b->branches = branches;
  drop_graph(b);
  return 0;
}

extern "C" void fc_batch_destroy(fc_batch* b) {
  if (b == nullptr) return;
  release(b);
  delete b;
}
