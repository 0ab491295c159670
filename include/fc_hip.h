/* fc_hip.h -- C ABI of the MI355X-native cross-sectional analysis library
 * (libfc_hip.so, gfx950). Plain pointers and sizes only; every entry point
 * returns an int status:
 *     0            success
 *     negative     contract error (the Python layer raises ValueError):
 *                    -1 null pointer or a dimension < 1
 *                    -2 N > 2^24 (rank ops) / T*N > INT32_MAX (correlations)
 *                    -3 T*N > INT32_MAX (rank ops) / N*N > INT32_MAX
 *                       (stock_corr) / F > 128 (factor_corr)
 *                    -4 min_valid < 2 (rolling_ic) / correlation value
 *                       domain violated (max|x| > 1e150 or min nonzero
 *                       |x| < 1e-150 over valid cells, stock_corr f64)
 *     positive     HIP runtime error code (the Python layer raises
 *                  RuntimeError); hipErrorInvalidConfiguration (9) and
 *                  hipErrorLaunchOutOfResources (701) keep their CUDA
 *                  numbers, so the parameter_scan group whitelist holds.
 * The status convention is the reference's (src/factor_cuda_pybind.cpp:117-123,
 * src/cross_sectional_rank.cu:233-235, src/rolling_ic.cu:333-335,
 * src/stock_corr.cu:440-442, src/factor_corr.cu:396-398).
 *
 * Host-pointer entry points are host-blocking: results are in the output
 * buffers on return. Device-pointer entry points (suffix _dev) take inputs
 * already resident in HBM, enqueue on the given stream (a hipStream_t passed
 * as void*, NULL = default stream) and return without synchronising.
 */
#ifndef FC_HIP_H_
#define FC_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC_GROUP_NOT_EXECUTED (-100)

/* Correlation precision tiers. */
#define FC_PREC_CONTRACT 0 /* FP64 sums + cancellation fallback (|dr| <= 1e-12 contract) */
#define FC_PREC_FP32 1     /* plain FP32 MFMA contraction, FP64 finalize */
#define FC_PREC_BF16X3 2   /* bf16 hi/lo split MFMA + INT8 MFMA counts, FP64 finalize */

/* stock_corr dispatch diagnostics; replaces StockCorrRunStats
 * (src/stock_corr.cuh:47-50). selected_path: 0 general (masked) path,
 * 1 fully-valid path; fallback_count: cells recomputed by the exact
 * two-pass fallback. */
typedef struct fc_stock_corr_stats {
  int selected_path;
  int fallback_count;
} fc_stock_corr_stats;

/* Allocation accounting of every device buffer the library allocates;
 * replaces factor_cuda::MemTracker (src/mem_tracker.h:35-77). */
typedef struct fc_mem_stats {
  int64_t live_bytes;
  int64_t peak_bytes;
  int64_t n_alloc;
  int64_t n_free;
} fc_mem_stats;

/* Opaque cache of device buffers reused across calls of the same shape and
 * device; replaces cs_rank_workspace / rolling_ic_workspace
 * (src/cross_sectional_rank.cuh:38-64, src/rolling_ic.cuh:33-60). Not
 * thread-safe: calls sharing one workspace must be serialised. */
typedef struct fc_workspace fc_workspace;

int fc_workspace_create(fc_workspace** out);
/* Frees every cached buffer; idempotent; safe on NULL. Replaces
 * cs_rank_workspace_clear / rolling_ic_workspace_clear. */
void fc_workspace_clear(fc_workspace* ws);
void fc_workspace_destroy(fc_workspace* ws);

/* ---- host-pointer drop-in drivers ------------------------------------ */

/* Stable ordinal rank of each row of a (T,N) float32 panel; valid =
 * finite && (mask == NULL || mask != 0); invalid cells -> quiet NaN
 * 0x7fc00000. Replaces cs_rank_gpu (src/cross_sectional_rank.cuh:90-99). */
int fc_cs_rank_f32(const float* h_X, const uint8_t* h_mask, int T, int N, int descending,
                   float* h_out, fc_workspace* ws);

/* The four rank variants of one panel from one launch. Group order:
 * 0 ascending-masked, 1 ascending-unmasked, 2 descending-masked,
 * 3 descending-unmasked. h_active (NULL = all) selects groups; skipped groups
 * report FC_GROUP_NOT_EXECUTED. h_mask may be NULL only when no masked group
 * is active. h_time_ms / h_time_gpu_ms (optional) receive each active
 * group's equal share of the one launch's wall / device time. Replaces
 * parameter_scan_gpu (src/parameter_scan.cuh:183-189). */
int fc_parameter_scan_f32(const float* h_X, const uint8_t* h_mask, int T, int N,
                          float* const h_out[4], int group_status[4], double* h_time_ms,
                          float* h_time_gpu_ms, const int* h_active, fc_workspace* ws);

/* Per-date Spearman IC of F vs R over the pairwise-valid cells (finite on
 * both sides and both masks set), stable ordinal ranks; NaN when fewer than
 * min_valid cells or either side is constant there. value_bytes 4 (float32
 * inputs) or 8 (float64). Optional h_rank_f / h_rank_r (T,N) receive the
 * ordinal ranks, 0 outside the pairwise-valid set. Replaces rolling_ic_gpu
 * (src/rolling_ic.cuh:102-108). */
int fc_rolling_ic(const void* h_F, const void* h_R, int value_bytes, const uint8_t* h_fmask,
                  const uint8_t* h_rmask, int T, int N, int min_valid, double* h_ic,
                  double* h_rank_f, double* h_rank_r, fc_workspace* ws);

/* (N,N) pairwise-complete Pearson matrix over the T rows of a (T,N) panel.
 * precision: FC_PREC_*. Replaces stock_corr_gpu (src/stock_corr.cuh:66-68). */
int fc_stock_corr(const void* h_X, int value_bytes, const uint8_t* h_mask, int T, int N,
                  int precision, double* h_out, fc_stock_corr_stats* stats);

/* (F,F) pooled pairwise-complete Pearson matrix of a (T,N,F) panel over its
 * T*N rows; mask is (T,N). Replaces factor_corr_gpu
 * (src/factor_corr.cuh:46-49). */
int fc_factor_corr(const void* h_F3, int value_bytes, const uint8_t* h_mask, int T, int N,
                   int F, int precision, double* h_out);

/* ---- device-pointer entry points (inputs resident in HBM) ------------- */

/* Batched rank of F planes in one launch. Element (f,t,i) of X at
 * d_X[f*plane_stride + t*row_stride + i*elem_stride] (elements); the mask
 * (f,t,i) at d_mask[f*mask_plane_stride + t*N + i] (mask_plane_stride 0 =
 * one shared (T,N) mask). d_out[g] (NULL = skip) is dense [F][T][N]. Rows
 * longer than fc_rank_max_n(4) are sorted with buffers this call allocates
 * and frees, so such a call waits for its stream before returning and cannot
 * be captured into a graph. */
int fc_rank_scan_dev(const float* d_X, int F, int T, int N, long long plane_stride,
                     long long row_stride, long long elem_stride, const uint8_t* d_mask,
                     long long mask_plane_stride, float* d_out0, float* d_out1, float* d_out2,
                     float* d_out3, void* stream);

/* Per-date order of one forward-return panel (T,N): the ordinal position and
 * tie group of every valid return (finite && (d_rmask == NULL || mask != 0)).
 * Kept in ws (4*T*N + 4*T bytes, allocated on first use) for any number of
 * later fc_rank_ic_ordered_dev calls with the same T and N, so one return
 * panel is sorted once however many factor batches are scored against it. */
int fc_fwd_order_dev(const void* d_R, int value_bytes, const uint8_t* d_rmask, int T, int N,
                     fc_workspace* ws, void* stream);

/* Fused rank + IC of F factor planes against the return order held in ws:
 * d_ic [F][T] Spearman IC on each factor's pairwise-valid set (finite factor,
 * factor mask, valid return); d_rank_out (optional, dense [F][T][N]) the
 * ascending masked cs_rank of each factor plane. Factor element (f,t,i) at
 * d_X[f*plane_stride + t*row_stride + i*elem_stride]; factor mask at
 * d_fmask[f*fmask_plane_stride + t*N + i] (NULL = all set). Returns -1 if ws
 * holds no return order of this T and N. */
int fc_rank_ic_ordered_dev(const void* d_X, int value_bytes, int F, int T, int N,
                           long long plane_stride, long long row_stride, long long elem_stride,
                           const uint8_t* d_fmask, long long fmask_plane_stride, int min_valid,
                           double* d_ic, float* d_rank_out, fc_workspace* ws, void* stream);

/* fc_fwd_order_dev followed by fc_rank_ic_ordered_dev (R and X share
 * value_bytes). Rows longer than fc_rank_max_n(value_bytes) take the
 * long-row sort, whose buffers the first call of a shape allocates in ws
 * (that call is not capture-safe; later calls of the shape allocate
 * nothing). */
int fc_rank_ic_dev(const void* d_X, int value_bytes, int F, int T, int N, long long plane_stride,
                   long long row_stride, long long elem_stride, const uint8_t* d_fmask,
                   long long fmask_plane_stride, const void* d_R, const uint8_t* d_rmask,
                   int min_valid, double* d_ic, float* d_rank_out, fc_workspace* ws,
                   void* stream);

/* fc_rank_ic_dev with no workspace and no allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The return order lives in a static device array of
 * fc_rank_ic_noalloc_scratch_bytes() (32 MiB) that each GPU holds until the
 * process ends and that every call on that GPU shares: calls that overlap on
 * one GPU race. Arguments and results as fc_rank_ic_dev; returns -3 (size cap)
 * when 4*T*N + 4*T bytes (rounded up to 256) exceed the array or
 * N > fc_rank_max_n(value_bytes). Calls no hipMalloc/hipFree. */
int fc_rank_ic_dev_noalloc(const void* d_X, int value_bytes, int F, int T, int N,
                           long long plane_stride, long long row_stride, long long elem_stride,
                           const uint8_t* d_fmask, long long fmask_plane_stride, const void* d_R,
                           const uint8_t* d_rmask, int min_valid, double* d_ic,
                           float* d_rank_out, void* stream);
/* Size in bytes of the static array behind fc_rank_ic_dev_noalloc. */
long long fc_rank_ic_noalloc_scratch_bytes(void);

/* Device-resident stock_corr: d_X (T,N) float32, d_mask (T,N) or NULL,
 * d_out (N,N) float64. ws provides the scratch planes. */
int fc_stock_corr_dev(const float* d_X, const uint8_t* d_mask, int T, int N, int precision,
                      double* d_out, fc_workspace* ws, void* stream);

/* Device-resident factor_corr: d_F3 (T,N,F) float32, d_mask (T,N) or NULL,
 * d_out (F,F) float64. */
int fc_factor_corr_dev(const float* d_F3, const uint8_t* d_mask, int T, int N, int F,
                       int precision, double* d_out, fc_workspace* ws, void* stream);

/* Device-resident stock_corr with the checks of fc_stock_corr: d_X (T,N) is
 * float32 (value_bytes 4) or float64 (8). It synchronises `stream` once to
 * read back the N column statistics: a float64 panel outside the value
 * domain returns -4, and a fully valid panel takes the bf16x3 full-row path,
 * so results are bitwise those of fc_stock_corr on the same panel. stats
 * (optional) as in fc_stock_corr. Replaces stock_corr_gpu
 * (src/stock_corr.cuh:66-68) for inputs already in HBM. */
int fc_stock_corr_dev_ex(const void* d_X, int value_bytes, const uint8_t* d_mask, int T, int N,
                         int precision, double* d_out, fc_stock_corr_stats* stats,
                         fc_workspace* ws, void* stream);

/* fc_factor_corr_dev for a float32 (value_bytes 4) or float64 (8) panel;
 * bitwise the result of fc_factor_corr on the same panel. Never
 * synchronises. Replaces factor_corr_gpu (src/factor_corr.cuh:46-49) for
 * inputs already in HBM. */
int fc_factor_corr_dev_ex(const void* d_F3, int value_bytes, const uint8_t* d_mask, int T, int N,
                          int F, int precision, double* d_out, fc_workspace* ws, void* stream);

/* ---- batched multi-panel schedule -------------------------------------
 * No counterpart in the reference, whose bindings take one panel per call
 * (src/factor_cuda_pybind.cpp:507-535, src/factor_corr_pybind.cpp:105-107);
 * it runs the per-panel device paths above for many panels at once. */

/* One panel of a batch; every pointer is device memory owned by the caller
 * and must stay valid (and unmodified while a run is in flight) for the
 * life of the batch. */
typedef struct fc_panel {
  const float* X;        /* (T,N,F) float32, factor fastest */
  const uint8_t* mask;   /* (T,N) factor mask, or NULL */
  const float* R;        /* (T,N) float32 forward returns */
  const uint8_t* rmask;  /* (T,N) return mask, or NULL */
  int T, N, F;
  int min_valid;         /* rolling_ic min_valid (>= 2) */
  float* rank_out;       /* [F][T][N] ascending masked cs_rank, or NULL */
  double* ic;            /* [F][T] Spearman IC */
  double* corr;          /* [F][F] factor_corr, or NULL to skip */
} fc_panel;

typedef struct fc_batch fc_batch;

/* Builds a batch over n_panels panels: validates them (status codes as the
 * single-panel entry points), allocates every per-panel workspace with one
 * eager run (so later runs allocate nothing), and orders panels largest
 * first. corr_precision: FC_PREC_* for the factor_corr part. Switched off:
 * after validation it returns hipErrorNotSupported (801) and *out stays NULL
 * (runs of the schedule crashed the host process; no cause is known). */
int fc_batch_create(const fc_panel* panels, int n_panels, int corr_precision, fc_batch** out);
/* Runs every panel: the first call (and the first after tuning) captures the
 * schedule into a hipGraph on an internal stream; each call launches that
 * graph on `stream` (NULL = default stream) and returns without waiting.
 * fc_batch_set_tuning and fc_batch_destroy wait for the last launch before
 * they destroy the graph. */
int fc_batch_run(fc_batch* b, void* stream);
/* Enqueues one panel's work on `stream` outside the graph (for per-panel
 * latency). */
int fc_batch_run_panel(fc_batch* b, int panel, void* stream);
/* Autotuner over the (segment schedule, branch count) knobs: switched off.
 * Returns -1 on invalid arguments, otherwise hipErrorNotSupported (801)
 * and leaves the knobs unchanged (runs of it crashed the host process with
 * no cause found); choose the knobs with fc_batch_set_tuning. */
int fc_batch_tune(fc_batch* b, int reps, void* stream);
/* Current knobs: sched 0 = dates grouped per XCD, 1 = factor-major, 2 =
 * date-major; branches = concurrent graph branches (1..4). */
int fc_batch_get_tuning(const fc_batch* b, int* sched, int* branches);
int fc_batch_set_tuning(fc_batch* b, int sched, int branches);
void fc_batch_destroy(fc_batch* b);

/* ---- accounting / info ------------------------------------------------ */
void fc_mem_get_stats(fc_mem_stats* out);
void fc_mem_reset_peak(void);

/* Device-memory plan of a host-pointer driver. The exact bytes (256-byte
 * granules, as fc_mem_stats counts them) that one call allocates in a new
 * workspace: the high-water mark over the call's own buffers. Replaces the
 * reference's per-op theoretical formulas
 * (benchmarks/results/poc3_calibration_v1.md:8-13); every host driver checks
 * the bytes it would add against the device's free HBM before allocating
 * and returns 2 (hipErrorOutOfMemory) without allocating when they do not
 * fit, in place of the reference's fixed 7,676 MiB budget
 * (benchmarks/memory_budget_v1.py:47). Negative = contract code for the
 * arguments; -1 also when the plan cannot be sized (rows beyond
 * fc_rank_max_n need rocPRIM's sort-storage size, a query that fails
 * without a device). Otherwise pure host arithmetic: needs no device. */
#define FC_OP_CS_RANK 0        /* fc_cs_rank_f32 */
#define FC_OP_PARAMETER_SCAN 1 /* fc_parameter_scan_f32 */
#define FC_OP_ROLLING_IC 2     /* fc_rolling_ic */
#define FC_OP_STOCK_CORR 3     /* fc_stock_corr */
#define FC_OP_FACTOR_CORR 4    /* fc_factor_corr */
#define FC_PLAN_MASK 1         /* mask (rolling_ic: factor mask) passed */
#define FC_PLAN_RMASK 2        /* rolling_ic: return mask passed */
#define FC_PLAN_RANKS 4        /* rolling_ic: rank planes requested */
#define FC_PLAN_GROUPS(g) ((g) << 8) /* parameter_scan: active groups, bit per group; 0 = all */
long long fc_op_device_bytes(int op, int T, int N, int F, int value_bytes, int precision,
                             int flags);

/* Free and total device memory of the current device (hipMemGetInfo). */
int fc_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Largest N the one-workgroup-per-row ranking kernels take for 4- or 8-byte
 * values on the current device (0 on error). Longer rows, up to the 2^24
 * cap, go through a stable rocPRIM radix sort (device-wide over (row, key)
 * for float rows, segmented for double rows) in every rank
 * entry point except fc_fwd_order_dev / fc_rank_ic_ordered_dev, which
 * return hipErrorLaunchOutOfResources for them. */
int fc_rank_max_n(int value_bytes);
const char* fc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FC_HIP_H_ */
