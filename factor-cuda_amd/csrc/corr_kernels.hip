// Pairwise-complete Pearson correlations on gfx950 (stock_corr, factor_corr).
//
// Both ops reduce to the six masked sums of every column pair over the rows
// where both columns are valid (finite && mask):
//   n = sum m_i m_j, Sx = sum x_i m_j, Sy = sum m_i x_j, Sxy = sum x_i x_j,
//   Sxx = sum x_i^2 m_j, Syy = sum m_i x_j^2
// i.e. matrix contractions over the row axis, run on MFMA, followed by an
// FP64 finalize r = (Sxy - Sx Sy/n) / sqrt(Sxx - Sx^2/n) / sqrt(Syy - Sy^2/n)
// in the reference's order ((sxy/sqrt(sxx))/sqrt(syy),
// src/stock_corr_impl.cuh:73-78). Before the contraction each column is
// shifted by its own valid-cell mean and scaled (Pearson is invariant to a
// per-column positive affine map), which removes the one-pass cancellation:
//   contract tier  FP64 MFMA (16x16x4), scale = a power of two (exact);
//   fp32 tier      FP32 MFMA (32x32x2) on standardised columns;
//   bf16x3 tier    standardised z split into bf16 hi + lo, products
//                  hi*hi + hi*lo + lo*hi (and hi*m + lo*m for the masked sums)
//                  on bf16 MFMA (32x32x16, FP32 accumulate), pair counts on
//                  INT8 MFMA (32x32x32, exact INT32).
// Cells whose sums still cancel (the reference's ratio-3 test,
// src/stock_corr_impl.cuh:102-132, on the raw sums in the contract tier),
// whose pair-restricted variance is near zero, or whose |r| > 1 are
// recomputed exactly by a serial two-pass with Kahan-compensated means
// (the HG-2 reference algorithm, tests/test_adapter_v1.py:96-124).
#include "corr_internal.h"
#include "fc_device.h"

#include <climits>
#include <type_traits>

namespace fcc {

using fcd::lane_id;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr double kNaN = __builtin_nan("");
constexpr double kCancelRatio = 3.0;  // src/stock_corr_impl.cuh ratio-3 gate
constexpr double kMaxAmp = 16.0;     // fp32 / bf16x3 error-amplification gate
// bf16x3 drops the lo*lo term of every product (relative error ~2^-17 each).
// Over many rows those errors average out, but over a handful they do not,
// and the rows of a small pairwise set sit far from the column shift: a
// bf16x3 cell over fewer rows than this goes to the exact pass.
constexpr double kBf16x3MinN = 16.0;

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even (finite f)
  const uint32_t u = __float_as_uint(f);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

template <typename V>
__device__ __forceinline__ bool finite_v(V v) {
  return __builtin_isfinite(v);
}

size_t stock_plane_bytes(int tier, int N, int T, int value_bytes) {
  const size_t Np = static_cast<size_t>(stock_rows_padded(N));
  const size_t Tp = static_cast<size_t>(stock_t_padded(T));
  // original values for the exact pass, in the input's precision
  const size_t xt = Np * Tp * static_cast<size_t>(value_bytes == 8 ? 8 : 4);
  if (tier == kTierContract) return 2 * Np * Tp * 8 + xt;
  if (tier == kTierFp32) return 2 * Np * Tp * 4 + xt;
  return 5 * Np * Tp * 2 + Np * Tp + xt;
}
int stock_tile(int tier) { return tier == kTierContract ? 32 : 64; }

// ---- finalize ----------------------------------------------------------------
struct Fin {
  double r;
  bool fb;
};

__device__ __forceinline__ double safe_pearson(double cov, double vx, double vy) {
  if (!(vx > 0.0 && vy > 0.0) || !__builtin_isfinite(vx) || !__builtin_isfinite(vy) ||
      !__builtin_isfinite(cov))
    return kNaN;
  const double r = (cov / sqrt(vx)) / sqrt(vy);
  return __builtin_isfinite(r) ? r : kNaN;
}

// Sums are of the shifted/scaled columns z = (x - mean) * scale.
// Contract-tier exact-pass triggers mirror the reference op, because on
// biased data the reference's answer is its compensated two-pass, not the
// exact Pearson of the data:
//   MIRROR = true  (stock_corr, src/stock_corr_impl.cuh:102-132): ratio-3
//                  cancellation on the shifted sums and on the reconstructed
//                  raw one-pass sums;
//   MIRROR = false (factor_corr, src/factor_corr.cu:147-173): bias =
//                  max(max|x| / sd_x, max|y| / sd_y) > 1e8, with max|x| the
//                  column's (ColStat.amax, an upper bound of the pair's), plus
//                  an FP64 amplification guard (kContractAmp).
constexpr double kContractAmp = 4.0;
constexpr double kBiasTrigger = 1e8;
template <int TIER, bool MIRROR = true>
__device__ __forceinline__ Fin finalize_cell(double n, double sx, double sy, double sxy, double sxx,
                                             double syy, const ColStat& ci, const ColStat& cj) {
  Fin o{kNaN, false};
  if (n < 2.0) return o;
  if (ci.minv == ci.maxv || cj.minv == cj.maxv) return o;  // constant column: constant on any subset
  const double cov = sxy - sx * sy / n;
  const double vx = sxx - sx * sx / n;
  const double vy = syy - sy * sy / n;
  bool fix;
  if (TIER == kTierContract && MIRROR) {
    fix = (fabs(sx * sy / n) > kCancelRatio * fabs(cov)) ||
          (fabs(sx * sx / n) > kCancelRatio * fabs(vx)) ||
          (fabs(sy * sy / n) > kCancelRatio * fabs(vy));
  } else {
    // Standardised sums carry an absolute error ~eps * magnitude; r's error is
    // that times the amplification below. Recompute exactly when it exceeds
    // kMaxAmp (absolute r error stays well inside 1e-5).
    const double sv = sqrt(fabs(vx) * fabs(vy));
    const double amp = fmax((fabs(sxy) + fabs(sx * sy) / n) / sv,
                            fmax((fabs(sxx) + sx * sx / n) / fabs(vx), (fabs(syy) + sy * sy / n) / fabs(vy)));
    fix = !(amp <= (TIER == kTierContract ? kContractAmp : kMaxAmp));
  }
  if (TIER == kTierBf16x3 && n < kBf16x3MinN) fix = true;
  const double tol = (TIER == kTierContract) ? 1e-10 : (TIER == kTierFp32 ? 1e-5 : 1e-3);
  const bool cx = fabs(vx) <= tol * (fmax(fabs(sxx), sx * sx / n) + 1e-300);
  const bool cy = fabs(vy) <= tol * (fmax(fabs(syy), sy * sy / n) + 1e-300);
  const double r = safe_pearson(cov, vx, vy);
  if (!__builtin_isfinite(r) || fabs(r) > 1.0) fix = true;
  if (TIER == kTierContract && !MIRROR) {
    // max|x| / sd with sd from the scaled centred sum: sd = sqrt(vx / n) / scale
    const double bx = ci.amax * ci.scale * sqrt(n / fabs(vx));
    const double by = cj.amax * cj.scale * sqrt(n / fabs(vy));
    fix = fix || !(fmax(bx, by) <= kBiasTrigger);
  }
  if (TIER == kTierContract && MIRROR) {
    // the reference's gate on the raw (unshifted) one-pass sums
    const double si = ci.scale, sj = cj.scale;
    const double rsx = sx / si + n * ci.mean;
    const double rsy = sy / sj + n * cj.mean;
    const double rcov = cov / (si * sj), rvx = vx / (si * si), rvy = vy / (sj * sj);
    fix = fix || (fabs(rsx * rsy / n) > kCancelRatio * fabs(rcov)) ||
          (fabs(rsx * rsx / n) > kCancelRatio * fabs(rvx)) ||
          (fabs(rsy * rsy / n) > kCancelRatio * fabs(rvy));
  }
  if (fix || cx || cy) {
    o.fb = true;
    return o;
  }
  o.r = r;
  return o;
}

// Fast tiers (fp32, bf16x3): the sums arrive in fp32, so the gates are
// evaluated in fp32 and only r itself in fp64. Everything is scaled by n to
// avoid divisions: c = n*sxy - sx*sy, vx = n*sxx - sx^2, vy = n*syy - sy^2,
// r = c / sqrt(vx * vy). A cell goes to the exact pass when the fp32 rounding
// error of its sums is amplified more than kMaxAmp (cancellation in c or a
// variance), when a pair-restricted variance is near zero relative to its
// terms, or when r is not finite or |r| > 1. `konst`: either column is
// constant (NaN directly). Branch-free: every lane evaluates every term.
template <int TIER>
__device__ __forceinline__ Fin finalize_fast(float n, float sx, float sy, float sxy, float sxx,
                                             float syy, bool konst) {
  const bool ok = (n >= 2.0f) & !konst;
  const float pxy = fabsf(sx * sy), pxx = sx * sx, pyy = sy * sy;
  const float nxy = n * fabsf(sxy), nxx = n * fabsf(sxx), nyy = n * fabsf(syy);
  const float ax = fabsf(n * sxx - pxx), ay = fabsf(n * syy - pyy);
  constexpr float tol = (TIER == kTierFp32) ? 1e-5f : 1e-3f;
  constexpr float K = static_cast<float>(kMaxAmp);
  const float a1 = nxy + pxy;
  // bitwise | keeps the gate branch-free
  bool fix = !(a1 * a1 <= (K * K) * (ax * ay)) | !(nxx + pxx <= K * ax) | !(nyy + pyy <= K * ay) |
             (ax <= tol * (fmaxf(nxx, pxx) + 1e-30f)) | (ay <= tol * (fmaxf(nyy, pyy) + 1e-30f));
  const double dn = n, dsx = sx, dsy = sy;
  const double c = dn * static_cast<double>(sxy) - dsx * dsy;
  const double vx = dn * static_cast<double>(sxx) - dsx * dsx;
  const double vy = dn * static_cast<double>(syy) - dsy * dsy;
  // 1/sqrt(vx*vy): fp32 hardware estimate (1 ulp) refined by one Newton step
  // in fp64 (relative error ~1e-14); p <= 0 or p out of fp32 range gives a
  // non-finite r and the cell goes to the exact pass
  const double pv = vx * vy;
  const double y0 = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(pv)));
  const double y1 = y0 * (1.5 - 0.5 * pv * y0 * y0);
  const double r = c * y1;
  if (TIER == kTierBf16x3) fix = fix | (n < static_cast<float>(kBf16x3MinN));
  fix = fix | !__builtin_isfinite(r) | (fabs(r) > 1.0) | !(pv > 0.0) |
        !(pv < 1e36);
  Fin o;
  o.fb = ok && fix;
  o.r = (ok && !fix) ? r : kNaN;
  return o;
}

// ---- stock_corr: column statistics + operand planes ------------------------
// (1) partial statistics per (64-column group, 128-row chunk): wave w of the
// block takes 32 consecutive rows, lane = column; each lane's count, mean and
// centred sum of squares over its rows, combined across waves and chunks with
// Chan's pairwise update in a fixed order (deterministic).
struct ColPart {
  double n, mean, m2, minv, maxv, amax, aminnz, pad;
};

__device__ __forceinline__ void chan_combine(double& n, double& mean, double& m2, double nb,
                                             double meanb, double m2b) {
  if (nb <= 0.0) return;
  if (n <= 0.0) {
    n = nb;
    mean = meanb;
    m2 = m2b;
    return;
  }
  const double tot = n + nb;
  const double d = meanb - mean;
  mean = mean + d * (nb / tot);
  m2 = m2 + m2b + d * d * (n * nb / tot);
  n = tot;
}

constexpr int kColRows = 128;  // rows per partial-statistics chunk

template <typename V>
__global__ __launch_bounds__(256) void stock_colpart_kernel(const V* __restrict__ X,
                                                           const uint8_t* __restrict__ mask, int T,
                                                           int N, ColPart* __restrict__ parts) {
  __shared__ ColPart wp[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = blockIdx.x * 64 + lane;
  const int tc = blockIdx.y;
  const int t0 = tc * kColRows + w * 32;
  double v[32];
  bool ok[32];
  double n = 0.0, sum = 0.0, mn = __builtin_inf(), mx = -__builtin_inf(), am = 0.0, az = 1e308;
  // every load issued before any is consumed: cells outside the panel read
  // element 0 (always in bounds) and are masked off afterwards, so no load
  // waits behind a branch of its own
  V xr[32];
  uint8_t mr[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const int t = t0 + k;
    const size_t o = (col < N && t < T) ? static_cast<size_t>(t) * N + col : 0;
    xr[k] = X[o];
  }
  if (mask != nullptr) {
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const int t = t0 + k;
      const size_t o = (col < N && t < T) ? static_cast<size_t>(t) * N + col : 0;
      mr[k] = mask[o];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 32; ++k) mr[k] = 1;
  }
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const int t = t0 + k;
    ok[k] = (col < N && t < T) && finite_v(xr[k]) && mr[k] != 0;
    v[k] = (col < N && t < T) ? static_cast<double>(xr[k]) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (ok[k]) {
      n += 1.0;
      sum += v[k];
      mn = fmin(mn, v[k]);
      mx = fmax(mx, v[k]);
      const double a = fabs(v[k]);
      am = fmax(am, a);
      if (a != 0.0) az = fmin(az, a);
    }
  }
  const double mean = n > 0.0 ? sum / n : 0.0;
  double m2 = 0.0;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (ok[k]) {
      const double d = v[k] - mean;
      m2 += d * d;
    }
  }
  wp[w][lane] = ColPart{n, mean, m2, mn, mx, am, az, 0.0};
  __syncthreads();
  if (w == 0 && col < N) {
    ColPart c = wp[0][lane];
    for (int k = 1; k < 4; ++k) {
      const ColPart b = wp[k][lane];
      chan_combine(c.n, c.mean, c.m2, b.n, b.mean, b.m2);
      c.minv = fmin(c.minv, b.minv);
      c.maxv = fmax(c.maxv, b.maxv);
      c.amax = fmax(c.amax, b.amax);
      c.aminnz = fmin(c.aminnz, b.aminnz);
    }
    parts[static_cast<size_t>(tc) * N + col] = c;
  }
}

// (2) per column: combine the chunk partials in chunk order, derive shift,
// scale and the diagonal value.
template <int TIER>
__global__ __launch_bounds__(256) void stock_colstat_kernel(const ColPart* __restrict__ parts,
                                                           int nchunks, int N,
                                                           ColStat* __restrict__ cs,
                                                           uint32_t* __restrict__ fb_count) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  // the tile kernels' fallback counter starts at 0 (a store here instead of
  // a memset node, so the call's sequence can be captured)
  if (col == 0 && fb_count != nullptr) *fb_count = 0u;
  if (col >= N) return;
  double n = 0.0, mean = 0.0, m2 = 0.0, mn = __builtin_inf(), mx = -__builtin_inf(), am = 0.0,
         az = 1e308;
  for (int c = 0; c < nchunks; ++c) {
    const ColPart b = parts[static_cast<size_t>(c) * N + col];
    chan_combine(n, mean, m2, b.n, b.mean, b.m2);
    mn = fmin(mn, b.minv);
    mx = fmax(mx, b.maxv);
    am = fmax(am, b.amax);
    az = fmin(az, b.aminnz);
  }
  double scale = 1.0;
  if (n > 0.0 && m2 > 0.0 && __builtin_isfinite(m2)) {
    const double sd = sqrt(m2 / n);
    scale = (TIER == kTierContract) ? ldexp(1.0, -ilogb(sd)) : 1.0 / sd;
  }
  ColStat c;
  c.n = n;
  c.mean = mean;
  c.scale = scale;
  c.sxx = m2;
  c.minv = mn;
  c.maxv = mx;
  c.amax = am;
  c.aminnz = az;
  c.diag = (n >= 2.0 && mn < mx && m2 > 0.0 && __builtin_isfinite(m2)) ? 1.0 : kNaN;
  for (int k = 0; k < 7; ++k) c.pad[k] = 0.0;
  cs[col] = c;
}

// (3) operand planes, one 64-row x 64-column tile per block, transposed
// through LDS so each stock's series is contiguous in every plane; also the
// original values (NaN where invalid, input precision) for the exact
// fallback pass. z is staged at the precision of its plane (FP64 for the
// contract tier, FP32 otherwise, where the planes round it from FP32).
template <typename V, int TIER>
__global__ __launch_bounds__(256) void stock_planes_kernel(const V* __restrict__ X,
                                                          const uint8_t* __restrict__ mask, int T,
                                                          int N, const ColStat* __restrict__ cs,
                                                          StockPlanes P) {
  using ZT = typename std::conditional<TIER == kTierContract, double, float>::type;
  __shared__ ZT zt[64][65];
  __shared__ V xt[64][65];
  __shared__ uint8_t vt[64][65];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 64, tb = blockIdx.y * 64;
  const int col = c0 + lane;
  double mean = 0.0, scale = 1.0;
  if (col < N) {
    mean = cs[col].mean;
    scale = cs[col].scale;
  }
  // all 16 rows of this thread loaded before any is used (out-of-panel
  // cells read element 0 and are masked off)
  V xr[16];
  uint8_t mr[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int t = tb + w + 4 * q;
    const size_t o = (col < N && t < T) ? static_cast<size_t>(t) * N + col : 0;
    xr[q] = X[o];
  }
  if (mask != nullptr) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int t = tb + w + 4 * q;
      const size_t o = (col < N && t < T) ? static_cast<size_t>(t) * N + col : 0;
      mr[q] = mask[o];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) mr[q] = 1;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = w + 4 * q;
    const int t = tb + r;
    const bool ok = (col < N && t < T) && finite_v(xr[q]) && mr[q] != 0;
    double z = 0.0;
    V xo = static_cast<V>(kNaN);
    if (ok) {
      xo = xr[q];
      z = (static_cast<double>(xr[q]) - mean) * scale;
    }
    zt[lane][r] = static_cast<ZT>(z);
    xt[lane][r] = xo;
    vt[lane][r] = ok ? 1 : 0;
  }
  __syncthreads();
  const int Tp = P.Tp;
  for (int c = w; c < 64; c += 4) {
    const size_t po = static_cast<size_t>(c0 + c) * Tp + tb + lane;
    const bool ok = vt[c][lane] != 0;
    static_cast<V*>(P.xt)[po] = xt[c][lane];
    if (TIER == kTierContract) {
      static_cast<double*>(P.z)[po] = zt[c][lane];
      static_cast<double*>(P.m)[po] = ok ? 1.0 : 0.0;
    } else if (TIER == kTierFp32) {
      static_cast<float*>(P.z)[po] = static_cast<float>(zt[c][lane]);
      static_cast<float*>(P.m)[po] = ok ? 1.0f : 0.0f;
    } else {
      const float z = static_cast<float>(zt[c][lane]);
      const uint16_t hi = f2bf(z);
      const uint16_t lo = f2bf(z - bf2f(hi));
      static_cast<uint16_t*>(P.z)[po] = hi;
      static_cast<uint16_t*>(P.lo)[po] = lo;
      // the square of the stored value hi + lo, split exactly as square_split
      const float z0 = bf2f(hi) + bf2f(lo);
      const __bf16 q0 = static_cast<__bf16>(z0 * z0);
      static_cast<__bf16*>(P.qh)[po] = q0;
      static_cast<__bf16*>(P.ql)[po] = static_cast<__bf16>(__builtin_fmaf(z0, z0, -static_cast<float>(q0)));
      static_cast<uint16_t*>(P.m)[po] = ok ? static_cast<uint16_t>(0x3f80) : static_cast<uint16_t>(0);
      P.m8[po] = ok ? 1 : 0;
    }
  }
}

// (4) fully-valid path only: sum z and sum z^2 of the stored bf16 hi+lo
// values per column (one wave per column, fixed butterfly order).
__global__ __launch_bounds__(256) void stock_fullsums_kernel(StockPlanes P, int N,
                                                            ColStat* __restrict__ cs) {
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= N) return;
  const uint16_t* HI = static_cast<const uint16_t*>(P.z) + static_cast<size_t>(col) * P.Tp;
  const uint16_t* LO = static_cast<const uint16_t*>(P.lo) + static_cast<size_t>(col) * P.Tp;
  double s1 = 0.0, s2 = 0.0;
  for (int t = lane; t < P.Tp; t += 64) {
    const double z = static_cast<double>(bf2f(HI[t]) + bf2f(LO[t]));
    s1 += z;
    s2 += z * z;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if (lane == 0) {
    cs[col].pad[0] = s1;
    cs[col].pad[1] = s2;
  }
}

// ---- stock_corr: tile scheduling helpers -------------------------------------
// Upper-triangle tile index b -> (I, J), I <= J < nt, row-major over I.
__device__ __forceinline__ void tri_tile(int b, int nt, int& I, int& J) {
  // rows I have (nt - I) tiles; start(I) = I*nt - I(I-1)/2
  double disc = (2.0 * nt + 1.0) * (2.0 * nt + 1.0) - 8.0 * b;
  int i = static_cast<int>(((2.0 * nt + 1.0) - sqrt(disc)) * 0.5);
  if (i < 0) i = 0;
  while (i > 0 && (i * nt - i * (i - 1) / 2) > b) --i;
  while (((i + 1) * nt - (i + 1) * i / 2) <= b) ++i;
  I = i;
  J = i + (b - (i * nt - i * (i - 1) / 2));
}

// Writes one finished cell; the mirror half goes through the wave's LDS
// transpose buffer (mt, SS x (SS+1) doubles) so both halves are coalesced.
template <int SS>
__device__ __forceinline__ void stage_cell(double* mt, int il, int jl, double v) {
  mt[jl * (SS + 1) + il] = v;
}

template <int TIER, int SS>
__device__ __forceinline__ void emit_cell(int i, int j, int il, int jl, int N, double n, double sx,
                                          double sy, double sxy, double sxx, double syy,
                                          const ColStat* __restrict__ cs, double* __restrict__ out,
                                          double* mt, uint32_t* fbl, uint32_t* fbc, uint32_t fbcap) {
  double v = kNaN;
  if (i < N && j < N && i < j) {
    const Fin f = finalize_cell<TIER>(n, sx, sy, sxy, sxx, syy, cs[i], cs[j]);
    v = f.r;
    if (f.fb) {
      const uint32_t q = atomicAdd(fbc, 1u);
      if (q < fbcap) {
        fbl[2 * q] = static_cast<uint32_t>(i);
        fbl[2 * q + 1] = static_cast<uint32_t>(j);
      }
    }
    out[static_cast<size_t>(i) * N + j] = v;
  } else if (i < N && i == j) {
    out[static_cast<size_t>(i) * N + i] = cs[i].diag;
  }
  stage_cell<SS>(mt, il, jl, v);
}

template <int TIER, int SS>
__device__ __forceinline__ void emit_fast(int i, int j, int il, int jl, int N, float n, float sx,
                                          float sy, float sxy, float sxx, float syy, bool konst,
                                          const ColStat* __restrict__ cs, double* __restrict__ out,
                                          double* mt, uint32_t* fbl, uint32_t* fbc, uint32_t fbcap) {
  const bool cell = (i < N) & (j < N) & (i < j);
  const Fin f = finalize_fast<TIER>(n, sx, sy, sxy, sxx, syy, konst);
  const bool fb = cell & f.fb;
  const double v = cell ? f.r : kNaN;
  // fallback cells of this wave appended with one atomic (rare: uniform skip)
  const uint64_t m = __ballot(fb);
  if (m != 0) {
    const int lead = __builtin_ctzll(m);
    uint32_t base = 0;
    if (lane_id() == lead) base = atomicAdd(fbc, static_cast<uint32_t>(__builtin_popcountll(m)));
    base = __shfl(base, lead);
    const uint32_t q = base + static_cast<uint32_t>(
                                  __builtin_popcountll(m & ((1ull << lane_id()) - 1ull)));
    if (fb && q < fbcap) {
      fbl[2 * q] = static_cast<uint32_t>(i);
      fbl[2 * q + 1] = static_cast<uint32_t>(j);
    }
  }
  if (cell) out[static_cast<size_t>(i) * N + j] = v;
  if (i == j && i < N) out[static_cast<size_t>(i) * N + i] = cs[i].diag;
  stage_cell<SS>(mt, il, jl, v);
}

// Constant-column flag of column c (false past N).
__device__ __forceinline__ bool col_const(const ColStat* __restrict__ cs, int c, int N) {
  return c < N && cs[c].minv == cs[c].maxv;
}

// Bit il set iff column rb + il (il < 32) is constant; lanes 0..31 load.
__device__ __forceinline__ uint32_t row_const_mask(const ColStat* __restrict__ cs, int rb, int N) {
  const int lane = lane_id();
  const bool c = lane < 32 && col_const(cs, rb + lane, N);
  return static_cast<uint32_t>(__ballot(c));
}

// Mirror store: row jl of the staged block -> out[(cb+jl)*N + rb + il] for
// i < j cells only.
template <int SS>
__device__ __forceinline__ void mirror_store(const double* mt, int rb, int cb, int N,
                                             double* __restrict__ out) {
  const int lane = lane_id();
  constexpr int rows_per_iter = 64 / SS;
  for (int jl0 = 0; jl0 < SS; jl0 += rows_per_iter) {
    const int jl = jl0 + lane / SS;
    const int il = lane % SS;
    const int i = rb + il, j = cb + jl;
    if (i < N && j < N && i < j) out[static_cast<size_t>(j) * N + i] = mt[jl * (SS + 1) + il];
  }
}

// ---- stock_corr: bf16x3 tile kernel (64x64 tile, 4 waves of 32x32) --------
__device__ __forceinline__ bf16x8 ld_bf8(const uint16_t* p) {
  const u16x8 v = *reinterpret_cast<const u16x8*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Two floats to packed bf16 (round to nearest even, one v_cvt_pk_bf16_f32),
// and the two bf16 values back as floats from the packed bits.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ __forceinline__ float bf16_lo_f(uint32_t p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float bf16_hi_f(uint32_t p) { return __uint_as_float(p & 0xffff0000u); }

// x^2 as an unevaluated bf16 pair: x = hi + lo (exact in fp32), q = fl(x*x),
// qh = bf16(q) (round to nearest even), ql = bf16(x*x - qh) with the
// residual taken by one fma from the exact square.
__device__ __forceinline__ void square_split(const u16x8& hi, const u16x8& lo, bf16x8& qh,
                                             bf16x8& ql) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 h = __builtin_bit_cast(u32x4, hi), l = __builtin_bit_cast(u32x4, lo);
  u32x4 a, b;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float z0 = __uint_as_float(h[e] << 16) + __uint_as_float(l[e] << 16);
    const float z1 = __uint_as_float(h[e] & 0xffff0000u) + __uint_as_float(l[e] & 0xffff0000u);
    a[e] = cvt_pk_bf16(z0 * z0, z1 * z1);
    b[e] = cvt_pk_bf16(__builtin_fmaf(z0, z0, -bf16_lo_f(a[e])), __builtin_fmaf(z1, z1, -bf16_hi_f(a[e])));
  }
  qh = __builtin_bit_cast(bf16x8, a);
  ql = __builtin_bit_cast(bf16x8, b);
}

template <int FULL>
__global__ __launch_bounds__(256) void stock_tile_bf16x3_kernel(StockPlanes P,
                                                                const ColStat* __restrict__ cs,
                                                                int T, int N, double* __restrict__ out,
                                                                uint32_t* fbl, uint32_t* fbc,
                                                                uint32_t fbcap) {
  __shared__ double mtb[4][32 * 33];
  const int nt = P.Np / 64;
  int I, J;
  tri_tile(blockIdx.x, nt, I, J);
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const int wi = w >> 1, wj = w & 1;
  if (I == J && wi > wj) return;  // below-diagonal quadrant of a diagonal tile
  const int rb = I * 64 + wi * 32, cb = J * 64 + wj * 32;
  const int Tp = P.Tp;
  const uint16_t* HI = static_cast<const uint16_t*>(P.z);
  const uint16_t* LO = static_cast<const uint16_t*>(P.lo);
  const uint16_t* MB = static_cast<const uint16_t*>(P.m);
  const size_t aoff = static_cast<size_t>(rb + (lane & 31)) * Tp;
  const size_t boff = static_cast<size_t>(cb + (lane & 31)) * Tp;
  const int kh = 8 * (lane >> 5);
  const int kq = 16 * (lane >> 5);
  f32x16 sxy = {}, sx = {}, sy = {}, sxx = {}, syy = {};
  i32x16 cnt = {};
  for (int k0 = 0; k0 < Tp; k0 += 32) {
#pragma unroll
    for (int hlf = 0; hlf < 2; ++hlf) {
      const int kk = k0 + 16 * hlf + kh;
      const u16x8 ah = *reinterpret_cast<const u16x8*>(HI + aoff + kk);
      const u16x8 al = *reinterpret_cast<const u16x8*>(LO + aoff + kk);
      const u16x8 bh = *reinterpret_cast<const u16x8*>(HI + boff + kk);
      const u16x8 bl = *reinterpret_cast<const u16x8*>(LO + boff + kk);
      const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
      const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
      sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bh, sxy, 0, 0, 0);
      sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bl, sxy, 0, 0, 0);
      sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bh, sxy, 0, 0, 0);
      if (!FULL) {
        const bf16x8 Am = ld_bf8(MB + aoff + kk);
        const bf16x8 Bm = ld_bf8(MB + boff + kk);
        bf16x8 Aqh, Aql, Bqh, Bql;
        square_split(ah, al, Aqh, Aql);
        square_split(bh, bl, Bqh, Bql);
        sx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bm, sx, 0, 0, 0);
        sx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bm, sx, 0, 0, 0);
        sy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bh, sy, 0, 0, 0);
        sy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bl, sy, 0, 0, 0);
        sxx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aqh, Bm, sxx, 0, 0, 0);
        sxx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aql, Bm, sxx, 0, 0, 0);
        syy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bqh, syy, 0, 0, 0);
        syy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bql, syy, 0, 0, 0);
      }
    }
    if (!FULL) {
      const i32x4 a8 = *reinterpret_cast<const i32x4*>(P.m8 + aoff + k0 + kq);
      const i32x4 b8 = *reinterpret_cast<const i32x4*>(P.m8 + boff + k0 + kq);
      cnt = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, b8, cnt, 0, 0, 0);
    }
  }
  double* mt = mtb[w];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int il = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int jl = lane & 31;
    const int i = rb + il, j = cb + jl;
    double n, vx, vy, vxx, vyy;
    if (FULL) {
      n = static_cast<double>(T);
      vx = 0.0;
      vy = 0.0;
      vxx = n;
      vyy = n;
      if (i < N && j < N) {
        // fully valid: sums over all T rows of standardised columns
        vx = cs[i].pad[0];
        vy = cs[j].pad[0];
        vxx = cs[i].pad[1];
        vyy = cs[j].pad[1];
      }
    } else {
      n = static_cast<double>(cnt[r]);
      vx = sx[r];
      vy = sy[r];
      vxx = sxx[r];
      vyy = syy[r];
    }
    emit_cell<kTierBf16x3, 32>(i, j, il, jl, N, n, vx, vy, static_cast<double>(sxy[r]), vxx, vyy, cs,
                               out, mt, fbl, fbc, fbcap);
  }
  __builtin_amdgcn_wave_barrier();
  mirror_store<32>(mt, rb, cb, N, out);
}

// ---- stock_corr: bf16x3 tile kernel, 64x64 tiles on an LDS DMA ring -------
// Each workgroup (4 waves) owns a 64 (rows i) x 64 (cols j) tile; wave w
// computes the 32x32 block of A rows 32(w & 1).. against B columns
// 32(w >> 1).. (96 accumulators: five f32 sums and one i32 count). The T axis
// runs in stages of 32 rows through a 2-deep LDS ring of 28 KB stages, so
// two workgroups share a CU and each SIMD holds one wave of each: their stage
// barriers and first fragment reads are independent, and one's waits overlap
// the other's MFMAs. Stages are filled by global_load_lds_dwordx4 (each wave
// instruction copies 1 KB straight into LDS, no staging registers), issued
// one stage ahead; a wave waits for its own copies with vmcnt and the block
// barrier publishes them.
//
// Stage layout (bytes): A hi, lo planes [64][32] bf16 (4 KB each), A mask
// int8 [64][32] (2 KB), B hi, lo, x^2 hi, x^2 lo planes [64][32] bf16 (4 KB
// each), B mask int8 [64][32] (2 KB) = 28 KB. B brings its squares from the
// x^2 planes; each wave splits the squares of its A rows in registers with
// the same formula. The bf16 0/1 mask operands are expanded from the int8
// masks. Within a row, 16-byte chunks are XOR-swizzled with row bits so the
// fragment reads of 16 consecutive rows hit distinct banks: bf16 rows (64 B)
// hold chunk c at slot c ^ ((r >> 2) & 3); int8 rows (32 B) at slot
// c ^ ((r >> 3) & 1).
constexpr int kRK = 32;  // T rows per stage

// One 16-byte global -> LDS copy per lane, to LDS byte address m0 + 16*lane.
// The copy is invisible to the compiler's wait-count tracking: callers wait
// with s_waitcnt vmcnt and a barrier before reading the destination.
__device__ __forceinline__ void lds_dma16(const void* src, uint32_t lds_addr) {
  const uint32_t a = __builtin_amdgcn_readfirstlane(lds_addr);
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(src), "s"(a)
               : "memory", "m0");
}

// Eight int8 0/1 mask bytes -> eight bf16 0.0/1.0 (0x3f80): each byte pair
// b0, b1 placed at bits 0 and 16 (v_perm), then times 0x3f80.
__device__ __forceinline__ bf16x8 mask_bf16(uint64_t m8) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const uint32_t w0 = static_cast<uint32_t>(m8), w1 = static_cast<uint32_t>(m8 >> 32);
  u32x4 r;
  r[0] = __umul24(__builtin_amdgcn_perm(0u, w0, 0x0c010c00u), 0x3f80u);
  r[1] = __umul24(__builtin_amdgcn_perm(0u, w0, 0x0c030c02u), 0x3f80u);
  r[2] = __umul24(__builtin_amdgcn_perm(0u, w1, 0x0c010c00u), 0x3f80u);
  r[3] = __umul24(__builtin_amdgcn_perm(0u, w1, 0x0c030c02u), 0x3f80u);
  return __builtin_bit_cast(bf16x8, r);
}

// The eleven bf16 products of one 32x32 block: sxy = hi.hi + hi.lo + lo.hi,
// sx = (hi + lo).m, sy = m.(hi + lo), sxx = (x^2 hi + lo).m, syy = m.(y^2 hi + lo).
__device__ __forceinline__ void stock_block_mfma(const bf16x8& Ah, const bf16x8& Al,
                                                 const bf16x8& Am, const bf16x8& Aqh,
                                                 const bf16x8& Aql, const u16x8& bh,
                                                 const u16x8& bl, const u16x8& bqh,
                                                 const u16x8& bql, const bf16x8& Bm, f32x16& sxy,
                                                 f32x16& sx, f32x16& sy, f32x16& sxx, f32x16& syy) {
  const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
  const bf16x8 Bqh = __builtin_bit_cast(bf16x8, bqh), Bql = __builtin_bit_cast(bf16x8, bql);
  sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bh, sxy, 0, 0, 0);
  sx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bm, sx, 0, 0, 0);
  sy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bh, sy, 0, 0, 0);
  sxx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aqh, Bm, sxx, 0, 0, 0);
  syy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bqh, syy, 0, 0, 0);
  sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ah, Bl, sxy, 0, 0, 0);
  sx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bm, sx, 0, 0, 0);
  sy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bl, sy, 0, 0, 0);
  sxx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aql, Bm, sxx, 0, 0, 0);
  syy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Am, Bql, syy, 0, 0, 0);
  sxy = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Al, Bh, sxy, 0, 0, 0);
}

// One 16-row sub-step for one 32-column B half.
struct StockFrag1 {
  u16x8 ah, al, bh, bl, bqh, bql;
  uint64_t am, bm;
};

__device__ __forceinline__ void stock_frag1_mfma(const StockFrag1& F, f32x16& sxy, f32x16& sx,
                                                 f32x16& sy, f32x16& sxx, f32x16& syy) {
  const bf16x8 Ah = __builtin_bit_cast(bf16x8, F.ah), Al = __builtin_bit_cast(bf16x8, F.al);
  const bf16x8 Am = mask_bf16(F.am);
  bf16x8 Aqh, Aql;
  square_split(F.ah, F.al, Aqh, Aql);
  stock_block_mfma(Ah, Al, Am, Aqh, Aql, F.bh, F.bl, F.bqh, F.bql, mask_bf16(F.bm), sxy, sx, sy,
                   sxx, syy);
}

constexpr int kRI = 64, kRJ = 64, kRStages = 2;
constexpr int kOffAh = 0, kOffAl = 4096, kOffA8 = 8192;
constexpr int kOffBh = 10240, kOffBl = 14336, kOffBqh = 18432, kOffBql = 22528, kOffB8 = 26624;
constexpr int kRStageBytes = 28672;
constexpr int kRUnits = 28;  // 1 KB copy units per stage, 7 per wave

// Copy unit u of a stage: 0..7 A hi, lo (16 rows each), 8..9 A int8 mask
// (32 rows), 10..25 B hi, lo, x^2 hi, x^2 lo, 26..27 B int8 mask.
__device__ __forceinline__ void ring_unit(const StockPlanes& P, int u, int ra, int rb, int lane,
                                            const unsigned char*& src, uint32_t& dst, int& esz) {
  const size_t Tp = static_cast<size_t>(P.Tp);
  if (u < 8 || (u >= 10 && u < 26)) {
    const bool a = u < 8;
    const int v = a ? u : u - 10;
    const int pl = v >> 2;  // 0 hi, 1 lo, 2 x^2 hi, 3 x^2 lo
    const int row = (v & 3) * 16 + (lane >> 2);
    const int c = (lane & 3) ^ ((lane >> 4) & 3);
    const uint16_t* base =
        static_cast<const uint16_t*>(pl == 0 ? P.z : (pl == 1 ? P.lo : (pl == 2 ? P.qh : P.ql)));
    src = reinterpret_cast<const unsigned char*>(base + static_cast<size_t>((a ? ra : rb) + row) * Tp +
                                                 c * 8);
    dst = static_cast<uint32_t>((a ? kOffAh : kOffBh) + pl * 4096 + (v & 3) * 1024);
    esz = 2;
  } else {
    const bool a = u < 10;
    const int v = a ? u - 8 : u - 26;
    const int row = v * 32 + (lane >> 1);
    const int c = (lane & 1) ^ ((lane >> 4) & 1);
    src = reinterpret_cast<const unsigned char*>(P.m8 + static_cast<size_t>((a ? ra : rb) + row) * Tp +
                                                 c * 16);
    dst = static_cast<uint32_t>((a ? kOffA8 : kOffB8) + v * 1024);
    esz = 1;
  }
}

struct RingUnits {
  const unsigned char* src[7];
  uint32_t dst[7];
  int esz[7];
};

__device__ __forceinline__ void ring_issue(const RingUnits& U, int stage, uint32_t ring) {
  const uint32_t base = ring + static_cast<uint32_t>((stage & 1) * kRStageBytes);
  const int k0 = stage * kRK;
#pragma unroll
  for (int q = 0; q < 7; ++q)
    lds_dma16(U.src[q] + static_cast<size_t>(k0) * U.esz[q], base + U.dst[q]);
}

__device__ __forceinline__ void stock_frag1_load(const unsigned char* S, int arow, int brow,
                                                   int ofs, int ma, int mb, StockFrag1& F) {
  F.ah = *reinterpret_cast<const u16x8*>(S + kOffAh + arow + ofs);
  F.al = *reinterpret_cast<const u16x8*>(S + kOffAl + arow + ofs);
  F.bh = *reinterpret_cast<const u16x8*>(S + kOffBh + brow + ofs);
  F.bl = *reinterpret_cast<const u16x8*>(S + kOffBl + brow + ofs);
  F.bqh = *reinterpret_cast<const u16x8*>(S + kOffBqh + brow + ofs);
  F.bql = *reinterpret_cast<const u16x8*>(S + kOffBql + brow + ofs);
  F.am = *reinterpret_cast<const uint64_t*>(S + ma);
  F.bm = *reinterpret_cast<const uint64_t*>(S + mb);
}

__global__ __launch_bounds__(256, 2) void stock_tile_bf16x3_ring_kernel(StockPlanes P,
                                                                    const ColStat* __restrict__ cs,
                                                                    int N, int ntiles,
                                                                    double* __restrict__ out,
                                                                    uint32_t* fbl, uint32_t* fbc,
                                                                    uint32_t fbcap) {
  __shared__ __align__(16) unsigned char ring[kRStages * kRStageBytes];
  const int per_xcd = (ntiles + 7) / 8;
  const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;  // whole workgroup: no barrier is left waiting
  const int nt = P.Np / kRJ;
  int I, J;
  tri_tile(tile, nt, I, J);
  const int lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ab = w & 1, bh = w >> 1;
  const int ra = I * kRI, rb = J * kRJ;
  // 32-row A block 2I+ab against B block 2J+bh; the block below the diagonal
  // of a diagonal tile is computed but not written
  const bool act = 2 * I + ab <= 2 * J + bh;
  const uint32_t ring_lds = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(ring));
  const uint32_t ring_u = __builtin_amdgcn_readfirstlane(ring_lds);
  const int nst = P.Tp / kRK;
  RingUnits U;
#pragma unroll
  for (int q = 0; q < 7; ++q) ring_unit(P, w + 4 * q, ra, rb, lane, U.src[q], U.dst[q], U.esz[q]);
  ring_issue(U, 0, ring_u);

  f32x16 sxy = {}, sx = {}, sy = {}, sxx = {}, syy = {};
  i32x16 cnt = {};
  const int r = lane & 31;
  const int swz = (r >> 2) & 3;
  const int swz8 = (r >> 3) & 1;
  const int arow = (ab * 32 + r) * 64, brow = (bh * 32 + r) * 64;
  const int a8 = kOffA8 + (ab * 32 + r) * 32 + (((lane >> 5) ^ swz8) * 16);
  const int b8 = kOffB8 + (bh * 32 + r) * 32 + (((lane >> 5) ^ swz8) * 16);
  const int ofs0 = ((lane >> 5) ^ swz) * 16, ofs1 = ((2 + (lane >> 5)) ^ swz) * 16;
  const int mh = 8 * (lane >> 5);
  const int ma0 = kOffA8 + (ab * 32 + r) * 32 + swz8 * 16 + mh;
  const int ma1 = kOffA8 + (ab * 32 + r) * 32 + (1 ^ swz8) * 16 + mh;
  const int mb0 = kOffB8 + (bh * 32 + r) * 32 + swz8 * 16 + mh;
  const int mb1 = kOffB8 + (bh * 32 + r) * 32 + (1 ^ swz8) * 16 + mh;

  for (int st = 0; st < nst; ++st) {
    // stage st landed for this wave (nothing later is in flight yet), then
    // for every wave; after the barrier the slot of stage st-1 is free
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    const unsigned char* S = ring + (st & 1) * kRStageBytes;
    StockFrag1 F0, F1;
    stock_frag1_load(S, arow, brow, ofs0, ma0, mb0, F0);
    if (st + 1 < nst) ring_issue(U, st + 1, ring_u);
    stock_frag1_load(S, arow, brow, ofs1, ma1, mb1, F1);
    stock_frag1_mfma(F0, sxy, sx, sy, sxx, syy);
    const i32x4 A8 = *reinterpret_cast<const i32x4*>(S + a8);
    const i32x4 B8 = *reinterpret_cast<const i32x4*>(S + b8);
    stock_frag1_mfma(F1, sxy, sx, sy, sxx, syy);
    cnt = __builtin_amdgcn_mfma_i32_32x32x32_i8(A8, B8, cnt, 0, 0, 0);
  }
  __syncthreads();  // the ring becomes the mirror transpose buffers
  double* mt = reinterpret_cast<double*>(ring) + w * 32 * 33;
  const int rbw = ra + ab * 32, cbw = rb + bh * 32;
  const uint32_t rowc = row_const_mask(cs, rbw, N);
  const int jl = lane & 31;
  if (act) {
    const bool colc = col_const(cs, cbw + jl, N);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int il = 8 * (k >> 2) + 4 * (lane >> 5) + (k & 3);
      emit_fast<kTierBf16x3, 32>(rbw + il, cbw + jl, il, jl, N, static_cast<float>(cnt[k]), sx[k],
                                 sy[k], sxy[k], sxx[k], syy[k], colc || ((rowc >> il) & 1u), cs,
                                 out, mt, fbl, fbc, fbcap);
    }
    __builtin_amdgcn_wave_barrier();
    mirror_store<32>(mt, rbw, cbw, N, out);
  }
}

// ---- stock_corr: fp32 tile kernel (64x64 tile, 4 waves of 32x32) ---------
// FP32 MFMA 32x32x2: lane half h holds rows k = t0 + 4h + m of the 8-row
// chunk, so MFMA m contracts rows {t0+m, t0+4+m} (the same row permutation on
// both operands).
__global__ __launch_bounds__(256) void stock_tile_fp32_kernel(StockPlanes P,
                                                             const ColStat* __restrict__ cs,
                                                             int T, int N, double* __restrict__ out,
                                                             uint32_t* fbl, uint32_t* fbc,
                                                             uint32_t fbcap) {
  __shared__ double mtb[4][32 * 33];
  const int nt = P.Np / 64;
  int I, J;
  tri_tile(blockIdx.x, nt, I, J);
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const int wi = w >> 1, wj = w & 1;
  if (I == J && wi > wj) return;
  const int rb = I * 64 + wi * 32, cb = J * 64 + wj * 32;
  const int Tp = P.Tp;
  const float* Z = static_cast<const float*>(P.z);
  const float* M = static_cast<const float*>(P.m);
  const size_t aoff = static_cast<size_t>(rb + (lane & 31)) * Tp + 4 * (lane >> 5);
  const size_t boff = static_cast<size_t>(cb + (lane & 31)) * Tp + 4 * (lane >> 5);
  f32x16 sxy = {}, sx = {}, sy = {}, sxx = {}, syy = {}, cn = {};
  for (int k0 = 0; k0 < Tp; k0 += 8) {
    const f32x4 za = *reinterpret_cast<const f32x4*>(Z + aoff + k0);
    const f32x4 ma = *reinterpret_cast<const f32x4*>(M + aoff + k0);
    const f32x4 zb = *reinterpret_cast<const f32x4*>(Z + boff + k0);
    const f32x4 mb = *reinterpret_cast<const f32x4*>(M + boff + k0);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const float a = za[m], b = zb[m], am = ma[m], bm = mb[m];
      sxy = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, sxy, 0, 0, 0);
      sx = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bm, sx, 0, 0, 0);
      sy = __builtin_amdgcn_mfma_f32_32x32x2f32(am, b, sy, 0, 0, 0);
      sxx = __builtin_amdgcn_mfma_f32_32x32x2f32(a * a, bm, sxx, 0, 0, 0);
      syy = __builtin_amdgcn_mfma_f32_32x32x2f32(am, b * b, syy, 0, 0, 0);
      cn = __builtin_amdgcn_mfma_f32_32x32x2f32(am, bm, cn, 0, 0, 0);
    }
  }
  double* mt = mtb[w];
  const uint32_t rowc = row_const_mask(cs, rb, N);
  const bool colc = col_const(cs, cb + (lane & 31), N);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int il = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int jl = lane & 31;
    emit_fast<kTierFp32, 32>(rb + il, cb + jl, il, jl, N, cn[r], sx[r], sy[r], sxy[r], sxx[r],
                             syy[r], colc || ((rowc >> il) & 1u), cs, out, mt, fbl, fbc, fbcap);
  }
  __builtin_amdgcn_wave_barrier();
  mirror_store<32>(mt, rb, cb, N, out);
}

// ---- stock_corr: contract tile kernel (32x32 tile, 4 waves of 16x16) -------
// FP64 MFMA 16x16x4: lane group g = lane/16 holds rows t0 + 4g + m of the
// 16-row chunk, so MFMA m contracts rows {t0 + 4g + m : g = 0..3}.
__global__ __launch_bounds__(256) void stock_tile_f64_kernel(StockPlanes P,
                                                            const ColStat* __restrict__ cs, int T,
                                                            int N, double* __restrict__ out,
                                                            uint32_t* fbl, uint32_t* fbc,
                                                            uint32_t fbcap) {
  __shared__ double mtb[4][16 * 17];
  const int nt = P.Np / 32;
  int I, J;
  tri_tile(blockIdx.x, nt, I, J);
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const int wi = w >> 1, wj = w & 1;
  if (I == J && wi > wj) return;
  const int rb = I * 32 + wi * 16, cb = J * 32 + wj * 16;
  const int Tp = P.Tp;
  const double* Z = static_cast<const double*>(P.z);
  const double* M = static_cast<const double*>(P.m);
  const size_t aoff = static_cast<size_t>(rb + (lane & 15)) * Tp + 4 * (lane >> 4);
  const size_t boff = static_cast<size_t>(cb + (lane & 15)) * Tp + 4 * (lane >> 4);
  f64x4 sxy = {}, sx = {}, sy = {}, sxx = {}, syy = {}, cn = {};
  for (int k0 = 0; k0 < Tp; k0 += 16) {
    double za[4], ma[4], zb[4], mb[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      za[m] = Z[aoff + k0 + m];
      ma[m] = M[aoff + k0 + m];
      zb[m] = Z[boff + k0 + m];
      mb[m] = M[boff + k0 + m];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      sxy = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m], zb[m], sxy, 0, 0, 0);
      sx = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m], mb[m], sx, 0, 0, 0);
      sy = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], zb[m], sy, 0, 0, 0);
      sxx = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m] * za[m], mb[m], sxx, 0, 0, 0);
      syy = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], zb[m] * zb[m], syy, 0, 0, 0);
      cn = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], mb[m], cn, 0, 0, 0);
    }
  }
  double* mt = mtb[w];
  // FP64 16x16 MFMA output: register r of lane l holds row 4r + l/16, col l%16
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int il = 4 * r + (lane >> 4);
    const int jl = lane & 15;
    emit_cell<kTierContract, 16>(rb + il, cb + jl, il, jl, N, cn[r], sx[r], sy[r], sxy[r], sxx[r],
                                 syy[r], cs, out, mt, fbl, fbc, fbcap);
  }
  __builtin_amdgcn_wave_barrier();
  mirror_store<16>(mt, rb, cb, N, out);
}

// ---- exact fallback ------------------------------------------------------------
// Serial two-pass over the pair-valid rows in row order with Kahan-compensated
// sums (means and centred sums), exact-constant operand -> NaN; the HG-2
// reference algorithm. FP contraction is off so every step rounds as the
// reference's float64 ops do.
struct Kahan {
  double s, c;
  __device__ __forceinline__ void add(double x) {
#pragma clang fp contract(off)
    const double y = x - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

template <typename V>
__device__ double pair_two_pass(const V* __restrict__ xi, const V* __restrict__ xj, long long stride,
                                const uint8_t* __restrict__ mi, const uint8_t* __restrict__ mj,
                                long long mstride, long long R) {
#pragma clang fp contract(off)
  Kahan ax{0.0, 0.0}, ay{0.0, 0.0};
  double n = 0.0, mnx = __builtin_inf(), mxx = -__builtin_inf(), mny = __builtin_inf(),
         mxy = -__builtin_inf();
  for (long long r = 0; r < R; ++r) {
    const V a = xi[r * stride], b = xj[r * stride];
    if (!finite_v(a) || !finite_v(b)) continue;
    if (mi != nullptr && (mi[r * mstride] == 0 || mj[r * mstride] == 0)) continue;
    const double x = static_cast<double>(a), y = static_cast<double>(b);
    n += 1.0;
    ax.add(x);
    ay.add(y);
    mnx = fmin(mnx, x);
    mxx = fmax(mxx, x);
    mny = fmin(mny, y);
    mxy = fmax(mxy, y);
  }
  if (n < 2.0 || mnx == mxx || mny == mxy) return kNaN;
  const double mx = ax.s / n, my = ay.s / n;
  Kahan kxx{0.0, 0.0}, kyy{0.0, 0.0}, kxy{0.0, 0.0};
  for (long long r = 0; r < R; ++r) {
    const V a = xi[r * stride], b = xj[r * stride];
    if (!finite_v(a) || !finite_v(b)) continue;
    if (mi != nullptr && (mi[r * mstride] == 0 || mj[r * mstride] == 0)) continue;
    const double dx = static_cast<double>(a) - mx, dy = static_cast<double>(b) - my;
    const double pxx = dx * dx, pyy = dy * dy, pxy = dx * dy;
    kxx.add(pxx);
    kyy.add(pyy);
    kxy.add(pxy);
  }
  const double sxx = kxx.s, syy = kyy.s, sxy = kxy.s;
  if (!(sxx > 0.0 && syy > 0.0) || !__builtin_isfinite(sxx) || !__builtin_isfinite(syy)) return kNaN;
  const double r = (sxy / sqrt(sxx)) / sqrt(syy);
  return __builtin_isfinite(r) ? r : kNaN;
}

// Same algorithm over two contiguous rows (NaN = invalid), 8 values at a time
// so the next block's loads are in flight while the serial compensated chains
// of the current block run.
template <typename X>
__device__ double pair_two_pass_rows(const X* __restrict__ xi, const X* __restrict__ xj, int T) {
#pragma clang fp contract(off)
  Kahan ax{0.0, 0.0}, ay{0.0, 0.0};
  double n = 0.0, mnx = __builtin_inf(), mxx = -__builtin_inf(), mny = __builtin_inf(),
         mxy = -__builtin_inf();
  const int T8 = T & ~7;
  double a[8], b[8];
  for (int t0 = 0; t0 < T8; t0 += 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = xi[t0 + k];
      b[k] = xj[t0 + k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (__builtin_isfinite(a[k]) && __builtin_isfinite(b[k])) {
        n += 1.0;
        ax.add(a[k]);
        ay.add(b[k]);
        mnx = fmin(mnx, a[k]);
        mxx = fmax(mxx, a[k]);
        mny = fmin(mny, b[k]);
        mxy = fmax(mxy, b[k]);
      }
    }
  }
  for (int t = T8; t < T; ++t) {
    const double x = xi[t], y = xj[t];
    if (__builtin_isfinite(x) && __builtin_isfinite(y)) {
      n += 1.0;
      ax.add(x);
      ay.add(y);
      mnx = fmin(mnx, x);
      mxx = fmax(mxx, x);
      mny = fmin(mny, y);
      mxy = fmax(mxy, y);
    }
  }
  if (n < 2.0 || mnx == mxx || mny == mxy) return kNaN;
  const double mx = ax.s / n, my = ay.s / n;
  Kahan kxx{0.0, 0.0}, kyy{0.0, 0.0}, kxy{0.0, 0.0};
  for (int t0 = 0; t0 < T8; t0 += 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[k] = xi[t0 + k];
      b[k] = xj[t0 + k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (__builtin_isfinite(a[k]) && __builtin_isfinite(b[k])) {
        const double dx = a[k] - mx, dy = b[k] - my;
        const double pxx = dx * dx, pyy = dy * dy, pxy = dx * dy;
        kxx.add(pxx);
        kyy.add(pyy);
        kxy.add(pxy);
      }
    }
  }
  for (int t = T8; t < T; ++t) {
    const double x = xi[t], y = xj[t];
    if (__builtin_isfinite(x) && __builtin_isfinite(y)) {
      const double dx = x - mx, dy = y - my;
      const double pxx = dx * dx, pyy = dy * dy, pxy = dx * dy;
      kxx.add(pxx);
      kyy.add(pyy);
      kxy.add(pxy);
    }
  }
  const double sxx = kxx.s, syy = kyy.s, sxy = kxy.s;
  if (!(sxx > 0.0 && syy > 0.0) || !__builtin_isfinite(sxx) || !__builtin_isfinite(syy)) return kNaN;
  const double r = (sxy / sqrt(sxx)) / sqrt(syy);
  return __builtin_isfinite(r) ? r : kNaN;
}

// One thread per flagged cell, streaming the two contiguous original-value
// rows of the transposed plane (NaN marks an invalid cell).
template <typename X>
__global__ __launch_bounds__(256) void stock_fallback_kernel(const X* __restrict__ xt, int Tp,
                                                            int T, int N,
                                                            const uint32_t* __restrict__ fbl,
                                                            const uint32_t* __restrict__ fbc,
                                                            uint32_t fbcap, double* __restrict__ out) {
  const uint32_t nfb = min(*fbc, fbcap);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nfb; q += gridDim.x * blockDim.x) {
    const int i = static_cast<int>(fbl[2 * q]), j = static_cast<int>(fbl[2 * q + 1]);
    const double r = pair_two_pass_rows(xt + static_cast<size_t>(i) * Tp,
                                        xt + static_cast<size_t>(j) * Tp, T);
    out[static_cast<size_t>(i) * N + j] = r;
    out[static_cast<size_t>(j) * N + i] = r;
  }
}

// ---- launchers (stock) -----------------------------------------------------------
hipError_t launch_stock_colprep(const void* X, int value_bytes, const uint8_t* mask, int T, int N,
                                int tier, ColStat* cs, ColPart* parts, const StockPlanes& P,
                                uint32_t* fb_count, hipStream_t s) {
  const unsigned cg = static_cast<unsigned>(P.Np / 64);
  const int nch = (T + kColRows - 1) / kColRows;
  const dim3 block(256);
  if (value_bytes == 8)
    hipLaunchKernelGGL(stock_colpart_kernel<double>, dim3(cg, nch), block, 0, s,
                       static_cast<const double*>(X), mask, T, N, parts);
  else
    hipLaunchKernelGGL(stock_colpart_kernel<float>, dim3(cg, nch), block, 0, s,
                       static_cast<const float*>(X), mask, T, N, parts);
  const dim3 g2((N + 255) / 256);
  if (tier == kTierContract) hipLaunchKernelGGL(stock_colstat_kernel<kTierContract>, g2, block, 0, s, parts, nch, N, cs, fb_count);
  else if (tier == kTierFp32) hipLaunchKernelGGL(stock_colstat_kernel<kTierFp32>, g2, block, 0, s, parts, nch, N, cs, fb_count);
  else hipLaunchKernelGGL(stock_colstat_kernel<kTierBf16x3>, g2, block, 0, s, parts, nch, N, cs, fb_count);
  const dim3 g3(cg, static_cast<unsigned>(P.Tp / 64));
#define FC_PLANES(V, TIER)                                                                        \
  hipLaunchKernelGGL((stock_planes_kernel<V, TIER>), g3, block, 0, s, static_cast<const V*>(X), mask, \
                     T, N, cs, P)
  if (value_bytes == 8) {
    if (tier == kTierContract) FC_PLANES(double, kTierContract);
    else if (tier == kTierFp32) FC_PLANES(double, kTierFp32);
    else FC_PLANES(double, kTierBf16x3);
  } else {
    if (tier == kTierContract) FC_PLANES(float, kTierContract);
    else if (tier == kTierFp32) FC_PLANES(float, kTierFp32);
    else FC_PLANES(float, kTierBf16x3);
  }
#undef FC_PLANES
  return hipGetLastError();
}

size_t stock_colpart_bytes(int T, int N) {
  return static_cast<size_t>((T + kColRows - 1) / kColRows) * N * sizeof(ColPart);
}

hipError_t launch_stock_fullsums(const StockPlanes& P, int N, ColStat* cs, hipStream_t s) {
  hipLaunchKernelGGL(stock_fullsums_kernel, dim3((N + 3) / 4), dim3(256), 0, s, P, N, cs);
  return hipGetLastError();
}

hipError_t launch_stock_tiles(int tier, int full, const StockPlanes& P, const ColStat* cs, int T,
                              int N, double* out, uint32_t* fb_list, uint32_t* fb_count,
                              uint32_t fb_cap, hipStream_t s) {
  const int ts = stock_tile(tier);
  const unsigned nt = static_cast<unsigned>(P.Np / ts);
  const dim3 grid(nt * (nt + 1) / 2);
  const dim3 block(256);
  if (tier == kTierContract) {
    hipLaunchKernelGGL(stock_tile_f64_kernel, grid, block, 0, s, P, cs, T, N, out, fb_list, fb_count,
                       fb_cap);
  } else if (tier == kTierFp32) {
    hipLaunchKernelGGL(stock_tile_fp32_kernel, grid, block, 0, s, P, cs, T, N, out, fb_list,
                       fb_count, fb_cap);
  } else if (full) {
    hipLaunchKernelGGL(stock_tile_bf16x3_kernel<1>, grid, block, 0, s, P, cs, T, N, out, fb_list,
                       fb_count, fb_cap);
  } else {
    const int nq = P.Np / kRJ;
    const int ntiles = nq * (nq + 1) / 2;
    const unsigned nblk = 8u * static_cast<unsigned>((ntiles + 7) / 8);
    hipLaunchKernelGGL(stock_tile_bf16x3_ring_kernel, dim3(nblk), dim3(256), 0, s, P, cs, N, ntiles,
                       out, fb_list, fb_count, fb_cap);
  }
  return hipGetLastError();
}

hipError_t launch_stock_fallback(const StockPlanes& P, int T, int N, const uint32_t* fb_list,
                                 const uint32_t* fb_count, uint32_t fb_cap, double* out,
                                 hipStream_t s) {
  if (P.xt_bytes == 8)
    hipLaunchKernelGGL(stock_fallback_kernel<double>, dim3(256), dim3(64), 0, s,
                       static_cast<const double*>(P.xt), P.Tp, T, N, fb_list, fb_count, fb_cap, out);
  else
    hipLaunchKernelGGL(stock_fallback_kernel<float>, dim3(256), dim3(64), 0, s,
                       static_cast<const float*>(P.xt), P.Tp, T, N, fb_list, fb_count, fb_cap, out);
  return hipGetLastError();
}

// ==== factor_corr ==============================================================
// Rows r = t*N + i of the (T*N, F) panel (factor index fastest); row valid per
// factor = finite && mask[r]. A pilot kernel takes each factor's shift and
// scale from a strided row sample (any shift/positive scale leaves Pearson
// unchanged; they only condition the sums). The main kernel is persistent:
// each workgroup walks row chunks, loads a chunk with coalesced reads,
// stages the transformed operands transposed in LDS ([factor][row]) and runs
// every upper 16x16 factor-block pair on MFMA, accumulating in registers.
// Per-wave partial sums go to scratch; the final kernel adds them in a fixed
// order in FP64 and finalizes; flagged cells get the exact serial two-pass.

constexpr int kFThreads = 512;
constexpr int kFWaves = kFThreads / 64;

struct FPlan {
  int F, Fp, nb, npairs;
  int RC;        // rows per chunk
  int nchunks;
  int maxp;      // pair slots per wave
  int ppass;     // pairs per pass = waves * maxp
  int npass;
  int groups;    // workgroups of the main kernel
};

__host__ __device__ inline void pair_blocks(int p, int nb, int& bi, int& bj) {
  int b = p, i = 0;
  while (b >= nb - i) {
    b -= nb - i;
    ++i;
  }
  bi = i;
  bj = i + b;
}

// wave w of a pass holding npp pairs: slot s -> pair (or -1), k-split
__host__ __device__ inline int wave_pair(int w, int s, int npp, int maxp, int& rsplit, int& kmod) {
  if (npp >= kFWaves) {
    rsplit = 1;
    kmod = 0;
    const int p = w + s * kFWaves;
    return (s < maxp && p < npp) ? p : -1;
  }
  rsplit = kFWaves / npp;
  kmod = w / npp;
  if (s != 0 || w >= npp * rsplit) return -1;
  return w % npp;
}

__host__ __device__ inline int tier_unit(int tier) { return tier == kTierBf16x3 ? 64 : 16; }

static FPlan make_plan(int tier, long long R, int F) {
  FPlan pl;
  pl.F = F;
  pl.nb = (F + 15) / 16;
  pl.Fp = pl.nb * 16;
  pl.npairs = pl.nb * (pl.nb + 1) / 2;
  // rows per chunk: as many 64-row groups as keep <= 16 elements per thread
  // (<= 8 for FP64) and the LDS staging within budget
  const int ept = (tier == kTierContract) ? 8 : 16;
  int rc = (ept * kFThreads / pl.Fp) / 64 * 64;
  rc = rc < 64 ? 64 : (rc > 512 ? 512 : rc);
  pl.RC = rc;
  pl.nchunks = static_cast<int>((R + rc - 1) / rc);
  pl.maxp = (tier == kTierContract) ? 2 : 5;
  pl.ppass = kFWaves * pl.maxp;
  pl.npass = (pl.npairs + pl.ppass - 1) / pl.ppass;
  int g = 512;
  if (g > pl.nchunks) g = pl.nchunks;
  pl.groups = g < 1 ? 1 : g;
  return pl;
}

static size_t f_lds_bytes(int tier, const FPlan& pl) {
  if (tier == kTierBf16x3) return static_cast<size_t>(pl.Fp) * (pl.RC + 8) * 2 * 3 + static_cast<size_t>(pl.Fp) * (pl.RC + 16);
  if (tier == kTierFp32) return static_cast<size_t>(pl.Fp) * (pl.RC + 4) * 4 * 2;
  return static_cast<size_t>(pl.Fp) * (pl.RC + 2) * 8 * 2;
}

// per-wave partials of the general (six-sum) path
static size_t f_partial_count(const FPlan& pl) {
  return static_cast<size_t>(pl.npass) * pl.groups * kFWaves * pl.maxp * 6 * 256;
}

// ---- uniform-mask (Gram) path plan --------------------------------------------
// When every masked cell is finite, the valid rows of every factor pair are the
// masked rows, so the six pair sums reduce to the masked Gram matrix
// G = sum_r m_r z_r z_r^T of the transformed rows, the column sums
// s = sum_r m_r z_r and the count n = sum_r m_r.
constexpr int kGWaves = 4;    // waves per Gram workgroup
using fcd::kBufWord3;

struct GPlan {
  int F, BW, nb, npairs, G, VPL;
  int RI;        // rows per wave iteration = (64 / BW) * VPL
  int nblk;      // workgroups; workgroup b owns rows [b*RB, (b+1)*RB)
  long long RB;  // multiple of RI
  int PB;        // doubles per workgroup partial (Gram blocks, column sums, count)
  int cs_off, cnt_off, am_off;
};

static GPlan make_gplan(int tier, long long R, int F) {
  GPlan g;
  g.F = F;
  if (tier == kTierContract || F <= 16) {
    g.BW = 16;
  } else {
    g.BW = 32;
  }
  g.nb = (F + g.BW - 1) / g.BW;
  g.npairs = g.nb * (g.nb + 1) / 2;
  g.G = (tier == kTierContract && g.nb >= 6) ? 2 : 1;
  if (tier == kTierContract)
    g.VPL = g.nb <= 2 ? 16 : (g.nb <= 4 ? 8 : 4);
  else
    g.VPL = g.nb == 1 ? 16 : 8;
  g.RI = (64 / g.BW) * g.VPL;
  const long long iters = (R + g.RI - 1) / g.RI;
  const int want = (g.nb <= 2) ? 1024 : 512;
  g.nblk = static_cast<int>(iters < want ? (iters < 1 ? 1 : iters) : want);
  long long ipb = (iters + g.nblk - 1) / g.nblk;
  // a workgroup's rows are one buffer resource: under 2^31 bytes of 8-byte values
  while (ipb * g.RI * F * 8 >= (1ll << 31)) {
    g.nblk *= 2;
    ipb = (iters + g.nblk - 1) / g.nblk;
  }
  g.RB = ipb * g.RI;
  g.cs_off = g.npairs * g.BW * g.BW;
  g.cnt_off = g.cs_off + g.nb * g.BW;
  g.am_off = g.cnt_off + 1;  // per-factor max |x| over valid cells (max-reduced)
  g.PB = (g.am_off + g.nb * g.BW + 31) / 32 * 32;
  return g;
}

// Exact-pass splits: flagged pair q is cut into K row ranges, K chosen from
// the flag count so that small counts still fill the chip; items <= max(2048,
// 8 * cells).
constexpr int kFbStride = 16;  // doubles per fallback item record
__host__ __device__ inline int fb_splits(uint32_t nfb) {
  const uint32_t k = nfb ? 2048u / nfb : 128u;
  return static_cast<int>(k < 8u ? 8u : (k > 128u ? 128u : k));
}
static size_t fb_items_cap(size_t cells) { return cells * 8 > 2048 ? cells * 8 : 2048; }

struct FLayout {
  size_t ss, gpart, stage, gen, fbl, fbp, total;
};
static FLayout factor_layout(int tier, long long R, int F) {
  const FPlan pl = make_plan(tier, R, F);
  const GPlan gp = make_gplan(tier, R, F);
  const size_t cells = static_cast<size_t>(F) * (F + 1) / 2;
  auto al = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
  FLayout L;
  size_t o = 256;  // flags: [0] fallback count, [1] non-finite masked cell seen
  L.ss = o; o = al(o + static_cast<size_t>(F) * 16);
  L.gpart = o; o = al(o + static_cast<size_t>(gp.nblk) * gp.PB * 8);
  L.stage = o; o = al(o + static_cast<size_t>(gp.PB) * 8);
  L.gen = o; o = al(o + f_partial_count(pl) * 8);
  L.fbl = o; o = al(o + cells * 8);
  L.fbp = o; o = al(o + fb_items_cap(cells) * kFbStride * 8);
  L.total = o;
  return L;
}

size_t factor_scratch_bytes(int tier, long long R, int F, int* n_groups) {
  if (n_groups) *n_groups = make_gplan(tier, R, F).nblk;
  return factor_layout(tier, R, F).total;
}

template <typename V, int TIER>
__global__ __launch_bounds__(256) void factor_pilot_kernel(FactorArgs a, double* __restrict__ ss,
                                                           uint32_t* __restrict__ flags) {
  __shared__ double rd[8][2];
  const int f = blockIdx.x;
  if (f == 0 && threadIdx.x == 0) {
    flags[0] = 0u;
    flags[1] = 0u;
  }
  // a sample of up to 4096 rows, 16 per thread held in registers so that
  // every load is in flight at once: every row of a panel of at most 4096
  // rows, else 256 runs of 16 consecutive rows spread evenly over the panel
  // (runs keep the sample to 256 address-translation pages; single rows
  // spread as far apart miss the translation cache on nearly every load)
  constexpr int kPer = 16;
  const V* X = static_cast<const V*>(a.x);
  const long long S = a.R < 256 * kPer ? a.R : 256 * kPer;
  const long long gap = a.R / 256;  // >= 16 when R > 4096
  auto row_of = [&](long long k) { return a.R <= 256 * kPer ? k : (k >> 4) * gap + (k & 15); };
  double v[kPer];
  bool ok[kPer];
  V xr[kPer];
  uint8_t mr[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const long long k = threadIdx.x + 256ll * u;
    xr[u] = X[row_of(k < S ? k : 0) * a.F + f];
  }
  if (a.mask != nullptr) {  // uniform: the mask loads stay in flight together
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const long long k = threadIdx.x + 256ll * u;
      mr[u] = a.mask[row_of(k < S ? k : 0)];
    }
  } else {
#pragma unroll
    for (int u = 0; u < kPer; ++u) mr[u] = 1;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const long long k = threadIdx.x + 256ll * u;
    ok[u] = (k < S) && mr[u] != 0 && finite_v(xr[u]);
    v[u] = static_cast<double>(xr[u]);
  }
  double n = 0.0, sum = 0.0;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    n += ok[u] ? 1.0 : 0.0;
    sum += ok[u] ? v[u] : 0.0;
  }
  // fixed-order reductions: xor butterfly per wave, then the 4 waves in order
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    n += __shfl_xor(n, o, 64);
    sum += __shfl_xor(sum, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    rd[wv][0] = n;
    rd[wv][1] = sum;
  }
  __syncthreads();
  const double N_ = (rd[0][0] + rd[1][0]) + (rd[2][0] + rd[3][0]);
  const double mean = N_ > 0.0 ? ((rd[0][1] + rd[1][1]) + (rd[2][1] + rd[3][1])) / N_ : 0.0;
  double q = 0.0;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const double d = v[u] - mean;
    q += ok[u] ? d * d : 0.0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) q += __shfl_xor(q, o, 64);
  if ((threadIdx.x & 63) == 0) rd[4 + wv][0] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double Q = (rd[4][0] + rd[5][0]) + (rd[6][0] + rd[7][0]);
    double scale = 1.0;
    if (N_ > 0.0 && Q > 0.0 && __builtin_isfinite(Q)) {
      const double sd = sqrt(Q / N_);
      scale = (TIER == kTierContract) ? ldexp(1.0, -ilogb(sd)) : 1.0 / sd;
    }
    ss[2 * f] = mean;
    ss[2 * f + 1] = scale;
  }
}

template <int TIER> struct FAcc;
template <> struct FAcc<kTierBf16x3> { typedef f32x4 T; };
template <> struct FAcc<kTierFp32> { typedef f32x4 T; };
template <> struct FAcc<kTierContract> { typedef f64x4 T; };

template <typename V, int TIER, int MAXP>
__global__ __launch_bounds__(kFThreads) void factor_main_kernel(FactorArgs a, const double* __restrict__ ss,
                                                                FPlan pl, int pass, double* __restrict__ partial,
                                                                const uint32_t* __restrict__ flags) {
  extern __shared__ __align__(16) unsigned char smem[];
  // general six-sum path: only when a masked cell is not finite
  if (flags[1] == 0u) return;
  typedef typename FAcc<TIER>::T Acc;
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  const int F = pl.F, Fp = pl.Fp, RC = pl.RC;
  const int p0 = pass * pl.ppass;
  const int npp = min(pl.ppass, pl.npairs - p0);
  int rsplit = 1, kmod = 0;
  int pr[MAXP];
#pragma unroll
  for (int s = 0; s < MAXP; ++s) pr[s] = wave_pair(w, s, npp, MAXP, rsplit, kmod);
  Acc acc[MAXP][6];
  i32x4 cnt8[MAXP];
#pragma unroll
  for (int s = 0; s < MAXP; ++s) {
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[s][q] = Acc{};
    cnt8[s] = i32x4{};
  }
  // LDS staging
  const int RS = (TIER == kTierBf16x3) ? RC + 8 : (TIER == kTierFp32 ? RC + 4 : RC + 2);
  const int RS8 = RC + 16;
  unsigned char* zbase = smem;
  const V* X = static_cast<const V*>(a.x);
  const int unit = tier_unit(TIER);
  const int nunits = RC / unit;
  for (int c = blockIdx.x; c < pl.nchunks; c += gridDim.x) {
    const long long r0 = static_cast<long long>(c) * RC;
    const long long nel = static_cast<long long>(RC) * F;
    for (int e = threadIdx.x; e < nel; e += kFThreads) {
      const int row = e / F;
      const int f = e - row * F;
      const long long r = r0 + row;
      double z = 0.0;
      bool ok = false;
      if (r < a.R) {
        const V v = X[r * F + f];
        ok = finite_v(v) && (a.mask == nullptr || a.mask[r] != 0);
        if (ok) z = (static_cast<double>(v) - ss[2 * f]) * ss[2 * f + 1];
      }
      if constexpr (TIER == kTierBf16x3) {
        uint16_t* hiT = reinterpret_cast<uint16_t*>(zbase);
        uint16_t* loT = hiT + static_cast<size_t>(Fp) * RS;
        uint16_t* mT = loT + static_cast<size_t>(Fp) * RS;
        int8_t* m8T = reinterpret_cast<int8_t*>(mT + static_cast<size_t>(Fp) * RS);
        const float zf = static_cast<float>(z);
        const uint16_t hi = f2bf(zf);
        hiT[f * RS + row] = hi;
        loT[f * RS + row] = f2bf(zf - bf2f(hi));
        mT[f * RS + row] = ok ? static_cast<uint16_t>(0x3f80) : static_cast<uint16_t>(0);
        m8T[f * RS8 + row] = ok ? 1 : 0;
      } else if constexpr (TIER == kTierFp32) {
        float* zT = reinterpret_cast<float*>(zbase);
        float* mT = zT + static_cast<size_t>(Fp) * RS;
        zT[f * RS + row] = static_cast<float>(z);
        mT[f * RS + row] = ok ? 1.0f : 0.0f;
      } else {
        double* zT = reinterpret_cast<double*>(zbase);
        double* mT = zT + static_cast<size_t>(Fp) * RS;
        zT[f * RS + row] = z;
        mT[f * RS + row] = ok ? 1.0 : 0.0;
      }
    }
    // padded factor rows F..Fp-1 stay zero: cleared once per chunk
    for (int e = threadIdx.x; e < (Fp - F) * RC; e += kFThreads) {
      const int f = F + e / RC, row = e % RC;
      if constexpr (TIER == kTierBf16x3) {
        uint16_t* hiT = reinterpret_cast<uint16_t*>(zbase);
        uint16_t* loT = hiT + static_cast<size_t>(Fp) * RS;
        uint16_t* mT = loT + static_cast<size_t>(Fp) * RS;
        int8_t* m8T = reinterpret_cast<int8_t*>(mT + static_cast<size_t>(Fp) * RS);
        hiT[f * RS + row] = 0;
        loT[f * RS + row] = 0;
        mT[f * RS + row] = 0;
        m8T[f * RS8 + row] = 0;
      } else if constexpr (TIER == kTierFp32) {
        float* zT = reinterpret_cast<float*>(zbase);
        zT[f * RS + row] = 0.0f;
        zT[static_cast<size_t>(Fp) * RS + f * RS + row] = 0.0f;
      } else {
        double* zT = reinterpret_cast<double*>(zbase);
        zT[f * RS + row] = 0.0;
        zT[static_cast<size_t>(Fp) * RS + f * RS + row] = 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < MAXP; ++s) {
      if (pr[s] < 0) continue;
      int bi, bj;
      pair_blocks(p0 + pr[s], pl.nb, bi, bj);
      const int fa = bi * 16 + (lane & 15), fb = bj * 16 + (lane & 15);
      for (int u = kmod; u < nunits; u += rsplit) {
        const int k0 = u * unit;
        if constexpr (TIER == kTierBf16x3) {
          const uint16_t* hiT = reinterpret_cast<const uint16_t*>(zbase);
          const uint16_t* loT = hiT + static_cast<size_t>(Fp) * RS;
          const uint16_t* mT = loT + static_cast<size_t>(Fp) * RS;
          const int8_t* m8T = reinterpret_cast<const int8_t*>(mT + static_cast<size_t>(Fp) * RS);
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int kk = k0 + 32 * h + 8 * (lane >> 4);
            const u16x8 ah = *reinterpret_cast<const u16x8*>(hiT + fa * RS + kk);
            const u16x8 al = *reinterpret_cast<const u16x8*>(loT + fa * RS + kk);
            const u16x8 bh = *reinterpret_cast<const u16x8*>(hiT + fb * RS + kk);
            const u16x8 bl = *reinterpret_cast<const u16x8*>(loT + fb * RS + kk);
            const bf16x8 Ah = __builtin_bit_cast(bf16x8, ah), Al = __builtin_bit_cast(bf16x8, al);
            const bf16x8 Bh = __builtin_bit_cast(bf16x8, bh), Bl = __builtin_bit_cast(bf16x8, bl);
            const bf16x8 Am = ld_bf8(mT + fa * RS + kk), Bm = ld_bf8(mT + fb * RS + kk);
            bf16x8 Aqh, Aql, Bqh, Bql;
            square_split(ah, al, Aqh, Aql);
            square_split(bh, bl, Bqh, Bql);
            acc[s][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc[s][0], 0, 0, 0);
            acc[s][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc[s][0], 0, 0, 0);
            acc[s][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc[s][0], 0, 0, 0);
            acc[s][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bm, acc[s][1], 0, 0, 0);
            acc[s][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bm, acc[s][1], 0, 0, 0);
            acc[s][2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bh, acc[s][2], 0, 0, 0);
            acc[s][2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bl, acc[s][2], 0, 0, 0);
            acc[s][3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Aqh, Bm, acc[s][3], 0, 0, 0);
            acc[s][3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Aql, Bm, acc[s][3], 0, 0, 0);
            acc[s][4] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bqh, acc[s][4], 0, 0, 0);
            acc[s][4] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am, Bql, acc[s][4], 0, 0, 0);
          }
          const int k8 = k0 + 16 * (lane >> 4);
          const i32x4 a8 = *reinterpret_cast<const i32x4*>(m8T + fa * RS8 + k8);
          const i32x4 b8 = *reinterpret_cast<const i32x4*>(m8T + fb * RS8 + k8);
          cnt8[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a8, b8, cnt8[s], 0, 0, 0);
        } else if constexpr (TIER == kTierFp32) {
          const float* zT = reinterpret_cast<const float*>(zbase);
          const float* mT = zT + static_cast<size_t>(Fp) * RS;
          const int kk = k0 + 4 * (lane >> 4);
          const f32x4 za = *reinterpret_cast<const f32x4*>(zT + fa * RS + kk);
          const f32x4 ma = *reinterpret_cast<const f32x4*>(mT + fa * RS + kk);
          const f32x4 zb = *reinterpret_cast<const f32x4*>(zT + fb * RS + kk);
          const f32x4 mb = *reinterpret_cast<const f32x4*>(mT + fb * RS + kk);
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            acc[s][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(za[m], zb[m], acc[s][0], 0, 0, 0);
            acc[s][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(za[m], mb[m], acc[s][1], 0, 0, 0);
            acc[s][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ma[m], zb[m], acc[s][2], 0, 0, 0);
            acc[s][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(za[m] * za[m], mb[m], acc[s][3], 0, 0, 0);
            acc[s][4] = __builtin_amdgcn_mfma_f32_16x16x4f32(ma[m], zb[m] * zb[m], acc[s][4], 0, 0, 0);
            acc[s][5] = __builtin_amdgcn_mfma_f32_16x16x4f32(ma[m], mb[m], acc[s][5], 0, 0, 0);
          }
        } else {
          const double* zT = reinterpret_cast<const double*>(zbase);
          const double* mT = zT + static_cast<size_t>(Fp) * RS;
          const int kk = k0 + 4 * (lane >> 4);
          double za[4], ma[4], zb[4], mb[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            za[m] = zT[fa * RS + kk + m];
            ma[m] = mT[fa * RS + kk + m];
            zb[m] = zT[fb * RS + kk + m];
            mb[m] = mT[fb * RS + kk + m];
          }
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m], zb[m], acc[s][0], 0, 0, 0);
            acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m], mb[m], acc[s][1], 0, 0, 0);
            acc[s][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], zb[m], acc[s][2], 0, 0, 0);
            acc[s][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(za[m] * za[m], mb[m], acc[s][3], 0, 0, 0);
            acc[s][4] = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], zb[m] * zb[m], acc[s][4], 0, 0, 0);
            acc[s][5] = __builtin_amdgcn_mfma_f64_16x16x4f64(ma[m], mb[m], acc[s][5], 0, 0, 0);
          }
        }
      }
    }
    __syncthreads();
  }
  // per-wave partials: [pass][group][wave][slot][sum][256]
#pragma unroll
  for (int s = 0; s < MAXP; ++s) {
    const size_t base =
        ((((static_cast<size_t>(pass) * pl.groups + blockIdx.x) * kFWaves + w) * MAXP + s) * 6) * 256;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v = static_cast<double>(acc[s][q][r]);
        if constexpr (TIER == kTierBf16x3) if (q == 5) v = static_cast<double>(cnt8[s][r]);
        partial[base + q * 256 + lane * 4 + r] = v;
      }
    }
  }
}

// ---- diagonal -------------------------------------------------------------------
// A diagonal cell is 1.0 iff the column's self-correlation is finite: at least
// two valid cells and a non-zero variance. A variance small against its terms
// (possibly an exactly constant column) is left to the exact pass.
template <int TIER>
__device__ __forceinline__ Fin finalize_diag(double n, double sx, double sxx) {
  Fin o{kNaN, false};
  if (n < 2.0) return o;
  const double tol = (TIER == kTierContract) ? 1e-10 : (TIER == kTierFp32 ? 1e-5 : 1e-3);
  const double p = sx * sx / n;
  const double vx = sxx - p;
  if (!(vx > tol * (fmax(fabs(sxx), p) + 1e-300)) || !__builtin_isfinite(vx)) {
    o.fb = true;
    return o;
  }
  o.r = 1.0;
  return o;
}

__device__ __forceinline__ void cell_ij(int cell, int F, int& i, int& j) {
  int ii = 0, rem = cell;
  while (rem >= F - ii) {
    rem -= F - ii;
    ++ii;
  }
  i = ii;
  j = ii + rem;
}

__device__ __forceinline__ void push_fallback(uint32_t* flags, uint32_t* fbl, int i, int j) {
  const uint32_t q = atomicAdd(&flags[0], 1u);
  fbl[2 * q] = static_cast<uint32_t>(i);
  fbl[2 * q + 1] = static_cast<uint32_t>(j);
}

// One wave per upper cell (i <= j) of the general path: lanes take the
// (workgroup, wave) partials round-robin, a fixed xor butterfly adds the lane
// sums, lane 0 finalizes; flagged cells are listed for the exact pass.
template <int TIER>
__global__ __launch_bounds__(256) void factor_final_kernel(FPlan pl, const double* __restrict__ ss,
                                                           const double* __restrict__ partial,
                                                           const double* __restrict__ amax,
                                                           double* __restrict__ out, uint32_t* fbl,
                                                           uint32_t* flags) {
  if (flags[1] == 0u) return;
  const int F = pl.F;
  const int cell = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = lane_id();
  const int ncell = F * (F + 1) / 2;
  if (cell >= ncell) return;
  int i, j;
  cell_ij(cell, F, i, j);
  const int bi = i / 16, bj = j / 16;
  // global pair index of (bi, bj)
  int p = 0;
  for (int k = 0; k < bi; ++k) p += pl.nb - k;
  p += bj - bi;
  const int pass = p / pl.ppass, pl_local = p - pass * pl.ppass;
  const int npp = min(pl.ppass, pl.npairs - pass * pl.ppass);
  const int ii = i - bi * 16, jj = j - bj * 16;
  // accumulator slot of (ii, jj): FP32 16x16 output rows are 4*(l/16) + r,
  // FP64 16x16 output rows are 4r + l/16
  const int ln = (TIER == kTierContract) ? 16 * (ii & 3) + jj : 16 * (ii >> 2) + jj;
  const int rr = (TIER == kTierContract) ? (ii >> 2) : (ii & 3);
  // waves holding this pair (see wave_pair): one (wave, slot) when the pass
  // has at least one pair per wave, else waves pl_local + h*npp, h < rsplit
  const int nh = (npp >= kFWaves) ? 1 : kFWaves / npp;
  double sum[6] = {0, 0, 0, 0, 0, 0};
  const int items = pl.groups * nh;
  for (int it = lane; it < items; it += 64) {
    const int g = it / nh, h = it - (it / nh) * nh;
    const int hw = (npp >= kFWaves) ? pl_local % kFWaves : pl_local + h * npp;
    const int hs = (npp >= kFWaves) ? pl_local / kFWaves : 0;
    const size_t base =
        ((((static_cast<size_t>(pass) * pl.groups + g) * kFWaves + hw) * pl.maxp + hs) * 6) * 256;
#pragma unroll
    for (int q = 0; q < 6; ++q) sum[q] += partial[base + q * 256 + ln * 4 + rr];
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sum[q] += __shfl_xor(sum[q], o, 64);
  }
  if (lane != 0) return;
  Fin fi;
  if (i == j) {
    fi = finalize_diag<TIER>(sum[5], sum[1], sum[3]);
  } else {
    ColStat ci, cj;
    ci.mean = ss[2 * i];
    ci.scale = ss[2 * i + 1];
    cj.mean = ss[2 * j];
    cj.scale = ss[2 * j + 1];
    ci.amax = amax[i];
    cj.amax = amax[j];
    ci.minv = cj.minv = -__builtin_inf();  // column extremes are not tracked: the
    ci.maxv = cj.maxv = __builtin_inf();   // exact pass decides constant operands
    fi = finalize_cell<TIER, false>(sum[5], sum[1], sum[2], sum[0], sum[3], sum[4], ci, cj);
  }
  if (fi.fb) {
    push_fallback(flags, fbl, i, j);
    return;
  }
  out[static_cast<size_t>(i) * F + j] = fi.r;
  out[static_cast<size_t>(j) * F + i] = fi.r;
}

// ---- factor_corr: uniform-mask (Gram) path ------------------------------------
// One pass over the row-major (R, F) panel. MFMA operand fragments are loaded
// straight from global memory: lane l of a wave owns factor b*BW + l%BW of
// every factor block b and, in each iteration of RI = RS*VPL rows (RS = 64/BW
// row slots), the rows RS*s + l/BW for s < VPL. Any bijection of rows onto
// MFMA k-slots that A and B share leaves sum_k A[i][k] B[k][j] unchanged, so
// this one map feeds every MFMA shape used:
//   contract  BW=16  v_mfma_f64_16x16x4f64 (one row per lane per MFMA)
//   fp32      BW=16  v_mfma_f32_16x16x4f32, BW=32 v_mfma_f32_32x32x2f32
//   bf16x3    BW=16  v_mfma_f32_16x16x32_bf16, BW=32 v_mfma_f32_32x32x16_bf16
//             (8 rows per lane per MFMA; z = hi + lo, hi*hi + hi*lo + lo*hi)
// The waves of a workgroup split its row iterations (and, with G = 2, the
// factor-block pairs) and are combined through LDS in FP64 in a fixed order;
// workgroup partials are then reduced in a fixed order. A masked
// non-finite cell raises flags[1], and the general six-sum path takes over.

template <int I, int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
  if constexpr (I < N) {
    fn(std::integral_constant<int, I>{});
    static_for<I + 1, N>(fn);
  }
}

// Balanced split of the row blocks into G pair groups: row blocks taken in
// the order 0, nb-1, 1, nb-2, ... join group 0 while it holds under half of
// the pairs.
__host__ __device__ constexpr int rb_group(int NB, int G, int bi) {
  if (G == 1) return 0;
  const int total = NB * (NB + 1) / 2;
  int cum = 0;
  for (int k = 0; k < NB; ++k) {
    const int b = (k % 2 == 0) ? k / 2 : NB - 1 - k / 2;
    const int grp = (2 * cum < total) ? 0 : 1;
    if (b == bi) return grp;
    cum += NB - b;
  }
  return 0;
}
__host__ __device__ constexpr int slot_of(int NB, int G, int bi, int bj) {
  int s = 0;
  for (int b = 0; b < bi; ++b)
    if (rb_group(NB, G, b) == rb_group(NB, G, bi)) s += NB - b;
  return s + (bj - bi);
}
__host__ __device__ constexpr int group_slots(int NB, int G, int grp) {
  int c = 0;
  for (int b = 0; b < NB; ++b)
    if (rb_group(NB, G, b) == grp) c += NB - b;
  return c;
}
__host__ __device__ constexpr int max_slots(int NB, int G) {
  int m = 0;
  for (int grp = 0; grp < G; ++grp) m = group_slots(NB, G, grp) > m ? group_slots(NB, G, grp) : m;
  return m;
}
__host__ __device__ constexpr int pair_index(int NB, int bi, int bj) {
  return bi * NB - bi * (bi - 1) / 2 + (bj - bi);
}
// pair index held in slot p of group grp, -1 if none
__host__ __device__ constexpr int slot_pair(int NB, int G, int grp, int p) {
  for (int bi = 0; bi < NB; ++bi)
    for (int bj = bi; bj < NB; ++bj)
      if (rb_group(NB, G, bi) == grp && slot_of(NB, G, bi, bj) == p) return pair_index(NB, bi, bj);
  return -1;
}

// MFMA shape per tier and block width: accumulator type, operand type, and
// the (row, col) of accumulator register r of lane l.
template <int TIER, int BW> struct GramMfma;
template <> struct GramMfma<kTierContract, 16> {
  typedef f64x4 Acc;
  typedef double Frag;
  static constexpr int kRegs = 4;
  __device__ static int row(int l, int r) { return 4 * r + (l >> 4); }
  __device__ static int col(int l, int) { return l & 15; }
  __device__ static Acc mma(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct GramMfma<kTierFp32, 16> {
  typedef f32x4 Acc;
  typedef float Frag;
  static constexpr int kRegs = 4;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static int col(int l, int) { return l & 15; }
  __device__ static Acc mma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <> struct GramMfma<kTierFp32, 32> {
  typedef f32x16 Acc;
  typedef float Frag;
  static constexpr int kRegs = 16;
  __device__ static int row(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
  __device__ static int col(int l, int) { return l & 31; }
  __device__ static Acc mma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
};
template <> struct GramMfma<kTierBf16x3, 16> {
  typedef f32x4 Acc;
  typedef float Frag;
  static constexpr int kRegs = 4;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
  __device__ static int col(int l, int) { return l & 15; }
  __device__ static Acc mma(bf16x8 a, bf16x8 b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <> struct GramMfma<kTierBf16x3, 32> {
  typedef f32x16 Acc;
  typedef float Frag;
  static constexpr int kRegs = 16;
  __device__ static int row(int l, int r) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }
  __device__ static int col(int l, int) { return l & 31; }
  __device__ static Acc mma(bf16x8 a, bf16x8 b, Acc c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

template <typename V, int TIER, int BW, int NB, int G, int VPL>
__global__ __launch_bounds__(kGWaves * 64) void factor_gram_kernel(FactorArgs a, const double* __restrict__ ss,
                                                                   GPlan gp, double* __restrict__ part,
                                                                   uint32_t* __restrict__ flags) {
  typedef GramMfma<TIER, BW> M;
  typedef typename M::Acc Acc;
  typedef typename M::Frag Frag;
  // transform arithmetic: FP64 in the contract tier (exact power-of-two scale,
  // shift rounded at 2^-53) and for float64 input, FP32 otherwise
  typedef typename std::conditional<(TIER == kTierContract) || (sizeof(V) == 8), double, float>::type W;
  constexpr int RS = 64 / BW;
  constexpr int RI = RS * VPL;
  constexpr int MW = (RI + 63) / 64;  // mask words (one row per lane) per iteration
  constexpr int MAXP = max_slots(NB, G);
  constexpr int NRW = kGWaves / G;
  static_assert(TIER != kTierBf16x3 || VPL % 8 == 0, "bf16 MFMA takes 8 rows per lane");
  __shared__ double tl[kGWaves * BW * BW];
  __shared__ double tcs[kGWaves * NB * BW];
  __shared__ double tam[kGWaves * NB * BW];
  __shared__ double tcnt[kGWaves];
  const int lane = lane_id();
  const int w = fcd::wave_id();  // wave-uniform: the row pointers below stay scalar
  const int g = w % G, rw = w / G;
  const int fl = lane % BW, q = lane / BW;
  const int F = a.F;
  W sh[NB], sc[NB];
  bool fin[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int f = b * BW + fl;
    fin[b] = f < F;
    sh[b] = static_cast<W>(fin[b] ? ss[2 * f] : 0.0);
    sc[b] = static_cast<W>(fin[b] ? ss[2 * f + 1] : 0.0);
  }
  Acc acc[MAXP];
#pragma unroll
  for (int p = 0; p < MAXP; ++p) acc[p] = Acc{};
  // FP64 contract tier on float input: the power-of-two column scales are
  // applied to the finished sums instead of to every item (exact either way:
  // the unscaled terms and sums stay normal in FP64), the max |x| is taken in
  // float (exact), and a non-finite masked item shows in the column sum
  constexpr bool kLateScale = TIER == kTierContract && sizeof(V) == 4;
  W csum[NB], amax[NB];
  float amf[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    csum[b] = W(0);
    amax[b] = W(0);
    amf[b] = 0.0f;
  }
  uint32_t cnt = 0;  // wave-uniform: masked rows seen by this wave
  bool bad = false;
  const long long rbeg = static_cast<long long>(blockIdx.x) * gp.RB;
  const long long rend = min(a.R, rbeg + gp.RB);
  const V* X = static_cast<const V*>(a.x);

  // The rows [r0, rend) of the panel and of the mask as buffer resources
  // whose base advances with the wave's iteration (scalar arithmetic), so a
  // lane's offsets stay fixed: 32-bit byte offsets from row r0 (the plan
  // keeps a workgroup's rows under 2^31 bytes), and a load at or past rend
  // returns 0 by the hardware range check (which covers the lane offset;
  // the scalar offset is left 0). Such rows count as unmasked, so their
  // values never enter a sum.
  const uint32_t vb = static_cast<uint32_t>(sizeof(V));
  auto xres = [&](long long r0) {
    const long long nr = rend > r0 ? rend - r0 : 0;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<V*>(X + r0 * F), 0, static_cast<int>(nr * F * vb),
                                             kBufWord3);
  };
  auto mres = [&](long long r0) {
    const long long nr = rend > r0 ? rend - r0 : 0;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(a.mask != nullptr ? a.mask + r0 : a.mask), 0,
                                             a.mask != nullptr ? static_cast<int>(nr) : 0, kBufWord3);
  };
  // lane offsets of the values of rows RS*s + q, factor b*BW + fl; a lane
  // past the last factor gets an offset beyond every buffer (the range
  // check reads it 0), and its shift and scale are 0, so its items are +0
  // with no test
  uint32_t xoff[NB][VPL];
#pragma unroll
  for (int s2 = 0; s2 < VPL; ++s2) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      xoff[b][s2] = fin[b] ? (static_cast<uint32_t>(RS * s2 + q) * static_cast<uint32_t>(F) +
                              static_cast<uint32_t>(b * BW + fl)) * vb
                           : 0xFFFFFF00u;
  }
  // Issue every load of the iteration at r0: the mask byte of row r0 + lane
  // (lanes < RI) and the values of rows r0 + RS*s + q.
  auto load = [&](long long r0, V (&xv)[NB][VPL], uint32_t (&mv)[MW]) {
    const __amdgpu_buffer_rsrc_t xbuf = xres(r0);
#pragma unroll
    for (int u = 0; u < MW; ++u) {
      const int o = 64 * u + lane;
      uint32_t m = 0u;
      if (a.mask == nullptr) {
        m = ((o < RI) && (r0 + o < rend)) ? 1u : 0u;
      } else if (o < RI) {
        const __amdgpu_buffer_rsrc_t mbuf = mres(r0);
        m = static_cast<uint32_t>(__builtin_amdgcn_raw_buffer_load_b8(mbuf, static_cast<uint32_t>(o), 0, 0));
      }
      mv[u] = m;
    }
#pragma unroll
    for (int s2 = 0; s2 < VPL; ++s2) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if constexpr (sizeof(V) == 8)
          xv[b][s2] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(xbuf, xoff[b][s2], 0, 0));
        else
          xv[b][s2] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(xbuf, xoff[b][s2], 0, 0));
      }
    }
  };

  auto compute = [&](const V (&xv)[NB][VPL], const uint32_t (&mv)[MW]) {
    uint64_t mb[MW];
#pragma unroll
    for (int u = 0; u < MW; ++u) {
      mb[u] = fcd::ballot(mv[u] != 0u);
      cnt += static_cast<uint32_t>(__popcll(mb[u]));
    }
    // bit RS*s2 + q of the ballot: this lane's rows, shifted down by q once
    // (q < RS, so the bit stays in word (RS*s2) >> 6)
    uint64_t sel[MW];
#pragma unroll
    for (int u = 0; u < MW; ++u) sel[u] = mb[u] >> q;
    if constexpr (NB == 1) {
#pragma unroll
      for (int u = 0; u < MW; ++u) sel[u] &= fin[0] ? ~0ull : 0ull;
    }
    Frag zf[NB][VPL];
#pragma unroll
    for (int s2 = 0; s2 < VPL; ++s2) {
      if constexpr (TIER != kTierContract) {
        // the row's mask bit as 0 / all-ones, applied to z's bits: an item
        // outside the mask becomes +0. A non-finite item inside it makes z,
        // and so the column sum, non-finite; that sum raises the flag.
        const int kb = RS * s2;  // (s2 unrolled: constant)
        const uint32_t half = static_cast<uint32_t>(sel[kb >> 6] >> (kb & 32));
        const uint32_t mm = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(half), kb & 31, 1));
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const uint32_t mk = mm;
          const W z0 = (static_cast<W>(xv[b][s2]) - sh[b]) * sc[b];
          W z;
          if constexpr (sizeof(W) == 8)
            z = __builtin_bit_cast(W, __builtin_bit_cast(uint64_t, z0) & ((static_cast<uint64_t>(mk) << 32) | mk));
          else
            z = __builtin_bit_cast(W, __builtin_bit_cast(uint32_t, z0) & mk);
          csum[b] += z;
          zf[b][s2] = static_cast<Frag>(z);
        }
      } else if constexpr (kLateScale) {
        const bool m = (sel[(RS * s2) >> 6] >> ((RS * s2) & 63)) & 1ull;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const bool v = m;
          const float x = xv[b][s2];
          amf[b] = fmaxf(amf[b], v ? fabsf(x) : 0.0f);
          const double z = v ? static_cast<double>(x) - sh[b] : 0.0;
          csum[b] += z;
          zf[b][s2] = z;
        }
      } else {
        const bool m = (sel[(RS * s2) >> 6] >> ((RS * s2) & 63)) & 1ull;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const bool v = m && fin[b];
          const W x = static_cast<W>(xv[b][s2]);
          const bool okv = v && __builtin_isfinite(x);
          bad |= v && !okv;
          const W z = okv ? (x - sh[b]) * sc[b] : W(0);
          csum[b] += z;
          amax[b] = okv ? fmax(amax[b], fabs(x)) : amax[b];
          zf[b][s2] = static_cast<Frag>(z);
        }
      }
    }
    if constexpr (TIER == kTierBf16x3) {
#pragma unroll
      for (int t = 0; t < VPL / 8; ++t) {
        bf16x8 hi[NB], lo[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
          u32x4 hw, lw;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float z0 = zf[b][8 * t + 2 * e], z1 = zf[b][8 * t + 2 * e + 1];
            hw[e] = cvt_pk_bf16(z0, z1);
            lw[e] = cvt_pk_bf16(z0 - bf16_lo_f(hw[e]), z1 - bf16_hi_f(hw[e]));
          }
          hi[b] = __builtin_bit_cast(bf16x8, hw);
          lo[b] = __builtin_bit_cast(bf16x8, lw);
        }
        static_for<0, NB>([&](auto BI) {
          constexpr int bi = decltype(BI)::value;
          if (rb_group(NB, G, bi) != g) return;
          static_for<bi, NB>([&](auto BJ) {
            constexpr int bj = decltype(BJ)::value;
            constexpr int p = slot_of(NB, G, bi, bj);
            acc[p] = M::mma(hi[bi], hi[bj], acc[p]);
            acc[p] = M::mma(hi[bi], lo[bj], acc[p]);
            acc[p] = M::mma(lo[bi], hi[bj], acc[p]);
          });
        });
      }
    } else {
#pragma unroll
      for (int s2 = 0; s2 < VPL; ++s2) {
        static_for<0, NB>([&](auto BI) {
          constexpr int bi = decltype(BI)::value;
          if (rb_group(NB, G, bi) != g) return;
          static_for<bi, NB>([&](auto BJ) {
            constexpr int bj = decltype(BJ)::value;
            constexpr int p = slot_of(NB, G, bi, bj);
            acc[p] = M::mma(zf[bi][s2], zf[bj][s2], acc[p]);
          });
        });
      }
    }
  };

  // software pipeline: the next iteration's loads are in flight while the
  // current one is transformed and contracted
  const long long step = static_cast<long long>(NRW) * RI;
  long long r0 = rbeg + static_cast<long long>(rw) * RI;
  V xa[NB][VPL];
  uint32_t ma[MW];
  if (r0 < rend) load(r0, xa, ma);
  if constexpr (NB == 1 && TIER != kTierContract) {
    // two register sets in turn (no copies between iterations)
    V xb[NB][VPL];
    uint32_t mbv[MW];
    for (; r0 < rend; r0 += 2 * step) {
      load(r0 + step, xb, mbv);
      // keep the next tile's loads issued ahead of this tile's work (the
      // scheduler would otherwise sink them below it and expose their latency)
      __builtin_amdgcn_sched_barrier(0);
      compute(xa, ma);
      if (r0 + step >= rend) break;
      load(r0 + 2 * step, xa, ma);
      __builtin_amdgcn_sched_barrier(0);
      compute(xb, mbv);
    }
  } else {
    for (; r0 < rend; r0 += step) {
      V xb[NB][VPL];
      uint32_t mbv[MW];
      load(r0 + step, xb, mbv);
      __builtin_amdgcn_sched_barrier(0);
      compute(xa, ma);
#pragma unroll
      for (int u = 0; u < MW; ++u) ma[u] = mbv[u];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int s2 = 0; s2 < VPL; ++s2) xa[b][s2] = xb[b][s2];
      }
    }
  }
  if constexpr (TIER != kTierContract || kLateScale) {
#pragma unroll
    for (int b = 0; b < NB; ++b) bad |= !__builtin_isfinite(csum[b]);
  }
  if constexpr (kLateScale) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      csum[b] *= sc[b];
      amax[b] = static_cast<W>(amf[b]);
    }
  }
  // (kLateScale) the scale of accumulator element (row, col) of this wave's
  // pair slot p: column factor b*BW + fl is this lane's, the row factor's
  // comes from ss
  auto acc_scale = [&](int p, int r) {
    int bi = 0, bj = 0;
#pragma unroll
    for (int x = 0; x < NB; ++x) {
#pragma unroll
      for (int y = x; y < NB; ++y) {
        if (rb_group(NB, G, x) == g && slot_of(NB, G, x, y) == p) {
          bi = x;
          bj = y;
        }
      }
    }
    const int fr = bi * BW + M::row(lane, r), fc = bj * BW + M::col(lane, r);
    return (fr < F ? ss[2 * fr + 1] : 0.0) * (fc < F ? ss[2 * fc + 1] : 0.0);
  };
  if (fcd::ballot(bad) != 0ull && lane == 0) atomicOr(&flags[1], 1u);
  // column sums and count over the RS lanes that share a factor (fixed butterfly)
  double csd[NB], amd[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    csd[b] = static_cast<double>(csum[b]);
    amd[b] = static_cast<double>(amax[b]);
  }
#pragma unroll
  for (int o = BW; o < 64; o <<= 1) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      csd[b] += __shfl_xor(csd[b], o, 64);
      amd[b] = fmax(amd[b], __shfl_xor(amd[b], o, 64));
    }
  }
  if (g == 0 && q == 0) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      tcs[rw * NB * BW + b * BW + fl] = csd[b];
      tam[rw * NB * BW + b * BW + fl] = amd[b];
    }
    if (fl == 0) tcnt[rw] = static_cast<double>(cnt);
  }
  double* pb = part + static_cast<size_t>(blockIdx.x) * gp.PB;
  static_for<0, MAXP>([&](auto P) {
    constexpr int p = decltype(P)::value;
#pragma unroll
    for (int r = 0; r < M::kRegs; ++r)
      tl[w * BW * BW + M::row(lane, r) * BW + M::col(lane, r)] =
          kLateScale ? static_cast<double>(acc[p][r]) * acc_scale(p, r) : static_cast<double>(acc[p][r]);
    __syncthreads();
    static_for<0, G>([&](auto GG) {
      constexpr int gg = decltype(GG)::value;
      constexpr int pid = slot_pair(NB, G, gg, p);
      if constexpr (pid >= 0) {
        for (int e = threadIdx.x; e < BW * BW; e += kGWaves * 64) {
          double sum = 0.0;
#pragma unroll
          for (int ww = gg; ww < kGWaves; ww += G) sum += tl[ww * BW * BW + e];
          pb[pid * BW * BW + e] = sum;
        }
      }
    });
    __syncthreads();
  });
  for (int e = threadIdx.x; e < NB * BW; e += kGWaves * 64) {
    double sum = 0.0, mx = 0.0;
#pragma unroll
    for (int r = 0; r < NRW; ++r) {
      sum += tcs[r * NB * BW + e];
      mx = fmax(mx, tam[r * NB * BW + e]);
    }
    pb[gp.cs_off + e] = sum;
    pb[gp.am_off + e] = mx;
  }
  if (threadIdx.x == 0) {
    double c = 0.0;
#pragma unroll
    for (int r = 0; r < NRW; ++r) c += tcnt[r];
    pb[gp.cnt_off] = c;
  }
}

// Fixed-order partial reduce, one wave per element: lane l adds the partials
// of workgroups l, l+64, ... in order, then a fixed xor butterfly.
// Elements from am_off on are maxima and are max-reduced.
__global__ __launch_bounds__(256) void factor_gram_reduce(const double* __restrict__ part, int nblk, int PB,
                                                          int am_off, double* __restrict__ red) {
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = lane_id();
  if (e >= PB) return;
  const bool mx = e >= am_off;
  double s = 0.0;
#pragma unroll 4
  for (int b = lane; b < nblk; b += 64) {
    const double v = part[static_cast<size_t>(b) * PB + e];
    s = mx ? fmax(s, v) : s + v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(s, o, 64);
    s = mx ? fmax(s, v) : s + v;
  }
  if (lane == 0) red[e] = s;
}

// One thread per upper cell: reduced sums, finalize (reference gates),
// both halves written; flagged cells listed for the exact pass. Writes
// nothing when a masked non-finite cell sent the call to the general path.
template <int TIER>
__global__ __launch_bounds__(256) void factor_gram_final(GPlan gp, const double* __restrict__ ss,
                                                         const double* __restrict__ stage,
                                                         double* __restrict__ out, uint32_t* fbl,
                                                         uint32_t* flags) {
  // the path flag is read with the sums and tested only before the writes,
  // so this kernel waits out one memory round trip, not two
  const uint32_t general = flags[1];
  const int F = gp.F;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  const int ncell = F * (F + 1) / 2;
  if (cell >= ncell) return;
  int i, j;
  cell_ij(cell, F, i, j);
  auto S = [&](int e) { return stage[e]; };
  auto gram = [&](int x, int y) {  // x <= y
    const int bx = x / gp.BW, by = y / gp.BW;
    const int p = bx * gp.nb - bx * (bx - 1) / 2 + (by - bx);
    return S(p * gp.BW * gp.BW + (x % gp.BW) * gp.BW + (y % gp.BW));
  };
  const double n = S(gp.cnt_off);
  const double sx = S(gp.cs_off + i), sxx = gram(i, i);
  Fin fi;
  if (i == j) {
    fi = finalize_diag<TIER>(n, sx, sxx);
  } else {
    const double sy = S(gp.cs_off + j), syy = gram(j, j), sxy = gram(i, j);
    ColStat ci, cj;
    ci.mean = ss[2 * i];
    ci.scale = ss[2 * i + 1];
    cj.mean = ss[2 * j];
    cj.scale = ss[2 * j + 1];
    ci.amax = S(gp.am_off + i);
    cj.amax = S(gp.am_off + j);
    ci.minv = cj.minv = -__builtin_inf();  // column extremes are not tracked: the
    ci.maxv = cj.maxv = __builtin_inf();   // exact pass decides constant operands
    fi = finalize_cell<TIER, false>(n, sx, sy, sxy, sxx, syy, ci, cj);
  }
  if (general != 0u) return;  // a masked non-finite cell: the general path decides
  if (fi.fb) {
    push_fallback(flags, fbl, i, j);
    return;
  }
  out[static_cast<size_t>(i) * F + j] = fi.r;
  out[static_cast<size_t>(j) * F + i] = fi.r;
}

// ---- exact pass for flagged factor pairs ------------------------------------
// Over the rows valid for both factors: double-double (error-free TwoSum)
// sums, so each mean is the correctly rounded sum / n, then compensated
// centred sums, then r in safe_pearson order. Flagged pair q is split into
// K = fb_splits(nfb) row ranges (items); per-item records are merged in split
// order, so the result does not depend on the list order.
struct DD {
  double hi, lo;
};
__device__ __forceinline__ void dd_add(DD& a, double x) {
#pragma clang fp contract(off)
  const double s = a.hi + x;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (x - bb);
  const double lo = a.lo + e;
  const double h = s + lo;
  a.lo = lo - (h - s);
  a.hi = h;
}
__device__ __forceinline__ void dd_merge(DD& a, const DD& b) {
  dd_add(a, b.hi);
  dd_add(a, b.lo);
}

// record of item it: [0] n, [1,2] sum x, [3,4] sum y, [5..8] min/max x, min/max y,
// [9..14] centred sums xx, yy, xy (hi, lo)
template <typename V>
__global__ __launch_bounds__(256) void factor_fb_sums(FactorArgs a, const uint32_t* __restrict__ fbl,
                                                      const uint32_t* __restrict__ flags,
                                                      double* __restrict__ fbp) {
#pragma clang fp contract(off)
  __shared__ DD sh[256][2];
  __shared__ double mm[256][5];
  const uint32_t nfb = flags[0];
  if (nfb == 0u) return;
  const int K = fb_splits(nfb);
  const long long items = static_cast<long long>(nfb) * K;
  const int F = a.F;
  const V* X = static_cast<const V*>(a.x);
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const long long qd = it / K;
    const int k = static_cast<int>(it - qd * K);
    const int i = static_cast<int>(fbl[2 * qd]), j = static_cast<int>(fbl[2 * qd + 1]);
    const long long r0 = a.R * k / K, r1 = a.R * (k + 1) / K;
    DD sx{0.0, 0.0}, sy{0.0, 0.0};
    double n = 0.0, mnx = __builtin_inf(), mxx = -__builtin_inf(), mny = __builtin_inf(),
           mxy = -__builtin_inf();
    for (long long r = r0 + threadIdx.x; r < r1; r += 256) {
      const V vx = X[r * F + i], vy = X[r * F + j];
      if (!finite_v(vx) || !finite_v(vy) || (a.mask != nullptr && a.mask[r] == 0)) continue;
      const double x = static_cast<double>(vx), y = static_cast<double>(vy);
      n += 1.0;
      dd_add(sx, x);
      dd_add(sy, y);
      mnx = fmin(mnx, x);
      mxx = fmax(mxx, x);
      mny = fmin(mny, y);
      mxy = fmax(mxy, y);
    }
    sh[threadIdx.x][0] = sx;
    sh[threadIdx.x][1] = sy;
    mm[threadIdx.x][0] = n;
    mm[threadIdx.x][1] = mnx;
    mm[threadIdx.x][2] = mxx;
    mm[threadIdx.x][3] = mny;
    mm[threadIdx.x][4] = mxy;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        dd_merge(sh[threadIdx.x][0], sh[threadIdx.x + o][0]);
        dd_merge(sh[threadIdx.x][1], sh[threadIdx.x + o][1]);
        mm[threadIdx.x][0] += mm[threadIdx.x + o][0];
        mm[threadIdx.x][1] = fmin(mm[threadIdx.x][1], mm[threadIdx.x + o][1]);
        mm[threadIdx.x][2] = fmax(mm[threadIdx.x][2], mm[threadIdx.x + o][2]);
        mm[threadIdx.x][3] = fmin(mm[threadIdx.x][3], mm[threadIdx.x + o][3]);
        mm[threadIdx.x][4] = fmax(mm[threadIdx.x][4], mm[threadIdx.x + o][4]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double* rec = fbp + it * kFbStride;
      rec[0] = mm[0][0];
      rec[1] = sh[0][0].hi;
      rec[2] = sh[0][0].lo;
      rec[3] = sh[0][1].hi;
      rec[4] = sh[0][1].lo;
      rec[5] = mm[0][1];
      rec[6] = mm[0][2];
      rec[7] = mm[0][3];
      rec[8] = mm[0][4];
    }
    __syncthreads();
  }
}

// merged first-pass record of pair q: n, means, degenerate flag
__device__ __forceinline__ void fb_merge_first(const double* __restrict__ fbp, long long q, int K,
                                               double& n, double& mx, double& my, bool& degenerate) {
#pragma clang fp contract(off)
  DD sx{0.0, 0.0}, sy{0.0, 0.0};
  double mnx = __builtin_inf(), mxx = -__builtin_inf(), mny = __builtin_inf(), mxy = -__builtin_inf();
  n = 0.0;
  for (int k = 0; k < K; ++k) {
    const double* rec = fbp + (q * K + k) * kFbStride;
    n += rec[0];
    dd_merge(sx, DD{rec[1], rec[2]});
    dd_merge(sy, DD{rec[3], rec[4]});
    mnx = fmin(mnx, rec[5]);
    mxx = fmax(mxx, rec[6]);
    mny = fmin(mny, rec[7]);
    mxy = fmax(mxy, rec[8]);
  }
  degenerate = (n < 2.0) || (mnx == mxx) || (mny == mxy);
  mx = (sx.hi + sx.lo) / n;
  my = (sy.hi + sy.lo) / n;
}

template <typename V>
__global__ __launch_bounds__(256) void factor_fb_centred(FactorArgs a, const uint32_t* __restrict__ fbl,
                                                         const uint32_t* __restrict__ flags,
                                                         double* __restrict__ fbp) {
#pragma clang fp contract(off)
  __shared__ DD sh[256][3];
  __shared__ double bc[3];
  const uint32_t nfb = flags[0];
  if (nfb == 0u) return;
  const int K = fb_splits(nfb);
  const long long items = static_cast<long long>(nfb) * K;
  const int F = a.F;
  const V* X = static_cast<const V*>(a.x);
  for (long long it = blockIdx.x; it < items; it += gridDim.x) {
    const long long qd = it / K;
    const int k = static_cast<int>(it - qd * K);
    const int i = static_cast<int>(fbl[2 * qd]), j = static_cast<int>(fbl[2 * qd + 1]);
    if (threadIdx.x == 0) {
      double n, mx, my;
      bool dg;
      fb_merge_first(fbp, qd, K, n, mx, my, dg);
      bc[0] = mx;
      bc[1] = my;
      bc[2] = dg ? 1.0 : 0.0;
    }
    __syncthreads();
    const double mx = bc[0], my = bc[1];
    const bool degenerate = bc[2] != 0.0;
    const long long r0 = a.R * k / K, r1 = a.R * (k + 1) / K;
    DD cxx{0.0, 0.0}, cyy{0.0, 0.0}, cxy{0.0, 0.0};
    if (!degenerate) {
      for (long long r = r0 + threadIdx.x; r < r1; r += 256) {
        const V vx = X[r * F + i], vy = X[r * F + j];
        if (!finite_v(vx) || !finite_v(vy) || (a.mask != nullptr && a.mask[r] == 0)) continue;
        const double dx = static_cast<double>(vx) - mx, dy = static_cast<double>(vy) - my;
        const double pxx = dx * dx, pyy = dy * dy, pxy = dx * dy;
        dd_add(cxx, pxx);
        dd_add(cyy, pyy);
        dd_add(cxy, pxy);
      }
    }
    sh[threadIdx.x][0] = cxx;
    sh[threadIdx.x][1] = cyy;
    sh[threadIdx.x][2] = cxy;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        dd_merge(sh[threadIdx.x][0], sh[threadIdx.x + o][0]);
        dd_merge(sh[threadIdx.x][1], sh[threadIdx.x + o][1]);
        dd_merge(sh[threadIdx.x][2], sh[threadIdx.x + o][2]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double* rec = fbp + it * kFbStride;
      rec[9] = sh[0][0].hi;
      rec[10] = sh[0][0].lo;
      rec[11] = sh[0][1].hi;
      rec[12] = sh[0][1].lo;
      rec[13] = sh[0][2].hi;
      rec[14] = sh[0][2].lo;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void factor_fb_final(int F, const uint32_t* __restrict__ fbl,
                                                       const uint32_t* __restrict__ flags,
                                                       const double* __restrict__ fbp,
                                                       double* __restrict__ out) {
#pragma clang fp contract(off)
  const uint32_t nfb = flags[0];
  const long long qd = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (qd >= nfb) return;
  const int K = fb_splits(nfb);
  const int i = static_cast<int>(fbl[2 * qd]), j = static_cast<int>(fbl[2 * qd + 1]);
  double n, mx, my;
  bool degenerate;
  fb_merge_first(fbp, qd, K, n, mx, my, degenerate);
  double r = kNaN;
  if (!degenerate) {
    DD cxx{0.0, 0.0}, cyy{0.0, 0.0}, cxy{0.0, 0.0};
    for (int k = 0; k < K; ++k) {
      const double* rec = fbp + (qd * K + k) * kFbStride;
      dd_merge(cxx, DD{rec[9], rec[10]});
      dd_merge(cyy, DD{rec[11], rec[12]});
      dd_merge(cxy, DD{rec[13], rec[14]});
    }
    const double sxx = cxx.hi + cxx.lo, syy = cyy.hi + cyy.lo, sxy = cxy.hi + cxy.lo;
    if (sxx > 0.0 && syy > 0.0 && __builtin_isfinite(sxx) && __builtin_isfinite(syy)) {
      r = (sxy / sqrt(sxx)) / sqrt(syy);
      if (!__builtin_isfinite(r)) r = kNaN;
    }
  }
  if (i == j) r = __builtin_isfinite(r) ? 1.0 : kNaN;
  out[static_cast<size_t>(i) * F + j] = r;
  out[static_cast<size_t>(j) * F + i] = r;
}

template <typename V, int TIER, int BW, int NB, int G, int VPL>
static hipError_t gram_launch(const GPlan& gp, const FactorArgs& a, const double* ss, double* part,
                              uint32_t* flags, hipStream_t s) {
  hipLaunchKernelGGL((factor_gram_kernel<V, TIER, BW, NB, G, VPL>), dim3(gp.nblk), dim3(kGWaves * 64), 0,
                     s, a, ss, gp, part, flags);
  return hipGetLastError();
}

// (BW, nb, G, VPL) instantiations; make_gplan picks one of these
template <typename V, int TIER>
static hipError_t gram_dispatch(const GPlan& gp, const FactorArgs& a, const double* ss, double* part,
                                uint32_t* flags, hipStream_t s) {
  if constexpr (TIER == kTierContract) {
    switch (gp.nb) {
      case 1: return gram_launch<V, TIER, 16, 1, 1, 16>(gp, a, ss, part, flags, s);
      case 2: return gram_launch<V, TIER, 16, 2, 1, 16>(gp, a, ss, part, flags, s);
      case 3: return gram_launch<V, TIER, 16, 3, 1, 8>(gp, a, ss, part, flags, s);
      case 4: return gram_launch<V, TIER, 16, 4, 1, 8>(gp, a, ss, part, flags, s);
      case 5: return gram_launch<V, TIER, 16, 5, 1, 4>(gp, a, ss, part, flags, s);
      case 6: return gram_launch<V, TIER, 16, 6, 2, 4>(gp, a, ss, part, flags, s);
      case 7: return gram_launch<V, TIER, 16, 7, 2, 4>(gp, a, ss, part, flags, s);
      case 8: return gram_launch<V, TIER, 16, 8, 2, 4>(gp, a, ss, part, flags, s);
      default: return hipErrorInvalidValue;
    }
  } else {
    if (gp.BW == 16) return gram_launch<V, TIER, 16, 1, 1, 16>(gp, a, ss, part, flags, s);
    switch (gp.nb) {
      case 1: return gram_launch<V, TIER, 32, 1, 1, 16>(gp, a, ss, part, flags, s);
      case 2: return gram_launch<V, TIER, 32, 2, 1, 8>(gp, a, ss, part, flags, s);
      case 3: return gram_launch<V, TIER, 32, 3, 1, 8>(gp, a, ss, part, flags, s);
      case 4: return gram_launch<V, TIER, 32, 4, 1, 8>(gp, a, ss, part, flags, s);
      default: return hipErrorInvalidValue;
    }
  }
}

template <typename V, int TIER>
static hipError_t factor_launch_t(const FactorArgs& a, void* scratch, double* out, hipStream_t s) {
  const FPlan pl = make_plan(TIER, a.R, a.F);
  const GPlan gp = make_gplan(TIER, a.R, a.F);
  const FLayout L = factor_layout(TIER, a.R, a.F);
  char* base = static_cast<char*>(scratch);
  uint32_t* flags = reinterpret_cast<uint32_t*>(base);
  double* ss = reinterpret_cast<double*>(base + L.ss);
  double* gpart = reinterpret_cast<double*>(base + L.gpart);
  double* stage = reinterpret_cast<double*>(base + L.stage);
  double* partial = reinterpret_cast<double*>(base + L.gen);
  uint32_t* fbl = reinterpret_cast<uint32_t*>(base + L.fbl);
  double* fbp = reinterpret_cast<double*>(base + L.fbp);
  const int ncell = a.F * (a.F + 1) / 2;
  // the pilot also clears the flags (no memset node, so the sequence can be captured)
  hipLaunchKernelGGL((factor_pilot_kernel<V, TIER>), dim3(a.F), dim3(256), 0, s, a, ss, flags);
  hipError_t e = gram_dispatch<V, TIER>(gp, a, ss, gpart, flags, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(factor_gram_reduce, dim3((gp.PB + 3) / 4), dim3(256), 0, s, gpart, gp.nblk, gp.PB,
                     gp.am_off, stage);
  hipLaunchKernelGGL(factor_gram_final<TIER>, dim3((ncell + 255) / 256), dim3(256), 0, s, gp, ss, stage,
                     out, fbl, flags);
  // general six-sum path: every kernel returns at once unless flags[1] is set
  const size_t lds = f_lds_bytes(TIER, pl);
  constexpr int MAXP = (TIER == kTierContract) ? 2 : 5;
  auto k = factor_main_kernel<V, TIER, MAXP>;
  if (lds > 65536) {
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  for (int pass = 0; pass < pl.npass; ++pass)
    hipLaunchKernelGGL(k, dim3(pl.groups), dim3(kFThreads), lds, s, a, ss, pl, pass, partial, flags);
  hipLaunchKernelGGL(factor_final_kernel<TIER>, dim3((ncell + 3) / 4), dim3(256), 0, s, pl, ss, partial,
                     stage + gp.am_off, out, fbl, flags);
  // exact pass over the flagged pairs of either path
  hipLaunchKernelGGL(factor_fb_sums<V>, dim3(1024), dim3(256), 0, s, a, fbl, flags, fbp);
  hipLaunchKernelGGL(factor_fb_centred<V>, dim3(1024), dim3(256), 0, s, a, fbl, flags, fbp);
  hipLaunchKernelGGL(factor_fb_final, dim3((ncell + 255) / 256), dim3(256), 0, s, a.F, fbl, flags, fbp,
                     out);
  return hipGetLastError();
}

hipError_t launch_factor_corr(int tier, const FactorArgs& a, void* scratch, size_t, double* out,
                              hipStream_t s) {
  if (a.value_bytes == 8) {
    if (tier == kTierContract) return factor_launch_t<double, kTierContract>(a, scratch, out, s);
    if (tier == kTierFp32) return factor_launch_t<double, kTierFp32>(a, scratch, out, s);
    return factor_launch_t<double, kTierBf16x3>(a, scratch, out, s);
  }
  if (tier == kTierContract) return factor_launch_t<float, kTierContract>(a, scratch, out, s);
  if (tier == kTierFp32) return factor_launch_t<float, kTierFp32>(a, scratch, out, s);
  return factor_launch_t<float, kTierBf16x3>(a, scratch, out, s);
}

}  // namespace fcc
