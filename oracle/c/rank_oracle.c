/* Plain-C restatement of the reference ranking semantics, a fast checker
 * for full-size panels (test infrastructure; never linked into the product).
 *
 * Ranks: stable ordinal rank over the participating cells of each row,
 * ties by ascending column index, descending = ascending of the negated
 * values (fc/_cpu_core.py:57-78; CLAUDE.md:63; the reference C++ selfcheck
 * sorts with std::stable_sort, poc/poc3_cs_rank_selfcheck.cu:48-70). Sorting
 * (key, column) pairs with a total order on both is the same permutation.
 * IC: Spearman of ordinal ranks on the pairwise-valid set
 * (fc/_cpu_core.py:271-293) with the exact integer sum S = sum rf*rr and the
 * centred sums sxy = S - m(m+1)^2/4, sxx = syy = m(m^2-1)/12, which the
 * reference's two-pass reduction (src/rolling_ic.cu:232-304) also obtains
 * exactly because every partial sum is a multiple of 1/4 below 2^53.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t key; uint32_t idx; } kv_t;

static int kv_cmp(const void* a, const void* b) {
  const kv_t* x = (const kv_t*)a;
  const kv_t* y = (const kv_t*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

static uint64_t key_f32(float v) {
  uint32_t b;
  if (v == 0.0f) v = 0.0f;
  memcpy(&b, &v, 4);
  if (v == 0.0f) b = 0u;
  return (b & 0x80000000u) ? (uint64_t)(~b) : (uint64_t)(b | 0x80000000u);
}

static uint64_t key_f64(double v) {
  uint64_t b;
  memcpy(&b, &v, 8);
  if (v == 0.0) b = 0ull;
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

/* out[t*N+i] = rank or quiet NaN; mask may be NULL. */
void oracle_rank_rows(const float* X, const uint8_t* mask, int T, int N, int descending,
                      float* out) {
  kv_t* buf = (kv_t*)malloc(sizeof(kv_t) * (size_t)N);
  const uint32_t qnan_bits = 0x7fc00000u;
  float qnan;
  memcpy(&qnan, &qnan_bits, 4);
  for (int t = 0; t < T; ++t) {
    const float* x = X + (size_t)t * N;
    float* o = out + (size_t)t * N;
    int k = 0;
    for (int i = 0; i < N; ++i) {
      o[i] = qnan;
      float v = descending ? -x[i] : x[i];
      if (!isfinite(v)) continue;
      if (mask && !mask[(size_t)t * N + i]) continue;
      buf[k].key = key_f32(v);
      buf[k].idx = (uint32_t)i;
      ++k;
    }
    qsort(buf, (size_t)k, sizeof(kv_t), kv_cmp);
    for (int p = 0; p < k; ++p) o[buf[p].idx] = (float)(p + 1);
  }
  free(buf);
}

/* rank of every valid cell of one row among the valid cells (1-based). */
static void row_ranks(const void* x, int is64, const uint8_t* valid, int N, kv_t* buf,
                      int64_t* rk) {
  int k = 0;
  for (int i = 0; i < N; ++i) {
    if (!valid[i]) continue;
    buf[k].key = is64 ? key_f64(((const double*)x)[i]) : key_f32(((const float*)x)[i]);
    buf[k].idx = (uint32_t)i;
    ++k;
  }
  qsort(buf, (size_t)k, sizeof(kv_t), kv_cmp);
  for (int p = 0; p < k; ++p) rk[buf[p].idx] = p + 1;
}

/* Spearman IC per row; is64 selects double inputs. */
void oracle_rolling_ic(const void* F, const void* R, int is64, const uint8_t* fmask,
                       const uint8_t* rmask, int T, int N, int min_valid, double* ic) {
  kv_t* buf = (kv_t*)malloc(sizeof(kv_t) * (size_t)N);
  uint8_t* valid = (uint8_t*)malloc((size_t)N);
  int64_t* rf = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  int64_t* rr = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  const size_t vs = is64 ? 8 : 4;
  for (int t = 0; t < T; ++t) {
    const char* f = (const char*)F + (size_t)t * N * vs;
    const char* r = (const char*)R + (size_t)t * N * vs;
    int64_t m = 0;
    double fmin = INFINITY, fmax = -INFINITY, rmin = INFINITY, rmax = -INFINITY;
    for (int i = 0; i < N; ++i) {
      double fv = is64 ? ((const double*)f)[i] : (double)((const float*)f)[i];
      double rv = is64 ? ((const double*)r)[i] : (double)((const float*)r)[i];
      int ok = isfinite(fv) && isfinite(rv);
      if (fmask && !fmask[(size_t)t * N + i]) ok = 0;
      if (rmask && !rmask[(size_t)t * N + i]) ok = 0;
      valid[i] = (uint8_t)ok;
      if (ok) {
        ++m;
        if (fv < fmin) fmin = fv;
        if (fv > fmax) fmax = fv;
        if (rv < rmin) rmin = rv;
        if (rv > rmax) rmax = rv;
      }
    }
    double out = NAN;
    if (m >= min_valid && m >= 2 && fmin != fmax && rmin != rmax) {
      row_ranks(f, is64, valid, N, buf, rf);
      row_ranks(r, is64, valid, N, buf, rr);
      /* 128-bit: at the 2^24 row cap m^3 and S reach ~2^72 */
      __int128 S = 0;
      for (int i = 0; i < N; ++i)
        if (valid[i]) S += (__int128)rf[i] * rr[i];
      const __int128 mm = m;
      const __int128 A = 4 * S - mm * (mm + 1) * (mm + 1);
      const __int128 B = mm * (mm * mm - 1);
      const double sxy = (double)A * 0.25;
      const double sxx = (double)B / 12.0;
      out = (sxy / sqrt(sxx)) / sqrt(sxx);
    }
    ic[t] = out;
  }
  free(buf);
  free(valid);
  free(rf);
  free(rr);
}
