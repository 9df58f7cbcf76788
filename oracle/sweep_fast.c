/*
 * sweep_fast.c -- the oracle's sweep written for speed: the CPU baseline that bench.py times.
 * TEST INFRASTRUCTURE ONLY, like oracle.c: never linked into, loaded by or called from the product path.
 *
 * Same operator as orc_sweep in oracle.c (and so the reference's check semantics restated there,
 * src/2D/check_correct.hpp:64-100 and src/1D/check_correct.hpp:15-38, in both compositions):
 *   out[p] = sum_a w[a] * in[(p - off(a)) mod n],   T steps with ping-pong buffers.
 * What differs is only how a step is evaluated:
 *   - each output is accumulated in a register over all non-zero taps (taps in row-major order, one
 *     FMA each) and stored once, instead of one pass over the output row per tap;
 *   - outputs go in blocks of 32 along the fastest axis, so the tap loop runs as 8 independent AVX2
 *     FMA chains (intrinsics: compilers interchange the plain-C form of this loop) over contiguous
 *     loads, then blocks of 4; the periodic wrap is taken only at the two ends of a row;
 *   - OpenMP splits the grid into work items of one row's 512-point chunk, ordered strip by strip
 *     down the rows (L1 reuse of the source rows), so a 1D grid spreads over all threads as well.
 * The explicit fma() calls make the result depend on the tap order only: it equals orc_sweep to a few
 * units of rounding per step (tests/test_oracle.py pins the two together at 1e-13 over many steps,
 * shapes, radii and both conventions).
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { kBlock = 32, kChunk = 512 };

typedef struct {
  int32_t row;  /* index into the item's source-row pointers: a0 * kw1 + a1 */
  int32_t pad;
  int64_t d;    /* axis-2 offset: out[i2] reads src[i2 - d] */
  double w;
} Tap;

static inline int64_t wrapi(int64_t i, int64_t n) {
  if (i >= 0 && i < n) return i;
  if (i < 0 && i >= -n) return i + n;
  if (i >= n && i < 2 * n) return i - n;
  i %= n;
  return i < 0 ? i + n : i;
}

static inline int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }
static inline int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* Outputs o[i2 .. i2 + 32) of an interior stretch: 8 AVX2 accumulators, one FMA per tap each. */
static inline void block32(const double** src, const Tap* taps, int ntap, int64_t i2, double* o) {
  __m256d a0 = _mm256_setzero_pd(), a1 = a0, a2 = a0, a3 = a0, a4 = a0, a5 = a0, a6 = a0, a7 = a0;
  for (int t = 0; t < ntap; t++) {
    const double* s = src[taps[t].row] + (i2 - taps[t].d);
    const __m256d wv = _mm256_set1_pd(taps[t].w);
    a0 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s), a0);
    a1 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 4), a1);
    a2 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 8), a2);
    a3 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 12), a3);
    a4 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 16), a4);
    a5 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 20), a5);
    a6 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 24), a6);
    a7 = _mm256_fmadd_pd(wv, _mm256_loadu_pd(s + 28), a7);
  }
  double* d = o + i2;
  _mm256_storeu_pd(d, a0);
  _mm256_storeu_pd(d + 4, a1);
  _mm256_storeu_pd(d + 8, a2);
  _mm256_storeu_pd(d + 12, a3);
  _mm256_storeu_pd(d + 16, a4);
  _mm256_storeu_pd(d + 20, a5);
  _mm256_storeu_pd(d + 24, a6);
  _mm256_storeu_pd(d + 28, a7);
}

/* One step; returns -1 if a thread could not allocate its source-row pointers ((2r+1)^2 of them, on the
 * heap: a large radius would not fit a thread's stack). */
static int step_fast(const double* in, double* out, const int64_t n[3], const int kw[3], const int64_t off0[3],
                     const Tap* taps, int ntap, int64_t dlo, int64_t dhi) {
  const int64_t n0 = n[0], n1 = n[1], n2 = n[2];
  const int64_t nch = (n2 + kChunk - 1) / kChunk;
  const int64_t items = n0 * n1 * nch;
  const int nrows = kw[0] * kw[1];
  int failed = 0;
#pragma omp parallel reduction(| : failed)
  {
    const double** src = (const double**)malloc(sizeof(const double*) * (size_t)nrows);
    if (!src) failed = 1;
#pragma omp for schedule(static)
    for (int64_t it = 0; it < items; it++) {
      if (!src) continue;
      /* column strips of kChunk points, each walked down the rows: the source row segments of one
       * output row are mostly those of the previous one, still in L1 */
      const int64_t c = it / (n0 * n1), row = it - c * (n0 * n1);
      const int64_t i0 = row / n1, i1 = row - i0 * n1;
      for (int a0 = 0; a0 < kw[0]; a0++) {
        const int64_t s0 = wrapi(i0 - (a0 + off0[0]), n0);
        for (int a1 = 0; a1 < kw[1]; a1++)
          src[a0 * kw[1] + a1] = in + (s0 * n1 + wrapi(i1 - (a1 + off0[1]), n1)) * n2;
      }
      double* o = out + row * n2;
      const int64_t b = c * kChunk, e = min64(b + kChunk, n2);
      /* outputs whose every read i2 - d (d in [dlo, dhi]) lies in [0, n2) need no wrap */
      int64_t lo = max64(b, max64(dhi, 0)), hi = min64(e, n2 + min64(dlo, 0));
      if (hi < lo) lo = hi = e;
      for (int64_t i2 = b; i2 < e; i2++) {
        if (i2 == lo) {
          /* interior: blocks of 32 outputs; the last one is moved back to end at hi (the outputs it
           * repeats get the same values), so only stretches shorter than a block go 4 or 1 at a time */
          for (; i2 + kBlock <= hi; i2 += kBlock) block32(src, taps, ntap, i2, o);
          if (i2 < hi && hi - lo >= kBlock) {
            block32(src, taps, ntap, hi - kBlock, o);
            i2 = hi;
          }
          for (; i2 + 4 <= hi; i2 += 4) {
            __m256d a0 = _mm256_setzero_pd();
            for (int t = 0; t < ntap; t++)
              a0 = _mm256_fmadd_pd(_mm256_set1_pd(taps[t].w), _mm256_loadu_pd(src[taps[t].row] + (i2 - taps[t].d)), a0);
            _mm256_storeu_pd(o + i2, a0);
          }
          if (i2 >= e) break;
        }
        const int wraps = i2 < lo || i2 >= hi;
        double acc = 0.0;
        for (int t = 0; t < ntap; t++) {
          const int64_t j = wraps ? wrapi(i2 - taps[t].d, n2) : i2 - taps[t].d;
          acc = fma(taps[t].w, src[taps[t].row][j], acc);
        }
        o[i2] = acc;
      }
    }
    free(src);
  }
  return failed ? -1 : 0;
}

/*
 * orc_sweep_fast: T steps of the periodic stencil, arguments and return value as orc_sweep (dims 1..3,
 * sizes slowest first, radius r, weights (2r+1)^dims row-major, convention 0 = ref, 1 = centred; `in`
 * and `out` may alias; threads <= 0 keeps the OpenMP default).
 */
int orc_sweep_fast(int dims, const int64_t* sizes, int radius, const double* weights, int convention,
                   int64_t steps, const double* in, double* out, int threads) {
  if (dims < 1 || dims > 3 || radius < 0 || steps < 0 || !sizes || !weights || !in || !out) return -1;
  int64_t n[3] = {1, 1, 1};
  int kw[3] = {1, 1, 1};
  int64_t off0[3] = {0, 0, 0};
  const int K = 2 * radius + 1;
  for (int d = 0; d < dims; d++) {
    n[3 - dims + d] = sizes[d];
    kw[3 - dims + d] = K;
    if (sizes[d] < 1) return -1;
  }
  const int64_t total = n[0] * n[1] * n[2];
  const int ntaps = kw[0] * kw[1] * kw[2];
  if (convention != 0)
    for (int d = 0; d < 3; d++)
      if (kw[d] > 1) off0[d] = -radius;
  /* non-zero taps in row-major order; the centred convention flips the weights (as orc_sweep) */
  Tap* taps = (Tap*)malloc(sizeof(Tap) * (size_t)ntaps);
  if (!taps) return -1;
  int nt = 0;
  int64_t dlo = 0, dhi = 0;
  for (int a = 0; a < ntaps; a++) {
    const double wv = convention == 0 ? weights[a] : weights[ntaps - 1 - a];
    if (wv == 0.0) continue;
    const int a2 = a % kw[2], a01 = a / kw[2];
    taps[nt].row = a01;
    taps[nt].pad = 0;
    taps[nt].d = a2 + off0[2];
    taps[nt].w = wv;
    if (nt == 0 || taps[nt].d < dlo) dlo = taps[nt].d;
    if (nt == 0 || taps[nt].d > dhi) dhi = taps[nt].d;
    nt++;
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  if (steps == 0) {
    if (out != in) memcpy(out, in, sizeof(double) * (size_t)total);
    free(taps);
    return 0;
  }
  double* tmp = (double*)malloc(sizeof(double) * (size_t)total);
  double* keep = NULL;
  if (!tmp) { free(taps); return -1; }
  const double* src = in;
  if (in == out) {
    keep = (double*)malloc(sizeof(double) * (size_t)total);
    if (!keep) { free(taps); free(tmp); return -1; }
    memcpy(keep, in, sizeof(double) * (size_t)total);
    src = keep;
  }
  int rc = 0;
  for (int64_t s = 0; s < steps && rc == 0; s++) {
    double* dst = ((steps - 1 - s) % 2 == 0) ? out : tmp;
    rc = step_fast(src, dst, n, kw, off0, taps, nt, dlo, dhi);
    src = dst;
  }
  free(tmp);
  free(keep);
  free(taps);
  return rc;
}
