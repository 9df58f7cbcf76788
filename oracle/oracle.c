/*
 * oracle.c -- FP64 direct periodic stencil sweep.  TEST INFRASTRUCTURE ONLY: this file is the
 * reference libffstencil is tested against and is never linked into, loaded by or called from the product path.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
 *
 * What it restates (KevinWu2017/FlashFFTStencil):
 *   - src/1D/check_correct.hpp:15-38 `stencil1D`: centred periodic correlation
 *       out[i] = sum_{k=-h..h} in[(i+k) mod n] * kernel[k+h]
 *     -> orc_sweep(..., convention = 1 (centred), ...), any number of dims.
 *   - src/2D/check_correct.hpp:17-31 `rotateKernel180`, :53-61 `getElement` (periodic wrap) and
 *     :64-100 `stencil2D` (centred correlation with the rotated kernel, written at (i+h, j+h) mod n),
 *     composed at src/2D/2d_main.cu:190-192, and src/1D/1d_main.cu:220-223 (reverse kernel, stencil1D,
 *     rotate right by h).  Both compositions equal
 *       out[p] = sum_{a in [0,K)^d} w[a] * in[(p - a) mod n]
 *     -> orc_sweep(..., convention = 0 (ref), ...).
 *   The reference checks one step; the sweep repeats the step T times with ping-pong buffers
 *   (multi-step semantics of the fused FFT path, src/1D/1d_main.cu:45-59 and assets/intro.png).
 *   The reference has no 3D check (src/3D/check_correct.hpp is a copy of the 2D one); the 3D sweep is
 *   the same formula with a third axis.
 *
 * Pinning (tests/test_oracle.py): against the reference's own check_correct.hpp compiled from its
 * sources (oracle/_ref, 1D and square 2D, both compositions), against an independent numpy FFT
 * convolution, the Heat-1D binomial closed form C(2T,j)/4^T, delta-kernel rolls, mass conservation
 * and Fourier-mode eigenvalues.
 *
 * Also here: SplitMix64 uniform[0,1) generation, restated from the published SplitMix64 algorithm
 * (Steele, Lea, Flood 2014), used by the tests and the bench to build seeded inputs independently of
 * the library's own generator.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

uint64_t orc_splitmix64_next(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_fill_uniform(uint64_t seed, double* out, int64_t n) {
  uint64_t s = seed;
  for (int64_t i = 0; i < n; i++) out[i] = (double)(orc_splitmix64_next(&s) >> 11) * 0x1.0p-53;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static inline int64_t wrap(int64_t i, int64_t n) {
  i %= n;
  return i < 0 ? i + n : i;
}

/* One step in the ref convention on a 3D-shaped grid (n0, n1, n2); lower dims use n = 1 axes and a
 * 1-wide tap range on them.  out[p] = sum_a w[a] in[p - off(a)], with off(a) = a - lo per axis where
 * lo = 0 (ref) or -r (centred, after flipping the weights). */
static void step3(const double* in, double* out, const int64_t n[3], const int kw[3], const int64_t off0[3],
                  const double* w) {
  const int64_t n0 = n[0], n1 = n[1], n2 = n[2];
#pragma omp parallel for schedule(static)
  for (int64_t i0 = 0; i0 < n0; i0++) {
    double* orow = out + i0 * n1 * n2;
    for (int64_t i1 = 0; i1 < n1; i1++) {
      double* o = orow + i1 * n2;
      for (int64_t i2 = 0; i2 < n2; i2++) o[i2] = 0.0;
      for (int a0 = 0; a0 < kw[0]; a0++) {
        const int64_t s0 = wrap(i0 - (a0 + off0[0]), n0);
        for (int a1 = 0; a1 < kw[1]; a1++) {
          const int64_t s1 = wrap(i1 - (a1 + off0[1]), n1);
          const double* src = in + (s0 * n1 + s1) * n2;
          for (int a2 = 0; a2 < kw[2]; a2++) {
            const double wv = w[(a0 * kw[1] + a1) * kw[2] + a2];
            if (wv == 0.0) continue;
            const int64_t d = a2 + off0[2]; /* o[i2] += wv * src[(i2 - d) mod n2] */
            if (n2 > 2 * (d < 0 ? -d : d) + 1) {
              const int64_t lo = d > 0 ? d : 0, hi = d < 0 ? n2 + d : n2; /* no-wrap range */
              for (int64_t i2 = 0; i2 < lo; i2++) o[i2] += wv * src[wrap(i2 - d, n2)];
              for (int64_t i2 = lo; i2 < hi; i2++) o[i2] += wv * src[i2 - d];
              for (int64_t i2 = hi; i2 < n2; i2++) o[i2] += wv * src[wrap(i2 - d, n2)];
            } else {
              for (int64_t i2 = 0; i2 < n2; i2++) o[i2] += wv * src[wrap(i2 - d, n2)];
            }
          }
        }
      }
    }
  }
}

/*
 * T steps of the periodic stencil.  dims 1..3, sizes n[0..dims-1] (slowest first), radius r,
 * weights (2r+1)^dims row-major, convention 0 = ref, 1 = centred.  `in` and `out` may alias.
 * threads <= 0 keeps the OpenMP default.  Returns 0 on success, -1 on bad arguments or allocation.
 */
int orc_sweep(int dims, const int64_t* sizes, int radius, const double* weights, int convention,
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
  int64_t total = n[0] * n[1] * n[2];
  int64_t ntaps = (int64_t)kw[0] * kw[1] * kw[2];
  double* w = (double*)malloc(sizeof(double) * ntaps);
  if (!w) return -1;
  if (convention == 0) {
    memcpy(w, weights, sizeof(double) * ntaps);
  } else {
    /* centred: out[p] = sum_o w[o+r] in[p+o] = sum_a w'[a] in[p - (a - r)] with w'[a] = w[2r - a]
     * (flip along every used axis); offsets per used axis = -r. */
    for (int64_t i = 0; i < ntaps; i++) w[i] = weights[ntaps - 1 - i];
    for (int d = 0; d < 3; d++)
      if (kw[d] > 1) off0[d] = -radius;
  }
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  if (steps == 0) {
    if (out != in) memcpy(out, in, sizeof(double) * total);
    free(w);
    return 0;
  }
  double* tmp = (double*)malloc(sizeof(double) * total);
  double* tmp2 = NULL;
  if (!tmp) { free(w); return -1; }
  const double* src = in;
  if (in == out) { /* keep `in` readable while writing */
    tmp2 = (double*)malloc(sizeof(double) * total);
    if (!tmp2) { free(w); free(tmp); return -1; }
    memcpy(tmp2, in, sizeof(double) * total);
    src = tmp2;
  }
  /* ping-pong so that the last step lands in `out` */
  for (int64_t s = 0; s < steps; s++) {
    double* dst = ((steps - 1 - s) % 2 == 0) ? out : tmp;
    step3(src, dst, n, kw, off0, w);
    src = dst;
  }
  free(w);
  free(tmp);
  free(tmp2);
  return 0;
}
