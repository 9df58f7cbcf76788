/* oracle_check.c -- runs the oracle (oracle/oracle.c, the strict FP64 sweep, and
 * oracle/sweep_fast.c, the timed CPU baseline) under AddressSanitizer and UndefinedBehaviorSanitizer over
 * the shapes the tests use: 1D/2D/3D, ragged and tiny grids (sides smaller than the stencil, so every
 * wrap branch runs; rows shorter and longer than the fast sweep's blocks and strips), both conventions,
 * T = 0, in == out aliasing and rejected arguments.  Each strict-sweep run is checked for mass conservation
 * (sum out = sum in * (sum w)^T, exact up to rounding for periodic grids) and each fast run against the
 * strict sweep's result.  Built and run by tests/test_sanitizers.py. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int orc_sweep(int dims, const int64_t* sizes, int radius, const double* weights, int convention, int64_t steps,
              const double* in, double* out, int threads);
int orc_sweep_fast(int dims, const int64_t* sizes, int radius, const double* weights, int convention,
                   int64_t steps, const double* in, double* out, int threads);
void orc_fill_uniform(uint64_t seed, double* out, int64_t n);

static int failures = 0;

static void one(int dims, const int64_t* sizes, int radius, int convention, int64_t steps, int alias) {
  int64_t n = 1, ntaps = 1;
  for (int d = 0; d < dims; d++) {
    n *= sizes[d];
    ntaps *= 2 * radius + 1;
  }
  double* w = malloc(sizeof(double) * ntaps);
  double* x = malloc(sizeof(double) * n);
  double* y = malloc(sizeof(double) * n);
  orc_fill_uniform(11 + (uint64_t)ntaps, w, ntaps);
  orc_fill_uniform(5 + (uint64_t)n, x, n);
  double sw = 0.0, sx = 0.0, sy = 0.0;
  for (int64_t i = 0; i < ntaps; i++) sw += w[i];
  for (int64_t i = 0; i < n; i++) sx += x[i];
  double* z = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) z[i] = x[i];
  double* dst = alias ? x : y;
  if (orc_sweep(dims, sizes, radius, w, convention, steps, x, dst, 2) != 0) {
    fprintf(stderr, "sweep failed: dims %d radius %d\n", dims, radius);
    failures++;
  }
  for (int64_t i = 0; i < n; i++) sy += dst[i];
  /* the fast sweep: in place when the strict sweep ran in place, else into a separate buffer */
  double* zo = alias ? z : malloc(sizeof(double) * n);
  if (orc_sweep_fast(dims, sizes, radius, w, convention, steps, z, zo, 2) != 0) {
    fprintf(stderr, "fast sweep failed: dims %d radius %d\n", dims, radius);
    failures++;
  }
  double dmax = 0.0, rmax = 0.0;
  for (int64_t i = 0; i < n; i++) {
    dmax = fmax(dmax, fabs(zo[i] - dst[i]));
    rmax = fmax(rmax, fabs(dst[i]));
  }
  if (dmax > 1e-13 * rmax) {
    fprintf(stderr, "fast vs strict sweep: dims %d radius %d conv %d: %.3g\n", dims, radius, convention, dmax / rmax);
    failures++;
  }
  if (!alias) free(zo);
  free(z);
  const double want = sx * pow(sw, (double)steps);
  if (fabs(sy - want) > 1e-11 * fabs(want) * (1 + steps)) {
    fprintf(stderr, "mass: dims %d radius %d conv %d T %lld: %.17g vs %.17g\n", dims, radius, convention,
            (long long)steps, sy, want);
    failures++;
  }
  free(w);
  free(x);
  free(y);
}

int main(void) {
  const int64_t s1[][1] = {{1}, {2}, {5}, {7}, {240}, {1001}, {1300}};
  const int64_t s2[][2] = {{1, 1}, {1, 19}, {5, 7}, {37, 53}, {3, 150}, {2, 600}};
  const int64_t s3[][3] = {{1, 1, 1}, {3, 2, 5}, {11, 13, 17}};
  for (int conv = 0; conv < 2; conv++)
    for (int r = 0; r <= 3; r++) {
      for (size_t i = 0; i < sizeof s1 / sizeof s1[0]; i++) one(1, s1[i], r, conv, 3, i & 1);
      for (size_t i = 0; i < sizeof s2 / sizeof s2[0]; i++) one(2, s2[i], r, conv, 2, i & 1);
      if (r <= 1)
        for (size_t i = 0; i < sizeof s3 / sizeof s3[0]; i++) one(3, s3[i], r, conv, 2, i & 1);
    }
  one(2, s2[3], 2, 0, 0, 0); /* T = 0: a copy */
  one(2, s2[3], 2, 0, 0, 1);
  /* rejected arguments: an error code, no access */
  const int64_t bad[2] = {0, 5};
  double w[9] = {0};
  double x[5] = {0};
  if (orc_sweep(2, bad, 1, w, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep(0, s2[2], 1, w, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep(2, s2[2], -1, w, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep(2, s2[2], 1, NULL, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep_fast(2, bad, 1, w, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep_fast(4, s2[2], 1, w, 0, 1, x, x, 1) != -1) failures++;
  if (orc_sweep_fast(2, s2[2], 1, w, 0, -1, x, x, 1) != -1) failures++;
  if (failures) {
    fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  printf("oracle checks passed\n");
  return 0;
}
