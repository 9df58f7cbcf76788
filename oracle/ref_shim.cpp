// ref_shim.cpp -- C-linkage wrapper around the reference's own CPU check functions, so the tests can
// call them through ctypes.  TEST INFRASTRUCTURE ONLY (oracle side).  The reference header is not
// copied: oracle/Makefile force-includes it from the reference checkout with `-include`, and the
// result lands in oracle/_ref/ (not committed).
//   FFS_REF_1D: src/1D/check_correct.hpp  -- stencil1D (:15-38)
//   FFS_REF_2D: src/2D/check_correct.hpp  -- rotateKernel180 (:17-31), convertTo2D (:35-48), stencil2D (:64-100)
#include <algorithm>
#include <vector>

extern "C" {
#if defined(FFS_REF_1D)
// One centred step exactly as stencil1D computes it.
int ref_stencil1d(const double* in, int n, const double* kernel, int k, double* out) {
  stencil1D(in, n, kernel, k, out);
  return 0;
}
// One step of the reference 1D main program's dormant check composition (src/1D/1d_main.cu:220-223): reverse the
// kernel, run stencil1D, rotate the result right by k/2.
int ref_step1d_driver(const double* in, int n, const double* kernel, int k, double* out) {
  std::vector<double> rk(kernel, kernel + k);
  std::reverse(rk.begin(), rk.end());
  std::vector<double> res(n);
  stencil1D(in, n, rk.data(), k, res.data());
  std::rotate(res.rbegin(), res.rbegin() + (k / 2), res.rend());
  std::copy(res.begin(), res.end(), out);
  return 0;
}
#elif defined(FFS_REF_2D)
// One step of the reference 2D main program's dormant check composition (src/2D/2d_main.cu:190-192):
// stencil2D(convertTo2D(in), rotateKernel180(kernel)) on a square width x width grid.
int ref_step2d_driver(const double* in, int width, const double* kernel, int k, double* out) {
  std::vector<double> kv(kernel, kernel + k * k);
  std::vector<std::vector<double>> rk = rotateKernel180(kv);
  std::vector<std::vector<double>> grid = convertTo2D(const_cast<double*>(in), width);
  std::vector<double> res = stencil2D(grid, rk);
  std::copy(res.begin(), res.end(), out);
  return 0;
}
// stencil2D with the kernel as given (no rotation): centred correlation written at (i+h, j+h).
int ref_stencil2d_raw(const double* in, int width, const double* kernel, int k, double* out) {
  std::vector<std::vector<double>> kk(k, std::vector<double>(k));
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) kk[i][j] = kernel[i * k + j];
  std::vector<std::vector<double>> grid = convertTo2D(const_cast<double*>(in), width);
  std::vector<double> res = stencil2D(grid, kk);
  std::copy(res.begin(), res.end(), out);
  return 0;
}
#endif
}
