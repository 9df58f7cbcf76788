// ffs_driver.cpp -- the drop-in `1d` / `2d` / `3d` programs (built three times with FFS_DRIVER_DIMS=1,2,3).
//
// Interface kept from the reference drivers (src/1D/1d_main.cu:23-38, src/2D/2d_main.cu:30-47,
// src/3D/3d_main.cu:26-42, README.md:44-58):
//     <prog> shape input_size time_iteration_size [options]
// printing one line
//     FlashFFTStencil, <tag>, <d1>, <d2>, <d3>, <T>, <ms>, <GStencil/s>
// (src/1D/1d_main.cu:195, src/2D/2d_main.cu:185, src/3D/3d_main.cu:185).  Usage errors print the help
// text and exit 1, as the reference does; an unknown shape exits 1 as well.
// Differences, all documented in the help text:
//   * `input_size` may be one side (square/cube grid, the reference's form) or one value per dimension;
//     any size is accepted (the reference rejects sizes not divisible by its valid width, and the 3D
//     reference ignores the size and runs 768^3);
//   * the grid really advances T steps (the reference relaunches on the same input);
//   * 1D reports points x T / time (the reference multiplies by the fused radius, 1d_main.cu:195);
//   * `--custom FILE` (or `-` for stdin) reads (2r+1)^d weights, row-major, in the ref convention;
//   * `--output FILE` writes the grid after the T steps (raw FP64, row-major, native byte order);
//   * `--help` prints the help and exits 0.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/ffstencil.h"

#ifndef FFS_DRIVER_DIMS
#error "build with -DFFS_DRIVER_DIMS=1, 2 or 3"
#endif

static const int kDims = FFS_DRIVER_DIMS;

static void print_help() {
  const char* shapes = kDims == 1 ? "Heat-1D / 1D5P / 1D7P"
                       : kDims == 2 ? "Heat-2D / Box2D9P / Star2D9P / Box2D25P"
                                    : "Heat-3D / Box3D27P (Box3d27P)";
  std::printf(
      "Program name: FlashFFTStencil-%dD (MI355X)\n"
      "Usage: %dd [stencil-shape] [input_size] [time_step] [options]\n"
      "Stencil-shape: %s\n"
      "input_size: one side (the grid is a square/cube of it) or %d sides, slowest axis first\n"
      "Options:\n"
      "  --help               print this text\n"
      "  --custom FILE|-      read (2r+1)^%d stencil weights (row-major, ref convention) from FILE or stdin;\n"
      "                       r is the shape's radius\n"
      "  --precision P        fp64 (default) | bf16 (mixed: BF16 split operands) | f16 (mixed: FP16 split\n"
      "                       operands, 2D) | f32 (mixed: FP32 operands; compensated DFT matrix in 3D)\n"
      "  --convention C       ref (default, out[p] = sum_a w[a] in[p-a]) | centred (out[p] = sum_o w[o+r] in[p+o])\n"
      "  --algorithm A        fft (default: tiled FFT stencil) | direct (FP64 stencil sum, temporal blocking) |\n"
      "                       auto (direct where offered)\n"
      "  --mapping M          linear (default: hardware workgroup order) | xcd (each XCD takes a contiguous\n"
      "                       band of tiles)\n"
      "  --seed S             seed of the weights and the initial grid (default 0x5EED0000; the grid is\n"
      "                       ffs_fill_uniform(S + 1), the weights ffs_stencil_named(shape, S))\n"
      "  --output FILE        write the grid after the T steps to FILE (raw FP64, row-major)\n"
      "  --fuse t             fused time depth (default automatic)\n"
      "  --no-graph           plain launches instead of a hipGraph replay\n"
      "  --tune [CACHE]       time the admissible tile/depth choices first (cached in CACHE, a JSON-lines file)\n",
      kDims, kDims, shapes, kDims, kDims);
}

static bool parse_i64(const char* s, long long* v) {
  char* end = nullptr;
  long long x = std::strtoll(s, &end, 10);
  if (!s[0] || *end) return false;
  *v = x;
  return true;
}

#define CHECK_FFS(call)                                                              \
  do {                                                                               \
    ffs_status st_ = (call);                                                         \
    if (st_ != FFS_OK) {                                                             \
      std::fprintf(stderr, "error: %s: %s\n", #call, ffs_last_error());              \
      return st_ == FFS_ERR_INVALID_ARG ? 1 : 2;                                     \
    }                                                                                \
  } while (0)

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  std::string custom, output, prec = "fp64", conv = "ref", algo = "fft", mapping = "linear";
  unsigned long long seed = 0x5EED0000ULL;
  int fuse = 0, use_graph = 1;
  bool tune = false;
  std::string tune_cache;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto need = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", what);
        print_help();
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "--help" || a == "-h") {
      print_help();
      return 0;
    } else if (a == "--custom") {
      custom = need("--custom");
    } else if (a == "--output") {
      output = need("--output");
    } else if (a == "--precision") {
      prec = need("--precision");
    } else if (a == "--convention") {
      conv = need("--convention");
    } else if (a == "--algorithm") {
      algo = need("--algorithm");
    } else if (a == "--mapping") {
      mapping = need("--mapping");
    } else if (a == "--seed") {
      seed = std::strtoull(need("--seed"), nullptr, 0);
    } else if (a == "--fuse") {
      fuse = std::atoi(need("--fuse"));
    } else if (a == "--no-graph") {
      use_graph = 0;
    } else if (a == "--tune") {
      tune = true;
      if (i + 1 < argc && argv[i + 1][0] != '-' && std::strchr(argv[i + 1], '.') != nullptr) tune_cache = argv[++i];
    } else if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      print_help();
      return 1;
    } else {
      pos.push_back(a);
    }
  }
  if (!(pos.size() == 3 || pos.size() == (size_t)(2 + kDims))) {
    print_help();
    return 1;
  }
  const std::string shape = pos[0];
  long long sizes[3] = {1, 1, 1};
  long long T = 0;
  if (pos.size() == 3) {
    long long w;
    if (!parse_i64(pos[1].c_str(), &w)) { print_help(); return 1; }
    for (int d = 0; d < kDims; d++) sizes[d] = w;
  } else {
    for (int d = 0; d < kDims; d++)
      if (!parse_i64(pos[1 + d].c_str(), &sizes[d])) { print_help(); return 1; }
  }
  if (!parse_i64(pos.back().c_str(), &T)) { print_help(); return 1; }

  int32_t dims = 0, radius = 0;
  char tag[32];
  std::vector<double> w(7 * 7 * 7);
  if (ffs_stencil_named(shape.c_str(), seed, &dims, &radius, w.data(), (int32_t)w.size(), tag, sizeof tag) != FFS_OK ||
      dims != kDims) {
    std::cout << "FlashFFTStencil can support stencil calculations for more shapes; '" << shape
              << "' is not one of the named " << kDims << "D shapes." << std::endl;
    return 1;
  }
  const int K = 2 * radius + 1;
  int ntaps = 1;
  for (int d = 0; d < dims; d++) ntaps *= K;
  w.resize(ntaps);
  if (!custom.empty()) {
    std::string text;
    if (custom == "-") {
      std::stringstream ss;
      ss << std::cin.rdbuf();
      text = ss.str();
    } else {
      std::ifstream f(custom);
      if (!f) {
        std::fprintf(stderr, "cannot open %s\n", custom.c_str());
        return 1;
      }
      std::stringstream ss;
      ss << f.rdbuf();
      text = ss.str();
    }
    std::istringstream is(text);
    std::vector<double> cw;
    double x;
    while (is >> x) cw.push_back(x);
    if ((int)cw.size() != ntaps) {
      std::fprintf(stderr, "--custom: %s needs %d weights, read %zu\n", shape.c_str(), ntaps, cw.size());
      return 1;
    }
    w = cw;
  }

  ffs_problem p;
  ffs_problem_init(&p);
  p.dims = dims;
  p.radius = radius;
  for (int d = 0; d < dims; d++) p.sizes[d] = sizes[d];
  p.weights = w.data();
  p.steps = T;
  p.precision = prec == "fp64"   ? FFS_PREC_FP64
                : prec == "bf16" ? FFS_PREC_MIXED_BF16
                : prec == "f16"  ? FFS_PREC_MIXED_F16
                : prec == "f32"  ? FFS_PREC_MIXED_F32
                                 : -1;
  p.convention = conv == "ref" ? FFS_CONV_REF : conv == "centred" || conv == "centered" ? FFS_CONV_CENTRED : -1;
  p.algorithm = algo == "fft" ? FFS_ALGO_FFT : algo == "direct" ? FFS_ALGO_DIRECT : algo == "auto" ? FFS_ALGO_AUTO : -1;
  p.mapping = mapping == "xcd" ? FFS_MAP_XCD : mapping == "linear" ? FFS_MAP_LINEAR : -1;
  p.fuse = fuse;
  p.use_graph = use_graph;

  if (tune) CHECK_FFS(ffs_autotune(&p, tune_cache.empty() ? nullptr : tune_cache.c_str(), nullptr));
  ffs_plan* plan = nullptr;
  CHECK_FFS(ffs_plan_create(&p, &plan));
  long long npts = 1;
  for (int d = 0; d < dims; d++) npts *= sizes[d];
  std::vector<double> h(npts);
  CHECK_FFS(ffs_fill_uniform(seed + 1, h.data(), npts));
  double *din = nullptr, *dout = nullptr;
  if (hipMalloc(&din, sizeof(double) * npts) != hipSuccess || hipMalloc(&dout, sizeof(double) * npts) != hipSuccess) {
    std::fprintf(stderr, "device allocation failed\n");
    return 2;
  }
  (void)hipMemcpy(din, h.data(), sizeof(double) * npts, hipMemcpyHostToDevice);
  CHECK_FFS(ffs_run(plan, din, dout, nullptr));  // warm-up (captures the graph)
  (void)hipDeviceSynchronize();
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, nullptr);
  CHECK_FFS(ffs_run(plan, din, dout, nullptr));
  (void)hipEventRecord(e1, nullptr);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  char line[256];
  int64_t s64[3] = {sizes[0], sizes[1], sizes[2]};
  CHECK_FFS(ffs_format_csv(tag, dims, s64, T, ms, line, sizeof line));
  std::cout << line << std::endl;
  if (!output.empty()) {
    if (hipMemcpy(h.data(), dout, sizeof(double) * npts, hipMemcpyDeviceToHost) != hipSuccess) {
      std::fprintf(stderr, "copy of the result failed\n");
      return 2;
    }
    std::FILE* f = std::fopen(output.c_str(), "wb");
    if (!f || std::fwrite(h.data(), sizeof(double), (size_t)npts, f) != (size_t)npts) {
      std::fprintf(stderr, "cannot write %s\n", output.c_str());
      if (f) std::fclose(f);
      return 1;
    }
    std::fclose(f);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(din);
  (void)hipFree(dout);
  ffs_plan_destroy(plan);
  return 0;
}
