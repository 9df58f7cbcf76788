// host_abi_check.cpp -- drives every host-only path of libffstencil (no GPU needed) for a run under
// AddressSanitizer and UndefinedBehaviorSanitizer: argument validation of plan and batch creation (it
// runs before any HIP call), the named-stencil table with small and exact capacities, the SplitMix64
// fill, the reference CSV line with short buffers, the status strings.  Built and run by
// tests/test_sanitizers.py; exits non-zero on the first unexpected status, and the sanitizers abort on
// any memory or UB error.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ffstencil.h"

static int failures = 0;

#define EXPECT(cond)                                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond); \
      failures++;                                                      \
    }                                                                  \
  } while (0)

static void named_stencils() {
  const char* names[] = {"Heat-1D", "1D5P", "1D7P", "Heat-2D", "Box2D9P", "Star2D9P", "Box2D25P", "Heat-3D",
                         "Box3D27P", "Box3d27P"};
  for (const char* n : names) {
    int32_t dims = 0, r = 0;
    // exact capacity: (2r+1)^dims weights, tag of 16
    std::vector<double> w(343, -1.0);
    char tag[16];
    EXPECT(ffs_stencil_named(n, 7, &dims, &r, w.data(), 343, tag, sizeof tag) == FFS_OK);
    int ntaps = 1;
    for (int d = 0; d < dims; d++) ntaps *= 2 * r + 1;
    std::vector<double> exact(ntaps);
    EXPECT(ffs_stencil_named(n, 7, &dims, &r, exact.data(), ntaps, nullptr, 0) == FFS_OK);
    double s = 0.0;
    for (double v : exact) s += v;
    EXPECT(std::fabs(s - 1.0) < 1e-12);
    EXPECT(std::strlen(tag) < sizeof tag);
    // one weight too few and a tag buffer too small are rejected, never overrun
    if (ntaps > 1) EXPECT(ffs_stencil_named(n, 7, &dims, &r, exact.data(), ntaps - 1, nullptr, 0) != FFS_OK);
    char tiny[2];
    EXPECT(ffs_stencil_named(n, 7, &dims, &r, exact.data(), ntaps, tiny, sizeof tiny) != FFS_OK);
  }
  int32_t dims, r;
  double w[9];
  EXPECT(ffs_stencil_named("NoSuchShape", 1, &dims, &r, w, 9, nullptr, 0) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_stencil_named(nullptr, 1, &dims, &r, w, 9, nullptr, 0) == FFS_ERR_INVALID_ARG);
  EXPECT(std::strlen(ffs_last_error()) > 0);
}

static void fill_and_csv() {
  std::vector<double> v(1000);
  EXPECT(ffs_fill_uniform(42, v.data(), 1000) == FFS_OK);
  for (double x : v) EXPECT(x >= 0.0 && x < 1.0);
  EXPECT(ffs_fill_uniform(42, v.data(), 0) == FFS_OK);
  EXPECT(ffs_fill_uniform(42, nullptr, 5) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_fill_uniform(42, v.data(), -1) == FFS_ERR_INVALID_ARG);
  const int64_t sizes[3] = {8192, 8192, 1};
  char line[128];
  EXPECT(ffs_format_csv("box_2d2r", 2, sizes, 256, 22.5, line, sizeof line) == FFS_OK);
  const char* want = "FlashFFTStencil, box_2d2r, 1, 8192, 8192, 256,";
  EXPECT(std::strncmp(line, want, std::strlen(want)) == 0);
  for (int cap = 0; cap < 60; cap += 7) {  // short buffers: an error, never an overrun
    std::vector<char> b(cap > 0 ? cap : 1);
    const ffs_status st = ffs_format_csv("box_2d2r", 2, sizes, 256, 22.5, cap > 0 ? b.data() : nullptr, cap);
    EXPECT(st != FFS_OK);
  }
  EXPECT(ffs_format_csv(nullptr, 2, sizes, 256, 1.0, line, sizeof line) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_format_csv("x", 4, sizes, 256, 1.0, line, sizeof line) == FFS_ERR_INVALID_ARG);
}

static ffs_problem base_problem(std::vector<double>& w) {
  ffs_problem p;
  ffs_problem_init(&p);
  w.assign(9, 1.0 / 9.0);
  p.dims = 2;
  p.radius = 1;
  p.sizes[0] = 64;
  p.sizes[1] = 48;
  p.weights = w.data();
  p.steps = 4;
  return p;
}

static void validation() {
  std::vector<double> w;
  ffs_plan* plan = reinterpret_cast<ffs_plan*>(0x1);
  EXPECT(ffs_plan_create(nullptr, &plan) == FFS_ERR_INVALID_ARG && plan == nullptr);
  ffs_problem p = base_problem(w);
  EXPECT(ffs_plan_create(&p, nullptr) == FFS_ERR_INVALID_ARG);
  struct Bad {
    const char* what;
    void (*edit)(ffs_problem&);
  };
  const Bad bad[] = {
      {"dims 0", [](ffs_problem& q) { q.dims = 0; }},
      {"dims 4", [](ffs_problem& q) { q.dims = 4; }},
      {"radius -1", [](ffs_problem& q) { q.radius = -1; }},
      {"size 0", [](ffs_problem& q) { q.sizes[1] = 0; }},
      {"size -5", [](ffs_problem& q) { q.sizes[0] = -5; }},
      {"steps -1", [](ffs_problem& q) { q.steps = -1; }},
      {"weights NULL", [](ffs_problem& q) { q.weights = nullptr; }},
      {"precision 9", [](ffs_problem& q) { q.precision = 9; }},
      {"convention 7", [](ffs_problem& q) { q.convention = 7; }},
      {"algorithm 5", [](ffs_problem& q) { q.algorithm = 5; }},
      {"mapping 4", [](ffs_problem& q) { q.mapping = 4; }},
      {"fuse -2", [](ffs_problem& q) { q.fuse = -2; }},
      {"tile 8", [](ffs_problem& q) { q.tile = 8; }},
      {"device -1", [](ffs_problem& q) { q.device = -1; }},
      {"flags 0x80", [](ffs_problem& q) { q.flags = 0x80; }},
      {"bf16x2", [](ffs_problem& q) { q.precision = FFS_PREC_MIXED_BF16X2; }},
      {"f16 in 3D", [](ffs_problem& q) {
         static std::vector<double> w3(27, 1.0 / 27.0);
         q.dims = 3;
         q.sizes[2] = 16;
         q.weights = w3.data();
         q.precision = FFS_PREC_MIXED_F16;
       }},
      {"f16 in 1D", [](ffs_problem& q) {
         static std::vector<double> w1(3, 1.0 / 3.0);
         q.dims = 1;
         q.radius = 1;
         q.weights = w1.data();
         q.precision = FFS_PREC_MIXED_F16;
       }},
      {"fuse too deep", [](ffs_problem& q) { q.fuse = 40; q.tile = 16; }},
  };
  for (const Bad& b : bad) {
    ffs_problem q = base_problem(w);
    b.edit(q);
    plan = reinterpret_cast<ffs_plan*>(0x1);
    const ffs_status st = ffs_plan_create(&q, &plan);
    if (st == FFS_OK || plan != nullptr) {
      std::fprintf(stderr, "validation accepted: %s\n", b.what);
      failures++;
    }
    if (plan) ffs_plan_destroy(plan);
    EXPECT(std::strlen(ffs_last_error()) > 0);
  }
  // batch creation: argument checks before any device work
  ffs_batch* batch = nullptr;
  EXPECT(ffs_batch_create(nullptr, 1, &batch) != FFS_OK);
  ffs_problem probs[2] = {base_problem(w), base_problem(w)};
  EXPECT(ffs_batch_create(probs, 0, &batch) != FFS_OK);
  EXPECT(ffs_batch_create(probs, -3, &batch) != FFS_OK);
  probs[1].precision = FFS_PREC_MIXED_F32;
  EXPECT(ffs_batch_create(probs, 2, &batch) != FFS_OK);
  EXPECT(batch == nullptr);
  // entry points given NULL plans
  ffs_plan_info info;
  EXPECT(ffs_plan_get_info(nullptr, &info) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_run(nullptr, nullptr, nullptr, nullptr) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_run_block(nullptr, nullptr, nullptr, 0, nullptr) == FFS_ERR_INVALID_ARG);
  EXPECT(ffs_plan_destroy(nullptr) == FFS_OK);
  EXPECT(ffs_batch_destroy(nullptr) == FFS_OK);
  EXPECT(ffs_problem_init(nullptr) == FFS_ERR_INVALID_ARG);
}

int main() {
  EXPECT(std::string(ffs_version()).size() > 0);
  named_stencils();
  fill_and_csv();
  validation();
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("host ABI checks passed\n");
  return 0;
}
