// ffs_api.cpp -- host side of libffstencil: validation, named stencils, plan creation, the fused
// time-block loop and its hipGraph replay.  See include/ffstencil.h for the contract.
#include "../../include/ffstencil.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ffs_internal.h"

namespace {

thread_local std::string g_err;

ffs_status fail(ffs_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

#define FFS_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return fail(FFS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

struct NamedStencil {
  const char* name;
  const char* alias;
  const char* tag;  // reference CSV tag (src/1D/1d_main.cu:188-193, src/2D/2d_main.cu:180-183, src/3D/3d_main.cu:180-183)
  int dims, radius;
  bool star;  // support = axes through the centre only
};

const NamedStencil kNamed[] = {
    {"Heat-1D", nullptr, "1d1r", 1, 1, false},      {"1D5P", nullptr, "1d2r", 1, 2, false},
    {"1D7P", nullptr, "1d3r", 1, 3, false},         {"Heat-2D", nullptr, "star_2d1r", 2, 1, true},
    {"Box2D9P", nullptr, "box_2d1r", 2, 1, false},  {"Star2D9P", nullptr, "star_2d2r", 2, 2, true},
    {"Box2D25P", nullptr, "box_2d2r", 2, 2, false}, {"Heat-3D", nullptr, "star_3d1r", 3, 1, true},
    {"Box3D27P", "Box3d27P", "box_3d1r", 3, 1, false},
};

}  // namespace

struct ffs_plan {
  ffs_problem prob{};
  std::vector<double> weights;
  int kind = 0;
  bool direct = false;  // FFS_ALGO_DIRECT path (no transform tables)
  bool star = false;    // direct path: taps off the axes are all zero
  int dev = 0;
  int num_cus = 256;
  int32_t P[3] = {1, 1, 1};      // DFT length per axis (1D: 240 on axis "1")
  int32_t L[3] = {16, 16, 16};   // per-axis DFT lengths of the operand tables
  int32_t V[3] = {1, 1, 1};
  int64_t t = 0, t_tail = 0, nblocks = 0;
  int64_t tiles[3] = {1, 1, 1};
  int64_t pairs = 0;
  int32_t h_main[3] = {0, 0, 0}, h_tail[3] = {0, 0, 0};
  void* d_opsi = nullptr;  // 32-point family with V <= 16: inverse operands of the main and tail depths
  double* d_ops = nullptr;
  uint32_t* d_opsm = nullptr;  // mixed-form operand table
  double* d_H = nullptr;       // [main | tail], each 2*len doubles
  float* d_Hf = nullptr;       // mixed forms: [main | tail] FP32 copies, each len (re, im) pairs
  double dc_main = 1.0, dc_tail = 1.0;  // K(0)^t and K(0)^t_tail
  int32_t scale_main = 0, scale_tail = 0;  // FP16 form: scale exponents of the main and tail depths
  double* d_taps = nullptr;
  double* d_scratch = nullptr;  // ping-pong grid
  int64_t npoints = 0;
  int64_t hlen = 0;
  hipStream_t cap_stream = nullptr;
  // Captured block chains for the last few (d_in, d_out) pairs, least recently used replaced first.
  struct GraphSlot {
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const double* in = nullptr;
    double* out = nullptr;
    uint64_t used = 0;
  };
  static constexpr int kGraphSlots = 4;
  GraphSlot graphs[kGraphSlots];
  uint64_t graph_clock = 0;
};

struct ffs_batch {
  std::vector<ffs_plan*> plans;
  int32_t precision = 0;
  int dev = 0;
  int num_cus = 256;
  int64_t rounds = 0;
  ffsk::TileArgs* d_items = nullptr;  // [round][grid]
  int64_t* d_prefix = nullptr;        // [round][grid + 1]
  std::vector<int32_t> active;        // grids with a block in each round
  std::vector<const double*> g_in;
  std::vector<double*> g_out;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

namespace {

constexpr int64_t kGraphMinLaunches = 4;  // chains up to this length are launched without a graph

ffs_status validate(const ffs_problem* p, int64_t* npoints) {
  if (!p) return fail(FFS_ERR_INVALID_ARG, "problem is NULL");
  if (p->dims < 1 || p->dims > 3) return fail(FFS_ERR_INVALID_ARG, "dims must be 1, 2 or 3 (got %d)", p->dims);
  if (p->radius < 0) return fail(FFS_ERR_INVALID_ARG, "radius must be >= 0 (got %d)", p->radius);
  if (p->steps < 0) return fail(FFS_ERR_INVALID_ARG, "steps must be >= 0 (got %lld)", (long long)p->steps);
  if (!p->weights) return fail(FFS_ERR_INVALID_ARG, "weights pointer is NULL");
  if (p->precision < FFS_PREC_FP64 || p->precision > FFS_PREC_MIXED_F16)
    return fail(FFS_ERR_INVALID_ARG, "unknown precision %d", p->precision);
  if (p->precision == FFS_PREC_MIXED_BF16X2)
    return fail(FFS_ERR_UNSUPPORTED,
                "the 2-term BF16 form is withdrawn (it misses 1e-6 before the field has smoothed); use "
                "FFS_PREC_MIXED_F16 or FFS_PREC_MIXED_BF16");
  if (p->convention != FFS_CONV_REF && p->convention != FFS_CONV_CENTRED)
    return fail(FFS_ERR_INVALID_ARG, "unknown convention %d", p->convention);
  if (p->fuse < 0) return fail(FFS_ERR_INVALID_ARG, "fuse must be >= 0 (got %d)", p->fuse);
  if (p->algorithm < FFS_ALGO_FFT || p->algorithm > FFS_ALGO_AUTO)
    return fail(FFS_ERR_INVALID_ARG, "unknown algorithm %d", p->algorithm);
  if (p->mapping < FFS_MAP_AUTO || p->mapping > FFS_MAP_LINEAR)
    return fail(FFS_ERR_INVALID_ARG, "unknown mapping %d", p->mapping);
  if (p->flags & ~FFS_FLAG_FULL_STAGES) return fail(FFS_ERR_INVALID_ARG, "unknown flags 0x%x", p->flags);
  int64_t n = 1;
  for (int d = 0; d < p->dims; d++) {
    if (p->sizes[d] < 1) return fail(FFS_ERR_INVALID_ARG, "size[%d] must be >= 1 (got %lld)", d, (long long)p->sizes[d]);
    if (n > (int64_t(1) << 40) / p->sizes[d]) return fail(FFS_ERR_INVALID_ARG, "grid too large");
    n *= p->sizes[d];
  }
  const int K = 2 * p->radius + 1;
  int64_t ntaps = 1;
  for (int d = 0; d < p->dims; d++) ntaps *= K;
  for (int64_t i = 0; i < ntaps; i++)
    if (!std::isfinite(p->weights[i])) return fail(FFS_ERR_INVALID_ARG, "weight %lld is not finite", (long long)i);
  *npoints = n;
  return FFS_OK;
}

// Split T into blocks of depth t (balanced when t was chosen automatically) plus a tail.
void set_blocks(ffs_plan* pl, int64_t tbest, bool forced) {
  const int64_t T = pl->prob.steps;
  if (T == 0) {
    pl->t = tbest;
    pl->nblocks = 0;
    pl->t_tail = 0;
    return;
  }
  if (tbest > T) tbest = T;
  int64_t nb = (T + tbest - 1) / tbest;
  if (!forced) {  // same number of blocks, shallowest depth that still covers T
    tbest = (T + nb - 1) / nb;
    nb = (T + tbest - 1) / tbest;
  }
  pl->t = tbest;
  pl->nblocks = nb;
  const int64_t tail = T - (nb - 1) * tbest;
  pl->t_tail = (tail == tbest) ? 0 : tail;
}

// Taps off the axes all zero (Heat-2D, Star2D9P, Heat-3D and custom weights of that form)?
bool star_support(const ffs_plan* pl) {
  const ffs_problem& p = pl->prob;
  const int r = p.radius, K = 2 * r + 1;
  if (p.dims == 1 || r == 0) return false;
  for (int64_t i = 0; i < (int64_t)pl->weights.size(); i++) {
    int off = 0;
    int64_t q = i;
    for (int d = 0; d < p.dims; d++, q /= K) off += (q % K) != r;
    if (off > 1 && pl->weights[i] != 0.0) return false;
  }
  return true;
}

// Direct path: fused depth per launch.  Per point-step cost max(compute, memory): FMAs x window/valid area
// x pipeline fill along axis 0 at ~24e12 FP64 FMA/s, against 16 B x window/valid area / t at an effective
// 2.5e12 B/s (fitted to MI355X timings of t = 1..6: the row stream is latency- rather than
// bandwidth-limited, so deeper fusion pays longer than the 8 TB/s peak suggests).
ffs_status choose_config_direct(ffs_plan* pl) {
  const ffs_problem& p = pl->prob;
  const int r = p.radius;
  pl->star = star_support(pl);
  const int tmax = ffsk::direct_max_depth(p.dims, r, pl->star);
  if (tmax < 1) return fail(FFS_ERR_UNSUPPORTED, "direct algorithm: radius %d not offered in %dD", r, p.dims);
  pl->direct = true;
  pl->kind = -1;
  int64_t tbest;
  if (p.fuse > 0) {
    if (p.fuse > tmax)
      return fail(FFS_ERR_UNSUPPORTED, "direct algorithm: fuse %d too deep for radius %d in %dD (max %d)", p.fuse, r,
                  p.dims, tmax);
    tbest = p.fuse;
  } else if (p.dims == 1) {
    tbest = tmax;
  } else {
    const int K = 2 * r + 1;
    double taps = p.dims == 2 ? K * K : K * K * K;
    if (pl->star) taps = 1 + 2 * r * p.dims;
    double best = 1e300;
    tbest = 1;
    for (int t = 1; t <= tmax; t++) {
      const double area = p.dims == 2
                              ? (double)ffsk::kDir2Cols / (ffsk::kDir2Cols - 2 * r * t)
                              : (double)ffsk::kDir3Rows * ffsk::kDir3Cols /
                                    ((ffsk::kDir3Rows - 2 * r * t) * (ffsk::kDir3Cols - 2 * r * t));
      const double fill = 1.0 + (2.0 * r * t + t - 1) / 256.0;
      const double cost = std::max(taps * area * fill / 24e12, 16.0 * area / t / 2.5e12);
      if (cost < best * (1 - 1e-9)) {
        best = cost;
        tbest = t;
      }
    }
  }
  set_blocks(pl, tbest, p.fuse > 0);
  return FFS_OK;
}

// Choose tile family, fused depth and valid widths.
ffs_status choose_config(ffs_plan* pl) {
  const ffs_problem& p = pl->prob;
  const int r = p.radius;
  if (p.algorithm == FFS_ALGO_DIRECT && p.precision != FFS_PREC_FP64)
    return fail(FFS_ERR_UNSUPPORTED, "the direct algorithm computes in FP64: use FFS_PREC_FP64");
  if (p.algorithm == FFS_ALGO_DIRECT) return choose_config_direct(pl);
  // AUTO: direct wherever it is offered (MI355X, GStencils/s direct vs FP64 FFT: Box2D25P 8192^2 756 vs 152,
  // Box3D27P 512^3 386 vs 121, 1D7P 2^20 1478 vs 1013, Heat-1D 2^26 5667 vs 4120).
  if (p.algorithm == FFS_ALGO_AUTO && ffsk::direct_max_depth(p.dims, r, star_support(pl)) > 0) {
    pl->prob.precision = FFS_PREC_FP64;
    return choose_config_direct(pl);
  }
  if (p.precision == FFS_PREC_MIXED_F16 && p.dims != 2)
    return fail(FFS_ERR_UNSUPPORTED, "the FP16 mixed form is offered for 2D grids (32-point tiles) only");
  if (p.dims == 1) {
    if (p.tile != 0 && p.tile != 240)
      return fail(FFS_ERR_UNSUPPORTED, "1D tile must be 240 (15x16 prime-factor window), got %d", p.tile);
    pl->kind = ffsk::KIND_1DPFA;
    pl->P[1] = 240;
    pl->L[0] = 15;
    pl->L[1] = 16;
  } else if (p.dims == 2) {
    const bool bf = p.precision == FFS_PREC_MIXED_BF16;
    const bool f16 = p.precision == FFS_PREC_MIXED_F16;
    const bool fp64 = p.precision == FFS_PREC_FP64;
    const int P = p.tile != 0 ? p.tile : (bf || f16 ? 32 : 16);
    if (!((P == 16 && !f16) || (P == 32 && (bf || f16 || fp64))))
      return fail(FFS_ERR_UNSUPPORTED,
                  "2D tile %d not available for this precision (16 or 32 with FP64 and the BF16 forms, 32 with FP16, "
                  "16 with FP32)", P);
    pl->kind = P == 32 ? ffsk::KIND_2D32 : ffsk::KIND_2D16;
    pl->P[0] = pl->P[1] = P;
    pl->L[0] = pl->L[1] = P;
  } else {
    if (p.tile != 0 && p.tile != 16) return fail(FFS_ERR_UNSUPPORTED, "3D tile must be 16, got %d", p.tile);
    pl->kind = ffsk::KIND_3D16;
    pl->P[0] = pl->P[1] = pl->P[2] = 16;
  }
  const int vmin = p.dims == 1 ? 16 : 2;
  const int dd = p.dims == 1 ? 1 : p.dims;
  // Minimise the modelled time of the whole run: nblocks x (launch + pairs x pair cost / 1024 SIMDs),
  // with the depth rebalanced as set_blocks does (same block count, shallowest depth; the tail block
  // shares the main block's tiling, so it costs the same).  The cost of a pair is max(compute, hbm):
  // compute fixed by the tile family and precision (cycles per SIMD, measured on MI355X), HBM 16 B per
  // valid point at ~3.26 B/cycle/SIMD (8 TB/s over 1024 SIMDs at 2.4 GHz); a launch adds a fixed ramp-up
  // and drain (~4 us), which decides the depth only for small grids.  Returns point-steps per modelled
  // cycle of the best (or the forced) depth of a family, -1 when none fits.
  auto search = [&](int kind, int Pm, int64_t& t_out) -> double {
    const int64_t tmax = r == 0 ? (int64_t)1 << 40 : (Pm - vmin) / (2 * r);
    if (tmax < 1 || (p.fuse > 0 && p.fuse > tmax)) return -1.0;
    double compute = p.precision == FFS_PREC_FP64 ? 4700.0 : (p.precision == FFS_PREC_MIXED_F32 ? 2600.0 : 2500.0);
    // 32-point pairs (full stages; SIMD cycles per pair from GRBM_GUI_ACTIVE of the reduced-stage kernels
    // on Box2D25P 8192^2, divided by their factor below): FP16 5550 / 0.63, BF16 8000 / 0.70
    if (kind == ffsk::KIND_2D32)
      compute = p.precision == FFS_PREC_MIXED_BF16 ? 12000.0
                : p.precision == FFS_PREC_MIXED_F16 ? 8800.0
                : p.precision == FFS_PREC_FP64     ? 31000.0  // 384 FP64 MFMAs per pair
                                                   : 9000.0;
    // 32-point family with V <= 16: reduced inverse stages, measured at 0.63 (FP16), 0.70 (BF16) and
    // 0.74 (FP64) of a full tile pair on MI355X
    const double cut = p.precision == FFS_PREC_FP64 ? 0.74 : 0.65;
    const double hbm_per_point = 2.0 * 16.0 / 3.26;
    const double launch = 8000.0;
    const int64_t T = std::max<int64_t>(p.steps, 1);
    double points = 1.0;
    for (int d = 0; d < p.dims; d++) points *= (double)p.sizes[d];
    double best = -1;
    const int64_t t0 = p.fuse > 0 ? p.fuse : 1, t1 = p.fuse > 0 ? p.fuse : std::min<int64_t>(tmax, 4096);
    for (int64_t t = t0; t <= t1; t++) {
      const int64_t nb = (T + t - 1) / t;
      const int64_t te = p.fuse > 0 ? t : (T + nb - 1) / nb;  // the depth set_blocks runs
      const int64_t V = Pm - 2 * r * te;
      const double Vd = std::pow((double)V, dd);
      double pairs = 1.0;  // tiles per axis, pairs along the last one
      for (int d = 0; d < p.dims; d++) {
        const double tiles = std::ceil((double)p.sizes[d] / (double)V);
        pairs *= d == p.dims - 1 ? std::ceil(tiles / 2.0) : tiles;
      }
      const bool out16 = kind == ffsk::KIND_2D32 && V <= 16 && !(p.flags & FFS_FLAG_FULL_STAGES);
      // 3D: 16 slices per pair; its FP32 form is the compensated one, two products per real GEMM
      const double c = compute * (p.dims == 3 ? (p.precision == FFS_PREC_MIXED_F32 ? 32.0 : 16.0) : 1.0) *
                       (out16 ? cut : 1.0);
      const double cycles = (double)nb * (launch + pairs * std::max(c, hbm_per_point * Vd) / 1024.0);
      const double score = points * (double)T / cycles;
      if (score > best * (1 + 1e-9)) {
        best = score;
        t_out = te;
      }
    }
    return best;
  };
  int64_t tbest = 1;
  double score = search(pl->kind, p.dims == 1 ? 240 : pl->P[0], tbest);
  if (p.dims == 2 && p.precision == FFS_PREC_FP64 && p.tile == 0) {  // FP64: 16- or 32-point tiles
    int64_t t32 = 1;
    const double s32 = search(ffsk::KIND_2D32, 32, t32);
    if (s32 > score * (1 + 1e-9)) {
      pl->kind = ffsk::KIND_2D32;
      pl->P[0] = pl->P[1] = pl->L[0] = pl->L[1] = 32;
      score = s32;
      tbest = t32;
    }
  }
  const int Pmain = p.dims == 1 ? 240 : pl->P[0];
  if (score < 0) {
    if (p.fuse > 0)
      return fail(FFS_ERR_UNSUPPORTED, "fuse %d too deep for radius %d and tile %d", p.fuse, r, Pmain);
    return fail(FFS_ERR_UNSUPPORTED, "radius %d leaves no valid region in a %d-wide tile", r, Pmain);
  }
  set_blocks(pl, tbest, p.fuse > 0);
  const int hi = p.convention == FFS_CONV_REF ? 2 * r : r;  // largest tap offset of one step
  for (int d = 0; d < 3; d++) pl->V[d] = 1;
  if (p.dims == 1) {
    pl->V[1] = (int32_t)(240 - 2 * r * pl->t);
    pl->h_main[1] = (int32_t)(hi * pl->t);
    pl->h_tail[1] = (int32_t)(hi * pl->t_tail);
    pl->tiles[1] = (p.sizes[0] + pl->V[1] - 1) / pl->V[1];
    pl->pairs = (pl->tiles[1] + 1) / 2;
  } else {
    for (int d = 0; d < p.dims; d++) {
      pl->V[d] = (int32_t)(Pmain - 2 * r * pl->t);
      pl->h_main[d] = (int32_t)(hi * pl->t);
      pl->h_tail[d] = (int32_t)(hi * pl->t_tail);
      pl->tiles[d] = (p.sizes[d] + pl->V[d] - 1) / pl->V[d];
    }
    if (p.dims == 2) pl->pairs = pl->tiles[0] * ((pl->tiles[1] + 1) / 2);
    if (p.dims == 3) pl->pairs = pl->tiles[0] * pl->tiles[1] * ((pl->tiles[2] + 1) / 2);
  }
  return FFS_OK;
}

void drop_graph(ffs_plan::GraphSlot& g) {
  if (g.gexec) (void)hipGraphExecDestroy(g.gexec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g = ffs_plan::GraphSlot{};
}

void release(ffs_plan* pl) {
  if (!pl) return;
  for (ffs_plan::GraphSlot& g : pl->graphs) drop_graph(g);
  if (pl->cap_stream) (void)hipStreamDestroy(pl->cap_stream);
  if (pl->d_ops) (void)hipFree(pl->d_ops);
  if (pl->d_opsm) (void)hipFree(pl->d_opsm);
  if (pl->d_opsi) (void)hipFree(pl->d_opsi);
  if (pl->d_H) (void)hipFree(pl->d_H);
  if (pl->d_Hf) (void)hipFree(pl->d_Hf);
  if (pl->d_taps) (void)hipFree(pl->d_taps);
  if (pl->d_scratch) (void)hipFree(pl->d_scratch);
  delete pl;
}

// 32-point family: does the valid region fit one 16 x 16 block (reduced inverse stages)?  The main and
// the tail block share the tiling, so both are out16 or neither; their inverse operands differ by shift.
bool out16_block(const ffs_plan* pl) {
  return pl->kind == ffsk::KIND_2D32 && pl->V[0] <= 16 && pl->V[1] <= 16 &&
         !(pl->prob.flags & FFS_FLAG_FULL_STAGES);
}

// Wave-to-tile mapping of the FFT tile kernels: true = hardware order.  AUTO takes XCD-contiguous bands
// for the FP64 32-point kernel (measured 4% faster on Box2D25P 8192^2: 1073 vs 1116 us per block) and
// hardware order for the others (1-3% faster for the mixed forms).
bool linear_map(const ffs_plan* pl) {
  const ffs_problem& p = pl->prob;
  if (p.mapping != FFS_MAP_AUTO) return p.mapping == FFS_MAP_LINEAR;
  // XCD bands where they measured faster: the 32-point FP64 kernel (4%) and the 3D mixed kernel (0.8%);
  // the hardware order elsewhere (2D mixed: XCD bands 4% slower; 3D FP64: 0.5% slower).
  const bool xcd = (pl->kind == ffsk::KIND_2D32 && p.precision == FFS_PREC_FP64) ||
                   (pl->kind == ffsk::KIND_3D16 && p.precision != FFS_PREC_FP64);
  return !xcd;
}

ffsk::TileArgs make_args(const ffs_plan* pl, const double* src, double* dst, bool tail) {
  ffsk::TileArgs a{};
  a.src = src;
  a.dst = dst;
  a.ops = pl->d_ops;
  a.H = pl->d_H + (tail ? 2 * pl->hlen : 0);
  a.Hf = pl->d_Hf ? pl->d_Hf + (tail ? 2 * pl->hlen : 0) : nullptr;
  a.opsm = pl->d_opsm;
  a.unit = 1.0f;
  a.dc_gain = tail ? pl->dc_tail : pl->dc_main;
  const ffs_problem& p = pl->prob;
  if (p.dims == 1) {
    a.n0 = 1;
    a.n1 = p.sizes[0];
    a.n2 = 1;
  } else {
    a.n0 = p.sizes[0];
    a.n1 = p.sizes[1];
    a.n2 = p.dims == 3 ? p.sizes[2] : 1;
  }
  a.tiles0 = pl->tiles[0];
  a.tiles1 = pl->tiles[1];
  a.tiles2 = pl->tiles[2];
  a.pairs = pl->pairs;
  a.V0 = pl->V[0];
  a.V1 = pl->V[1];
  a.V2 = pl->V[2];
  const int32_t* h = tail ? pl->h_tail : pl->h_main;
  a.h0 = h[0];
  a.h1 = h[1];
  a.h2 = h[2];
  a.precision = p.precision;
  a.scale_exp = tail ? pl->scale_tail : pl->scale_main;
  a.linear_map = linear_map(pl) ? 1 : 0;
  a.out16 = out16_block(pl) ? 1 : 0;
  if (a.out16) {
    const size_t per = pl->prob.precision == FFS_PREC_FP64 ? sizeof(double) * ffsk::kInv32F64
                                                            : sizeof(uint32_t) * ffsk::kInv32Words;
    a.opsi = static_cast<const char*>(pl->d_opsi) + (tail ? per : 0);
  }
  return a;
}

// Direct path: launch geometry of one block of depth t (the window is fixed, the valid part shrinks
// with t).  About two workgroups per CU along axis 0, each streaming a segment of rows (planes).
ffsk::DirectArgs make_direct_args(const ffs_plan* pl, const double* src, double* dst, int64_t t) {
  const ffs_problem& p = pl->prob;
  const int r = p.radius, K = 2 * r + 1;
  ffsk::DirectArgs a{};
  a.src = src;
  a.dst = dst;
  a.n0 = p.sizes[0];
  a.n1 = p.dims > 1 ? p.sizes[1] : 1;
  a.n2 = p.dims > 2 ? p.sizes[2] : 1;
  a.steps = (int32_t)t;
  a.lo = p.convention == FFS_CONV_REF ? 2 * r : r;
  // window-order taps: g[e] multiplies u[i + e]; ref convention flips every axis (w[2r - e])
  const int64_t ntaps = (int64_t)pl->weights.size();
  for (int64_t i = 0; i < ntaps; i++) {
    int64_t src_i = i;
    if (p.convention == FFS_CONV_REF) {
      src_i = 0;
      int64_t q = i, mul = 1;
      for (int d = 0; d < p.dims; d++, q /= K, mul *= K) src_i += (2 * r - q % K) * mul;
    }
    a.g[i] = pl->weights[src_i];
  }
  auto cdiv = [](int64_t x, int64_t y) { return (x + y - 1) / y; };
  const int64_t shrink = 2 * (int64_t)r * t;
  if (p.dims == 1) {
    a.valid1 = (int32_t)(ffsk::kDir1Window - shrink);
    a.nseg = (int32_t)cdiv(a.n0, a.valid1);
    a.strips1 = a.strips2 = 1;
    a.seg = a.n0;
    return a;
  }
  int64_t strips;
  if (p.dims == 2) {
    a.valid1 = (int32_t)(ffsk::kDir2Cols - shrink);
    a.strips1 = (int32_t)cdiv(a.n1, a.valid1);
    a.strips2 = 1;
    strips = a.strips1;
  } else {
    a.valid1 = (int32_t)(ffsk::kDir3Rows - shrink);
    a.valid2 = (int32_t)(ffsk::kDir3Cols - shrink);
    a.strips1 = (int32_t)cdiv(a.n1, a.valid1);
    a.strips2 = (int32_t)cdiv(a.n2, a.valid2);
    strips = (int64_t)a.strips1 * a.strips2;
  }
  int64_t nseg = std::max<int64_t>(1, (2 * (int64_t)pl->num_cus + strips / 2) / strips);
  nseg = std::min<int64_t>(nseg, cdiv(a.n0, 16));
  a.seg = cdiv(a.n0, nseg);
  a.nseg = (int32_t)cdiv(a.n0, a.seg);
  return a;
}

hipError_t launch_one_block(ffs_plan* pl, const double* src, double* dst, bool tail, hipStream_t s) {
  if (pl->direct) {
    const ffsk::DirectArgs a = make_direct_args(pl, src, dst, tail ? pl->t_tail : pl->t);
    return ffsk::launch_direct(a, pl->prob.dims, pl->prob.radius, pl->star, s);
  }
  ffsk::TileArgs a = make_args(pl, src, dst, tail);
  return ffsk::launch_block(a, pl->kind, s, pl->num_cus);
}

// Enqueue the whole block chain d_in -> ... -> d_out on stream s.
ffs_status enqueue_chain(ffs_plan* pl, const double* d_in, double* d_out, hipStream_t s) {
  const int64_t nb = pl->nblocks;
  const double* src = d_in;
  for (int64_t b = 0; b < nb; b++) {
    double* dst = ((nb - 1 - b) % 2 == 0) ? d_out : pl->d_scratch;
    const bool tail = (b == nb - 1) && pl->t_tail > 0;
    hipError_t e = launch_one_block(pl, src, dst, tail, s);
    if (e != hipSuccess) return fail(FFS_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    src = dst;
  }
  return FFS_OK;
}

// ---- plan set-up steps of ffs_plan_create ----

// Validation, the weights and the planner's choice (host only, no HIP call).
ffs_status plan_host(const ffs_problem* p, ffs_plan* pl) {
  int64_t npoints = 0;
  ffs_status st = validate(p, &npoints);
  if (st != FFS_OK) return st;
  pl->prob = *p;
  const int K = 2 * p->radius + 1;
  int64_t ntaps = 1;
  for (int d = 0; d < p->dims; d++) ntaps *= K;
  pl->weights.assign(p->weights, p->weights + ntaps);
  pl->prob.weights = pl->weights.data();
  pl->npoints = npoints;
  return choose_config(pl);
}

// The plan's device exists and is a gfx950; makes it current and records its CU count (no allocation).
ffs_status plan_device(ffs_plan* pl) {
  int ndev = 0;
  const hipError_t ce = hipGetDeviceCount(&ndev);
  if (ce != hipSuccess || ndev <= 0)
    return fail(FFS_ERR_NO_DEVICE, "no HIP device visible (%s, count %d)", hipGetErrorString(ce), ndev);
  if (pl->prob.device < 0 || pl->prob.device >= ndev)
    return fail(FFS_ERR_INVALID_ARG, "device %d out of range (%d devices)", pl->prob.device, ndev);
  pl->dev = pl->prob.device;
  hipDeviceProp_t prop;
  if (hipSetDevice(pl->dev) != hipSuccess || hipGetDeviceProperties(&prop, pl->dev) != hipSuccess)
    return fail(FFS_ERR_HIP, "cannot query device %d", pl->dev);
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(FFS_ERR_NO_DEVICE, "device %d is %s; this library is built for gfx950 only", pl->dev, prop.gcnArchName);
  pl->num_cus = prop.multiProcessorCount;
  return FFS_OK;
}

// Host constants: the constant-mode gains K(0)^t and, for the FP16 form, the tile scale exponents.
ffs_status plan_constants(ffs_plan* pl) {
  if (pl->direct) return FFS_OK;
  const ffs_problem& p = pl->prob;
  const int K = 2 * p.radius + 1;
  pl->hlen = ffsk::spectrum_len(pl->kind);
  {  // gain of the constant mode, K(0)^t = (sum w)^t, by repeated squaring in FP64
    double sw = 0.0;
    for (double wv : pl->weights) sw += wv;
    auto ipow = [](double b, int64_t e) {
      double r = 1.0;
      for (; e > 0; e >>= 1, b *= b)
        if (e & 1) r *= b;
      return r;
    };
    pl->dc_main = ipow(sw, pl->t);
    pl->dc_tail = ipow(sw, pl->t_tail > 0 ? pl->t_tail : pl->t);
  }
  if (p.precision == FFS_PREC_MIXED_F16) {
    // largest |K(k)| over the 32 x 32 frequencies (direct DFT of the taps; |K| does not depend on the
    // convention), then scale_exp = floor(10 - log2 max(1, |K|^t)) for each depth
    double kmax = 0.0;
    const int P = 32;
    for (int k0 = 0; k0 < P; k0++)
      for (int k1 = 0; k1 < P; k1++) {
        double re = 0.0, im = 0.0;
        for (int a0 = 0; a0 < K; a0++)
          for (int a1 = 0; a1 < K; a1++) {
            const double ang = -2.0 * M_PI * (double)((a0 * k0 + a1 * k1) % P) / P;
            re += pl->weights[a0 * K + a1] * std::cos(ang);
            im += pl->weights[a0 * K + a1] * std::sin(ang);
          }
        kmax = std::max(kmax, std::sqrt(re * re + im * im));
      }
    auto expo = [&](int64_t t) {
      const double lg = kmax > 1.0 ? (double)t * std::log2(kmax) : 0.0;
      return std::floor(10.0 - lg);
    };
    const double em = expo(pl->t), et = expo(pl->t_tail > 0 ? pl->t_tail : pl->t);
    if (em < -100.0 || et < -100.0)
      return fail(FFS_ERR_UNSUPPORTED, "FP16 form: |K|^t = 2^%.0f is beyond the FP16 range; use another form",
                  10.0 - std::min(em, et));
    pl->scale_main = (int32_t)em;
    pl->scale_tail = (int32_t)et;
  }
  return FFS_OK;
}

// Bytes of each device table of a transform plan (0 = not used).
struct TableSizes {
  size_t ops, opsm, H, Hf, opsi, taps;
};
TableSizes table_sizes(const ffs_plan* pl) {
  TableSizes t{};
  if (pl->direct) return t;
  t.ops = sizeof(double) * std::max(ffsk::kOpsLen, ffsk::kOps32F64);
  t.opsm = pl->prob.precision != FFS_PREC_FP64 ? sizeof(uint32_t) * std::max(ffsk::kOpsMixedLen, ffsk::kOps32Words) : 0;
  t.H = sizeof(double) * 4 * (size_t)ffsk::spectrum_len(pl->kind);
  t.Hf = pl->prob.precision != FFS_PREC_FP64 ? sizeof(float) * 4 * (size_t)ffsk::spectrum_len(pl->kind) : 0;
  t.opsi = out16_block(pl) ? 2 * std::max(sizeof(double) * ffsk::kInv32F64, sizeof(uint32_t) * ffsk::kInv32Words) : 0;
  t.taps = sizeof(double) * pl->weights.size();
  return t;
}

// DFT operands and spectra into the plan's device tables, on stream s (the taps are already in d_taps).
hipError_t plan_tables(ffs_plan* pl, hipStream_t s) {
  const ffs_problem& p = pl->prob;
  const int K = 2 * p.radius + 1;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (x != hipSuccess && e == hipSuccess) e = x;
  };
  const int layout = p.precision == FFS_PREC_FP64 ? 0 : 1;
  if (out16_block(pl)) {  // inverse operands of window columns h .. h + 15, main and tail depth
    const int32_t hm = pl->h_main[1], ht = pl->t_tail > 0 ? pl->h_tail[1] : pl->h_main[1];
    if (p.precision == FFS_PREC_FP64) {
      double* o = static_cast<double*>(pl->d_opsi);
      chk(ffsk::gen_inverse32_f64(o, hm, s));
      chk(ffsk::gen_inverse32_f64(o + ffsk::kInv32F64, ht, s));
    } else {
      uint32_t* o = static_cast<uint32_t*>(pl->d_opsi);
      chk(ffsk::gen_inverse32(o, p.precision == FFS_PREC_MIXED_F16, hm, s));
      chk(ffsk::gen_inverse32(o + ffsk::kInv32Words, p.precision == FFS_PREC_MIXED_F16, ht, s));
    }
  }
  if (pl->kind == ffsk::KIND_2D32 && p.precision == FFS_PREC_FP64)
    chk(ffsk::gen_operands32_f64(pl->d_ops, s));
  else if (pl->kind == ffsk::KIND_2D32)
    chk(ffsk::gen_operands32(pl->d_opsm, p.precision == FFS_PREC_MIXED_F16, s));
  else
    chk(ffsk::gen_operands(pl->d_ops, pl->d_opsm, pl->L, s));
  chk(ffsk::gen_spectrum(pl->d_H, pl->kind, layout, pl->L, pl->d_taps, K, p.dims, p.convention, p.radius, pl->t, s));
  chk(ffsk::gen_spectrum(pl->d_H + 2 * pl->hlen, pl->kind, layout, pl->L, pl->d_taps, K, p.dims, p.convention,
                         p.radius, pl->t_tail > 0 ? pl->t_tail : pl->t, s));
  if (pl->d_Hf) {
    chk(ffsk::spectrum_to_f32(pl->d_H, pl->d_Hf, pl->hlen, s));
    chk(ffsk::spectrum_to_f32(pl->d_H + 2 * pl->hlen, pl->d_Hf + 2 * pl->hlen, pl->hlen, s));
  }
  return e;
}

// ---- device scratch of the allocation-free entry point (ffs_run_static) ----
// One static array per GPU: the plan tables in the first FFS_STATIC_TABLE_BYTES, the ping-pong grid in the
// rest.  In host code `scratch` is only a host-side stand-in for the symbol; its device address on each
// GPU comes from hipGetSymbolAddress, resolved once per device with that device current and kept in
// g_static_ptr, and only that address is handed to kernels.
constexpr int kStaticMaxDevices = 64;
std::mutex g_static_mu;
void* g_static_ptr[kStaticMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[FFS_STATIC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on device `dev` (which must be current).
ffs_status static_scratch(int dev, void** out) {
  if (dev < 0 || dev >= kStaticMaxDevices) return fail(FFS_ERR_UNSUPPORTED, "device %d: static scratch is kept for "
                                                                            "devices 0..%d", dev, kStaticMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FFS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    g_static_ptr[dev] = p;
  }
  *out = g_static_ptr[dev];
  return FFS_OK;
}

}  // namespace

extern "C" {

ffs_status ffs_problem_init(ffs_problem* p) {
  if (!p) return fail(FFS_ERR_INVALID_ARG, "problem is NULL");
  std::memset(p, 0, sizeof *p);
  p->dims = 2;
  p->radius = 1;
  p->sizes[0] = p->sizes[1] = p->sizes[2] = 1;
  p->steps = 1;
  p->precision = FFS_PREC_FP64;
  p->convention = FFS_CONV_REF;
  p->use_graph = 1;
  return FFS_OK;
}

ffs_status ffs_stencil_named(const char* shape, uint64_t seed, int32_t* dims, int32_t* radius, double* weights,
                             int32_t capacity, char* tag, int32_t tag_capacity) {
  if (!shape) return fail(FFS_ERR_INVALID_ARG, "shape is NULL");
  const NamedStencil* ns = nullptr;
  for (const NamedStencil& n : kNamed)
    if (std::strcmp(shape, n.name) == 0 || (n.alias && std::strcmp(shape, n.alias) == 0)) ns = &n;
  if (!ns) return fail(FFS_ERR_INVALID_ARG, "unknown stencil shape '%s'", shape);
  const int K = 2 * ns->radius + 1;
  int ntaps = 1;
  for (int d = 0; d < ns->dims; d++) ntaps *= K;
  if (dims) *dims = ns->dims;
  if (radius) *radius = ns->radius;
  if (tag) {
    if (tag_capacity < (int32_t)std::strlen(ns->tag) + 1) return fail(FFS_ERR_INVALID_ARG, "tag buffer too small");
    std::strcpy(tag, ns->tag);
  }
  if (!weights) return FFS_OK;
  if (capacity < ntaps) return fail(FFS_ERR_INVALID_ARG, "weights capacity %d < %d taps", capacity, ntaps);
  if (std::strcmp(ns->name, "Heat-1D") == 0) {
    weights[0] = 0.25;
    weights[1] = 0.5;
    weights[2] = 0.25;
    return FFS_OK;
  }
  uint64_t s = seed;
  double sum = 0.0;
  for (int i = 0; i < ntaps; i++) {
    int rem = i, off = 0;
    for (int d = 0; d < ns->dims; d++) {
      const int a = rem % K - ns->radius;
      rem /= K;
      if (a != 0) off++;
    }
    const bool on = !ns->star || off <= 1;
    weights[i] = on ? uniform01(s) : 0.0;
    sum += weights[i];
  }
  if (!(sum > 0.0)) return fail(FFS_ERR_INTERNAL, "degenerate weights");
  for (int i = 0; i < ntaps; i++) weights[i] /= sum;
  return FFS_OK;
}

ffs_status ffs_fill_uniform(uint64_t seed, double* host_out, int64_t n) {
  if (!host_out && n > 0) return fail(FFS_ERR_INVALID_ARG, "output is NULL");
  if (n < 0) return fail(FFS_ERR_INVALID_ARG, "negative count");
  uint64_t s = seed;
  for (int64_t i = 0; i < n; i++) host_out[i] = uniform01(s);
  return FFS_OK;
}

ffs_status ffs_fill_uniform_device(uint64_t seed, double* d_out, int64_t n, void* stream) {
  if (!d_out || n < 0) return fail(FFS_ERR_INVALID_ARG, "bad device buffer or size");
  if (n == 0) return FFS_OK;
  const hipError_t e = ffsk::fill_uniform_device(d_out, n, seed, (hipStream_t)stream);
  if (e != hipSuccess) return fail(FFS_ERR_HIP, "fill kernel launch failed: %s", hipGetErrorString(e));
  return FFS_OK;
}

ffs_status ffs_format_csv(const char* tag, int32_t dims, const int64_t* sizes, int64_t steps, double ms, char* buf,
                          int32_t capacity) {
  if (!tag || !sizes || !buf || capacity <= 0) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  if (dims < 1 || dims > 3) return fail(FFS_ERR_INVALID_ARG, "dims must be 1..3");
  long long d[3] = {1, 1, 1};
  double pts = 1.0;
  for (int i = 0; i < dims; i++) {
    d[3 - dims + i] = (long long)sizes[i];
    pts *= (double)sizes[i];
  }
  const double gst = ms > 0 ? pts * (double)steps / ms / 1e6 : 0.0;
  const int n = std::snprintf(buf, (size_t)capacity, "FlashFFTStencil, %s, %lld, %lld, %lld, %lld, %g, %g", tag, d[0],
                              d[1], d[2], (long long)steps, ms, gst);
  if (n < 0 || n >= capacity) return fail(FFS_ERR_INVALID_ARG, "buffer too small");
  return FFS_OK;
}

ffs_status ffs_plan_create(const ffs_problem* p, ffs_plan** out) {
  if (!out) return fail(FFS_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  ffs_plan* pl = new (std::nothrow) ffs_plan();
  if (!pl) return fail(FFS_ERR_OUT_OF_MEMORY, "host allocation failed");
  ffs_status st = plan_host(p, pl);
  if (st == FFS_OK) st = plan_device(pl);
  if (st == FFS_OK) st = plan_constants(pl);
  if (st != FFS_OK) {
    release(pl);
    return st;
  }
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) {
    if (x != hipSuccess && e == hipSuccess) e = x;
  };
  const TableSizes ts = table_sizes(pl);
  if (!pl->direct) {
    chk(hipMalloc(&pl->d_ops, ts.ops));
    if (ts.opsm) chk(hipMalloc(&pl->d_opsm, ts.opsm));
    chk(hipMalloc(&pl->d_H, ts.H));
    if (ts.Hf) chk(hipMalloc(&pl->d_Hf, ts.Hf));
    if (ts.opsi) chk(hipMalloc(&pl->d_opsi, ts.opsi));
    chk(hipMalloc(&pl->d_taps, ts.taps));
  }
  if (pl->nblocks >= 2) chk(hipMalloc(&pl->d_scratch, sizeof(double) * pl->npoints));
  if (e != hipSuccess) {
    release(pl);
    return fail(e == hipErrorOutOfMemory ? FFS_ERR_OUT_OF_MEMORY : FFS_ERR_HIP, "device allocation failed: %s",
                hipGetErrorString(e));
  }
  if (!pl->direct) {
    chk(hipMemcpy(pl->d_taps, pl->weights.data(), ts.taps, hipMemcpyHostToDevice));
    chk(plan_tables(pl, nullptr));
    chk(hipDeviceSynchronize());
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    release(pl);
    return fail(FFS_ERR_HIP, "plan set-up failed: %s", hipGetErrorString(e));
  }
  *out = pl;
  return FFS_OK;
}

ffs_status ffs_plan_get_info(const ffs_plan* pl, ffs_plan_info* info) {
  if (!pl || !info) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  std::memset(info, 0, sizeof *info);
  const ffs_problem& p = pl->prob;
  info->dims = p.dims;
  info->radius = p.radius;
  for (int d = 0; d < 3; d++) info->sizes[d] = d < p.dims ? p.sizes[d] : 1;
  info->steps = p.steps;
  info->precision = p.precision;
  info->convention = p.convention;
  info->algorithm = pl->direct ? FFS_ALGO_DIRECT : FFS_ALGO_FFT;
  info->mapping = linear_map(pl) ? FFS_MAP_LINEAR : FFS_MAP_XCD;
  if (pl->direct) {
    const ffsk::DirectArgs a = make_direct_args(pl, nullptr, nullptr, pl->t);
    const int r = p.radius, K = 2 * r + 1;
    const int64_t halo = 2 * (int64_t)r * pl->t;
    double taps = std::pow((double)K, p.dims);
    if (pl->star) taps = 1 + 2 * r * p.dims;
    for (int d = 0; d < 3; d++) info->tile[d] = info->valid[d] = 1;
    double computed;  // points the launch computes, all levels (loop rounding to 2r+1 rows ignored)
    if (p.dims == 1) {
      info->tile[0] = ffsk::kDir1Window;
      info->valid[0] = a.valid1;
      info->tiles = a.nseg;
      computed = (double)a.nseg * ffsk::kDir1Window * pl->t;
    } else {
      const double rows = (double)a.n0 + (double)a.nseg * (halo + pl->t - 1);
      info->tile[0] = (int32_t)(a.seg + halo);
      info->valid[0] = (int32_t)a.seg;
      if (p.dims == 2) {
        info->tile[1] = ffsk::kDir2Cols;
        info->valid[1] = a.valid1;
        info->tiles = (int64_t)a.nseg * a.strips1;
        computed = rows * a.strips1 * ffsk::kDir2Cols * pl->t;
      } else {
        info->tile[1] = ffsk::kDir3Rows;
        info->tile[2] = ffsk::kDir3Cols;
        info->valid[1] = a.valid1;
        info->valid[2] = a.valid2;
        info->tiles = (int64_t)a.nseg * a.strips1 * a.strips2;
        computed = rows * a.strips1 * a.strips2 * ffsk::kDir3Rows * ffsk::kDir3Cols * pl->t;
      }
    }
    info->fuse = (int32_t)pl->t;
    info->fuse_tail = (int32_t)pl->t_tail;
    info->blocks = info->launches = pl->nblocks;
    info->flops_per_block = 2.0 * taps * computed;
    info->bytes_per_block = 16.0 * (double)pl->npoints;
    return FFS_OK;
  }
  if (p.dims == 1) {
    info->tile[0] = 240;
    info->valid[0] = pl->V[1];
    info->tile[1] = info->tile[2] = 1;
    info->valid[1] = info->valid[2] = 1;
    info->tiles = pl->tiles[1];
  } else {
    info->tiles = 1;
    for (int d = 0; d < 3; d++) {
      info->tile[d] = d < p.dims ? pl->P[d] : 1;
      info->valid[d] = d < p.dims ? pl->V[d] : 1;
      if (d < p.dims) info->tiles *= pl->tiles[d];
    }
  }
  info->fuse = (int32_t)pl->t;
  info->fuse_tail = (int32_t)pl->t_tail;
  info->blocks = pl->nblocks;
  info->launches = pl->nblocks;
  // Matrix-core flops executed per tile pair: stages (2 per axis) x 16^(d-2) slices x real 16x16x16
  // GEMMs per complex product x MFMA flops per real GEMM.
  //   FP64: Gauss form, 3 GEMMs of 4 x v_mfma_f64_16x16x4f64 (2*16^3 flops each GEMM)
  //   BF16: 4 GEMMs of NTERMS x v_mfma_f32_16x16x32_bf16 (2*16*16*32 flops each instruction)
  //   FP32: 4 GEMMs of 4 x v_mfma_f32_16x16x4f32 (2*16^3 flops each GEMM); twice that in 3D, whose
  //         FP32 form adds the products with the matrix's low-order terms
  const double gemm = 2.0 * 16 * 16 * 16;
  const int dd = p.dims == 1 ? 2 : p.dims;
  double per_cplx = 3.0 * gemm;
  if (p.precision == FFS_PREC_MIXED_BF16) per_cplx = 4.0 * 3.0 * 2.0 * 16 * 16 * 32;
  if (p.precision == FFS_PREC_MIXED_F32) per_cplx = (p.dims == 3 ? 8.0 : 4.0) * gemm;
  double per_pair = 2.0 * dd * per_cplx * std::pow(16.0, dd - 2);
  // 3D: the last inverse stage (phase 3) runs only the V0 slices of the valid block, not all 16
  if (pl->kind == ffsk::KIND_3D16) per_pair = (5.0 * 16.0 + std::min(16, pl->V[0])) * per_cplx;
  // 32-point family: output blocks over the 4 stages (4 + 4 + 4 + 4, or 4 + 4 + 2 + 1 for an out16 main
  // block) x 4 real GEMMs x (6, 4 or 3) K=32 MFMAs
  if (pl->kind == ffsk::KIND_2D32 && p.precision == FFS_PREC_FP64)  // x 3 Gauss GEMMs x 8 K=4 chunks
    per_pair = (out16_block(pl) ? 11.0 : 16.0) * 3.0 * 8.0 * 2.0 * 16 * 16 * 4;
  else if (pl->kind == ffsk::KIND_2D32)
    per_pair = (out16_block(pl) ? 11.0 : 16.0) * 4.0 *
               (p.precision == FFS_PREC_MIXED_BF16 ? 6.0 : (p.precision == FFS_PREC_MIXED_F16 ? 3.0 : 4.0)) *
               2.0 * 16 * 16 * 32;
  info->flops_per_block = per_pair * (double)pl->pairs;
  info->bytes_per_block = 16.0 * (double)pl->npoints;
  return FFS_OK;
}

ffs_status ffs_plan_read_table(const ffs_plan* pl, int32_t which, void* host_out, int64_t capacity,
                               int64_t* bytes) {
  if (!pl) return fail(FFS_ERR_INVALID_ARG, "plan is NULL");
  if (pl->direct) return fail(FFS_ERR_UNSUPPORTED, "the direct algorithm has no transform tables");
  const TableSizes ts = table_sizes(pl);
  const void* src = nullptr;
  size_t n = 0;
  switch (which) {
    case FFS_TABLE_OPS: src = pl->d_ops; n = ts.ops; break;
    case FFS_TABLE_OPS_MIXED: src = pl->d_opsm; n = ts.opsm; break;
    case FFS_TABLE_SPECTRUM: src = pl->d_H; n = ts.H; break;
    case FFS_TABLE_INVERSE16: src = pl->d_opsi; n = ts.opsi; break;
    default: return fail(FFS_ERR_INVALID_ARG, "unknown table %d", which);
  }
  if (bytes) *bytes = (int64_t)n;
  if (!src || n == 0) return fail(FFS_ERR_UNSUPPORTED, "this plan has no table %d", which);
  if (!host_out) return FFS_OK;  // size query
  if (capacity < (int64_t)n) return fail(FFS_ERR_INVALID_ARG, "capacity %lld < %zu bytes", (long long)capacity, n);
  FFS_HIP(hipSetDevice(pl->dev));
  FFS_HIP(hipMemcpy(host_out, src, n, hipMemcpyDeviceToHost));
  return FFS_OK;
}

ffs_status ffs_plan_destroy(ffs_plan* pl) {
  if (!pl) return FFS_OK;
  (void)hipSetDevice(pl->dev);
  (void)hipDeviceSynchronize();
  release(pl);
  return FFS_OK;
}

ffs_status ffs_run(ffs_plan* pl, const double* d_in, double* d_out, void* stream) {
  if (!pl || !d_in || !d_out) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  if (d_in == d_out) return fail(FFS_ERR_INVALID_ARG, "d_in must not alias d_out");
  hipStream_t s = (hipStream_t)stream;
  FFS_HIP(hipSetDevice(pl->dev));
  if (pl->nblocks == 0) {
    FFS_HIP(hipMemcpyAsync(d_out, d_in, sizeof(double) * pl->npoints, hipMemcpyDeviceToDevice, s));
    return FFS_OK;
  }
  // Chains of a few launches go out as plain launches: replaying a graph costs ~9 us more per run on
  // MI355X than launching up to four kernels directly (measured on 1 to 4 block chains).
  if (!pl->prob.use_graph || pl->nblocks <= kGraphMinLaunches) return enqueue_chain(pl, d_in, d_out, s);
  ffs_plan::GraphSlot* slot = nullptr;
  for (ffs_plan::GraphSlot& g : pl->graphs)
    if (g.gexec && g.in == d_in && g.out == d_out) slot = &g;
  if (!slot) {  // capture this pair into the least recently used slot
    slot = &pl->graphs[0];
    for (ffs_plan::GraphSlot& g : pl->graphs)
      if (g.used < slot->used) slot = &g;
    drop_graph(*slot);
    FFS_HIP(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeThreadLocal));
    ffs_status st = enqueue_chain(pl, d_in, d_out, pl->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(pl->cap_stream, &g);
    if (st != FFS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (ee != hipSuccess) return fail(FFS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ee));
    slot->graph = g;
    const hipError_t ie = hipGraphInstantiate(&slot->gexec, slot->graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      drop_graph(*slot);
      return fail(FFS_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    }
    slot->in = d_in;
    slot->out = d_out;
  }
  slot->used = ++pl->graph_clock;
  FFS_HIP(hipGraphLaunch(slot->gexec, s));
  return FFS_OK;
}

ffs_status ffs_run_block(ffs_plan* pl, const double* d_in, double* d_out, int32_t tail, void* stream) {
  if (!pl || !d_in || !d_out) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  if (d_in == d_out) return fail(FFS_ERR_INVALID_ARG, "d_in must not alias d_out");
  if (pl->nblocks == 0) return fail(FFS_ERR_INVALID_ARG, "plan has no blocks (T = 0)");
  FFS_HIP(hipSetDevice(pl->dev));
  const bool use_tail = tail != 0 && pl->t_tail > 0;
  hipError_t e = launch_one_block(pl, d_in, d_out, use_tail, (hipStream_t)stream);
  if (e != hipSuccess) return fail(FFS_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  return FFS_OK;
}

ffs_status ffs_run_static(const ffs_problem* p, const double* d_in, double* d_out, void* stream) {
  if (!p || !d_in || !d_out) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  if (d_in == d_out) return fail(FFS_ERR_INVALID_ARG, "d_in must not alias d_out");
  // A host-side plan whose device tables and ping-pong grid are carved out of the static array; it owns
  // no device memory, so it is never passed to release().
  ffs_plan pl;
  ffs_status st = plan_host(p, &pl);
  if (st == FFS_OK) st = plan_device(&pl);
  if (st == FFS_OK) st = plan_constants(&pl);
  if (st != FFS_OK) return st;
  const TableSizes ts = table_sizes(&pl);
  size_t off = 0;
  auto carve = [&](size_t n) {
    const size_t at = off;
    off += (n + 255) & ~size_t(255);
    return at;
  };
  const size_t o_ops = carve(ts.ops), o_opsm = carve(ts.opsm), o_H = carve(ts.H), o_Hf = carve(ts.Hf),
               o_opsi = carve(ts.opsi), o_taps = carve(ts.taps);
  if (off > FFS_STATIC_TABLE_BYTES)
    return fail(FFS_ERR_INTERNAL, "plan tables need %zu bytes of the %llu reserved", off, FFS_STATIC_TABLE_BYTES);
  const size_t grid_bytes = pl.nblocks >= 2 ? sizeof(double) * (size_t)pl.npoints : 0;
  if (grid_bytes > FFS_STATIC_SCRATCH_BYTES - FFS_STATIC_TABLE_BYTES)
    return fail(FFS_ERR_UNSUPPORTED,
                "ffs_run_static: %lld points need %zu bytes of static scratch for the ping-pong grid, the limit is "
                "%llu (%llu points); use ffs_plan_create + ffs_run",
                (long long)pl.npoints, grid_bytes, FFS_STATIC_SCRATCH_BYTES - FFS_STATIC_TABLE_BYTES,
                (FFS_STATIC_SCRATCH_BYTES - FFS_STATIC_TABLE_BYTES) / sizeof(double));
  void* base = nullptr;
  st = static_scratch(pl.dev, &base);
  if (st != FFS_OK) return st;
  char* b = static_cast<char*>(base);
  if (!pl.direct) {
    pl.d_ops = reinterpret_cast<double*>(b + o_ops);
    pl.d_opsm = ts.opsm ? reinterpret_cast<uint32_t*>(b + o_opsm) : nullptr;
    pl.d_H = reinterpret_cast<double*>(b + o_H);
    pl.d_Hf = ts.Hf ? reinterpret_cast<float*>(b + o_Hf) : nullptr;
    pl.d_opsi = ts.opsi ? static_cast<void*>(b + o_opsi) : nullptr;
    pl.d_taps = reinterpret_cast<double*>(b + o_taps);
  }
  pl.d_scratch = grid_bytes ? reinterpret_cast<double*>(b + FFS_STATIC_TABLE_BYTES) : nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (!pl.direct) {
    hipError_t e = ffsk::put_taps(pl.d_taps, pl.weights.data(), (int64_t)pl.weights.size(), s);
    if (e == hipSuccess) e = plan_tables(&pl, s);
    if (e != hipSuccess) return fail(FFS_ERR_HIP, "table generation failed: %s", hipGetErrorString(e));
  }
  if (pl.nblocks == 0) {
    FFS_HIP(hipMemcpyAsync(d_out, d_in, sizeof(double) * pl.npoints, hipMemcpyDeviceToDevice, s));
    return FFS_OK;
  }
  return enqueue_chain(&pl, d_in, d_out, s);
}

ffs_status ffs_run_host(ffs_plan* pl, const double* h_in, double* h_out) {
  if (!pl || !h_in || !h_out) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  FFS_HIP(hipSetDevice(pl->dev));
  double *din = nullptr, *dout = nullptr;
  const size_t bytes = sizeof(double) * pl->npoints;
  hipError_t e = hipMalloc(&din, bytes);
  if (e == hipSuccess) e = hipMalloc(&dout, bytes);
  if (e != hipSuccess) {
    if (din) (void)hipFree(din);
    return fail(FFS_ERR_OUT_OF_MEMORY, "device allocation failed: %s", hipGetErrorString(e));
  }
  ffs_status st = FFS_OK;
  e = hipMemcpy(din, h_in, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    st = ffs_run(pl, din, dout, nullptr);
    if (st == FFS_OK) {
      e = hipStreamSynchronize(nullptr);
      if (e == hipSuccess) e = hipMemcpy(h_out, dout, bytes, hipMemcpyDeviceToHost);
    }
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  // a cached graph referenced the temporaries: drop it
  for (ffs_plan::GraphSlot& g : pl->graphs)
    if (g.gexec && g.in == din) drop_graph(g);
  if (st != FFS_OK) return st;
  if (e != hipSuccess) return fail(FFS_ERR_HIP, "run_host failed: %s", hipGetErrorString(e));
  return FFS_OK;
}

ffs_status ffs_autotune(ffs_problem* p, const char* cache_path, double* best_ms) {
  int64_t npoints = 0;
  ffs_status st = validate(p, &npoints);
  if (st != FFS_OK) return st;
  // The key holds everything the admissible configurations and their timings depend on: the problem's
  // shape, the weights themselves (star support bounds the direct depth, |K| the FP16 scale; FNV-1a over
  // their bytes), the flags and the device architecture.
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(FFS_ERR_NO_DEVICE, "no HIP device visible");
  if (p->device < 0 || p->device >= ndev) return fail(FFS_ERR_INVALID_ARG, "device %d out of range", p->device);
  hipDeviceProp_t prop;
  FFS_HIP(hipGetDeviceProperties(&prop, p->device));
  char arch[32];
  std::snprintf(arch, sizeof arch, "%.*s", 31, prop.gcnArchName);
  for (char* c = arch; *c; c++)
    if (*c == ':' || *c == '"' || *c == ' ') *c = '.';
  int64_t ntaps_key = 1;
  for (int d = 0; d < p->dims; d++) ntaps_key *= 2 * p->radius + 1;
  uint64_t wh = 0xcbf29ce484222325ULL;
  for (int64_t i = 0; i < ntaps_key; i++) {
    uint64_t bits;
    std::memcpy(&bits, &p->weights[i], sizeof bits);
    for (int k = 0; k < 8; k++, bits >>= 8) wh = (wh ^ (bits & 0xff)) * 0x100000001b3ULL;
  }
  char key[256];
  std::snprintf(key, sizeof key, "d%d_r%d_p%d_c%d_a%d_n%lldx%lldx%lld_T%lld_w%016llx_f%d_%s", p->dims, p->radius,
                p->precision, p->convention, p->algorithm, (long long)p->sizes[0],
                (long long)(p->dims > 1 ? p->sizes[1] : 1), (long long)(p->dims > 2 ? p->sizes[2] : 1),
                (long long)p->steps, (unsigned long long)wh, p->flags, arch);
  // host-only admissibility check of a configuration (the planner's own validation, no device work)
  auto admissible = [&](int algo, int tile, int fuse, int mapping) {
    ffs_plan probe;
    probe.prob = *p;
    probe.prob.algorithm = algo;
    if (algo == FFS_ALGO_DIRECT) probe.prob.precision = FFS_PREC_FP64;
    probe.prob.tile = algo == FFS_ALGO_DIRECT ? 0 : tile;
    probe.prob.fuse = fuse;
    probe.prob.mapping = mapping;
    if (validate(&probe.prob, &probe.npoints) != FFS_OK) return false;
    probe.weights.assign(p->weights, p->weights + ntaps_key);
    probe.prob.weights = probe.weights.data();
    return choose_config(&probe) == FFS_OK;
  };
  auto apply = [&](int algo, int tile, int fuse, int mapping) {
    p->algorithm = algo;
    if (algo == FFS_ALGO_DIRECT) p->precision = FFS_PREC_FP64;
    p->tile = algo == FFS_ALGO_DIRECT ? 0 : tile;
    p->fuse = fuse;
    p->mapping = mapping;
  };
  if (cache_path) {  // a line: {"key": "...", "algorithm": A, "tile": P, "fuse": t, "mapping": M, "ms": x}
    if (FILE* f = std::fopen(cache_path, "r")) {
      char line[512];
      int algo = -1, tile = -1, fuse = -1, mapping = 0;
      double ms = 0;
      while (std::fgets(line, sizeof line, f)) {
        char k[256];
        int al, tl, fu, mp;
        double m;
        if (std::sscanf(line,
                        " {\"key\": \"%255[^\"]\", \"algorithm\": %d, \"tile\": %d, \"fuse\": %d, \"mapping\": %d, \"ms\": %lf",
                        k, &al, &tl, &fu, &mp, &m) == 6 &&
            std::strcmp(k, key) == 0) {
          algo = al;
          tile = tl;
          fuse = fu;
          mapping = mp;
          ms = m;
        }
      }
      std::fclose(f);
      if (algo >= 0 && admissible(algo, tile, fuse, mapping)) {  // else: stale or foreign line, tune again
        apply(algo, tile, fuse, mapping);
        if (best_ms) *best_ms = ms;
        return FFS_OK;
      }
    }
  }
  struct Cand {
    int algo, tile, fuse, mapping;
  };
  std::vector<Cand> cands;
  const int64_t tcap = std::max<int64_t>(p->steps, 1);
  auto pick = [&](int64_t t, int64_t tmax) { return t <= 8 || (t & (t - 1)) == 0 || t == tmax; };
  if (p->algorithm != FFS_ALGO_DIRECT) {  // FFT candidates: every tile of the form, depths up to 8, then 2^k + max
    std::vector<int> tiles;
    if (p->dims == 1) tiles = {240};
    else if (p->dims == 3) tiles = {16};
    else if (p->precision == FFS_PREC_MIXED_BF16) tiles = {16, 32};
    else if (p->precision == FFS_PREC_MIXED_F16) tiles = {32};
    else if (p->precision == FFS_PREC_FP64) tiles = {16, 32};
    else tiles = {16};
    for (int tile : tiles) {
      const int vmin = p->dims == 1 ? 16 : 2;
      const int64_t tmax = p->radius == 0 ? 1 : std::max<int64_t>(1, (tile - vmin) / (2 * p->radius));
      for (int64_t t = 1; t <= std::min(tmax, tcap); t++)
        if (pick(t, tmax))
          for (int mp : {FFS_MAP_XCD, FFS_MAP_LINEAR}) cands.push_back({FFS_ALGO_FFT, tile, (int)t, mp});
    }
  }
  if (p->algorithm != FFS_ALGO_FFT) {  // direct candidates (FP64), when the radius is offered
    ffs_plan probe;
    probe.prob = *p;
    int64_t ntaps = 1;
    for (int d = 0; d < p->dims; d++) ntaps *= 2 * p->radius + 1;
    probe.weights.assign(p->weights, p->weights + ntaps);
    const int64_t tmax = ffsk::direct_max_depth(p->dims, p->radius, star_support(&probe));
    for (int64_t t = 1; t <= std::min(tmax, tcap); t++)
      if (pick(t, tmax)) cands.push_back({FFS_ALGO_DIRECT, 0, (int)t, FFS_MAP_AUTO});
  }
  double best = 1e300;
  Cand win{-1, 0, 0, 0};
  double *din = nullptr, *dout = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  for (const Cand& cd : cands) {
    ffs_problem c = *p;
    c.algorithm = cd.algo;
    if (cd.algo == FFS_ALGO_DIRECT) c.precision = FFS_PREC_FP64;
    c.tile = cd.tile;
    c.fuse = cd.fuse;
    c.mapping = cd.mapping;
    ffs_plan* pl = nullptr;
    if (ffs_plan_create(&c, &pl) != FFS_OK) continue;
    if (!din) {
      hipError_t e = hipMalloc(&din, sizeof(double) * npoints);
      if (e == hipSuccess) e = hipMalloc(&dout, sizeof(double) * npoints);
      // time on data like the caller's: an all-zero grid runs at a higher clock (less switching)
      if (e == hipSuccess) e = ffsk::fill_uniform_device(din, npoints, 0x5EED0000ULL, nullptr);
      if (e == hipSuccess) e = hipEventCreate(&e0);
      if (e == hipSuccess) e = hipEventCreate(&e1);
      if (e != hipSuccess) {
        ffs_plan_destroy(pl);
        cleanup();
        return fail(FFS_ERR_OUT_OF_MEMORY, "autotune scratch: %s", hipGetErrorString(e));
      }
    }
    double ms_best = 1e300;
    for (int rep = 0; rep < 3; rep++) {  // rep 0 captures the graph and warms up
      (void)hipEventRecord(e0, nullptr);
      st = ffs_run(pl, din, dout, nullptr);
      (void)hipEventRecord(e1, nullptr);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (st != FFS_OK) break;
      if (rep > 0) ms_best = std::min(ms_best, (double)ms);
    }
    ffs_plan_destroy(pl);
    if (st != FFS_OK) {
      cleanup();
      return st;
    }
    if (ms_best < best) {
      best = ms_best;
      win = cd;
    }
  }
  cleanup();
  if (win.algo < 0) return fail(FFS_ERR_UNSUPPORTED, "no admissible configuration");
  apply(win.algo, win.tile, win.fuse, win.mapping);
  if (best_ms) *best_ms = best;
  if (cache_path) {
    if (FILE* f = std::fopen(cache_path, "a")) {
      std::fprintf(f, "{\"key\": \"%s\", \"algorithm\": %d, \"tile\": %d, \"fuse\": %d, \"mapping\": %d, \"ms\": %.6f}\n",
                   key, win.algo, win.tile, win.fuse, win.mapping, best);
      std::fclose(f);
    }
  }
  return FFS_OK;
}

ffs_status ffs_batch_create(const ffs_problem* problems, int32_t count, ffs_batch** out) {
  if (!out) return fail(FFS_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  if (!problems || count <= 0) return fail(FFS_ERR_INVALID_ARG, "batch needs at least one problem");
  const int32_t prec = problems[0].precision;
  for (int32_t i = 0; i < count; i++) {
    if (problems[i].dims != 2) return fail(FFS_ERR_UNSUPPORTED, "batch problem %d: only 2D grids are batched", i);
    if (problems[i].algorithm != FFS_ALGO_FFT)
      return fail(FFS_ERR_UNSUPPORTED, "batch problem %d: batches run the FFT algorithm only", i);
    if (problems[i].precision != prec)
      return fail(FFS_ERR_INVALID_ARG, "batch problem %d: all problems must share one precision", i);
    if (prec != FFS_PREC_FP64 && prec != FFS_PREC_MIXED_BF16 && prec != FFS_PREC_MIXED_F16)
      return fail(FFS_ERR_UNSUPPORTED, "batches run the FP64, BF16 and FP16 forms (32-point tiles), not FP32");
    if (problems[i].tile != 0 && problems[i].tile != 32)
      return fail(FFS_ERR_UNSUPPORTED, "batch problem %d: tile must be 32", i);
    int64_t np = 0;
    ffs_status st = validate(&problems[i], &np);
    if (st != FFS_OK) return st;
  }
  ffs_batch* b = new (std::nothrow) ffs_batch();
  if (!b) return fail(FFS_ERR_OUT_OF_MEMORY, "host allocation failed");
  b->precision = prec;
  auto bail = [&](ffs_status st) {
    ffs_batch_destroy(b);
    return st;
  };
  for (int32_t i = 0; i < count; i++) {
    ffs_problem p = problems[i];
    p.tile = 32;
    ffs_plan* pl = nullptr;
    ffs_status st = ffs_plan_create(&p, &pl);
    if (st != FFS_OK) return bail(st);
    b->plans.push_back(pl);
    b->rounds = std::max(b->rounds, pl->nblocks);
  }
  b->dev = b->plans[0]->dev;
  b->num_cus = b->plans[0]->num_cus;
  b->active.assign((size_t)b->rounds, 0);
  hipError_t e = hipSuccess;
  if (b->rounds > 0) {
    e = hipMalloc(&b->d_items, sizeof(ffsk::TileArgs) * b->rounds * count);
    if (e == hipSuccess) e = hipMalloc(&b->d_prefix, sizeof(int64_t) * b->rounds * (count + 1));
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fail(FFS_ERR_HIP, "batch set-up failed: %s", hipGetErrorString(e));
    return bail(FFS_ERR_HIP);
  }
  *out = b;
  return FFS_OK;
}

ffs_status ffs_batch_get_info(const ffs_batch* b, int32_t index, ffs_plan_info* info) {
  if (!b || index < 0 || index >= (int32_t)b->plans.size()) return fail(FFS_ERR_INVALID_ARG, "bad batch index");
  return ffs_plan_get_info(b->plans[index], info);
}

ffs_status ffs_batch_run(ffs_batch* b, const double* const* d_in, double* const* d_out, void* stream) {
  if (!b || !d_in || !d_out) return fail(FFS_ERR_INVALID_ARG, "NULL argument");
  const int32_t n = (int32_t)b->plans.size();
  for (int32_t i = 0; i < n; i++) {
    if (!d_in[i] || !d_out[i]) return fail(FFS_ERR_INVALID_ARG, "grid %d: NULL pointer", i);
    if (d_in[i] == d_out[i]) return fail(FFS_ERR_INVALID_ARG, "grid %d: d_in must not alias d_out", i);
  }
  hipStream_t s = (hipStream_t)stream;
  FFS_HIP(hipSetDevice(b->dev));
  const bool same = b->gexec && std::equal(b->g_in.begin(), b->g_in.end(), d_in) &&
                    std::equal(b->g_out.begin(), b->g_out.end(), d_out) && (int32_t)b->g_in.size() == n;
  if (!same && b->rounds > 0) {
    // per-round work lists: grid g's block r reads its previous output and writes d_out or the plan's
    // ping-pong grid so that its last block lands in d_out
    std::vector<ffsk::TileArgs> items((size_t)(b->rounds * n));
    std::vector<int64_t> prefix((size_t)(b->rounds * (n + 1)));
    for (int64_t r = 0; r < b->rounds; r++) {
      int32_t k = 0;
      int64_t acc = 0;
      for (int32_t g = 0; g < n; g++) {
        ffs_plan* pl = b->plans[g];
        const int64_t nb = pl->nblocks;
        if (r >= nb) continue;
        const double* src = d_in[g];
        double* dst = nullptr;
        for (int64_t q = 0; q <= r; q++) {
          dst = ((nb - 1 - q) % 2 == 0) ? d_out[g] : pl->d_scratch;
          if (q < r) src = dst;
        }
        const bool tail = (r == nb - 1) && pl->t_tail > 0;
        items[(size_t)(r * n + k)] = make_args(pl, src, dst, tail);
        prefix[(size_t)(r * (n + 1) + k)] = acc;
        acc += pl->pairs;
        k++;
      }
      prefix[(size_t)(r * (n + 1) + k)] = acc;
      b->active[(size_t)r] = k;
    }
    // a replay of the previous graph may still be reading the work lists (on any stream): let it finish
    FFS_HIP(hipDeviceSynchronize());
    FFS_HIP(hipMemcpy(b->d_items, items.data(), sizeof(ffsk::TileArgs) * items.size(), hipMemcpyHostToDevice));
    FFS_HIP(hipMemcpy(b->d_prefix, prefix.data(), sizeof(int64_t) * prefix.size(), hipMemcpyHostToDevice));
    if (b->gexec) {
      FFS_HIP(hipGraphExecDestroy(b->gexec));
      b->gexec = nullptr;
    }
    if (b->graph) {
      FFS_HIP(hipGraphDestroy(b->graph));
      b->graph = nullptr;
    }
    FFS_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int64_t r = 0; r < b->rounds && le == hipSuccess; r++)
      le = ffsk::launch_batch(b->d_items + r * n, b->d_prefix + r * (n + 1), b->active[(size_t)r], b->precision,
                              b->cap_stream, b->num_cus);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(FFS_ERR_HIP, "batch capture failed: %s", hipGetErrorString(le != hipSuccess ? le : ee));
    }
    b->graph = g;
    FFS_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    b->g_in.assign(d_in, d_in + n);
    b->g_out.assign(d_out, d_out + n);
  }
  for (int32_t g = 0; g < n; g++)  // grids with T = 0 are copies
    if (b->plans[g]->nblocks == 0)
      FFS_HIP(hipMemcpyAsync(d_out[g], d_in[g], sizeof(double) * b->plans[g]->npoints, hipMemcpyDeviceToDevice, s));
  if (b->gexec) FFS_HIP(hipGraphLaunch(b->gexec, s));
  return FFS_OK;
}

ffs_status ffs_batch_destroy(ffs_batch* b) {
  if (!b) return FFS_OK;
  (void)hipSetDevice(b->dev);
  (void)hipDeviceSynchronize();
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_prefix) (void)hipFree(b->d_prefix);
  for (ffs_plan* pl : b->plans) release(pl);
  delete b;
  return FFS_OK;
}

ffs_status ffs_batch_autotune(ffs_problem* problems, int32_t count, const char* cache_path, double* best_ms) {
  if (!problems || count <= 0) return fail(FFS_ERR_INVALID_ARG, "batch needs at least one problem");
  // Depth classes: grids of one radius move together.  Base depth of each grid: the planner's own choice
  // on 32-point tiles (fuse = 0), or the caller's fuse when one is set.
  std::vector<int64_t> base((size_t)count), tmax((size_t)count);
  std::vector<int32_t> cls((size_t)count);
  std::vector<int32_t> radii;
  for (int32_t i = 0; i < count; i++) {
    ffs_problem q = problems[i];
    q.tile = 32;
    ffs_plan probe;
    ffs_status st = plan_host(&q, &probe);
    if (st != FFS_OK) return st;
    base[(size_t)i] = probe.t;
    const int r = q.radius;
    tmax[(size_t)i] = r == 0 ? 1 : std::max<int64_t>(1, std::min<int64_t>((32 - 2) / (2 * r), std::max<int64_t>(q.steps, 1)));
    auto it = std::find(radii.begin(), radii.end(), r);
    if (it == radii.end()) {
      radii.push_back(r);
      cls[(size_t)i] = (int32_t)radii.size() - 1;
    } else {
      cls[(size_t)i] = (int32_t)(it - radii.begin());
    }
  }
  // Key: every field the batch's timing depends on, hashed (FNV-1a) over all problems, and the device.
  uint64_t h = 0xcbf29ce484222325ULL;
  auto mix = [&](const void* data, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(data);
    for (size_t k = 0; k < n; k++) h = (h ^ c[k]) * 0x100000001b3ULL;
  };
  for (int32_t i = 0; i < count; i++) {
    const ffs_problem& q = problems[i];
    const int32_t f[7] = {q.dims, q.radius, q.precision, q.convention, q.flags, q.fuse, q.mapping};
    mix(f, sizeof f);
    mix(q.sizes, sizeof q.sizes);
    mix(&q.steps, sizeof q.steps);
    int64_t ntaps = 1;
    for (int d = 0; d < q.dims; d++) ntaps *= 2 * q.radius + 1;
    mix(q.weights, sizeof(double) * (size_t)ntaps);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(FFS_ERR_NO_DEVICE, "no HIP device visible");
  const int dev = problems[0].device;
  if (dev < 0 || dev >= ndev) return fail(FFS_ERR_INVALID_ARG, "device %d out of range", dev);
  hipDeviceProp_t prop;
  FFS_HIP(hipGetDeviceProperties(&prop, dev));
  char key[96];
  std::snprintf(key, sizeof key, "batch%d_%016llx_%.*s", count, (unsigned long long)h, 24, prop.gcnArchName);
  for (char* c = key; *c; c++)
    if (*c == ':' || *c == '"' || *c == ' ') *c = '.';
  auto apply = [&](const std::vector<int32_t>& fuse, int32_t mapping) {
    for (int32_t i = 0; i < count; i++) {
      problems[i].tile = 32;
      problems[i].algorithm = FFS_ALGO_FFT;
      problems[i].fuse = fuse[(size_t)i];
      problems[i].mapping = mapping;
    }
  };
  // a line: {"key": "...", "mapping": M, "ms": x, "fuse": [t0, t1, ...]}
  if (cache_path) {
    if (FILE* f = std::fopen(cache_path, "r")) {
      std::string line;
      std::vector<int32_t> got;
      int32_t mapping = 0;
      double ms = 0.0;
      bool found = false;
      for (int ch; (ch = std::fgetc(f)) != EOF || !line.empty();) {
        if (ch != '\n' && ch != EOF) {
          line.push_back((char)ch);
          continue;
        }
        char k[96];
        int mp = 0;
        double m = 0.0;
        int used = 0;
        if (std::sscanf(line.c_str(), " {\"key\": \"%95[^\"]\", \"mapping\": %d, \"ms\": %lf, \"fuse\": [%n", k, &mp, &m,
                        &used) == 3 && used > 0 && std::strcmp(k, key) == 0) {
          std::vector<int32_t> v;
          const char* c = line.c_str() + used;
          while (*c) {
            char* end = nullptr;
            const long x = std::strtol(c, &end, 10);
            if (end == c) break;
            v.push_back((int32_t)x);
            c = end;
            while (*c == ',' || *c == ' ') c++;
          }
          if ((int32_t)v.size() == count) {
            got = v;
            mapping = mp;
            ms = m;
            found = true;
          }
        }
        line.clear();
        if (ch == EOF) break;
      }
      std::fclose(f);
      bool ok = found;
      for (int32_t i = 0; ok && i < count; i++) ok = got[(size_t)i] >= 1 && got[(size_t)i] <= tmax[(size_t)i];
      if (ok) {  // else: stale or foreign line, tune again
        apply(got, mapping);
        if (best_ms) *best_ms = ms;
        return FFS_OK;
      }
    }
  }
  // Device grids for the timing runs (the caller's sizes, seeded data like the caller's).
  FFS_HIP(hipSetDevice(dev));
  std::vector<double*> din((size_t)count, nullptr), dout((size_t)count, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    for (double* d : din)
      if (d) (void)hipFree(d);
    for (double* d : dout)
      if (d) (void)hipFree(d);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int32_t i = 0; i < count && e == hipSuccess; i++) {
    int64_t n = 1;
    for (int d = 0; d < problems[i].dims; d++) n *= problems[i].sizes[d];
    e = hipMalloc(&din[(size_t)i], sizeof(double) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc(&dout[(size_t)i], sizeof(double) * (size_t)n);
    if (e == hipSuccess) e = ffsk::fill_uniform_device(din[(size_t)i], n, 0x5EED0000ULL + (uint64_t)i, nullptr);
  }
  if (e != hipSuccess) {
    cleanup();
    return fail(FFS_ERR_OUT_OF_MEMORY, "batch autotune grids: %s", hipGetErrorString(e));
  }
  // Time one configuration: a whole batch run (all rounds from its hipGraph), best of 3 after a warm-up.
  std::vector<ffs_problem> trial(problems, problems + count);
  auto timed = [&](const std::vector<int32_t>& fuse, int32_t mapping, double* ms_out) -> ffs_status {
    for (int32_t i = 0; i < count; i++) {
      trial[(size_t)i] = problems[i];
      trial[(size_t)i].tile = 32;
      trial[(size_t)i].algorithm = FFS_ALGO_FFT;
      trial[(size_t)i].fuse = fuse[(size_t)i];
      trial[(size_t)i].mapping = mapping;
    }
    ffs_batch* b = nullptr;
    ffs_status st = ffs_batch_create(trial.data(), count, &b);
    if (st != FFS_OK) return st;
    double best = 1e300;
    for (int rep = 0; rep < 4 && st == FFS_OK; rep++) {  // rep 0 captures the graph and warms up
      (void)hipEventRecord(e0, nullptr);
      st = ffs_batch_run(b, din.data(), dout.data(), nullptr);
      (void)hipEventRecord(e1, nullptr);
      if (hipEventSynchronize(e1) != hipSuccess) st = fail(FFS_ERR_HIP, "batch autotune: run failed");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0) best = std::min(best, (double)ms);
    }
    ffs_batch_destroy(b);
    *ms_out = best;
    return st;
  };
  std::vector<int32_t> cur((size_t)count);
  for (int32_t i = 0; i < count; i++)
    cur[(size_t)i] = (int32_t)std::max<int64_t>(1, std::min(base[(size_t)i], tmax[(size_t)i]));
  // Mapping: the one the caller set, else both orders are tried.
  std::vector<int32_t> maps;
  if (problems[0].mapping != FFS_MAP_AUTO) maps = {problems[0].mapping};
  else maps = {FFS_MAP_LINEAR, FFS_MAP_XCD};
  int32_t best_map = maps[0];
  double best = 1e300;
  for (int32_t mp : maps) {
    double ms = 0.0;
    ffs_status st = timed(cur, mp, &ms);
    if (st != FFS_OK) {
      cleanup();
      return st;
    }
    if (ms < best) {
      best = ms;
      best_map = mp;
    }
  }
  // Coordinate descent over the depth classes: shift every grid of one radius by -2 .. +2 (clamped to
  // its admissible range) and keep a shift that makes the whole batch faster.
  for (size_t c = 0; c < radii.size(); c++) {
    std::vector<int32_t> keep = cur;
    for (int delta : {-2, -1, 1, 2}) {
      std::vector<int32_t> cand = keep;
      bool changed = false;
      for (int32_t i = 0; i < count; i++) {
        if (cls[(size_t)i] != (int32_t)c) continue;
        const int64_t t = std::min<int64_t>(std::max<int64_t>(keep[(size_t)i] + delta, 1), tmax[(size_t)i]);
        changed |= t != keep[(size_t)i];
        cand[(size_t)i] = (int32_t)t;
      }
      if (!changed) continue;
      double ms = 0.0;
      if (timed(cand, best_map, &ms) == FFS_OK && ms < best) {
        best = ms;
        cur = cand;
      }
    }
  }
  cleanup();
  apply(cur, best_map);
  if (best_ms) *best_ms = best;
  if (cache_path) {
    if (FILE* f = std::fopen(cache_path, "a")) {
      std::fprintf(f, "{\"key\": \"%s\", \"mapping\": %d, \"ms\": %.6f, \"fuse\": [", key, best_map, best);
      for (int32_t i = 0; i < count; i++) std::fprintf(f, "%s%d", i ? ", " : "", cur[(size_t)i]);
      std::fprintf(f, "]}\n");
      std::fclose(f);
    }
  }
  return FFS_OK;
}

const char* ffs_last_error(void) { return g_err.c_str(); }
const char* ffs_version(void) { return FFS_VERSION_STRING; }

}  // extern "C"
