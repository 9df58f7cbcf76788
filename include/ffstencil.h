/*
 * ffstencil.h -- C ABI of libffstencil, the MI355X (gfx950) tiled FFT-stencil engine.
 *
 * A periodic, constant-coefficient stencil of radius r (2r+1 taps per axis) is applied T times to a
 * compact FP64 grid of 1, 2 or 3 dimensions.  The grid is cut into tiles; every tile window (valid
 * region plus an r*t halo, gathered with periodic wrap) is transformed by DFT-as-GEMM on the matrix
 * cores, multiplied by the stencil spectrum raised to the fused depth t, transformed back, and only
 * the valid region is stored (overlap-save).  T steps run as floor(T/t) fused blocks plus one tail
 * block, ping-ponged through HBM and replayed from a hipGraph.
 *
 * All pointers are plain host or device pointers; sizes are element counts; no HIP or torch types
 * appear in the signatures (a stream is passed as `void*` holding a hipStream_t, NULL = null stream).
 * Functions never exit the process: they return an ffs_status and leave a message for
 * ffs_last_error() (thread-local).
 *
 * Reference interfaces replaced (KevinWu2017/FlashFFTStencil):
 *   - the per-dimension `CreatePlan(double* k, int K, bool print)` that fills __constant__ DFT
 *     matrices and the kernel spectrum (src/1D/create_fft_pfa_plan.cu:19, src/2D/create_fft_pfa_plan.cu:17,
 *     src/3D/create_fft_pfa_plan.cu:17)                                    -> ffs_plan_create
 *   - the time loop launching the tile kernel T (or T/t) times on the padded grid
 *     (src/1D/1d_main.cu:164-177, src/2D/2d_main.cu:150-170, src/3D/3d_main.cu:147-170) -> ffs_run
 *   - the shape table of the drivers (src/1D/1d_main.cu:44-65, src/2D/2d_main.cu:49-52,
 *     src/3D/3d_main.cu:44-47) and their rand() weights (src/1D/1d_main.cu:130-141,
 *     src/2D/2d_main.cu:113-126, src/3D/3d_main.cu:111-127)                 -> ffs_stencil_named
 *   - the rand() grid initialisation (src/1D/1d_main.cu:115-129, src/2D/2d_main.cu:92-112,
 *     src/3D/3d_main.cu:86-110)                                             -> ffs_fill_uniform
 *   - the CSV report line (src/1D/1d_main.cu:187-195, src/2D/2d_main.cu:179-185,
 *     src/3D/3d_main.cu:179-185)                                            -> ffs_format_csv
 */
#ifndef FFSTENCIL_H
#define FFSTENCIL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFS_VERSION_STRING "0.1.0"

typedef enum ffs_status {
  FFS_OK = 0,
  FFS_ERR_INVALID_ARG = 1,   /* bad shape, size, weights, T, pointer ... */
  FFS_ERR_UNSUPPORTED = 2,   /* valid request the library does not implement (e.g. tile too small) */
  FFS_ERR_OUT_OF_MEMORY = 3, /* device or host allocation failed */
  FFS_ERR_HIP = 4,           /* a HIP runtime call failed */
  FFS_ERR_NO_DEVICE = 5,     /* no gfx950 device visible */
  FFS_ERR_INTERNAL = 6
} ffs_status;

/* Arithmetic of the DFT contractions.  The grid state between fused blocks, the DFT matrices and
 * the spectrum power are FP64 in every form. */
typedef enum ffs_precision {
  FFS_PREC_FP64 = 0,       /* FP64 MFMA (v_mfma_f64_16x16x4f64) contractions: the drop-in form */
  FFS_PREC_MIXED_BF16 = 1,  /* 3-term BF16 split operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16) */
  FFS_PREC_MIXED_F32 = 2,   /* FP32 operands, FP32 accumulation (v_mfma_f32_16x16x4f32); in 3D the DFT
                               matrix is compensated (FP32 entries plus their low-order FP32 terms, two
                               products per real GEMM) */
  FFS_PREC_MIXED_BF16X2 = 3, /* withdrawn (FFS_ERR_UNSUPPORTED): a 2-term BF16 data split (16 bits) misses 1e-6
                                before the field has smoothed, and was not faster than the 3-term form */
  FFS_PREC_MIXED_F16 = 4     /* 2-term FP16 split operands (22 bits) with a per-tile power-of-2 scale, FP32
                                accumulation (v_mfma_f32_16x16x32_f16), 3 products per block; 2D only */
} ffs_precision;

/* Index convention of one step.
 *   REF:     out[p] = sum_{a in [0,2r]^d} w[a] * in[(p - a) mod n]   (true convolution anchored at the
 *            first tap; the reference's CPU check composition, src/2D/check_correct.hpp:64-100 with
 *            rotateKernel180 :17-31, and src/1D/1d_main.cu:220-223)
 *   CENTRED: out[p] = sum_{o in [-r,r]^d} w[o + r] * in[(p + o) mod n]   (src/1D/check_correct.hpp:15-38) */
typedef enum ffs_convention { FFS_CONV_REF = 0, FFS_CONV_CENTRED = 1 } ffs_convention;

/* How the T steps are computed.
 *   FFT:    the tiled FFT-stencil path (DFT-as-GEMM on the matrix cores, spectrum power K^t), in the
 *           requested precision.  The drop-in for the reference's kernels (src/2D/rfft_2d/2d_rfft_1_async.cu).
 *   DIRECT: the stencil sum itself in FP64 FMAs, t steps per launch through LDS/registers (temporal
 *           blocking); precision must be FFS_PREC_FP64.  No kernel counterpart in the reference (its
 *           CPU check, src/2D/check_correct.hpp:64-100, is the semantics): a fallback and a baseline.
 *           Radius <= 3 in 1D and 2D, <= 1 in 3D.
 *   AUTO:   DIRECT where it is offered, else FFT.  On MI355X the FP64 vector and FP64 matrix peaks are
 *           equal, so the direct sum, with several times fewer FP64 operations per point-step than the
 *           transform path at these radii, is the faster one (ffs_autotune with AUTO measures both).
 *           A DIRECT result is FP64 whatever precision was asked. */
typedef enum ffs_algorithm { FFS_ALGO_FFT = 0, FFS_ALGO_DIRECT = 1, FFS_ALGO_AUTO = 2 } ffs_algorithm;

/* Wave-to-tile mapping of the FFT tile kernels: which workgroup takes which tile pairs.
 *   XCD:         workgroups are renumbered so that each of the 8 XCDs takes one contiguous band of tile
 *                pairs (neighbouring windows share that XCD's L2);
 *   LINEAR:      hardware dispatch order (consecutive workgroups go to different XCDs);
 *   AUTO:        XCD for the 2D FP64 form on 32-point tiles (measured 4% faster on Box2D25P 8192^2),
 *                LINEAR otherwise (1-3% faster than XCD bands for the mixed 2D forms there).
 * The direct kernels ignore it.  No reference counterpart (the reference launches one warp per tile
 * pair in grid order). */
typedef enum ffs_mapping { FFS_MAP_AUTO = 0, FFS_MAP_XCD = 1, FFS_MAP_LINEAR = 2 } ffs_mapping;

/* ffs_problem.flags bits.
 *   FULL_STAGES: on 32-point tiles whose valid region fits one 16 x 16 block (V <= 16), compute every
 *                output block of the two inverse stages instead of only the kept one.  The results are
 *                bitwise the same (the kept values come from the same operands in the same order); it
 *                exists to check exactly that, and costs ~1.4x the tile time.  No reference counterpart. */
#define FFS_FLAG_FULL_STAGES 0x1

typedef struct ffs_problem {
  int32_t dims;          /* 1, 2 or 3 */
  int32_t radius;        /* r >= 0; taps per axis K = 2r+1 */
  int64_t sizes[3];      /* grid sides, slowest axis first; entries past `dims` are ignored */
  const double* weights; /* K^dims taps, row-major (slowest axis first); copied at plan creation */
  int64_t steps;         /* T >= 0 */
  int32_t precision;     /* ffs_precision */
  int32_t convention;    /* ffs_convention */
  int32_t tile;          /* DFT length per axis; 0 = automatic (1D: 240 = 16x15 prime-factor; 3D: 16; 2D: 32
                            for the FP16/BF16 forms, 16 or 32 for FP64 by the depth cost model, 16 for FP32) */
  int32_t fuse;          /* fused depth t; 0 = automatic */
  int32_t use_graph;     /* 1 = replay the block loop from a hipGraph (chains of more than 4 launches;
                            shorter chains launch directly, which is cheaper), 0 = plain launches */
  int32_t device;        /* HIP device ordinal */
  int32_t algorithm;     /* ffs_algorithm (default FFS_ALGO_FFT) */
  int32_t mapping;       /* ffs_mapping (default FFS_MAP_AUTO) */
  int32_t flags;         /* FFS_FLAG_* bits (default 0) */
  int32_t reserved;
} ffs_problem;

/* What a plan chose. */
typedef struct ffs_plan_info {
  int32_t dims, radius;
  int64_t sizes[3];
  int64_t steps;
  int32_t precision, convention;
  int32_t tile[3];        /* FFT: DFT length per axis; DIRECT: input window per workgroup (1 = unused axis) */
  int32_t valid[3];       /* output width per tile (workgroup) and axis: V = P - 2 r t */
  int32_t fuse;           /* t */
  int32_t fuse_tail;      /* T mod t (0 = no tail block) */
  int64_t blocks;         /* fused blocks per run, tail included */
  int64_t tiles;          /* tiles per block */
  int64_t launches;       /* kernel launches per run */
  double flops_per_block; /* flops one block executes: matrix-core flops (FFT; see the kernels' header
                             comment) or FP64 FMA flops including the halo recomputation (DIRECT) */
  double bytes_per_block; /* algorithmic HBM bytes of one block: 16 B per grid point */
  int32_t algorithm;      /* ffs_algorithm actually used (FFT or DIRECT) */
  int32_t mapping;        /* ffs_mapping used by the tile kernels (XCD or LINEAR) */
  int32_t reserved[2];
} ffs_plan_info;

typedef struct ffs_plan ffs_plan;

/* Fill `p` with defaults: dims 2, radius 1, sizes 1, T 1, FP64, REF, automatic tile/fuse, graph on. */
ffs_status ffs_problem_init(ffs_problem* p);

/* Named stencils.  `shape` is one of Heat-1D, 1D5P, 1D7P, Heat-2D, Box2D9P, Star2D9P, Box2D25P,
 * Heat-3D, Box3D27P (Box3d27P accepted).  Heat-1D is [0.25, 0.5, 0.25]; the others carry weights
 * drawn uniform[0,1) from `seed` on the stencil's support and normalised to sum 1 (zeros off the
 * support).  Writes dims, radius and (2r+1)^dims weights into `weights` (capacity in doubles).
 * `tag` (capacity >= 16) receives the reference's CSV tag (1d1r, box_2d1r, ...); may be NULL. */
ffs_status ffs_stencil_named(const char* shape, uint64_t seed, int32_t* dims, int32_t* radius,
                             double* weights, int32_t capacity, char* tag, int32_t tag_capacity);

/* Host helper: n doubles uniform in [0,1) from SplitMix64(seed), 53-bit mantissas. */
ffs_status ffs_fill_uniform(uint64_t seed, double* host_out, int64_t n);

/* Device twin of ffs_fill_uniform: the same n values written to device memory on `stream` (async).
 * No reference counterpart (the reference fills its grids with rand() on the host). */
ffs_status ffs_fill_uniform_device(uint64_t seed, double* d_out, int64_t n, void* stream);

/* Validate `p`, choose (P, t, V), allocate device tables and the ping-pong buffer, generate the DFT
 * operands and the FP64 spectrum power on the GPU.  Validation happens before any HIP call. */
ffs_status ffs_plan_create(const ffs_problem* p, ffs_plan** out);
ffs_status ffs_plan_get_info(const ffs_plan* plan, ffs_plan_info* info);

/* Diagnostics: copy one of the plan's device tables, as generated on the GPU, to host memory (for unit
 * checks of the DFT operands and the spectrum).  `bytes` receives the table's size; host_out = NULL only
 * queries it.  Layouts are internal (per-lane MFMA operand order, see the library sources):
 *   OPS:        FP64 DFT operands (16-point families: [axis][Fr, -(Fr+Fi), Fi-Fr][chunk][lane] doubles;
 *               32-point FP64: [R][chunk][column block][Fr, -(Fr+Fi), Fi-Fr][lane])
 *   OPS_MIXED:  split operands of the mixed forms (BF16/FP16 words, FP32 chunks)
 *   SPECTRUM:   [main | tail] x [re plane | im plane] of K^t / prod(P), in the tile family's order
 *   INVERSE16:  out16 inverse operands of the main and the tail depth
 * No reference counterpart (the reference's tables live in __constant__ memory, src/2D/create_fft_pfa_plan.cu:25-107). */
typedef enum ffs_table {
  FFS_TABLE_OPS = 0,
  FFS_TABLE_OPS_MIXED = 1,
  FFS_TABLE_SPECTRUM = 2,
  FFS_TABLE_INVERSE16 = 3
} ffs_table;
ffs_status ffs_plan_read_table(const ffs_plan* plan, int32_t which, void* host_out, int64_t capacity,
                               int64_t* bytes);
ffs_status ffs_plan_destroy(ffs_plan* plan);

/* Advance T steps: d_out = S^T(d_in).  Device pointers to compact row-major FP64 grids of the plan's
 * sizes; d_in is not modified and must not alias d_out.  Asynchronous on `stream`. */
ffs_status ffs_run(ffs_plan* plan, const double* d_in, double* d_out, void* stream);

/* Same as ffs_run on host buffers (allocates, copies, runs, synchronises). */
ffs_status ffs_run_host(ffs_plan* plan, const double* h_in, double* h_out);

/* Allocation-free run: d_out = S^T(d_in) for problem `p` in one call that allocates no device memory.
 * No counterpart in the reference; it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It plans on the host exactly as
 * ffs_plan_create does, writes the plan's DFT operands, spectra and taps into a static device array of
 * FFS_STATIC_SCRATCH_BYTES per GPU (held by every GPU the library runs on, from library load until the
 * process ends; its address is resolved once per device on the first call there), keeps the ping-pong
 * grid in the same array, and enqueues the T steps on `stream` as plain kernel launches (no graph, no
 * host-to-device copy, no hipMalloc/hipFree/hipMemset).  Results are bitwise those of ffs_plan_create +
 * ffs_run for the same problem.  Limit: the tables take at most FFS_STATIC_TABLE_BYTES, and a run of two
 * or more fused blocks needs 8 bytes per grid point of the rest (grids up to 8,257,536 points);
 * larger problems get FFS_ERR_UNSUPPORTED before any device work.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads) race, so serialise them
 * (e.g. one stream).  Nothing in the array needs clearing: each call writes every byte it reads. */
#define FFS_STATIC_SCRATCH_BYTES (64ull << 20)
#define FFS_STATIC_TABLE_BYTES (1ull << 20)
ffs_status ffs_run_static(const ffs_problem* p, const double* d_in, double* d_out, void* stream);

/* One fused block only (depth t, or the tail depth when `tail` != 0 and the plan has a tail),
 * d_out = S^t(d_in).  One kernel launch; used to time the tile kernel alone. */
ffs_status ffs_run_block(ffs_plan* plan, const double* d_in, double* d_out, int32_t tail, void* stream);

/* Autotuner: time the admissible (algorithm, tile, fused depth, wave-to-tile mapping) plans of `p` on
 * its device (full T-step runs on a scratch grid of p's sizes) and write the fastest into p->tile,
 * p->fuse, p->mapping and p->algorithm (FFS_ALGO_AUTO searches both algorithms; a DIRECT winner also
 * sets p->precision to FFS_PREC_FP64).  With a non-NULL `cache_path`, results are read from / appended to that JSON-lines
 * file keyed by (dims, radius, precision, convention, algorithm, sizes, T), so each configuration is
 * measured once.  best_ms (may be NULL) receives the winner's time per run.  No reference counterpart
 * (the reference hard-codes its tile). */
ffs_status ffs_autotune(ffs_problem* p, const char* cache_path, double* best_ms);

/* Batches of independent 2D grids (different sizes, stencils and T).  Block b of every grid that still
 * has blocks runs in ONE persistent launch over the concatenated tile-pair list of those grids; each
 * grid keeps its own fused depth, tail and spectrum.  All problems must be 2D and share one precision:
 * FP64, FP16 or one of the BF16 forms (32-point tiles; FP32 is not batched).  No reference counterpart (the reference runs one grid per process);
 * it is the multi-grid workload's entry point.  d_in / d_out are host arrays of `count` device
 * pointers; grid i advances problems[i].steps steps. */
typedef struct ffs_batch ffs_batch;
ffs_status ffs_batch_create(const ffs_problem* problems, int32_t count, ffs_batch** out);
ffs_status ffs_batch_run(ffs_batch* batch, const double* const* d_in, double* const* d_out, void* stream);
ffs_status ffs_batch_get_info(const ffs_batch* batch, int32_t index, ffs_plan_info* info);
ffs_status ffs_batch_destroy(ffs_batch* batch);

/* Batch autotuner: picks the fused depth of every problem and the batch's wave-to-tile mapping by timing
 * whole batch runs (every round of the batch, from its hipGraph, on seeded grids of the problems' sizes),
 * since a grid's blocks share launches with the other grids and its standalone timing would count launch
 * costs it does not pay in the batch.  Start: the planner's depth of each grid on 32-point tiles (or
 * the caller's fuse); then, for each radius in turn, every grid of that radius is shifted by -2..+2
 * depths (within its admissible range) and a shift is kept when the whole batch gets faster; both
 * mappings are tried unless problems[0].mapping is set.  Writes tile = 32, fuse, mapping and
 * algorithm = FFT into every problem, ready for ffs_batch_create.  With a non-NULL `cache_path` the result
 * is read from / appended to that JSON-lines file, keyed by a hash of all problems (sizes, T, radius,
 * weights, precision, convention, flags) and the device architecture.  best_ms (may be NULL) receives
 * the tuned batch's time per run.  No reference counterpart (the reference runs one grid per process). */
ffs_status ffs_batch_autotune(ffs_problem* problems, int32_t count, const char* cache_path, double* best_ms);

/* Reference-format report line:
 *   "FlashFFTStencil, <tag>, <d1>, <d2>, <d3>, <T>, <ms>, <GStencil/s>"
 * with leading 1s for unused dimensions and GStencil/s = points * T / ms / 1e6. */
ffs_status ffs_format_csv(const char* tag, int32_t dims, const int64_t* sizes, int64_t steps, double ms,
                          char* buf, int32_t capacity);

const char* ffs_last_error(void);
const char* ffs_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FFSTENCIL_H */
