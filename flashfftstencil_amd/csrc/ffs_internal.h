// ffs_internal.h -- types shared by the host side (ffs_api.cpp) and the gfx950 kernels (ffs_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ffsk {

// Tile families.  Every family runs its DFTs on 16x16 MFMA blocks.
enum TileKind : int32_t {
  KIND_2D16 = 0,   // 2D, 16x16 window
  KIND_1DPFA = 1,  // 1D, 240-sample window as a 15x16 prime-factor (Good-Thomas) matrix
  KIND_3D16 = 2,   // 3D, 16x16x16 window (one workgroup of 4 waves per tile pair)
  KIND_2D32 = 3,   // 2D, 32x32 window (mixed forms: the contraction fills K = 32 of the MFMA; FP64: 8 chunks)
};

// Operand-table layout in doubles (one table per plan, generated on the GPU in FP64):
//   DFT B-operands  [axis 0..2][arr: R = Fr, S = -(Fr+Fi), D = Fi-Fr][chunk 0..3][lane 0..63]
//   spectrum        [main|tail][...] laid out per tile family, see spectrum_len()
constexpr int kLanes = 64;
constexpr int kOpsPerAxis = 3 * 4 * kLanes;  // 768 doubles
constexpr int kOpsLen = 3 * kOpsPerAxis;     // 2304 doubles

// Mixed-form operand table, 32-bit words, generated on the GPU from the FP64 DFT matrices:
//   BF16 B operands [axis][arr: R = Fr, I = Fi, NI = -Fi][pairing g][word 0..3][lane]
//     element e of lane (j + 16q): row 4q + (e & 3), column j, term (e < 4 ? first : second) of
//     g = 0: (Fh | Fh), g = 1: (Fm | Fm), g = 2: (Fl | Fh), g = 3: (Fm | Fl)
//     where F = Fh + Fm + Fl is the 3-term BF16 split.
//   FP32 B operands [axis][arr][chunk c][lane] = F[4(lane/16) + c][lane % 16], rounded to FP32,
//   then the same layout again with the low-order terms fl32(F - fl32(F)) (the 3D FP32 form's
//   compensating products)
constexpr int kPairings = 4;
constexpr int kBf16WordsPerAxis = 3 * kPairings * 4 * kLanes;  // 3072
constexpr int kF32PerAxis = 3 * 4 * kLanes;            // 768
constexpr int kOpsMixedLen = 3 * kBf16WordsPerAxis + 2 * 3 * kF32PerAxis;
constexpr int kOpsF32Offset = 3 * kBf16WordsPerAxis;
constexpr int kOpsF32LoOffset = kOpsF32Offset + 3 * kF32PerAxis;

inline int64_t spectrum_len(int kind) {  // complex values per spectrum (one depth)
  return kind == KIND_3D16 ? 16 * 16 * 16 : (kind == KIND_2D32 ? 32 * 32 : 16 * 16);
}

// 32-point B operands (KIND_2D32): [arr: R, I, NI][column block Kb][term][word 0..3][lane]; terms h, m, l of
// the 3-term BF16 split, or h, l (and a zero term) of the 2-term FP16 split for the FP16 form;
// element e of lane (j + 16q) is row 16 (e >> 2) + 4q + (e & 3), column 16 Kb + j of that term.
constexpr int kOps32Words = 3 * 2 * 3 * 4 * kLanes;  // 4608

// 32-point FP64 B operands (KIND_2D32, FP64 form): [R][chunk v][column block Kb][arr: R, S, D][lane], lane
// (n + 16k) holding entry (row 16 R + 4 v + k, column 16 Kb + n) of Fr, -(Fr + Fi) or Fi - Fr (the sum is
// stored negated so that both output accumulators chain from the shared Gauss product).
constexpr int kOps32F64 = 2 * 4 * 2 * 3 * kLanes;  // 3072 doubles

struct TileArgs {
  const double* src;
  double* dst;
  const double* ops;  // DFT operands (doubles)
  const double* H;    // spectrum^t / (P0 P1 [P2]) for this block, family layout, [re plane | im plane]
  const float* Hf;    // mixed forms: the same spectrum rounded to FP32, (re, im) pairs (one 8-byte load each)
  const uint32_t* opsm;  // mixed-form operand table (kOpsMixedLen words)
  int64_t n0, n1, n2;       // grid sides (unused axes = 1); 1D uses n1
  int64_t tiles0, tiles1, tiles2;
  int64_t pairs;            // tile pairs in the block
  int32_t V0, V1, V2;       // valid widths
  int32_t h0, h1, h2;       // window start offset (valid region starts at h in the window)
  int32_t precision;
  int32_t scale_exp;  // FP16 form: the tile pair is scaled so that its largest fluctuation is below 2^scale_exp
  double dc_gain;  // K(0)^t: gain of the constant mode (mixed forms add mean * gain back in FP64)
  int32_t linear_map;  // 1: workgroups take tiles in hardware order; 0: XCD-contiguous bands
  // 32-point family: 1 when the valid region fits one 16 x 16 block (V0, V1 <= 16).  The two inverse
  // stages then compute only output block 0 (2 and 1 of 4 output blocks) with the inverse operands of
  // window columns h .. h + 15 (`opsi`), so output index m is window position h + m: every value kept is
  // the one the full stages would produce at that position, from the same operands (the batch kernels
  // read those columns from the forward operands they hold in LDS instead of `opsi`).
  int32_t out16;
  const void* opsi;  // out16: inverse operands (kInv32Words words or kInv32F64 doubles, see below)
  float unit;  // 1.0f: a multiplier the kernels must not see as a constant (the FP16 split's residual FMA)
};

// Direct (non-FFT) FP64 kernels with temporal blocking (ffs_direct.hip).
constexpr int kDir1Pts = 8;                    // 1D: points per thread, in registers
constexpr int kDir1Window = 256 * kDir1Pts;    // 1D: window per workgroup
constexpr int kDir1MaxDepth = 64;              // 1D: largest fused depth of one launch
constexpr int kDir2Cols = 512;                 // 2D: window columns per workgroup (2 per thread)
constexpr int kDir3Threads = 512;              // 3D: threads per workgroup
constexpr int kDir3PerThread = 2;              // 3D: points per thread along axis 1 (2 along axis 2)
constexpr int kDir3Cols = 64;                  // 3D: plane window along the contiguous axis 2 (32 threads x 2)
constexpr int kDir3Rows = kDir3Threads / (kDir3Cols / 2) * kDir3PerThread;  // along axis 1

struct DirectArgs {
  const double* src;
  double* dst;
  int64_t n0, n1, n2;        // grid sides (1D: n0 = n)
  int64_t seg;               // 2D/3D: output rows (planes) along axis 0 per workgroup
  int32_t nseg;              // 2D/3D: workgroups along axis 0; 1D: workgroups
  int32_t strips1, strips2;  // workgroups along axes 1 and 2
  int32_t valid1, valid2;    // output width per workgroup along axes 1 and 2 (1D: valid1)
  int32_t lo;                // per-step window offset: 2r (ref convention) or r (centred)
  int32_t steps;             // fused depth of this launch
  double g[49];              // taps in window order, g[e] multiplies u[i + e], e in [0, 2r]^d row-major
};

int direct_max_depth(int dims, int radius, bool star);  // 0 = radius not offered
hipError_t launch_direct(const DirectArgs& a, int dims, int radius, bool star, hipStream_t s);

// Launch helpers implemented in ffs_kernels.hip.
hipError_t gen_operands(double* ops, uint32_t* opsm, const int32_t* axis_len, hipStream_t s);
hipError_t gen_operands32(uint32_t* ops32, bool f16, hipStream_t s);
hipError_t gen_operands32_f64(double* ops, hipStream_t s);
// Inverse operands of an out16 block, output column j <- window column (j + shift) mod 32:
//   mixed forms: u4 per lane, [arr: R, I, NI][term 0..2][lane] (the kOps32Words entries of block Kb = 0);
//   FP64: [R][chunk v][arr: Fr, -(Fr + Fi), Fi - Fr][lane] doubles (the kOps32F64 entries of block Kb = 0).
constexpr int kInv32Words = 3 * 3 * 64 * 4;   // 2304 words
constexpr int kInv32F64 = 2 * 4 * 3 * 64;     // 1536 doubles
hipError_t gen_inverse32(uint32_t* out, bool f16, int32_t shift, hipStream_t s);
hipError_t gen_inverse32_f64(double* out, int32_t shift, hipStream_t s);
// dst[0 .. n) = w[0 .. n) from host memory through kernel arguments (chunks of kTapsPerLaunch), so
// that no host-to-device copy is enqueued (the allocation-free entry point runs under graph capture).
constexpr int kTapsPerLaunch = 256;
hipError_t put_taps(double* dst, const double* w, int64_t n, hipStream_t s);
// d[i] = output i of SplitMix64(seed) as a double in [0, 1) (the host ffs_fill_uniform sequence)
hipError_t fill_uniform_device(double* d, int64_t n, uint64_t seed, hipStream_t s);
// layout 0: FP64 accumulator order (row freq q + 4v); layout 1: FP32 accumulator order (row freq 4q + v)
hipError_t gen_spectrum(double* H, int kind, int layout, const int32_t* axis_len, const double* d_taps,
                        int32_t ntaps_per_axis, int32_t dims, int32_t convention, int32_t radius, int64_t t,
                        hipStream_t s);
// Hf[2e], Hf[2e + 1] = (float)H[e], (float)H[len + e] for e < len: the FP32 copy of one depth's spectrum
// that the mixed-form kernels read from global memory (the batch kernels and the 3D kernel).
hipError_t spectrum_to_f32(const double* H, float* Hf, int64_t len, hipStream_t s);
hipError_t launch_block(const TileArgs& a, int kind, hipStream_t s, int num_cus);
// One fused block of every grid of a batch (32-point family: FP64, BF16, FP16), items/prefix in device memory.
hipError_t launch_batch(const TileArgs* d_items, const int64_t* d_prefix, int32_t n, int32_t precision, hipStream_t s,
                        int num_cus);

}  // namespace ffsk
