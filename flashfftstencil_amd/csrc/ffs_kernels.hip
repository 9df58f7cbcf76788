// ffs_kernels.hip -- gfx950 kernels of the tiled FFT-stencil engine.
//
// One fused time block (depth t) is one launch.  Per tile pair (two real tiles carried as the real
// and imaginary parts of one complex tile, as the reference does at src/2D/rfft_2d/2d_rfft_1_async.cu:81-88):
//   gather the periodic window -> forward DFT per axis -> multiply by H = K^t/(prod P) -> inverse DFT
//   per axis -> store the valid region (overlap-save; the reference stores overlap-add with atomics,
//   src/2D/rfft_2d/2d_rfft_1_async.cu:299-364, which this design replaces).
//
// MFMA layout facts (v_mfma_f64_16x16x4f64, measured on gfx950):
//   A operand:  lane i + 16k holds A[i][k]          (16x4 block)
//   B operand:  lane j + 16k holds B[k][j]          (4x16 block)
//   D (4 regs): lane j + 16q, reg v holds D[q + 4v][j]
// Hence D's register v, read in lane j+16q, is chunk v (rows 4v..4v+3) of D seen as the A operand of
// D^T: the product D^T * F contracts D's row index and leaves D's column index as the new row index.
// Each stage therefore contracts the current row axis and transposes, so the four stages of a 2D tile
// (fwd axis0, fwd axis1, inv axis0, inv axis1) chain in registers with no data exchange.  This
// replaces the reference's permuted-row DFT copy for accumulator reuse (src/2D/create_fft_pfa_plan.cu:27-37).
//
// Complex GEMMs use the 3-multiplication (Gauss/Karatsuba) form, as the reference's 1D kernel does
// (src/1D/rfftstencil_pfa/rfft_8_pfa_fastcomplex.cu:84-101):
//   k1 = (Ar+Ai)Br, k2 = Ar(Bi-Br), k3 = Ai(Br+Bi);  Re = k1-k3, Im = k1+k2.
// The inverse uses conj(F) with the same operand tables and flipped combination signs.
#include "ffs_internal.h"

#include <algorithm>
#include <type_traits>

namespace ffsk {

typedef double d4 __attribute__((ext_vector_type(4)));

#define FFS_MFMA64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)

__device__ __forceinline__ int64_t wrap64(int64_t i, int64_t n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Cheap wrap for |i| < 2n (the common case), falling back to the modulo for tiny grids.
__device__ __forceinline__ int64_t wrap_near(int64_t i, int64_t n) {
  if (i >= 0 && i < n) return i;
  if (i < 0 && i >= -n) return i + n;
  if (i >= n && i < 2 * n) return i - n;
  return wrap64(i, n);
}

// Complex product (yr + i yi)(hr + i hi) with the rounding pinned (explicit FMAs), so every kernel
// that applies the spectrum produces bitwise the same result for the same inputs.
template <typename T>
__device__ __forceinline__ void cmul(T& yr, T& yi, T hr, T hi) {
  const T r = __builtin_fma(yr, hr, -(yi * hi));
  const T i = __builtin_fma(yr, hi, yi * hr);
  yr = r;
  yi = i;
}
__device__ __forceinline__ void cmul(float& yr, float& yi, float hr, float hi) {
  const float r = __builtin_fmaf(yr, hr, -(yi * hi));
  const float i = __builtin_fmaf(yr, hi, yi * hr);
  yr = r;
  yi = i;
}

// Output stores of the tile kernels are non-temporal: a block's output grid is read only by the next
// block (hundreds of MiB, beyond L2 and MALL), so its lines should not displace the input windows, whose
// halos the neighbouring pairs re-read (FP16 Box2D25P block +1.5%, 96-grid FP16 batch +2%; same values).
__device__ __forceinline__ void put_out(double* p, double v) { __builtin_nontemporal_store(v, p); }

// XCD-aware workgroup order: the hardware hands workgroup b to XCD (b mod 8); renumber so that each
// XCD receives one contiguous range of logical workgroups (neighbouring tiles share its L2).
__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t G) {
  const int64_t x = b & 7, per = G >> 3, rem = G & 7;
  const int64_t start = x * per + (x < rem ? x : rem);
  return start + (b >> 3);
}
// The plan's wave-to-tile mapping: the hardware order (default) or XCD-contiguous bands.
__device__ __forceinline__ int64_t block_order(int32_t linear, int64_t b, int64_t G) {
  return linear ? b : xcd_logical_block(b, G);
}

struct OpsAxis {
  double r[4], s[4], d[4];
};

__device__ __forceinline__ void load_ops(OpsAxis& o, const double* __restrict__ ops, int axis, int lane) {
  const double* base = ops + axis * kOpsPerAxis;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    o.r[c] = base[(0 * 4 + c) * kLanes + lane];
    o.s[c] = base[(1 * 4 + c) * kLanes + lane];
    o.d[c] = base[(2 * 4 + c) * kLanes + lane];
  }
}

// out = A^T-chain stage: A given as 4 chunks (A operand values), B = F (or conj F when INV).  Gauss form
// on the tables r = Fr, s = -(Fr + Fi), d = Fi - Fr: k1 = (ar + ai) r first, then both outputs chain from
// it (C input k1, result a new accumulator), so the stage needs no output additions:
//   forward  Re = k1 + ai s, Im = k1 + ar d;   inverse  Re = k1 + ai d, Im = k1 + ar s.
template <bool INV>
__device__ __forceinline__ void cstage(const double (&ar)[4], const double (&ai)[4], const OpsAxis& f, d4& outr,
                                       d4& outi) {
  d4 k1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 4; c++) k1 = FFS_MFMA64(ar[c] + ai[c], f.r[c], k1);
  d4 re = k1, im = k1;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    re = FFS_MFMA64(ai[c], INV ? f.d[c] : f.s[c], re);
    im = FFS_MFMA64(ar[c], INV ? f.s[c] : f.d[c], im);
  }
  outr = re;
  outi = im;
}

__device__ __forceinline__ void to_chunks(const d4& v, double (&c)[4]) {
  c[0] = v[0];
  c[1] = v[1];
  c[2] = v[2];
  c[3] = v[3];
}

// ---------------------------------------------------------------------------------------------
// 2D 16x16 and 1D 240-sample prime-factor tiles, FP64.  One wave per tile pair, grid-stride loop
// over tile pairs with the next pair's window prefetched into registers while the current pair is
// on the matrix cores.  Tile indexing is wave-uniform (scalar registers).
//
// 1D prime-factor map (Good-Thomas, 240 = 15 x 16 coprime; the reference uses 56 = 8 x 7,
// src/1D/rfftstencil_pfa/rfft_8_pfa_fastcomplex.cu:39-47 and src/1D/create_fft_pfa_plan.cu:171-178):
//   window sample n <-> (row n2 = n mod 15, col n1 = n mod 16),  n = (225 n1 + 16 n2) mod 240
//   spectrum index k = (15 k1 + 16 k0) mod 240 for (row freq k0 < 15, col freq k1)
// Row 15 of the 16x16 block is padding: loaded as zero, its DFT rows/cols are zero, never stored.
// ---------------------------------------------------------------------------------------------
struct PairCoord {
  int64_t r0;     // 2D: first output row of the pair (both tiles share it)
  int64_t orgA;   // 2D: first output column of tile A; 1D: first output sample of tile A
  int64_t orgB;
  bool hasB;
};

template <int KIND>
__device__ __forceinline__ PairCoord pair_coord(const TileArgs& a, int64_t tp) {
  PairCoord pc;
  if (KIND == KIND_2D16) {
    const int64_t pairs1 = (a.tiles1 + 1) >> 1;
    const int64_t T0 = tp / pairs1, T1 = (tp - T0 * pairs1) * 2;
    pc.hasB = (T1 + 1) < a.tiles1;
    pc.r0 = T0 * a.V0;
    pc.orgA = T1 * a.V1;
  } else {
    const int64_t T1 = tp * 2;
    pc.hasB = (T1 + 1) < a.tiles1;
    pc.r0 = 0;
    pc.orgA = T1 * a.V1;
  }
  pc.orgB = pc.orgA + a.V1;
  return pc;
}

// Load the pair's windows as stage-1 A operands: chunk c of lane (j + 16q) = window[4c + q][j].
template <int KIND>
__device__ __forceinline__ void gather_pair(const TileArgs& a, const PairCoord& pc, int q, int j, double (&xr)[4],
                                            double (&xi)[4]) {
  const double* __restrict__ src = a.src;
  if (KIND == KIND_2D16) {
    const int64_t rw = pc.r0 - a.h0, cwA = pc.orgA - a.h1, cwB = pc.orgB - a.h1;
    const bool interior = rw >= 0 && rw + 16 <= a.n0 && cwA >= 0 && (pc.hasB ? cwB : cwA) + 16 <= a.n1;
    if (interior) {  // wave-uniform: plain affine addressing
      const double* p = src + (rw + q) * a.n1 + cwA + j;
      const int64_t step = 4 * a.n1;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        xr[c] = p[c * step];
        xi[c] = pc.hasB ? p[c * step + a.V1] : 0.0;
      }
    } else {
      const int64_t cA = wrap_near(cwA + j, a.n1);
      const int64_t cB = wrap_near(cwB + j, a.n1);
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int64_t row = wrap_near(rw + 4 * c + q, a.n0) * a.n1;
        xr[c] = src[row + cA];
        xi[c] = pc.hasB ? src[row + cB] : 0.0;
      }
    }
  } else {
    const int64_t wA = pc.orgA - a.h1, wB = pc.orgB - a.h1;
    const bool interior = wA >= 0 && (pc.hasB ? wB : wA) + 240 <= a.n1;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int row = 4 * c + q;
      const int n = (225 * j + 16 * row) % 240;
      const bool pad = row >= 15;
      if (interior) {
        xr[c] = pad ? 0.0 : src[wA + n];
        xi[c] = (pad || !pc.hasB) ? 0.0 : src[wB + n];
      } else {
        xr[c] = pad ? 0.0 : src[wrap_near(wA + n, a.n1)];
        xi[c] = (pad || !pc.hasB) ? 0.0 : src[wrap_near(wB + n, a.n1)];
      }
    }
  }
}

template <int KIND>
__device__ __forceinline__ void store_pair(const TileArgs& a, const PairCoord& pc, int q, int j, const d4& yr,
                                           const d4& yi) {
  if (KIND == KIND_2D16) {
    const int64_t cA = pc.orgA + j - a.h1, cB = pc.orgB + j - a.h1;
    const bool colok = j >= a.h1 && j < a.h1 + a.V1;
    const bool okA = colok && cA < a.n1;
    const bool okB = colok && pc.hasB && cB < a.n1;
    double* base = a.dst + (pc.r0 + q - a.h0) * a.n1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = q + 4 * v;
      if (row >= a.h0 && row < a.h0 + a.V0 && pc.r0 + row - a.h0 < a.n0) {
        double* drow = base + (int64_t)(4 * v) * a.n1;
        if (okA) put_out(drow + cA, yr[v]);
        if (okB) put_out(drow + cB, yi[v]);
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = q + 4 * v;
      if (row < 15) {
        const int n = (225 * j + 16 * row) % 240;
        if (n >= a.h1 && n < a.h1 + a.V1) {
          const int64_t pA = pc.orgA + n - a.h1, pB = pc.orgB + n - a.h1;
          if (pA < a.n1) put_out(a.dst + pA, yr[v]);
          if (pc.hasB && pB < a.n1) put_out(a.dst + pB, yi[v]);
        }
      }
    }
  }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_tile16_f64(TileArgs a) {
  // The spectrum (256 complex, 4 KB) is staged in LDS once per workgroup and read per tile pair,
  // which keeps the 2D kind within 128 VGPRs (4 waves per SIMD).  The 1D kind also stages its second
  // (15-point) operand set there and reads it just before each of its two stages, and does not prefetch
  // the next pair: its waves take about one pair each, and at 158 VGPRs three workgroups per CU are
  // resident instead of two.
  constexpr bool ONE_D = KIND != KIND_2D16;
  __shared__ double2 sH[256];
  __shared__ double sO1[ONE_D ? kOpsPerAxis : 1];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) sH[i] = double2{a.H[i], a.H[256 + i]};
  if (ONE_D)
    for (int i = threadIdx.x; i < kOpsPerAxis; i += blockDim.x) sO1[i] = a.ops[kOpsPerAxis + i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(a.linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // 2D: both axes are 16-point DFTs and share one operand set; 1D: 15-point rows (from LDS at each use),
  // 16-point columns.
  OpsAxis f0;
  load_ops(f0, a.ops, 0, lane);
  auto stage1 = [&](const double (&ar)[4], const double (&ai)[4], d4& outr, d4& outi, auto inv) {
    if (ONE_D) {
      OpsAxis g;
      load_ops(g, sO1, 0, lane);
      cstage<decltype(inv)::value>(ar, ai, g, outr, outi);
    } else {
      cstage<decltype(inv)::value>(ar, ai, f0, outr, outi);
    }
  };

  int64_t tp = gw;
  if (tp >= a.pairs) return;
  PairCoord pc = pair_coord<KIND>(a, tp);
  double xr[4], xi[4];
  gather_pair<KIND>(a, pc, q, j, xr, xi);
  for (; tp < a.pairs; tp += W) {
    // 2D: prefetch the next pair while this one computes
    const int64_t tn = tp + W;
    PairCoord pn = pc;
    double nr[4] = {0.0, 0.0, 0.0, 0.0}, ni[4] = {0.0, 0.0, 0.0, 0.0};
    if (tn < a.pairs) {
      pn = pair_coord<KIND>(a, tn);
      if (!ONE_D) gather_pair<KIND>(a, pn, q, j, nr, ni);
    }

    d4 sr, si;
    cstage<false>(xr, xi, f0, sr, si);  // contract axis 0 (rows)
    double cr[4], ci[4];
    to_chunks(sr, cr);
    to_chunks(si, ci);
    stage1(cr, ci, sr, si, std::false_type{});  // contract axis 1 -> Y[k0 = q+4v][k1 = j]
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const double2 h = sH[v * kLanes + lane];
      double yr = sr[v], yi = si[v];
      cmul(yr, yi, h.x, h.y);
      sr[v] = yr;
      si[v] = yi;
    }
    to_chunks(sr, cr);
    to_chunks(si, ci);
    cstage<true>(cr, ci, f0, sr, si);  // inverse over k0
    to_chunks(sr, cr);
    to_chunks(si, ci);
    stage1(cr, ci, sr, si, std::true_type{});  // inverse over k1 -> X'[n0 = q+4v][n1 = j]
    store_pair<KIND>(a, pc, q, j, sr, si);

    pc = pn;
    if (ONE_D) {
      if (tn < a.pairs) gather_pair<KIND>(a, pc, q, j, xr, xi);
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) {
        xr[c] = nr[c];
        xi[c] = ni[c];
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// 2D 32x32 tiles, FP64.  A 32x32 matrix is held as 4 blocks (R, C) of 16x16 in FP64-accumulator order:
//   block b = 2R + C, lane (j + 16q), reg v  <->  M[16R + q + 4v][16C + j],
// so register v of block (R, I) is the A operand of contraction rows 16R + 4v .. + 3 (lane (i + 16k)
// holds M[16R + 4v + k][16I + i]) and a stage Out = Z^T F is, per output block (I, Kb), 8 chunks x 3
// Gauss GEMMs = 24 v_mfma_f64_16x16x4f64 into three accumulators, as in the 16-point kernel.  With OUT16
// (V <= 16) the inverse stages compute only output block 0 from the inverse operands of window columns
// h .. h + 15: blocks (I, 0), then (0, 0).  DFT operands (24 KB), the FP64 spectrum (16 KB) and the
// out16 inverse operands (12 KB) are staged in LDS per workgroup.
// ---------------------------------------------------------------------------------------------
struct Blk4d {
  d4 b[4];
};

// Output blocks (I, Kb), I < NI, Kb < NK, from a table holding KBS column blocks per chunk: the full
// operands (NI = NK = KBS = 2), or for the inverse stages of an out16 tile the shifted one-block table
// (KBS = 1): NI = 2, NK = 1, then NI = NK = 1.  Gauss form with the tables r = Fr, n = -(Fr + Fi),
// d = Fi - Fr: k1 = (zr + zi) r is accumulated first, and both outputs chain from it (the MFMA's C input
// is k1, its result a new accumulator), so no output additions are needed:
//   forward  Re = k1 + zi n, Im = k1 + zr d;   inverse  Re = k1 + zi d, Im = k1 + zr n.
template <bool INV, int NI, int NK, int KBS>
__device__ __forceinline__ void stage32_f64(const Blk4d& zr, const Blk4d& zi, const double* __restrict__ sF, int lane,
                                            Blk4d& outr, Blk4d& outi) {
#pragma unroll
  for (int I = 0; I < NI; I++) {
    double sum[2][4];
#pragma unroll
    for (int R = 0; R < 2; R++)
#pragma unroll
      for (int v = 0; v < 4; v++) sum[R][v] = zr.b[2 * R + I][v] + zi.b[2 * R + I][v];
#pragma unroll
    for (int Kb = 0; Kb < NK; Kb++) {
      d4 k1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int v = 0; v < 4; v++) k1 = FFS_MFMA64(sum[R][v], sF[(((R * 4 + v) * KBS + Kb) * 3) * kLanes + lane], k1);
      d4 re = k1, im = k1;
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const double* f = sF + (((R * 4 + v) * KBS + Kb) * 3) * kLanes + lane;
          const double fn = f[kLanes], fd = f[2 * kLanes];
          re = FFS_MFMA64(zi.b[2 * R + I][v], INV ? fd : fn, re);
          im = FFS_MFMA64(zr.b[2 * R + I][v], INV ? fn : fd, im);
        }
      outr.b[2 * I + Kb] = re;
      outi.b[2 * I + Kb] = im;
    }
  }
}

// Interior windows are read through a raw buffer view based at the window's first element: one 32-bit
// per-lane offset plus a wave-uniform offset per row, so no 64-bit address arithmetic per load. The flags
// word is the gfx9 raw-buffer setting (32-bit data format). A view with zero records reads as zeros, which
// stands in for the absent second tile of an odd pair. Offsets stay below 2^31 for n1 <= 2^22.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t window_rsrc(const double* base, bool live) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, live ? 0x7fffffff : 0, 0x00020000);
}

__device__ __forceinline__ double window_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

// Window values of the pair in block order: x.b[2R + C][v] = window[16R + q + 4v][16C + j].
__device__ __forceinline__ void gather32_f64(const TileArgs& a, const PairCoord& pc, int q, int j, Blk4d& xa,
                                             Blk4d& xb) {
  const double* __restrict__ src = a.src;
  const int64_t rw = pc.r0 - a.h0, cwA = pc.orgA - a.h1, cwB = pc.orgB - a.h1;
  const bool interior = rw >= 0 && rw + 32 <= a.n0 && cwA >= 0 && (pc.hasB ? cwB : cwA) + 32 <= a.n1 &&
                        a.n1 <= (1 << 22);
  if (interior) {
    const double* baseA = src + rw * a.n1 + cwA;
    const __amdgpu_buffer_rsrc_t ra = window_rsrc(baseA, true), rb = window_rsrc(baseA + a.V1, pc.hasB);
    const uint32_t voff = ((uint32_t)q * (uint32_t)a.n1 + (uint32_t)j) * 8u;
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t soff = ((uint32_t)(16 * (b >> 1) + 4 * v) * (uint32_t)a.n1 + 16 * (b & 1)) * 8u;
        xa.b[b][v] = window_load(ra, voff, soff);
        xb.b[b][v] = window_load(rb, voff, soff);
      }
  } else {
#pragma unroll
    for (int C = 0; C < 2; C++) {
      const int64_t cA = wrap_near(cwA + 16 * C + j, a.n1);
      const int64_t cB = wrap_near(cwB + 16 * C + j, a.n1);
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int64_t row = wrap_near(rw + 16 * R + q + 4 * v, a.n0) * a.n1;
          xa.b[2 * R + C][v] = src[row + cA];
          xb.b[2 * R + C][v] = pc.hasB ? src[row + cB] : 0.0;
        }
    }
  }
}

// Store the valid outputs of all four blocks (window rows 16R + q + 4v, columns 16C + j).
__device__ __forceinline__ void store32_f64(const TileArgs& a, const PairCoord& pc, int q, int j, const Blk4d& yr,
                                            const Blk4d& yi) {
#pragma unroll
  for (int C = 0; C < 2; C++) {
    const int wc = 16 * C + j;
    const int64_t cA = pc.orgA + wc - a.h1, cB = pc.orgB + wc - a.h1;
    const bool colok = wc >= a.h1 && wc < a.h1 + a.V1;
    const bool okA = colok && cA < a.n1;
    const bool okB = colok && pc.hasB && cB < a.n1;
#pragma unroll
    for (int R = 0; R < 2; R++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 16 * R + q + 4 * v;
        const int64_t y = pc.r0 + row - a.h0;
        if (row >= a.h0 && row < a.h0 + a.V0 && y < a.n0) {
          double* drow = a.dst + y * a.n1;
          if (okA) put_out(drow + cA, yr.b[2 * R + C][v]);
          if (okB) put_out(drow + cB, yi.b[2 * R + C][v]);
        }
      }
  }
}

// Store of an out16 tile pair: output block 0, lane (j + 16q), reg v = valid point (q + 4v, j) (window
// position (h0 + q + 4v, h1 + j)).
__device__ __forceinline__ void store16_f64(const TileArgs& a, const PairCoord& pc, int q, int j, const d4& yr,
                                            const d4& yi) {
  const int64_t cA = pc.orgA + j, cB = pc.orgB + j;
  const bool okA = j < a.V1 && cA < a.n1;
  const bool okB = j < a.V1 && pc.hasB && cB < a.n1;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = q + 4 * v;
    const int64_t y = pc.r0 + row;
    if (row < a.V0 && y < a.n0) {
      double* drow = a.dst + y * a.n1;
      if (okA) put_out(drow + cA, yr[v]);
      if (okB) put_out(drow + cB, yi[v]);
    }
  }
}

// One tile pair of the 32-point FP64 family; the spectrum comes from LDS (single-grid launch) or from
// the item's FP64 table in global memory (batch launch).
// The inverse operands of shift h (output column m <- window column (m + h) mod 32) are columns of the
// forward operand table: lane j + 16q of column block 0 of the shifted table holds what lane
// (c & 15) + 16q of column block c >> 4 holds in the forward one, c = (j + h) mod 32.  ShiftedView
// gives this lane's column block (`base`, read with KBS = 2 as block 0) and the lane to read there.
template <typename T>
struct ShiftedView {
  const T* base;
  int lane;
};
template <typename T>
__device__ __forceinline__ ShiftedView<T> shifted_view(const T* table, int h, int lane) {
  const int c = ((lane & 15) + h) & 31;
  return {table + (c >> 4) * 3 * kLanes, (lane & 48) | (c & 15)};
}

// IKBS: column blocks per chunk of the inverse-operand table sG (1: the shifted one-block table; 2: the
// forward operand table, read through a shifted_view whose lane is `ilane`).
template <bool H_FROM_LDS, bool OUT16, int IKBS = 1>
__device__ __forceinline__ void tile32_pair_f64(const TileArgs& a, int64_t tp, const double* __restrict__ sF,
                                                const double2* __restrict__ sH, const double* __restrict__ sG,
                                                int lane, int q, int j, int ilane = -1) {
  const PairCoord pc = pair_coord<KIND_2D16>(a, tp);
  Blk4d zr, zi, orr, oi;
  gather32_f64(a, pc, q, j, zr, zi);
  stage32_f64<false, 2, 2, 2>(zr, zi, sF, lane, orr, oi);
  stage32_f64<false, 2, 2, 2>(orr, oi, sF, lane, zr, zi);  // Y[k0][k1], block order
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int e = (b * 4 + v) * kLanes + lane;
      const double hr = H_FROM_LDS ? sH[e].x : a.H[e];
      const double hi = H_FROM_LDS ? sH[e].y : a.H[1024 + e];
      double yr = zr.b[b][v], yi = zi.b[b][v];
      cmul(yr, yi, hr, hi);
      zr.b[b][v] = yr;
      zi.b[b][v] = yi;
    }
  if (OUT16) {
    const int il = ilane < 0 ? lane : ilane;  // the lane whose inverse operands this lane reads
    stage32_f64<true, 2, 1, IKBS>(zr, zi, sG, il, orr, oi);  // blocks 0 and 2: window rows h0 ..
    stage32_f64<true, 1, 1, IKBS>(orr, oi, sG, il, zr, zi);  // block 0 = the valid region
    store16_f64(a, pc, q, j, zr.b[0], zi.b[0]);
  } else {
    stage32_f64<true, 2, 2, 2>(zr, zi, sF, lane, orr, oi);
    stage32_f64<true, 2, 2, 2>(orr, oi, sF, lane, zr, zi);  // X'[n0][n1], block order
    store32_f64(a, pc, q, j, zr, zi);
  }
}

template <bool OUT16>
__global__ __launch_bounds__(256, 2) void k_tile32_f64(TileArgs a) {
  __shared__ double sF[kOps32F64];           // 24 KB
  __shared__ double2 sH[4 * 4 * kLanes];     // spectrum [block][reg][lane] (re, im), 16 KB
  __shared__ double sG[OUT16 ? kInv32F64 : 1];  // out16 inverse operands, 12 KB
  for (int i = threadIdx.x; i < kOps32F64; i += blockDim.x) sF[i] = a.ops[i];
  if (OUT16)
    for (int i = threadIdx.x; i < kInv32F64; i += blockDim.x) sG[i] = static_cast<const double*>(a.opsi)[i];
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) sH[i] = double2{a.H[i], a.H[1024 + i]};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(a.linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int64_t tp = gw; tp < a.pairs; tp += W) tile32_pair_f64<true, OUT16>(a, tp, sF, sH, sG, lane, q, j);
}

// Multi-grid launch of the FP64 form (the k_batch32_* structure): the DFT operands, the same for every
// item, in LDS; each item's spectrum read per tile pair from its FP64 table (L2-resident).
// Work item idx of a batch round -> grid g with prefix[g] <= idx < prefix[g + 1].  A wave takes its items in
// increasing order, so the search starts after the grid of its previous item and most items stay in that
// grid (no dependent scalar loads); a change costs one binary search over the later grids.  g = -1 and
// next = 0 before the first item.
__device__ __forceinline__ int batch_grid(const int64_t* __restrict__ prefix, int32_t n, int64_t idx, int g,
                                          int64_t& next) {
  if (idx < next) return g;
  int lo = g + 1, hi = n;  // prefix[lo] = next <= idx
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid;
  }
  next = prefix[lo + 1];
  return lo;
}

__global__ __launch_bounds__(256, 2) void k_batch32_f64(const TileArgs* __restrict__ items,
                                                        const int64_t* __restrict__ prefix, int32_t n) {
  __shared__ double sF[kOps32F64];
  for (int i = threadIdx.x; i < kOps32F64; i += blockDim.x) sF[i] = items[0].ops[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(items[0].linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t total = prefix[n];
  int lo = -1;
  int64_t next = 0;
  for (int64_t idx = gw; idx < total; idx += W) {
    lo = batch_grid(prefix, n, idx, lo, next);
    const TileArgs a = items[lo];  // independent scalar loads, from the scalar cache while lo is unchanged
    if (a.out16) {
      const ShiftedView<double> v = shifted_view(sF, a.h1, lane);
      tile32_pair_f64<false, true, 2>(a, idx - prefix[lo], sF, nullptr, v.base, lane, q, j, v.lane);
    } else {
      tile32_pair_f64<false, false>(a, idx - prefix[lo], sF, nullptr, nullptr, lane, q, j);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Mixed-precision forms, 2D 16x16 and 1D prime-factor tiles.
//
// FP32 accumulator layout (v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x4f32, measured on gfx950):
//   D (4 regs): lane j + 16q, reg v holds D[4q + v][j]
//   bf16 A:     lane i + 16q, element e (0..7) holds A[i][8q + e]; B: lane j + 16q, element e holds B[8q + e][j]
//   f32 A/B:    as the FP64 form (lane i + 16k holds A[i][k], lane j + 16k holds B[k][j])
// The K = 32 contraction of the BF16 instruction carries 16 rows x 2 split terms: element e of lane
// group q is (row 4q + (e & 3), term e >> 2), so D's four registers feed the next stage's A operand
// (as D^T) without any exchange, exactly like the FP64 chain.  The data (FP32) is split into 3 BF16
// terms a = ah + am + al and the DFT matrix (FP64) into 3 terms F = Fh + Fm + Fl; each real 16x16x16
// GEMM keeps the six products of order <= 2^-16 in three MFMAs (two products per K = 32):
//   [ah | am] x [Fh ; Fh]  +  [ah | am] x [Fm ; Fm]  +  [ah | al] x [Fl ; Fh]
// accumulated in FP32 (~FP32 accuracy).  NTERMS = 2 keeps a 2-term data split (16 bits) with two MFMAs,
//   [ah | am] x [Fh ; Fh]  +  [ah | ah] x [Fm ; Fl],
// whose error (~4e-6 of the fluctuation per block) only meets 1e-6 once the field has smoothed.
// The FP32 form uses one v_mfma_f32_16x16x4f32 per 4-row chunk, with
// the DFT tables stored in the permuted chunk order (rows 4k + c) that the FP32 accumulator implies.
//
// Every tile window is taken relative to a DC offset c (FP64, the window's center sample): the FFT path
// runs on x - c and the store adds c * K(0)^t back in FP64 (a constant field is an eigenvector with
// eigenvalue K(0)).  This keeps the split/rounding errors relative to the fluctuation, not to the
// (large, slowly varying) level of the field, and is what holds the mixed forms below 1e-6 over
// hundreds of steps.  The grid state between blocks stays FP64 in HBM.  The complex products are the plain 4-GEMM
// form (no Gauss trick): these forms are VALU-bound, not MFMA-bound, and the 4-GEMM form needs fewer
// operand splits and no combination adds.
// ---------------------------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

#define FFS_MFMA_BF16(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)
#define FFS_MFMA_F32(a, b, c) __builtin_amdgcn_mfma_f32_16x16x4f32((a), (b), (c), 0, 0, 0)

// No inline assembly anywhere in these kernels.  On gfx950 a VALU read of a VGPR that an MFMA wrote, and
// an MFMA read (SrcA/B/C) of a VGPR that a VALU wrote, need software wait states (the hardware does not
// interlock them); the compiler's hazard recognizer inserts them, but it does not see an inline-asm
// statement as a VALU instruction, so an asm that consumes an MFMA result or produces an MFMA operand
// gets no wait states.  Every split below is therefore plain code the compiler schedules itself.

enum { PREC_BF16 = 1, PREC_F32 = 2, PREC_F16 = 4 };

template <int NTERMS>
struct OpsBF16 {
  bf8 r[NTERMS == 3 ? 3 : 2], i[NTERMS == 3 ? 3 : 2], ni[NTERMS == 3 ? 3 : 2];  // pairings used by the scheme
};
struct OpsF32 {
  float r[4], i[4], ni[4];  // [chunk]
};
struct OpsF32C {  // FP32 operands and their low-order terms, F = hi + lo to ~2^-48
  OpsF32 hi, lo;
};

__device__ __forceinline__ bf8 as_bf8(u4 w) { return __builtin_bit_cast(bf8, w); }

template <int NTERMS>
__device__ __forceinline__ void load_ops_mixed(OpsBF16<NTERMS>& o, const uint32_t* __restrict__ t, int axis,
                                               int lane) {
  const uint32_t* base = t + axis * kBf16WordsPerAxis;
  bf8* dst[3] = {o.r, o.i, o.ni};
  constexpr int ng = NTERMS == 3 ? 3 : 2;
#pragma unroll
  for (int arr = 0; arr < 3; arr++)
#pragma unroll
    for (int g = 0; g < ng; g++) {
      const int pg = NTERMS == 3 ? g : (g == 0 ? 0 : 3);
      u4 w;
#pragma unroll
      for (int k = 0; k < 4; k++) w[k] = base[((arr * kPairings + pg) * 4 + k) * kLanes + lane];
      dst[arr][g] = as_bf8(w);
    }
}

__device__ __forceinline__ void load_ops_f32(OpsF32& o, const uint32_t* __restrict__ t, int offset, int axis,
                                             int lane) {
  const float* base = reinterpret_cast<const float*>(t + offset) + axis * kF32PerAxis;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    o.r[c] = base[(0 * 4 + c) * kLanes + lane];
    o.i[c] = base[(1 * 4 + c) * kLanes + lane];
    o.ni[c] = base[(2 * 4 + c) * kLanes + lane];
  }
}
__device__ __forceinline__ void load_ops_mixed(OpsF32& o, const uint32_t* __restrict__ t, int axis, int lane) {
  load_ops_f32(o, t, kOpsF32Offset, axis, lane);
}
__device__ __forceinline__ void load_ops_mixed(OpsF32C& o, const uint32_t* __restrict__ t, int axis, int lane) {
  load_ops_f32(o.hi, t, kOpsF32Offset, axis, lane);
  load_ops_f32(o.lo, t, kOpsF32LoOffset, axis, lane);
}

typedef float f2 __attribute__((ext_vector_type(2)));
typedef __bf16 b2 __attribute__((ext_vector_type(2)));

// Two FP32 values -> one word of two BF16 (round to nearest even, one v_cvt_pk_bf16_f32); element 0
// (`a`) in the low half.
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, b2));
}
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// Split of four FP32 values (rows 4q..4q+3 of this lane's column) into the A operands of the scheme.
//   NTERMS = 3: A[0] = A[1] = [am | ah] (pairs with (Fh|Fh) and (Fm|Fm)), A[2] = [ah | al] (pairs with (Fl|Fh))
//   NTERMS = 2: A[0] = [am | ah] (with (Fh|Fh)), A[1] = [ah | ah] (with (Fm|Fl))
// Putting am before ah lets [am | ah] and [ah | al] share the ah registers.
template <int NTERMS>
__device__ __forceinline__ void split_a(const f4& x, bf8 (&A)[NTERMS]) {
  const uint32_t h01 = pk_bf16(x[0], x[1]), h23 = pk_bf16(x[2], x[3]);
  const float r0 = x[0] - bf_lo(h01), r1 = x[1] - bf_hi(h01), r2 = x[2] - bf_lo(h23), r3 = x[3] - bf_hi(h23);
  const uint32_t m01 = pk_bf16(r0, r1), m23 = pk_bf16(r2, r3);
  A[0] = as_bf8(u4{m01, m23, h01, h23});
  if (NTERMS == 3) {
    const uint32_t l01 = pk_bf16(r0 - bf_lo(m01), r1 - bf_hi(m01)), l23 = pk_bf16(r2 - bf_lo(m23), r3 - bf_hi(m23));
    A[1] = A[0];
    A[2] = as_bf8(u4{h01, h23, l01, l23});
  } else {
    A[1] = as_bf8(u4{h01, h23, h01, h23});
  }
}

// out = A^T F (forward) or A^T conj(F) (INV); A complex, given as this lane's 4 rows of A's column.
//   forward: Re = ar Fr + ai (-Fi), Im = ar Fi + ai Fr;  inverse: Re = ar Fr + ai Fi, Im = ar (-Fi) + ai Fr
// v_mfma_f32_16x16x32_bf16 does not round like a chain of FP32 FMAs (measured with single-wave probes):
// each output sums its 32 products in four groups of 8 (the K values one 16-lane group of A holds), every
// product truncated toward zero 24 bits below the group's largest exponent sum e(a) + e(b), and the
// accumulator C is truncated toward zero 24 bits below the largest over all 32; only the final sum is
// rounded to nearest.  An accumulator smaller than the products therefore loses its low bits, always
// toward zero.  The split's small products (pairings 1 and 2, about 2^-8 and 2^-16 of the result)
// are summed on their own, and the two dominant products (pairing 0: ah*Fh and am*Fh, real and
// imaginary input) each start from zero; the three partial sums are added in FP32.  Chained into
// one accumulator instead (small first), the truncation biased every stage towards zero: on the
// Box3D27P 512^3 x 32 sinusoid field 8.9e-7 max-rel with a mean error of -1.3e-7, against 5.9e-7 and
// -1.3e-8 this way.
template <bool INV, int NTERMS>
__device__ __forceinline__ void mstage(const f4& ar, const f4& ai, const OpsBF16<NTERMS>& f, f4& outr, f4& outi) {
  bf8 R[NTERMS], I[NTERMS];
  split_a<NTERMS>(ar, R);
  split_a<NTERMS>(ai, I);
  const f4 z = {0.f, 0.f, 0.f, 0.f};
  f4 re = z, im = z;
#pragma unroll
  for (int g = NTERMS - 1; g >= 1; g--) {
    re = FFS_MFMA_BF16(R[g], f.r[g], re);
    im = FFS_MFMA_BF16(R[g], INV ? f.ni[g] : f.i[g], im);
    re = FFS_MFMA_BF16(I[g], INV ? f.i[g] : f.ni[g], re);
    im = FFS_MFMA_BF16(I[g], f.r[g], im);
  }
  const f4 rr = FFS_MFMA_BF16(R[0], f.r[0], z);
  const f4 ri = FFS_MFMA_BF16(R[0], INV ? f.ni[0] : f.i[0], z);
  const f4 ir = FFS_MFMA_BF16(I[0], INV ? f.i[0] : f.ni[0], z);
  const f4 ii = FFS_MFMA_BF16(I[0], f.r[0], z);
  outr = (rr + ir) + re;
  outi = (ri + ii) + im;
}

template <bool INV, int C0 = 0, int C1 = 4>
__device__ __forceinline__ void mstage_f32(const f4& ar, const f4& ai, const OpsF32& f, f4& re, f4& im) {
#pragma unroll
  for (int c = C0; c < C1; c++) {
    re = FFS_MFMA_F32(ar[c], f.r[c], re);
    im = FFS_MFMA_F32(ar[c], INV ? f.ni[c] : f.i[c], im);
    re = FFS_MFMA_F32(ai[c], INV ? f.i[c] : f.ni[c], re);
    im = FFS_MFMA_F32(ai[c], f.r[c], im);
  }
}
// FP32 stage: v_mfma_f32_16x16x4f32 rounds every product and every addition (in sequence) to FP32, so
// chunks 0-1 and 2-3 (8 of the 16 K rows each) go to two accumulators, added at the end: 16 roundings
// in each chain instead of 32 in one.
template <bool INV, int NTERMS>
__device__ __forceinline__ void mstage(const f4& ar, const f4& ai, const OpsF32& f, f4& outr, f4& outi) {
  const f4 z = {0.f, 0.f, 0.f, 0.f};
  f4 re = z, im = z, re2 = z, im2 = z;
  mstage_f32<INV, 0, 2>(ar, ai, f, re, im);
  mstage_f32<INV, 2, 4>(ar, ai, f, re2, im2);
  outr = re + re2;
  outi = im + im2;
}
// Compensated FP32 stage: the products with the matrix's low-order terms first (about 2^-24 of the
// result, so their roundings are negligible), then those with its FP32-rounded entries.  The matrix then
// enters to ~2^-48 instead of FP32's 2^-25: its rounding is the same in every window and every block, so
// uncompensated it adds up coherently over a run.  What remains is the rounding along the accumulation
// chain, halved as in the FP32 stage above: the main products of chunks 0-1 and 2-3 go to two
// accumulators, added at the end.
template <bool INV, int NTERMS>
__device__ __forceinline__ void mstage(const f4& ar, const f4& ai, const OpsF32C& f, f4& outr, f4& outi) {
  const f4 z = {0.f, 0.f, 0.f, 0.f};
  f4 re = z, im = z, re2 = z, im2 = z;
  mstage_f32<INV>(ar, ai, f.lo, re, im);
  mstage_f32<INV, 0, 2>(ar, ai, f.hi, re, im);
  mstage_f32<INV, 2, 4>(ar, ai, f.hi, re2, im2);
  outr = re + re2;
  outi = im + im2;
}

// Load the pair's windows in the FP32-accumulator order: lane (j + 16q), element r = window[4q + r][j].
template <int KIND>
__device__ __forceinline__ void gather_pair_m(const TileArgs& a, const PairCoord& pc, int q, int j, double (&xr)[4],
                                              double (&xi)[4]) {
  const double* __restrict__ src = a.src;
  if (KIND == KIND_2D16) {
    const int64_t rw = pc.r0 - a.h0, cwA = pc.orgA - a.h1, cwB = pc.orgB - a.h1;
    const bool interior = rw >= 0 && rw + 16 <= a.n0 && cwA >= 0 && (pc.hasB ? cwB : cwA) + 16 <= a.n1;
    if (interior) {
      const double* p = src + (rw + 4 * q) * a.n1 + cwA + j;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        xr[r] = p[r * a.n1];
        xi[r] = pc.hasB ? p[r * a.n1 + a.V1] : 0.0;
      }
    } else {
      const int64_t cA = wrap_near(cwA + j, a.n1);
      const int64_t cB = wrap_near(cwB + j, a.n1);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t row = wrap_near(rw + 4 * q + r, a.n0) * a.n1;
        xr[r] = src[row + cA];
        xi[r] = pc.hasB ? src[row + cB] : 0.0;
      }
    }
  } else {
    const int64_t wA = pc.orgA - a.h1, wB = pc.orgB - a.h1;
    const bool interior = wA >= 0 && (pc.hasB ? wB : wA) + 240 <= a.n1;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = 4 * q + r;
      const int n = (225 * j + 16 * row) % 240;
      const bool pad = row >= 15;
      if (interior) {
        xr[r] = pad ? 0.0 : src[wA + n];
        xi[r] = (pad || !pc.hasB) ? 0.0 : src[wB + n];
      } else {
        xr[r] = pad ? 0.0 : src[wrap_near(wA + n, a.n1)];
        xi[r] = (pad || !pc.hasB) ? 0.0 : src[wrap_near(wB + n, a.n1)];
      }
    }
  }
}

template <int KIND>
__device__ __forceinline__ void store_pair_m(const TileArgs& a, const PairCoord& pc, int q, int j, const f4& yr,
                                             const f4& yi, double addA, double addB) {
  if (KIND == KIND_2D16) {
    const int64_t cA = pc.orgA + j - a.h1, cB = pc.orgB + j - a.h1;
    const bool colok = j >= a.h1 && j < a.h1 + a.V1;
    const bool okA = colok && cA < a.n1;
    const bool okB = colok && pc.hasB && cB < a.n1;
    double* base = a.dst + (pc.r0 + 4 * q - a.h0) * a.n1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = 4 * q + v;
      if (row >= a.h0 && row < a.h0 + a.V0 && pc.r0 + row - a.h0 < a.n0) {
        double* drow = base + (int64_t)v * a.n1;
        if (okA) put_out(drow + cA, (double)yr[v] + addA);
        if (okB) put_out(drow + cB, (double)yi[v] + addB);
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = 4 * q + v;
      if (row < 15) {
        const int n = (225 * j + 16 * row) % 240;
        if (n >= a.h1 && n < a.h1 + a.V1) {
          const int64_t pA = pc.orgA + n - a.h1, pB = pc.orgB + n - a.h1;
          if (pA < a.n1) put_out(a.dst + pA, (double)yr[v] + addA);
          if (pc.hasB && pB < a.n1) put_out(a.dst + pB, (double)yi[v] + addB);
        }
      }
    }
  }
}

// Wave-uniform copy of lane L's value.
__device__ __forceinline__ double lane_value(double v, int L) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, L), hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), L);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// One tile pair through the mixed pipeline.  The DC offset of each window is its center sample, which
// tracks the local mean of a smooth field and bounds the fluctuation by half the window's range
// otherwise.  2D: window (8, 8) = lane 8 + 16*2, element 0; 1D: window sample 120 = (row 0, col 8) =
// lane 8, element 0.
template <int KIND, int NTERMS, typename Ops>
__device__ __forceinline__ void mixed_pair(const TileArgs& a, const PairCoord& pc, int q, int j, const double (&xr)[4],
                                           const double (&xi)[4], const Ops& f0, const Ops& g1, const f4& hr,
                                           const f4& hi) {
  const int cl = KIND == KIND_2D16 ? 40 : 8;
  const double cA = lane_value(xr[0], cl), cB = lane_value(xi[0], cl);
  f4 ar, ai;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    ar[r] = (float)(xr[r] - cA);
    ai[r] = (float)(xi[r] - cB);
  }
  f4 sr, si;
  mstage<false, NTERMS>(ar, ai, f0, sr, si);
  mstage<false, NTERMS>(sr, si, g1, ar, ai);  // Y[k0 = 4q+v][k1 = j]
#pragma unroll
  for (int v = 0; v < 4; v++) {
    float yr = ar[v], yi = ai[v];
    cmul(yr, yi, hr[v], hi[v]);
    sr[v] = yr;
    si[v] = yi;
  }
  mstage<true, NTERMS>(sr, si, f0, ar, ai);
  mstage<true, NTERMS>(ar, ai, g1, sr, si);  // X'[n0 = 4q+v][n1 = j]
  store_pair_m<KIND>(a, pc, q, j, sr, si, cA * a.dc_gain, cB * a.dc_gain);
}

template <int KIND, int PREC, int NTERMS>
__global__ __launch_bounds__(256) void k_tile16_mixed(TileArgs a) {
  typedef typename std::conditional<PREC == PREC_BF16, OpsBF16<NTERMS>, OpsF32>::type Ops;
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(a.linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  Ops f0, f1;
  load_ops_mixed(f0, a.opsm, 0, lane);
  if (KIND != KIND_2D16) load_ops_mixed(f1, a.opsm, 1, lane);
  const Ops& g1 = (KIND == KIND_2D16) ? f0 : f1;
  f4 hr, hi;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    hr[v] = (float)a.H[v * kLanes + lane];
    hi[v] = (float)a.H[256 + v * kLanes + lane];
  }

  // Two register sets alternate (unrolled by 2): the next pair's window loads while this one computes.
  int64_t tp = gw;
  if (tp >= a.pairs) return;
  PairCoord pa = pair_coord<KIND>(a, tp), pb = pa;
  double ar[4], ai[4], br[4] = {0.0, 0.0, 0.0, 0.0}, bi[4] = {0.0, 0.0, 0.0, 0.0};
  gather_pair_m<KIND>(a, pa, q, j, ar, ai);
  while (true) {
    if (tp + W < a.pairs) {
      pb = pair_coord<KIND>(a, tp + W);
      gather_pair_m<KIND>(a, pb, q, j, br, bi);
    }
    mixed_pair<KIND, NTERMS>(a, pa, q, j, ar, ai, f0, g1, hr, hi);
    tp += W;
    if (tp >= a.pairs) break;
    if (tp + W < a.pairs) {
      pa = pair_coord<KIND>(a, tp + W);
      gather_pair_m<KIND>(a, pa, q, j, ar, ai);
    }
    mixed_pair<KIND, NTERMS>(a, pb, q, j, br, bi, f0, g1, hr, hi);
    tp += W;
    if (tp >= a.pairs) break;
  }
}

// ---------------------------------------------------------------------------------------------
// 2D 32x32 tiles, mixed BF16 form.
//
// A 32x32 matrix is held as 4 blocks (R, C) of 16x16 in FP32-accumulator order:
//   block b = 2R + C, lane (j + 16q), reg v  <->  M[16R + 4q + v][16C + j].
// A stage Out = Z^T F contracts all 32 rows of Z: output block (I, Kb) = sum_R Z(R, I)^T F(R, Kb).  The
// BF16 A operand of one MFMA spans K = 32 = both row blocks: element e of lane (i + 16q) is
// Z[16 (e >> 2) + 4q + (e & 3)][16 I + i], i.e. registers (e & 3) of blocks (0, I) and (1, I), so the
// chain again needs no exchange between lanes.  Each real 16x16 output block of a GEMM keeps the
// split products of order <= 2^-16: (hh, hm, mh, hl, mm, lh) for NTERMS = 3 (6 MFMAs), (hh, hm, mh, hl)
// for NTERMS = 2 (4 MFMAs).  DFT operands (18 KB) and the spectrum (8 KB) are staged once per
// workgroup in LDS.  With the 32-point window the fused depth can be 2-3x deeper than at 16 points
// for the same valid fraction, which cuts the split work and the HBM traffic per stencil step.
// ---------------------------------------------------------------------------------------------
struct Blk4 {
  f4 b[4];
};

template <int NTERMS>
__device__ __forceinline__ void split8(const f4& x0, const f4& x1, bf8& h, bf8& m, bf8& l) {
  const uint32_t h0 = pk_bf16(x0[0], x0[1]), h1 = pk_bf16(x0[2], x0[3]);
  const uint32_t h2 = pk_bf16(x1[0], x1[1]), h3 = pk_bf16(x1[2], x1[3]);
  const float r0 = x0[0] - bf_lo(h0), r1 = x0[1] - bf_hi(h0), r2 = x0[2] - bf_lo(h1), r3 = x0[3] - bf_hi(h1);
  const float r4 = x1[0] - bf_lo(h2), r5 = x1[1] - bf_hi(h2), r6 = x1[2] - bf_lo(h3), r7 = x1[3] - bf_hi(h3);
  const uint32_t m0 = pk_bf16(r0, r1), m1 = pk_bf16(r2, r3), m2 = pk_bf16(r4, r5), m3 = pk_bf16(r6, r7);
  h = as_bf8(u4{h0, h1, h2, h3});
  m = as_bf8(u4{m0, m1, m2, m3});
  if (NTERMS == 3) {
    const uint32_t l0 = pk_bf16(r0 - bf_lo(m0), r1 - bf_hi(m0)), l1 = pk_bf16(r2 - bf_lo(m1), r3 - bf_hi(m1));
    const uint32_t l2 = pk_bf16(r4 - bf_lo(m2), r5 - bf_hi(m2)), l3 = pk_bf16(r6 - bf_lo(m3), r7 - bf_hi(m3));
    l = as_bf8(u4{l0, l1, l2, l3});
  }
}

// acc += (ah + am + al) x (Fh + Fm + Fl) without the dominant product ah x Fh, i.e. the split products
// of order 2^-8 .. 2^-16 (added first, while the accumulator is still small).
template <int NTERMS>
__device__ __forceinline__ f4 prod32_small(const bf8& ah, const bf8& am, const bf8& al, const bf8 (&F)[3], f4 acc) {
  if (NTERMS == 3) {
    acc = FFS_MFMA_BF16(al, F[0], acc);
    acc = FFS_MFMA_BF16(am, F[1], acc);
  }
  acc = FFS_MFMA_BF16(ah, F[2], acc);
  acc = FFS_MFMA_BF16(am, F[0], acc);
  acc = FFS_MFMA_BF16(ah, F[1], acc);
  return acc;
}

// KBS: column blocks per arr in the table (2: the full operands; 1: the out16 inverse operands).
template <int KBS>
__device__ __forceinline__ void load_f32ops(const u4* __restrict__ sF, int arr, int Kb, int lane, bf8 (&F)[3]) {
#pragma unroll
  for (int t = 0; t < 3; t++) F[t] = as_bf8(sF[((arr * KBS + Kb) * 3 + t) * kLanes + lane]);
}

// Output blocks (I, Kb), I < NI, Kb < NK, from a table of KBS column blocks: all four from the full
// operands (NI = NK = KBS = 2); for the inverse stages of an out16 tile block Kb = 0 of the shifted
// one-block table (the valid window rows), NI = 2, NK = KBS = 1, then NI = NK = KBS = 1.
template <bool INV, int NTERMS, int NI = 2, int NK = 2, int KBS = 2>
__device__ __forceinline__ void stage32(const Blk4& zr, const Blk4& zi, const u4* __restrict__ sF, int lane, Blk4& outr,
                                        Blk4& outi) {
#pragma unroll
  for (int I = 0; I < NI; I++) {
    bf8 rh, rm, rl, ih, im, il;
    const f4 x0 = zr.b[I], x1 = zr.b[2 + I], y0 = zi.b[I], y1 = zi.b[2 + I];
    split8<NTERMS>(x0, x1, rh, rm, rl);
    split8<NTERMS>(y0, y1, ih, im, il);
#pragma unroll
    for (int Kb = 0; Kb < NK; Kb++) {
      bf8 Fr[3], Fi[3], Fn[3];
      load_f32ops<KBS>(sF, 0, Kb, lane, Fr);
      load_f32ops<KBS>(sF, 1, Kb, lane, Fi);
      load_f32ops<KBS>(sF, 2, Kb, lane, Fn);
      f4 re = {0.f, 0.f, 0.f, 0.f}, iv = re;
      // forward: Re = zr Fr + zi (-Fi), Im = zr Fi + zi Fr;  inverse: Re = zr Fr + zi Fi, Im = zr (-Fi) + zi Fr
      // small split products first, the two dominant ones last (two full-magnitude roundings per output)
      re = prod32_small<NTERMS>(rh, rm, rl, Fr, re);
      iv = prod32_small<NTERMS>(rh, rm, rl, INV ? Fn : Fi, iv);
      re = prod32_small<NTERMS>(ih, im, il, INV ? Fi : Fn, re);
      iv = prod32_small<NTERMS>(ih, im, il, Fr, iv);
      re = FFS_MFMA_BF16(rh, Fr[0], re);
      iv = FFS_MFMA_BF16(rh, INV ? Fn[0] : Fi[0], iv);
      re = FFS_MFMA_BF16(ih, INV ? Fi[0] : Fn[0], re);
      iv = FFS_MFMA_BF16(ih, Fr[0], iv);
      outr.b[2 * I + Kb] = re;
      outi.b[2 * I + Kb] = iv;
    }
  }
}

// Window values in block order: x[b * 4 + r] = window[16 (b >> 1) + 4q + r][16 (b & 1) + j].
__device__ __forceinline__ void gather32(const TileArgs& a, const PairCoord& pc, int q, int j, double (&xa)[16],
                                         double (&xb)[16]) {
  const double* __restrict__ src = a.src;
  const int64_t rw = pc.r0 - a.h0, cwA = pc.orgA - a.h1, cwB = pc.orgB - a.h1;
  const bool interior = rw >= 0 && rw + 32 <= a.n0 && cwA >= 0 && (pc.hasB ? cwB : cwA) + 32 <= a.n1 &&
                        a.n1 <= (1 << 22);
  if (interior) {
    const double* baseA = src + rw * a.n1 + cwA;
    const __amdgpu_buffer_rsrc_t ra = window_rsrc(baseA, true), rb = window_rsrc(baseA + a.V1, pc.hasB);
    const uint32_t voff = ((uint32_t)(4 * q) * (uint32_t)a.n1 + (uint32_t)j) * 8u;
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const uint32_t soff = ((uint32_t)(16 * (b >> 1) + r) * (uint32_t)a.n1 + 16 * (b & 1)) * 8u;
        xa[b * 4 + r] = window_load(ra, voff, soff);
        xb[b * 4 + r] = window_load(rb, voff, soff);
      }
  } else {
#pragma unroll
    for (int C = 0; C < 2; C++) {
      const int64_t cA = wrap_near(cwA + 16 * C + j, a.n1);
      const int64_t cB = wrap_near(cwB + 16 * C + j, a.n1);
#pragma unroll
      for (int R = 0; R < 2; R++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int64_t row = wrap_near(rw + 16 * R + 4 * q + r, a.n0) * a.n1;
          xa[(2 * R + C) * 4 + r] = src[row + cA];
          xb[(2 * R + C) * 4 + r] = pc.hasB ? src[row + cB] : 0.0;
        }
    }
  }
}

__device__ __forceinline__ void store32(const TileArgs& a, const PairCoord& pc, int q, int j, const Blk4& yr,
                                        const Blk4& yi, double addA, double addB, double scale = 1.0) {
#pragma unroll
  for (int C = 0; C < 2; C++) {
    const int wc = 16 * C + j;
    const int64_t cA = pc.orgA + wc - a.h1, cB = pc.orgB + wc - a.h1;
    const bool colok = wc >= a.h1 && wc < a.h1 + a.V1;
    const bool okA = colok && cA < a.n1;
    const bool okB = colok && pc.hasB && cB < a.n1;
#pragma unroll
    for (int R = 0; R < 2; R++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 16 * R + 4 * q + v;
        const int64_t y = pc.r0 + row - a.h0;
        if (row >= a.h0 && row < a.h0 + a.V0 && y < a.n0) {
          double* drow = a.dst + y * a.n1;
          if (okA) put_out(drow + cA, (double)yr.b[2 * R + C][v] * scale + addA);
          if (okB) put_out(drow + cB, (double)yi.b[2 * R + C][v] * scale + addB);
        }
      }
  }
}

// Valid region of an out16 tile pair: output block 0, lane (j + 16q), reg v = valid point (4q + v, j)
// (window position (h0 + 4q + v, h1 + j)).
__device__ __forceinline__ void store16(const TileArgs& a, const PairCoord& pc, int q, int j, const f4& yr,
                                        const f4& yi, double addA, double addB, double scale = 1.0) {
  const int64_t cA = pc.orgA + j, cB = pc.orgB + j;  // window column h1 + j
  const bool okA = j < a.V1 && cA < a.n1;
  const bool okB = j < a.V1 && pc.hasB && cB < a.n1;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int row = 4 * q + v;  // window row h0 + row
    const int64_t y = pc.r0 + row;
    if (row < a.V0 && y < a.n0) {
      double* drow = a.dst + y * a.n1;
      if (okA) put_out(drow + cA, (double)yr[v] * scale + addA);
      if (okB) put_out(drow + cB, (double)yi[v] * scale + addB);
    }
  }
}

// One tile pair of the 32-point family.  The spectrum comes from LDS (single-grid launch) or from the
// grid's FP64 table in global memory (batch launch, where every grid has its own stencil and depth).
template <int NTERMS, bool H_FROM_LDS, bool OUT16, int IKBS = 1>  // IKBS as for tile32_pair_f64
__device__ __forceinline__ void tile32_pair(const TileArgs& a, int64_t tp, const u4* __restrict__ sF,
                                            const f2* __restrict__ sH, const u4* __restrict__ sG, int lane, int q,
                                            int j, int ilane = -1) {
  const PairCoord pc = pair_coord<KIND_2D16>(a, tp);
  double xa[16], xb[16];
  gather32(a, pc, q, j, xa, xb);
  // DC offset: window center (16, 16) = block 3, lane 0, element 0
  const double cA = lane_value(xa[12], 0), cB = lane_value(xb[12], 0);
  Blk4 zr, zi, orr, oi;
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      zr.b[b][r] = (float)(xa[b * 4 + r] - cA);
      zi.b[b][r] = (float)(xb[b * 4 + r] - cB);
    }
  stage32<false, NTERMS>(zr, zi, sF, lane, orr, oi);
  stage32<false, NTERMS>(orr, oi, sF, lane, zr, zi);  // Y[k0][k1], block order
  // Keep the spectrum reads here: hoisted to the top of the pair they would hold 32 VGPRs through two
  // stages and cost the third wave per SIMD.
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int e = (b * 4 + v) * kLanes + lane;
      const f2 hv = H_FROM_LDS ? sH[e] : reinterpret_cast<const f2*>(a.Hf)[e];
      const float hr = hv[0], hi = hv[1];
      float yr = zr.b[b][v], yi = zi.b[b][v];
      cmul(yr, yi, hr, hi);
      zr.b[b][v] = yr;
      zi.b[b][v] = yi;
    }
  if (OUT16) {
    const int il = ilane < 0 ? lane : ilane;  // the lane whose inverse operands this lane reads
    stage32<true, NTERMS, 2, 1, IKBS>(zr, zi, sG, il, orr, oi);  // blocks 0 and 2: window rows h0 ..
    stage32<true, NTERMS, 1, 1, IKBS>(orr, oi, sG, il, zr, zi);  // block 0 = the valid region
    store16(a, pc, q, j, zr.b[0], zi.b[0], cA * a.dc_gain, cB * a.dc_gain);
  } else {
    stage32<true, NTERMS>(zr, zi, sF, lane, orr, oi);
    stage32<true, NTERMS>(orr, oi, sF, lane, zr, zi);  // X'[n0][n1], block order
    store32(a, pc, q, j, zr, zi, cA * a.dc_gain, cB * a.dc_gain);
  }
}

// ---------------------------------------------------------------------------------------------
// 2D 32x32 tiles, mixed FP16 form.  Each A operand is split into two FP16 terms (x = h + l, 11 + 11
// significand bits, residual <= 2^-23 |x|) and the DFT operands likewise; a real 16x16 output block keeps
// the three products hh, hl, lh (3 MFMAs instead of the BF16 form's 6; the dropped ll is ~2^-22 of the
// result), so the MFMA work and the split work are both about half.  FP16's range is what BF16 does not
// need: the tile pair is scaled by a power of two (exact) so that its largest fluctuation lies below
// 2^scale_exp, with scale_exp = 10 - log2 max(1, max|K^t|) chosen on the host; then every split operand
// stays below 32 sqrt(2) 2^10 < 65504 through the four stages, and values too small for FP16 normals
// sit ~2^-34 below the tile's largest and do not matter.  The store undoes the scale (exact) in FP64.
// ---------------------------------------------------------------------------------------------
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
#define FFS_MFMA_F16(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_f16((a), (b), (c), 0, 0, 0)

__device__ __forceinline__ h8 as_h8(u4 w) { return __builtin_bit_cast(h8, w); }
// Two FP32 values -> one word of two FP16 (round to nearest even, one v_cvt_pk_f16_f32).
__device__ __forceinline__ uint32_t pk_f16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, h2));
}

// Same element order as split8: word k of the operand holds elements (2k, 2k+1) of (x0 | x1); h = f16(x),
// l = f16(x - h).  x - h is exact in FP32 (h is x's own FP16 rounding), so l is x's residual rounded once.
// The residual is written as fma(x, unit, -h) with `unit` = 1 held in a register: the compiler then folds
// the FP16 -> FP32 conversion of h into a v_fma_mix (with a literal 1 it would fold the FMA into a
// subtraction and convert h separately), 4 instructions per word instead of 6.  h is bit-cast from the
// scalar word, never from the vector element hw[k] (a bit_cast of that element read element 0 here).
__device__ __forceinline__ void split8_f16(const f4& x0, const f4& x1, float unit, h8& h, h8& l) {
  u4 hw, lw;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const f4& x = k < 2 ? x0 : x1;
    const int e = 2 * (k & 1);
    const uint32_t hk = pk_f16(x[e], x[e + 1]);
    const h2 hh = __builtin_bit_cast(h2, hk);
    hw[k] = hk;
    const _Float16 l0 = (_Float16)__builtin_fmaf(x[e], unit, -(float)hh[0]);
    const _Float16 l1 = (_Float16)__builtin_fmaf(x[e + 1], unit, -(float)hh[1]);
    lw[k] = __builtin_bit_cast(uint32_t, h2{l0, l1});
  }
  h = as_h8(hw);
  l = as_h8(lw);
}

// split8_f16 of (x0 | x1) * sc for a power-of-two sc: h = f16(x sc), l = f16(fma(x, sc, -h)).  x sc is
// exact in FP32 (a power-of-two scale of a normal value) and so is x sc - h, so the words equal
// split8_f16 of the scaled values; the compiler selects v_fma_mix for the residual FMA.
__device__ __forceinline__ void split8_f16_scaled(const f4& x0, const f4& x1, float sc, h8& h, h8& l) {
  u4 hw, lw;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const f4& x = k < 2 ? x0 : x1;
    const int e = 2 * (k & 1);
    const _Float16 h0 = (_Float16)(x[e] * sc), h1 = (_Float16)(x[e + 1] * sc);
    const _Float16 l0 = (_Float16)__builtin_fmaf(x[e], sc, -(float)h0);
    const _Float16 l1 = (_Float16)__builtin_fmaf(x[e + 1], sc, -(float)h1);
    hw[k] = __builtin_bit_cast(uint32_t, h2{h0, h1});
    lw[k] = __builtin_bit_cast(uint32_t, h2{l0, l1});
  }
  h = as_h8(hw);
  l = as_h8(lw);
}

// Maximum of a non-negative v over the wave, in every lane: DPP exchanges inside each row of 16 lanes
// (quad swaps, then the half-row and row mirrors), then the four row maxima through readlane.  For
// non-negative floats the bit patterns order like unsigned integers, so the max is taken on those.
// Exact (a max), and no LDS round trip as a __shfl_xor butterfly would take.
__device__ __forceinline__ uint32_t dpp_umax(uint32_t x, int sel) {
  int y;
  switch (sel) {  // the DPP control must be a constant
    case 0: y = __builtin_amdgcn_update_dpp((int)x, (int)x, 0xB1, 0xF, 0xF, false); break;   // quad_perm [1,0,3,2]
    case 1: y = __builtin_amdgcn_update_dpp((int)x, (int)x, 0x4E, 0xF, 0xF, false); break;   // quad_perm [2,3,0,1]
    case 2: y = __builtin_amdgcn_update_dpp((int)x, (int)x, 0x141, 0xF, 0xF, false); break;  // row_half_mirror
    default: y = __builtin_amdgcn_update_dpp((int)x, (int)x, 0x140, 0xF, 0xF, false); break; // row_mirror
  }
  return max(x, (uint32_t)y);
}
__device__ __forceinline__ float wave_max_nonneg(float v) {
  uint32_t x = __float_as_uint(v);
  x = dpp_umax(x, 0);
  x = dpp_umax(x, 1);
  x = dpp_umax(x, 2);
  x = dpp_umax(x, 3);
  const uint32_t r0 = __builtin_amdgcn_readlane((int)x, 0), r1 = __builtin_amdgcn_readlane((int)x, 16);
  const uint32_t r2 = __builtin_amdgcn_readlane((int)x, 32), r3 = __builtin_amdgcn_readlane((int)x, 48);
  return __uint_as_float(max(max(r0, r1), max(r2, r3)));
}

// SCALED: the inputs are split as z * sc (the first forward stage, which applies the tile's scale).
// All NI input row blocks are split first (the inputs are dead afterwards, so the split operands take
// their registers), then each column block Kb of the operands is read from LDS once and used for the NI
// output blocks (I, Kb): half the operand reads of splitting and loading per (I, Kb), and four
// independent accumulator chains per operand load.  Every output block sees the same MFMAs in the same
// order as when computed alone.
template <bool INV, int NI = 2, int NK = 2, int KBS = 2, bool SCALED = false>
__device__ __forceinline__ void stage32_f16(const Blk4& zr, const Blk4& zi, const u4* __restrict__ sF, int lane,
                                            Blk4& outr, Blk4& outi, float sc, float unit) {
  h8 rh[NI], rl[NI], ih[NI], il[NI];
#pragma unroll
  for (int I = 0; I < NI; I++) {
    const f4 x0 = zr.b[I], x1 = zr.b[2 + I], y0 = zi.b[I], y1 = zi.b[2 + I];
    if (SCALED) {
      split8_f16_scaled(x0, x1, sc, rh[I], rl[I]);
      split8_f16_scaled(y0, y1, sc, ih[I], il[I]);
    } else {
      split8_f16(x0, x1, unit, rh[I], rl[I]);
      split8_f16(y0, y1, unit, ih[I], il[I]);
    }
  }
#pragma unroll
  for (int Kb = 0; Kb < NK; Kb++) {
    h8 Fr[2], Fi[2], Fn[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      Fr[t] = as_h8(sF[((0 * KBS + Kb) * 3 + t) * kLanes + lane]);
      Fi[t] = as_h8(sF[((1 * KBS + Kb) * 3 + t) * kLanes + lane]);
      Fn[t] = as_h8(sF[((2 * KBS + Kb) * 3 + t) * kLanes + lane]);
    }
    // forward: Re = zr Fr + zi (-Fi), Im = zr Fi + zi Fr;  inverse: Re = zr Fr + zi Fi, Im = zr (-Fi) + zi Fr
    const h8* Gi = INV ? Fn : Fi;  // multiplies zr into Im
    const h8* Gr = INV ? Fi : Fn;  // multiplies zi into Re
#pragma unroll
    for (int I = 0; I < NI; I++) {
      f4 re = {0.f, 0.f, 0.f, 0.f}, iv = re;
      re = FFS_MFMA_F16(rh[I], Fr[1], re);  // the cross products first, the two dominant ones last
      re = FFS_MFMA_F16(rl[I], Fr[0], re);
      iv = FFS_MFMA_F16(rh[I], Gi[1], iv);
      iv = FFS_MFMA_F16(rl[I], Gi[0], iv);
      re = FFS_MFMA_F16(ih[I], Gr[1], re);
      re = FFS_MFMA_F16(il[I], Gr[0], re);
      iv = FFS_MFMA_F16(ih[I], Fr[1], iv);
      iv = FFS_MFMA_F16(il[I], Fr[0], iv);
      re = FFS_MFMA_F16(rh[I], Fr[0], re);
      iv = FFS_MFMA_F16(rh[I], Gi[0], iv);
      re = FFS_MFMA_F16(ih[I], Gr[0], re);
      iv = FFS_MFMA_F16(ih[I], Fr[0], iv);
      outr.b[2 * I + Kb] = re;
      outi.b[2 * I + Kb] = iv;
    }
  }
}

template <bool H_FROM_LDS, bool OUT16, int IKBS = 1>  // IKBS as for tile32_pair_f64
__device__ __forceinline__ void tile32_pair_f16(const TileArgs& a, int64_t tp, const u4* __restrict__ sF,
                                                const f2* __restrict__ sH, const u4* __restrict__ sG, int lane,
                                                int q, int j, int ilane = -1) {
  const PairCoord pc = pair_coord<KIND_2D16>(a, tp);
  double xa[16], xb[16];
  // Wave priority 1 from the window loads to the first stage: a wave that starts a pair issues its loads
  // and takes in its window ahead of the other waves' stage work, so the waves of a SIMD drift out of
  // step instead of gathering together (1-2% on the single-grid and batch kernels, same results).
  __builtin_amdgcn_s_setprio(1);
  gather32(a, pc, q, j, xa, xb);
  const double cA = lane_value(xa[12], 0), cB = lane_value(xb[12], 0);
  Blk4 zr, zi, orr, oi;
  float m = 0.f;
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      zr.b[b][r] = (float)(xa[b * 4 + r] - cA);
      zi.b[b][r] = (float)(xb[b * 4 + r] - cB);
      m = fmaxf(m, fmaxf(fabsf(zr.b[b][r]), fabsf(zi.b[b][r])));
    }
  m = wave_max_nonneg(m);
  int k;
  (void)frexpf(m, &k);  // m < 2^k
  const int e = __builtin_amdgcn_readfirstlane(min(max(a.scale_exp - k, -120), 120));
  __builtin_amdgcn_s_setprio(0);
  stage32_f16<false, 2, 2, 2, true>(zr, zi, sF, lane, orr, oi, ldexpf(1.0f, e), a.unit);
  stage32_f16<false>(orr, oi, sF, lane, zr, zi, 1.0f, a.unit);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int ee = (b * 4 + v) * kLanes + lane;
      const f2 hv = H_FROM_LDS ? sH[ee] : reinterpret_cast<const f2*>(a.Hf)[ee];
      const float hr = hv[0], hi = hv[1];
      float yr = zr.b[b][v], yi = zi.b[b][v];
      cmul(yr, yi, hr, hi);
      zr.b[b][v] = yr;
      zi.b[b][v] = yi;
    }
  if (OUT16) {
    const int il = ilane < 0 ? lane : ilane;  // the lane whose inverse operands this lane reads
    stage32_f16<true, 2, 1, IKBS>(zr, zi, sG, il, orr, oi, 1.0f, a.unit);
    stage32_f16<true, 1, 1, IKBS>(orr, oi, sG, il, zr, zi, 1.0f, a.unit);
    store16(a, pc, q, j, zr.b[0], zi.b[0], cA * a.dc_gain, cB * a.dc_gain, ldexp(1.0, -e));
  } else {
    stage32_f16<true>(zr, zi, sF, lane, orr, oi, 1.0f, a.unit);
    stage32_f16<true>(orr, oi, sF, lane, zr, zi, 1.0f, a.unit);
    store32(a, pc, q, j, zr, zi, cA * a.dc_gain, cB * a.dc_gain, ldexp(1.0, -e));
  }
}

__device__ __forceinline__ void stage_ops32(u4* sF, const uint32_t* __restrict__ opsm) {
  for (int i = threadIdx.x; i < 3 * 2 * 3 * kLanes; i += blockDim.x) {
    const int idx = i / kLanes, ln = i % kLanes;
    u4 w;
#pragma unroll
    for (int k = 0; k < 4; k++) w[k] = opsm[(idx * 4 + k) * kLanes + ln];
    sF[i] = w;
  }
}

// Inverse operands of an out16 block (9 KB), already in u4 order.
template <bool OUT16>
__device__ __forceinline__ void stage_inv32(u4* sG, const void* opsi) {
  if (OUT16)
    for (int i = threadIdx.x; i < 3 * 3 * kLanes; i += blockDim.x) sG[i] = static_cast<const u4*>(opsi)[i];
}

template <int NTERMS, bool OUT16>
__global__ __launch_bounds__(256, 3) void k_tile32_bf16(TileArgs a) {
  __shared__ u4 sF[3 * 2 * 3 * kLanes];  // B operands, 18 KB
  __shared__ f2 sH[4 * 4 * kLanes];      // spectrum [block][reg][lane] (re, im), 8 KB
  __shared__ u4 sG[OUT16 ? 3 * 3 * kLanes : 1];  // out16 inverse operands, 9 KB
  stage_ops32(sF, a.opsm);
  stage_inv32<OUT16>(sG, a.opsi);
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) sH[i] = f2{(float)a.H[i], (float)a.H[1024 + i]};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(a.linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int64_t tp = gw; tp < a.pairs; tp += W) tile32_pair<NTERMS, true, OUT16>(a, tp, sF, sH, sG, lane, q, j);
}

// Multi-grid launch: one fused block of every grid still running, as one persistent launch over the
// concatenated work list.  Each item carries its own spectrum (an FP32 copy for the mixed forms, FP64 for
// the FP64 form); the DFT operands are staged in LDS and the kernels run 3 waves per SIMD (2 for FP64).
// items[g] describes grid g's block (source, destination, depth, spectrum); prefix[g] is the first work
// index of grid g (prefix[n] = total).  All items cost the same (one 32x32 tile pair), so waves take them
// in a fixed order (strided, or in chunks for the FP16 kernel): balanced and bitwise reproducible.
//
// Work-item walk of the FP16 batch kernel.  Each wave takes kBatchChunk consecutive tile pairs, then
// strides by the wave count times that: the chip-wide front of the launch is still one contiguous band of
// pairs (neighbouring windows share L2), and a wave usually stays in one grid for a whole chunk.  So it
// keeps that grid's FP32 spectrum in its own LDS slice, refilled when its grid changes: lane L fills
// entries L, L + 64, ..., the only ones the 32-point pairs read in lane L, so no barrier is needed
// (96-grid FP16 batch 918 -> 984 GStencils/s; with the plain stride the slice would be refilled for
// almost every pair, and one contiguous range per wave loses the L2 sharing, 6% slower).
constexpr int kBatchChunk = 16;
template <typename PairFn>
__device__ __forceinline__ void batch_walk(const TileArgs* __restrict__ items, const int64_t* __restrict__ prefix,
                                           int32_t n, f2* sHw, int lane, PairFn pair) {
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(items[0].linear_map, blockIdx.x, gridDim.x) * 4 +
                     __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t total = prefix[n];
  int lo = -1, cur = -1;
  int64_t next = 0;
  for (int64_t base = gw * kBatchChunk; base < total; base += W * kBatchChunk)
    for (int64_t idx = base; idx < base + kBatchChunk && idx < total; idx++) {
      lo = batch_grid(prefix, n, idx, lo, next);
      const TileArgs a = items[lo];  // independent scalar loads, from the scalar cache while lo is unchanged
      if (lo != cur) {
        const f2* Hf = reinterpret_cast<const f2*>(a.Hf);
#pragma unroll
        for (int k = 0; k < 16; k++) sHw[k * kLanes + lane] = Hf[k * kLanes + lane];
        cur = lo;
      }
      pair(a, idx - prefix[lo]);
    }
}

template <int NTERMS>
__global__ __launch_bounds__(256, 3) void k_batch32_bf16(const TileArgs* __restrict__ items,
                                                         const int64_t* __restrict__ prefix, int32_t n) {
  // The spectrum stays in L2 here: with the per-wave LDS slices of the FP16 batch kernel (batch_walk) this
  // kernel spills at 3 waves per SIMD and gains 0.4%.
  __shared__ u4 sF[3 * 2 * 3 * kLanes];
  stage_ops32(sF, items[0].opsm);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(items[0].linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t total = prefix[n];
  int lo = -1;
  int64_t next = 0;
  for (int64_t idx = gw; idx < total; idx += W) {
    lo = batch_grid(prefix, n, idx, lo, next);
    const TileArgs a = items[lo];  // independent scalar loads, from the scalar cache while lo is unchanged
    if (a.out16) {
      const ShiftedView<u4> v = shifted_view(sF, a.h1, lane);
      tile32_pair<NTERMS, false, true, 2>(a, idx - prefix[lo], sF, nullptr, v.base, lane, q, j, v.lane);
    } else {
      tile32_pair<NTERMS, false, false>(a, idx - prefix[lo], sF, nullptr, nullptr, lane, q, j);
    }
  }
}

// The FP16 form of the two kernels above (same structure, FP16 operand pairs).  The out16 variant fits
// 128 VGPRs and runs 4 waves per SIMD (measured 5% faster than 3).
template <bool OUT16>
__global__ __launch_bounds__(256, OUT16 ? 4 : 3) void k_tile32_f16(TileArgs a) {
  __shared__ u4 sF[3 * 2 * 3 * kLanes];
  __shared__ f2 sH[4 * 4 * kLanes];
  __shared__ u4 sG[OUT16 ? 3 * 3 * kLanes : 1];
  stage_ops32(sF, a.opsm);
  stage_inv32<OUT16>(sG, a.opsi);
  for (int i = threadIdx.x; i < 1024; i += blockDim.x) sH[i] = f2{(float)a.H[i], (float)a.H[1024 + i]};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = block_order(a.linear_map, blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int64_t tp = gw; tp < a.pairs; tp += W) tile32_pair_f16<true, OUT16>(a, tp, sF, sH, sG, lane, q, j);
}

__global__ __launch_bounds__(256, 3) void k_batch32_f16(const TileArgs* __restrict__ items,
                                                        const int64_t* __restrict__ prefix, int32_t n) {
  __shared__ u4 sF[3 * 2 * 3 * kLanes];   // B operands, 18 KB
  __shared__ f2 sHw[4][4 * 4 * kLanes];   // per-wave spectrum slices, 32 KB
  stage_ops32(sF, items[0].opsm);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const f2* sH = sHw[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
  batch_walk(items, prefix, n, sHw[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)], lane,
             [&](const TileArgs& a, int64_t tp) {
               if (a.out16) {
                 const ShiftedView<u4> v = shifted_view(sF, a.h1, lane);
                 tile32_pair_f16<true, true, 2>(a, tp, sF, sH, v.base, lane, q, j, v.lane);
               } else {
                 tile32_pair_f16<true, false>(a, tp, sF, sH, nullptr, lane, q, j);
               }
             });
}

hipError_t launch_batch(const TileArgs* d_items, const int64_t* d_prefix, int32_t n, int32_t precision, hipStream_t s,
                        int num_cus) {
  const dim3 g((unsigned)num_cus * (precision == 0 ? 2 : 3)), b(256);
  if (precision == 0)
    hipLaunchKernelGGL(k_batch32_f64, g, b, 0, s, d_items, d_prefix, n);
  else if (precision == PREC_BF16)
    hipLaunchKernelGGL(k_batch32_bf16<3>, g, b, 0, s, d_items, d_prefix, n);
  else if (precision == PREC_F16)
    hipLaunchKernelGGL(k_batch32_f16, g, b, 0, s, d_items, d_prefix, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// 3D 16x16x16 tiles.  One workgroup (4 waves) per tile pair; the complex window (tile A real, tile B
// imaginary) sits in LDS with a padded pitch, idx(x, y, z) = (16 x + y) * 17 + z.  Three phases:
//   1. wave w, slices a = 4w..4w+3: contract axis 2 then axis 1 in registers  -> Y[a][k2][k1]
//   2. wave w, k2 = 4w..4w+3: contract axis 0, multiply by H, inverse axis 1, inverse axis 0 -> W[k2][n1][n0]
//   3. wave w, slices n0 = 4w..4w+3: inverse axis 2; lanes now run along axis 2 -> coalesced store
// Every stage is the same accumulator-chained A^T F step as in 2D; only the re-slicing between phases
// goes through LDS.  rowmap(q, x) is the row a lane group q holds in register x: 4x + q for FP64
// accumulators, 4q + x for FP32 accumulators (the MFMA output orders measured on gfx950).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int lds3(int x, int y, int z) { return (16 * x + y) * 17 + z; }

template <int PREC>
__device__ __forceinline__ int rowmap(int q, int x) {
  return PREC == 0 ? 4 * x + q : 4 * q + x;
}

template <int PREC, int NTERMS>
struct Prec3;

template <int NTERMS>
struct Prec3<0, NTERMS> {  // FP64
  typedef double V;
  typedef d4 V4;
  typedef OpsAxis Ops;
  struct C2 {
    double re, im;
  };
  template <bool INV>
  static __device__ __forceinline__ void stage(const V4& ar, const V4& ai, const Ops& f, V4& outr, V4& outi) {
    double cr[4], ci[4];
    to_chunks(ar, cr);
    to_chunks(ai, ci);
    cstage<INV>(cr, ci, f, outr, outi);
  }
  static __device__ __forceinline__ void load_ops(Ops& f, const TileArgs& a, int lane) { ffsk::load_ops(f, a.ops, 0, lane); }
};

template <int NTERMS>
struct Prec3<PREC_BF16, NTERMS> {
  typedef float V;
  typedef f4 V4;
  typedef OpsBF16<NTERMS> Ops;
  struct C2 {
    float re, im;
  };
  template <bool INV>
  static __device__ __forceinline__ void stage(const V4& ar, const V4& ai, const Ops& f, V4& outr, V4& outi) {
    mstage<INV, NTERMS>(ar, ai, f, outr, outi);
  }
  static __device__ __forceinline__ void load_ops(Ops& f, const TileArgs& a, int lane) { load_ops_mixed(f, a.opsm, 0, lane); }
};

template <int NTERMS>
struct Prec3<PREC_F32, NTERMS> {  // compensated FP32 (two products per real GEMM)
  typedef float V;
  typedef f4 V4;
  typedef OpsF32C Ops;
  struct C2 {
    float re, im;
  };
  template <bool INV>
  static __device__ __forceinline__ void stage(const V4& ar, const V4& ai, const Ops& f, V4& outr, V4& outi) {
    mstage<INV, NTERMS>(ar, ai, f, outr, outi);
  }
  static __device__ __forceinline__ void load_ops(Ops& f, const TileArgs& a, int lane) { load_ops_mixed(f, a.opsm, 0, lane); }
};

// Geometry of 3D tile pair tp: tile origins (o*), window origins (w*), whether tile B exists and
// whether the window needs no periodic wrap.
struct Win3 {
  int64_t o0, o1, o2A, o2B, w0, w1, w2A, w2B;
  bool hasB, interior;
};

__device__ __forceinline__ Win3 win3(const TileArgs& a, int64_t tp) {
  Win3 g;
  const int64_t pairs2 = (a.tiles2 + 1) >> 1;
  const int64_t T0 = tp / (a.tiles1 * pairs2);
  const int64_t rem = tp - T0 * (a.tiles1 * pairs2);
  const int64_t T1 = rem / pairs2, T2 = (rem - T1 * pairs2) * 2;
  g.hasB = T2 + 1 < a.tiles2;
  g.o0 = T0 * a.V0;
  g.o1 = T1 * a.V1;
  g.o2A = T2 * a.V2;
  g.o2B = g.o2A + a.V2;
  g.w0 = g.o0 - a.h0;
  g.w1 = g.o1 - a.h1;
  g.w2A = g.o2A - a.h2;
  g.w2B = g.o2B - a.h2;
  g.interior = g.w0 >= 0 && g.w0 + 16 <= a.n0 && g.w1 >= 0 && g.w1 + 16 <= a.n1 && g.w2A >= 0 &&
               (g.hasB ? g.w2B : g.w2A) + 16 <= a.n2;
  return g;
}

// Window values of thread (y, z): slice x, row y, column z of tile A (va) and tile B (vb, 0 without B).
__device__ __forceinline__ void load3(const TileArgs& a, const Win3& g, int y, int z, double (&va)[16],
                                      double (&vb)[16]) {
  const int64_t r1 = g.interior ? g.w1 + y : wrap_near(g.w1 + y, a.n1);
  const int64_t cAz = g.interior ? g.w2A + z : wrap_near(g.w2A + z, a.n2);
  const int64_t cBz = g.interior ? g.w2B + z : wrap_near(g.w2B + z, a.n2);
#pragma unroll
  for (int x = 0; x < 16; x++) {
    const int64_t r0 = g.interior ? g.w0 + x : wrap_near(g.w0 + x, a.n0);
    const double* row = a.src + (r0 * a.n1 + r1) * a.n2;
    va[x] = row[cAz];
    vb[x] = g.hasB ? row[cBz] : 0.0;
  }
}

template <int PREC, int NTERMS>
__global__ __launch_bounds__(256) void k_tile3d(TileArgs a) {
  typedef Prec3<PREC, NTERMS> P;
  typedef typename P::V V;
  typedef typename P::V4 V4;
  typedef typename P::C2 C2;
  __shared__ C2 sX[16 * 16 * 17];
  // BF16 form: planes k2 = 0..8 of the FP32 spectrum in LDS (18 KB; with the window 52 KB, 3 workgroups
  // per CU).  Real taps make the spectrum Hermitian, H(-k mod 16) = conj H(k), so planes 9..15 are read
  // as conjugates of planes 7..1: the same FP32 values as the global table, without its L2 latency in
  // phase 2.  First read in phase 2, after the barriers of the window fill.  The compensated FP32 form
  // multiplies by the FP64 spectrum instead (in FP64, rounded once to FP32), from global memory.
  constexpr bool H64 = PREC == PREC_F32;
  constexpr bool HLDS = PREC != 0 && !H64;
  __shared__ f2 sH[HLDS ? 9 * 4 * kLanes : 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = lane >> 4, j = lane & 15;
  typename P::Ops f;
  P::load_ops(f, a, lane);
  const bool mixed = PREC != 0;
  if (HLDS)
    for (int i = threadIdx.x; i < 9 * 4 * kLanes; i += blockDim.x) sH[i] = reinterpret_cast<const f2*>(a.Hf)[i];
  // Phase loops: fully unrolled in the mixed forms (later slices' loads issue early, 1.5% faster), two
  // slices at a time in the FP64 form (full unrolling costs it 3% at its 2 waves per SIMD) and in phase 2
  // of the FP32 form (fully unrolled, its FP64 spectrum loads take it to 178 VGPRs, 2 waves per SIMD).
  constexpr int kSlices = PREC == 0 ? 2 : 4;
  constexpr int kSlices2 = H64 ? 2 : kSlices;

  // Window of the pair: every thread (y = tid / 16, z = tid % 16) loads column z of row y in all 16
  // slices x, 32 loads in flight.  With PREFETCH (FP64) the next pair's window is loaded right after
  // the current one is in LDS, so its latency hides behind the three compute phases (64 VGPRs, within
  // the 2 waves per SIMD the 70 KB of LDS allow anyway).
  constexpr bool PREFETCH = PREC == 0;
  const int fy = threadIdx.x >> 4, fz = threadIdx.x & 15;
  double va[16], vb[16];
  int64_t tp = block_order(a.linear_map, blockIdx.x, gridDim.x);
  Win3 g = win3(a, tp);
  if (tp < a.pairs) load3(a, g, fy, fz, va, vb);
  for (; tp < a.pairs; tp += gridDim.x) {
    if (!PREFETCH) load3(a, g, fy, fz, va, vb);
    const Win3 cur = g;
    const int64_t o0 = cur.o0, o1 = cur.o1, o2A = cur.o2A, o2B = cur.o2B;
    const bool hasB = cur.hasB;
    // DC offsets (mixed forms): the window's center sample (8, 8, 8), a wave-uniform load
    double cA = 0.0, cB = 0.0;
    if (mixed) {
      const int64_t base = (wrap_near(cur.w0 + 8, a.n0) * a.n1 + wrap_near(cur.w1 + 8, a.n1)) * a.n2;
      cA = a.src[base + wrap_near(cur.w2A + 8, a.n2)];
      cB = hasB ? a.src[base + wrap_near(cur.w2B + 8, a.n2)] : 0.0;
    }
#pragma unroll
    for (int x = 0; x < 16; x++) {
      C2 c;
      c.re = (V)(va[x] - cA);
      c.im = (V)(hasB ? vb[x] - cB : 0.0);
      sX[lds3(x, fy, fz)] = c;
    }
    __syncthreads();
    const int64_t tn = tp + gridDim.x;
    if (tn < a.pairs) {
      g = win3(a, tn);
      if (PREFETCH) load3(a, g, fy, fz, va, vb);
    }
    // ---- phase 1: slices a = 4w + s; contract axis 2, then axis 1 ----
#pragma unroll kSlices
    for (int s = 0; s < 4; s++) {
      const int sl = 4 * w + s;
      V4 ar, ai;
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const C2 c = sX[lds3(sl, j, rowmap<PREC>(q, x))];
        ar[x] = c.re;
        ai[x] = c.im;
      }
      V4 br, bi;
      P::template stage<false>(ar, ai, f, br, bi);  // [n1 = row][k2 = lane]
      P::template stage<false>(br, bi, f, ar, ai);  // [k2 = row][k1 = lane]
#pragma unroll
      for (int v = 0; v < 4; v++) {
        C2 c;
        c.re = ar[v];
        c.im = ai[v];
        sX[lds3(sl, rowmap<PREC>(q, v), j)] = c;
      }
    }
    __syncthreads();
    // ---- phase 2: k2 = 4w + s; all reads first (the writes reuse the same LDS) ----
    V4 pr[4], pi[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int k2 = 4 * w + s;
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const C2 c = sX[lds3(rowmap<PREC>(q, x), k2, j)];  // Y[k0 row][k2][k1 = lane]
        pr[s][x] = c.re;
        pi[s][x] = c.im;
      }
    }
    __syncthreads();
#pragma unroll kSlices2
    for (int s = 0; s < 4; s++) {
      const int k2 = 4 * w + s;
      V4 br, bi;
      P::template stage<false>(pr[s], pi[s], f, br, bi);  // [k1 = row][k0 = lane]
      // H entry (k2, v, lane): k0 = j, k1 = rowmap(q, v), k2
      const double* H = a.H + (int64_t)k2 * 4 * kLanes + lane;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        if (H64) {  // FP64 product with the FP64 spectrum, one rounding to FP32
          double yr = br[v], yi = bi[v];
          cmul(yr, yi, H[v * kLanes], H[4096 + v * kLanes]);
          br[v] = (V)yr;
          bi[v] = (V)yi;
          continue;
        }
        V hr, hi;
        if (mixed) {  // the FP32 copy, from LDS; k2 is wave-uniform, so is the branch
          if (k2 <= 8) {
            const f2 hv = sH[k2 * 4 * kLanes + v * kLanes + lane];
            hr = (V)hv[0];
            hi = (V)hv[1];
          } else {  // conj H at (-k0, -k1, 16 - k2) mod 16; k1 = 4q + v in the FP32 accumulator order
            const int k1c = (16 - rowmap<PREC>(q, v)) & 15, k0c = (16 - j) & 15;
            const f2 hv = sH[(16 - k2) * 4 * kLanes + (k1c & 3) * kLanes + 16 * (k1c >> 2) + k0c];
            hr = (V)hv[0];
            hi = -(V)hv[1];
          }
        } else {
          hr = (V)H[v * kLanes];
          hi = (V)H[4096 + v * kLanes];
        }
        V yr = br[v], yi = bi[v];
        cmul(yr, yi, hr, hi);
        br[v] = yr;
        bi[v] = yi;
      }
      V4 cr, ci;
      P::template stage<true>(br, bi, f, cr, ci);  // [k0 = row][n1 = lane]
      P::template stage<true>(cr, ci, f, br, bi);  // [n1 = row][n0 = lane]
#pragma unroll
      for (int v = 0; v < 4; v++) {
        C2 c;
        c.re = br[v];
        c.im = bi[v];
        sX[lds3(k2, rowmap<PREC>(q, v), j)] = c;  // W[k2][n1][n0]
      }
    }
    __syncthreads();
    // ---- phase 3: the slices n0 of the valid block only, h0 + w + 4s (V0 of the 16, fewer at the grid's
    // far edge, spread over the four waves: 3 each at V0 = 12); inverse axis 2; store ----
    const int nvalid0 = (int)(a.V0 < a.n0 - o0 ? a.V0 : a.n0 - o0);
#pragma unroll kSlices
    for (int s = 0; s < 4; s++) {
      const int i0 = w + 4 * s;
      if (i0 >= nvalid0) break;
      const int n0 = (int)a.h0 + i0;
      V4 ar, ai;
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const C2 c = sX[lds3(rowmap<PREC>(q, x), j, n0)];  // W[k2 row][n1 = lane][n0]
        ar[x] = c.re;
        ai[x] = c.im;
      }
      V4 br, bi;
      P::template stage<true>(ar, ai, f, br, bi);  // [n1 = row][n2 = lane]
      const int64_t y0 = o0 + i0;
      {
        const bool colok = j >= a.h2 && j < a.h2 + a.V2;
        const int64_t zA = o2A + j - a.h2, zB = o2B + j - a.h2;
        const bool okA = colok && zA < a.n2, okB = colok && hasB && zB < a.n2;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int n1 = rowmap<PREC>(q, v);
          const int64_t y1 = o1 + n1 - a.h1;
          if (n1 >= a.h1 && n1 < a.h1 + a.V1 && y1 < a.n1) {
            double* row = a.dst + (y0 * a.n1 + y1) * a.n2;
            if (okA) put_out(row + zA, (double)br[v] + cA * a.dc_gain);
            if (okB) put_out(row + zB, (double)bi[v] + cB * a.dc_gain);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------
// Plan-time generation (FP64 on the GPU).
// ---------------------------------------------------------------------------------------------
__global__ void k_gen_operands(double* ops, int32_t L0, int32_t L1, int32_t L2) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over [axis][chunk][lane] = 3*4*64
  if (idx >= 3 * 4 * kLanes) return;
  const int axis = idx / (4 * kLanes);
  const int c = (idx / kLanes) % 4;
  const int lane = idx % kLanes;
  const int L = axis == 0 ? L0 : (axis == 1 ? L1 : L2);
  const int row = 4 * c + (lane >> 4), col = lane & 15;
  double fr = 0.0, fi = 0.0;
  if (row < L && col < L) {
    const int m = (row * col) % L;  // exact angle reduction: theta = -2 pi m / L
    double s, co;
    sincospi(-2.0 * (double)m / (double)L, &s, &co);
    fr = co;
    fi = s;
  }
  double* base = ops + axis * kOpsPerAxis;
  base[(0 * 4 + c) * kLanes + lane] = fr;
  base[(1 * 4 + c) * kLanes + lane] = -(fr + fi);
  base[(2 * 4 + c) * kLanes + lane] = fi - fr;
}

// 3-term BF16 split of an FP64 value: x ~ hi + mid + lo, each rounded to nearest even.
__device__ __forceinline__ void split3_bf16(double x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)(float)x;
  const double r1 = x - (double)(float)h;
  m = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)m;
  l = (__bf16)(float)r2;
}

__device__ __forceinline__ void dft_entry(int row, int col, int L, double& fr, double& fi) {
  fr = 0.0;
  fi = 0.0;
  if (row < L && col < L) {
    const int m = (row * col) % L;
    double s, c;
    sincospi(-2.0 * (double)m / (double)L, &s, &c);
    fr = c;
    fi = s;
  }
}

// Mixed-form operand tables (layout in ffs_internal.h); one thread per (axis, arr, lane).
__global__ void k_gen_operands_mixed(uint32_t* t, int32_t L0, int32_t L1, int32_t L2) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 3 * 3 * kLanes) return;
  const int axis = idx / (3 * kLanes), arr = (idx / kLanes) % 3, lane = idx % kLanes;
  const int L = axis == 0 ? L0 : (axis == 1 ? L1 : L2);
  const int q = lane >> 4, j = lane & 15;
  // BF16 pairings: (h|h), (m|m), (l|h), (m|l)
  for (int g = 0; g < kPairings; g++) {
    for (int w = 0; w < 4; w++) {
      uint32_t word = 0;
      for (int half = 0; half < 2; half++) {
        const int e = 2 * w + half;
        double fr, fi;
        dft_entry(4 * q + (e & 3), j, L, fr, fi);
        const double v = arr == 0 ? fr : (arr == 1 ? fi : -fi);
        __bf16 h, m, l;
        split3_bf16(v, h, m, l);
        const __bf16 first[kPairings] = {h, m, l, m}, second[kPairings] = {h, m, h, l};
        const __bf16 pick = e < 4 ? first[g] : second[g];
        const uint32_t bits = (uint32_t)__builtin_bit_cast(uint16_t, pick);
        word |= bits << (16 * half);
      }
      t[axis * kBf16WordsPerAxis + ((arr * kPairings + g) * 4 + w) * kLanes + lane] = word;
    }
  }
  // FP32 chunks: chunk c of lane (j + 16q) = F[4q + c][j], and its low-order term F - fl32(F)
  float* tf = reinterpret_cast<float*>(t + kOpsF32Offset) + axis * kF32PerAxis;
  float* tl = reinterpret_cast<float*>(t + kOpsF32LoOffset) + axis * kF32PerAxis;
  for (int c = 0; c < 4; c++) {
    double fr, fi;
    dft_entry(4 * q + c, j, L, fr, fi);
    const double v = arr == 0 ? fr : (arr == 1 ? fi : -fi);
    const float hi = (float)v;
    tf[(arr * 4 + c) * kLanes + lane] = hi;
    tl[(arr * 4 + c) * kLanes + lane] = (float)(v - (double)hi);
  }
}

// 32-point B operands for KIND_2D32 (layout in ffs_internal.h); one thread per (arr, Kb, term, lane).
// BF16 forms: terms h, m, l of the 3-term split.  FP16 form: terms h, l of the 2-term split (term 2 = 0).
__global__ void k_gen_operands32(uint32_t* t, int32_t f16) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 3 * 2 * 3 * kLanes) return;
  const int lane = idx % kLanes, term = (idx / kLanes) % 3, Kb = (idx / (3 * kLanes)) % 2, arr = idx / (6 * kLanes);
  const int q = lane >> 4, j = lane & 15;
  for (int w = 0; w < 4; w++) {
    uint32_t word = 0;
    for (int half = 0; half < 2; half++) {
      const int e = 2 * w + half;
      double fr, fi;
      dft_entry(16 * (e >> 2) + 4 * q + (e & 3), 16 * Kb + j, 32, fr, fi);
      const double v = arr == 0 ? fr : (arr == 1 ? fi : -fi);
      uint16_t bits;
      if (f16) {
        const _Float16 h = (_Float16)(float)v;
        const _Float16 l = (_Float16)(float)(v - (double)h);
        bits = __builtin_bit_cast(uint16_t, term == 0 ? h : (term == 1 ? l : (_Float16)0.0f));
      } else {
        __bf16 h, m, l;
        split3_bf16(v, h, m, l);
        bits = __builtin_bit_cast(uint16_t, term == 0 ? h : (term == 1 ? m : l));
      }
      word |= (uint32_t)bits << (16 * half);
    }
    t[((((arr * 2 + Kb) * 3 + term) * 4) + w) * kLanes + lane] = word;
  }
}

// Inverse operands of an out16 block (layouts in ffs_internal.h): the Kb = 0 entries of the operand
// tables with output column j taken from window column (j + shift) mod 32.
__global__ void k_gen_inverse32(uint32_t* out, int32_t f16, int32_t shift) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 3 * 3 * kLanes) return;
  const int lane = idx % kLanes, term = (idx / kLanes) % 3, arr = idx / (3 * kLanes);
  const int q = lane >> 4, j = lane & 15;
  for (int w = 0; w < 4; w++) {
    uint32_t word = 0;
    for (int half = 0; half < 2; half++) {
      const int e = 2 * w + half;
      double fr, fi;
      dft_entry(16 * (e >> 2) + 4 * q + (e & 3), (j + shift) & 31, 32, fr, fi);
      const double v = arr == 0 ? fr : (arr == 1 ? fi : -fi);
      uint16_t bits;
      if (f16) {
        const _Float16 h = (_Float16)(float)v;
        const _Float16 l = (_Float16)(float)(v - (double)h);
        bits = __builtin_bit_cast(uint16_t, term == 0 ? h : (term == 1 ? l : (_Float16)0.0f));
      } else {
        __bf16 h, m, l;
        split3_bf16(v, h, m, l);
        bits = __builtin_bit_cast(uint16_t, term == 0 ? h : (term == 1 ? m : l));
      }
      word |= (uint32_t)bits << (16 * half);
    }
    out[idx * 4 + w] = word;
  }
}

__global__ void k_gen_inverse32_f64(double* out, int32_t shift) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kInv32F64) return;
  const int lane = idx % kLanes, arr = (idx / kLanes) % 3, v = (idx / (3 * kLanes)) % 4, R = idx / (12 * kLanes);
  double fr, fi;
  dft_entry(16 * R + 4 * v + (lane >> 4), ((lane & 15) + shift) & 31, 32, fr, fi);
  out[idx] = arr == 0 ? fr : (arr == 1 ? -(fr + fi) : fi - fr);
}

// 32-point FP64 B operands (layout kOps32F64): one thread per entry.
__global__ void k_gen_operands32_f64(double* t) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kOps32F64) return;
  const int lane = idx % kLanes, arr = (idx / kLanes) % 3, Kb = (idx / (3 * kLanes)) % 2;
  const int v = (idx / (6 * kLanes)) % 4, R = idx / (24 * kLanes);
  double fr, fi;
  dft_entry(16 * R + 4 * v + (lane >> 4), 16 * Kb + (lane & 15), 32, fr, fi);
  t[idx] = arr == 0 ? fr : (arr == 1 ? -(fr + fi) : fi - fr);
}

// K(k) = sum_taps w[a] exp(-2 pi i sum_ax (d_ax(a) k_ax mod L_ax)/L_ax),  d = a (ref) or r - a (centred);
// H = K^t / prod(L) by repeated squaring in complex FP64.  `kind` fixes which frequency sits at
// entry e of the family layout.
__global__ void k_gen_spectrum(double* H, int kind, int layout, int32_t L0, int32_t L1, int32_t L2,
                               const double* taps, int32_t K, int32_t dims, int32_t convention, int32_t radius,
                               int64_t t) {
  const int64_t len = kind == KIND_3D16 ? 4096 : (kind == KIND_2D32 ? 1024 : 256);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  const int lane = (int)(e % kLanes);
  const int q = lane >> 4;
  int j = lane & 15;
  int v, kb = 0;
  if (kind == KIND_3D16 || kind == KIND_2D32) {
    v = (int)((e / kLanes) % 4);
    kb = (int)(e / (4 * kLanes));
  } else {
    v = (int)(e / kLanes);
  }
  int krow = layout == 0 ? q + 4 * v : 4 * q + v;  // frequency on the row axis of the family layout
  if (kind == KIND_2D32) {  // block kb = 2 R + C: rows 16 R + (4q + v | q + 4v), columns 16 C + j
    krow = 16 * (kb >> 1) + (layout == 0 ? q + 4 * v : 4 * q + v);
    j += 16 * (kb & 1);
  }
  double kr = 0.0, ki = 0.0;
  bool zero = false;
  int kax[3] = {0, 0, 0};
  int Lax[3] = {1, 1, 1};
  int nax = 0;
  int64_t k1d = 0;
  if (kind == KIND_2D16 || kind == KIND_2D32) {
    kax[0] = krow; kax[1] = j; Lax[0] = L0; Lax[1] = L1; nax = 2;
    zero = krow >= L0 || j >= L1;
  } else if (kind == KIND_1DPFA) {
    zero = krow >= 15;
    k1d = (15 * j + 16 * krow) % 240;
    nax = 1;
  } else {  // 3D entry (k2 block, reg v, lane): k0 = lane % 16, k1 = rowmap(q, v), k2 = block
    kax[0] = j; kax[1] = krow; kax[2] = kb; Lax[0] = L0; Lax[1] = L1; Lax[2] = L2; nax = 3;
  }
  if (!zero) {
    int64_t ntaps = 1;
    for (int d = 0; d < dims; d++) ntaps *= K;
    for (int64_t tap = 0; tap < ntaps; tap++) {
      const double w = taps[tap];
      if (w == 0.0) continue;
      // tap multi-index, slowest axis first
      int64_t rem = tap;
      int aidx[3] = {0, 0, 0};
      for (int d = dims - 1; d >= 0; d--) {
        aidx[d] = (int)(rem % K);
        rem /= K;
      }
      double frac = 0.0;
      if (kind == KIND_1DPFA) {
        const int64_t dd = convention == 0 ? aidx[0] : radius - aidx[0];
        int64_t m = (dd * k1d) % 240;
        if (m < 0) m += 240;
        frac = (double)m / 240.0;
      } else {
        for (int d = 0; d < nax; d++) {
          const int64_t dd = convention == 0 ? aidx[d] : radius - aidx[d];
          int64_t m = (dd * kax[d]) % Lax[d];
          if (m < 0) m += Lax[d];
          frac += (double)m / (double)Lax[d];
        }
      }
      double s, c;
      sincospi(-2.0 * frac, &s, &c);
      kr += w * c;
      ki += w * s;
    }
    // K^t by repeated squaring
    double pr = 1.0, pi = 0.0, br = kr, bi = ki;
    for (int64_t n = t; n > 0; n >>= 1) {
      if (n & 1) {
        const double tr = pr * br - pi * bi;
        pi = pr * bi + pi * br;
        pr = tr;
      }
      const double tr = br * br - bi * bi;
      bi = 2.0 * br * bi;
      br = tr;
    }
    double norm = kind == KIND_1DPFA ? 240.0 : (double)L0 * (double)L1 * (kind == KIND_3D16 ? (double)L2 : 1.0);
    kr = pr / norm;
    ki = pi / norm;
  }
  H[e] = zero ? 0.0 : kr;
  H[len + e] = zero ? 0.0 : ki;
}

struct TapChunk {
  double w[kTapsPerLaunch];
  int32_t n;
};

__global__ void k_put_taps(double* dst, TapChunk c) {
  for (int i = threadIdx.x; i < c.n; i += blockDim.x) dst[i] = c.w[i];
}

hipError_t put_taps(double* dst, const double* w, int64_t n, hipStream_t s) {
  for (int64_t o = 0; o < n; o += kTapsPerLaunch) {
    TapChunk c{};
    c.n = (int32_t)std::min<int64_t>(kTapsPerLaunch, n - o);
    for (int i = 0; i < c.n; i++) c.w[i] = w[o + i];
    hipLaunchKernelGGL(k_put_taps, dim3(1), dim3(256), 0, s, dst + o, c);
  }
  return hipGetLastError();
}

hipError_t gen_operands(double* ops, uint32_t* opsm, const int32_t* axis_len, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_operands, dim3(3), dim3(256), 0, s, ops, axis_len[0], axis_len[1], axis_len[2]);
  if (opsm)
    hipLaunchKernelGGL(k_gen_operands_mixed, dim3(3), dim3(192), 0, s, opsm, axis_len[0], axis_len[1], axis_len[2]);
  return hipGetLastError();
}

// SplitMix64 stream on the device: element i is output i of the generator seeded with `seed` (the same
// values ffs_fill_uniform writes on the host), so timing runs see data like the caller's, not zeros.
__global__ void k_fill_uniform(double* d, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    d[i] = (double)(z >> 11) * 0x1.0p-53;
  }
}

hipError_t fill_uniform_device(double* d, int64_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_uniform, dim3(1024), dim3(256), 0, s, d, n, seed);
  return hipGetLastError();
}

hipError_t gen_inverse32(uint32_t* out, bool f16, int32_t shift, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_inverse32, dim3((3 * 3 * kLanes + 255) / 256), dim3(256), 0, s, out, (int32_t)f16, shift);
  return hipGetLastError();
}

hipError_t gen_inverse32_f64(double* out, int32_t shift, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_inverse32_f64, dim3((kInv32F64 + 255) / 256), dim3(256), 0, s, out, shift);
  return hipGetLastError();
}

hipError_t gen_operands32_f64(double* ops, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_operands32_f64, dim3((kOps32F64 + 255) / 256), dim3(256), 0, s, ops);
  return hipGetLastError();
}

hipError_t gen_operands32(uint32_t* ops32, bool f16, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_operands32, dim3((3 * 2 * 3 * kLanes + 255) / 256), dim3(256), 0, s, ops32, (int32_t)f16);
  return hipGetLastError();
}

hipError_t gen_spectrum(double* H, int kind, int layout, const int32_t* axis_len, const double* d_taps, int32_t K,
                        int32_t dims, int32_t convention, int32_t radius, int64_t t, hipStream_t s) {
  const int len = (int)spectrum_len(kind);
  hipLaunchKernelGGL(k_gen_spectrum, dim3((len + 255) / 256), dim3(256), 0, s, H, kind, layout, axis_len[0],
                     axis_len[1], axis_len[2], d_taps, K, dims, convention, radius, t);
  return hipGetLastError();
}

__global__ void k_spectrum_f32(const double* H, float* Hf, int64_t len) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < len) {
    Hf[2 * e] = (float)H[e];
    Hf[2 * e + 1] = (float)H[len + e];
  }
}

hipError_t spectrum_to_f32(const double* H, float* Hf, int64_t len, hipStream_t s) {
  hipLaunchKernelGGL(k_spectrum_f32, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, H, Hf, len);
  return hipGetLastError();
}

// The 16-point kernels (1D and 2D) stride persistently over their pairs.  They are launched with at most
// the workgroups that can be resident at once: workgroups beyond that would wait for a later round and
// start cold, while resident waves take further pairs (measured on the 1D kernel when it held 2
// workgroups per CU: Heat-1D 2^20 x 64, 17.3 -> 14.9 us).  Residency is read from the compiled kernel,
// once per kernel.
template <void (*K)(TileArgs)>
void launch16(const TileArgs& a, int64_t blocks, int num_cus, hipStream_t s) {
  static const int resident = [] {  // thread-safe one-time initialisation
    int nb = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, K, 256, 0);
    return (e == hipSuccess && nb >= 1) ? nb : 8;
  }();
  const int64_t nb = std::min<int64_t>(blocks, (int64_t)num_cus * resident);
  hipLaunchKernelGGL(K, dim3((unsigned)nb), dim3(256), 0, s, a);
}

hipError_t launch_block(const TileArgs& a, int kind, hipStream_t s, int num_cus) {
  if (a.pairs <= 0) return hipSuccess;
  if (kind == KIND_3D16) {  // one workgroup per tile pair, two resident per CU (68 KB LDS each for FP64)
    const int64_t blocks3 = std::min<int64_t>(a.pairs, (int64_t)num_cus * (a.precision == 0 ? 2 : 3));
    const dim3 g3((unsigned)blocks3), b3(256);
    if (a.precision == PREC_BF16)
      hipLaunchKernelGGL((k_tile3d<PREC_BF16, 3>), g3, b3, 0, s, a);
    else if (a.precision == PREC_F32)
      hipLaunchKernelGGL((k_tile3d<PREC_F32, 1>), g3, b3, 0, s, a);
    else if (a.precision == 0)
      hipLaunchKernelGGL((k_tile3d<0, 1>), g3, b3, 0, s, a);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  const int64_t waves_needed = a.pairs;
  int64_t blocks = (waves_needed + 3) / 4;
  // 8 workgroups (32 waves) per CU at most, grid-stride beyond; the 32-point kernels stage their tables
  // in LDS per workgroup and run 3 waves per SIMD (4 for the FP16 out16 kernel), so they launch exactly
  // the resident workgroups.
  const int64_t per_cu32 = (a.precision == PREC_F16 && a.out16) ? 4 : (a.precision == 0 ? 2 : 3);
  const int64_t cap = (int64_t)num_cus * (kind == KIND_2D32 ? per_cu32 : 8);
  if (blocks > cap) blocks = cap;
  const dim3 g((unsigned)blocks), b(256);
  if (kind == KIND_2D16) {
    if (a.precision == PREC_BF16)
      launch16<k_tile16_mixed<KIND_2D16, PREC_BF16, 3>>(a, blocks, num_cus, s);
    else if (a.precision == PREC_F32)
      launch16<k_tile16_mixed<KIND_2D16, PREC_F32, 1>>(a, blocks, num_cus, s);
    else
      launch16<k_tile16_f64<KIND_2D16>>(a, blocks, num_cus, s);
  } else if (kind == KIND_1DPFA) {
    if (a.precision == PREC_BF16)
      launch16<k_tile16_mixed<KIND_1DPFA, PREC_BF16, 3>>(a, blocks, num_cus, s);
    else if (a.precision == PREC_F32)
      launch16<k_tile16_mixed<KIND_1DPFA, PREC_F32, 1>>(a, blocks, num_cus, s);
    else
      launch16<k_tile16_f64<KIND_1DPFA>>(a, blocks, num_cus, s);
  } else if (kind == KIND_2D32) {
    if (a.precision == PREC_BF16 && a.out16)
      hipLaunchKernelGGL((k_tile32_bf16<3, true>), g, b, 0, s, a);
    else if (a.precision == PREC_BF16)
      hipLaunchKernelGGL((k_tile32_bf16<3, false>), g, b, 0, s, a);
    else if (a.precision == PREC_F16 && a.out16)
      hipLaunchKernelGGL(k_tile32_f16<true>, g, b, 0, s, a);
    else if (a.precision == PREC_F16)
      hipLaunchKernelGGL(k_tile32_f16<false>, g, b, 0, s, a);
    else if (a.precision == 0 && a.out16)
      hipLaunchKernelGGL(k_tile32_f64<true>, g, b, 0, s, a);
    else if (a.precision == 0)
      hipLaunchKernelGGL(k_tile32_f64<false>, g, b, 0, s, a);
    else
      return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ffsk
