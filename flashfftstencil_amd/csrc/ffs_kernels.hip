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

// XCD-aware workgroup order: the hardware hands workgroup b to XCD (b mod 8); renumber so that each
// XCD receives one contiguous range of logical workgroups (neighbouring tiles share its L2).
__device__ __forceinline__ int64_t xcd_logical_block(int64_t b, int64_t G) {
  const int64_t x = b & 7, per = G >> 3, rem = G & 7;
  const int64_t start = x * per + (x < rem ? x : rem);
  return start + (b >> 3);
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

// out = A^T-chain stage: A given as 4 chunks (A operand values), B = F (or conj F when INV).
template <bool INV>
__device__ __forceinline__ void cstage(const double (&ar)[4], const double (&ai)[4], const OpsAxis& f, d4& outr,
                                       d4& outi) {
  d4 k1 = {0.0, 0.0, 0.0, 0.0}, k2 = k1, k3 = k1;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    k1 = FFS_MFMA64(ar[c] + ai[c], f.r[c], k1);
    k2 = FFS_MFMA64(ar[c], INV ? f.s[c] : f.d[c], k2);
    k3 = FFS_MFMA64(ai[c], INV ? f.d[c] : f.s[c], k3);
  }
  if (!INV) {
    outr = k1 - k3;
    outi = k1 + k2;
  } else {
    outr = k1 + k3;
    outi = k1 - k2;
  }
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
        if (okA) drow[cA] = yr[v];
        if (okB) drow[cB] = yi[v];
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
          if (pA < a.n1) a.dst[pA] = yr[v];
          if (pc.hasB && pB < a.n1) a.dst[pB] = yi[v];
        }
      }
    }
  }
}

template <int KIND>
__global__ __launch_bounds__(256) void k_tile16_f64(TileArgs a) {
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, j = lane & 15;
  const int64_t W = (int64_t)gridDim.x * 4;
  const int64_t gw = xcd_logical_block(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // 2D: both axes are 16-point DFTs and share one operand set; 1D: 15-point rows, 16-point columns.
  OpsAxis f0, f1;
  load_ops(f0, a.ops, 0, lane);
  if (KIND != KIND_2D16) load_ops(f1, a.ops, 1, lane);
  const OpsAxis& g1 = (KIND == KIND_2D16) ? f0 : f1;
  d4 hr, hi;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    hr[v] = a.H[v * kLanes + lane];
    hi[v] = a.H[256 + v * kLanes + lane];
  }

  int64_t tp = gw;
  if (tp >= a.pairs) return;
  PairCoord pc = pair_coord<KIND>(a, tp);
  double xr[4], xi[4];
  gather_pair<KIND>(a, pc, q, j, xr, xi);
  for (; tp < a.pairs; tp += W) {
    // prefetch the next pair while this one computes
    const int64_t tn = tp + W;
    PairCoord pn = pc;
    double nr[4] = {0.0, 0.0, 0.0, 0.0}, ni[4] = {0.0, 0.0, 0.0, 0.0};
    if (tn < a.pairs) {
      pn = pair_coord<KIND>(a, tn);
      gather_pair<KIND>(a, pn, q, j, nr, ni);
    }

    d4 sr, si;
    cstage<false>(xr, xi, f0, sr, si);  // contract axis 0 (rows)
    double cr[4], ci[4];
    to_chunks(sr, cr);
    to_chunks(si, ci);
    cstage<false>(cr, ci, g1, sr, si);  // contract axis 1 -> Y[k0 = q+4v][k1 = j]
    {
      const d4 yr = sr, yi = si;
      sr = yr * hr - yi * hi;
      si = yr * hi + yi * hr;
    }
    to_chunks(sr, cr);
    to_chunks(si, ci);
    cstage<true>(cr, ci, f0, sr, si);  // inverse over k0
    to_chunks(sr, cr);
    to_chunks(si, ci);
    cstage<true>(cr, ci, g1, sr, si);  // inverse over k1 -> X'[n0 = q+4v][n1 = j]
    store_pair<KIND>(a, pc, q, j, sr, si);

    pc = pn;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      xr[c] = nr[c];
      xi[c] = ni[c];
    }
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
  base[(1 * 4 + c) * kLanes + lane] = fr + fi;
  base[(2 * 4 + c) * kLanes + lane] = fi - fr;
}

// K(k) = sum_taps w[a] exp(-2 pi i sum_ax (d_ax(a) k_ax mod L_ax)/L_ax),  d = a (ref) or r - a (centred);
// H = K^t / prod(L) by repeated squaring in complex FP64.  `kind` fixes which frequency sits at
// entry e of the family layout.
__global__ void k_gen_spectrum(double* H, int kind, int32_t L0, int32_t L1, int32_t L2, const double* taps,
                               int32_t K, int32_t dims, int32_t convention, int32_t radius, int64_t t) {
  const int64_t len = kind == KIND_3D16 ? 4096 : 256;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  const int lane = (int)(e % kLanes);
  const int q = lane >> 4, j = lane & 15;
  int v, kb = 0;
  if (kind == KIND_3D16) {
    v = (int)((e / kLanes) % 4);
    kb = (int)(e / (4 * kLanes));
  } else {
    v = (int)(e / kLanes);
  }
  const int krow = q + 4 * v;  // frequency on the row axis of the family layout
  double kr = 0.0, ki = 0.0;
  bool zero = false;
  int kax[3] = {0, 0, 0};
  int Lax[3] = {1, 1, 1};
  int nax = 0;
  int64_t k1d = 0;
  if (kind == KIND_2D16) {
    kax[0] = krow; kax[1] = j; Lax[0] = L0; Lax[1] = L1; nax = 2;
    zero = krow >= L0 || j >= L1;
  } else if (kind == KIND_1DPFA) {
    zero = krow >= 15;
    k1d = (15 * j + 16 * krow) % 240;
    nax = 1;
  } else {
    kax[0] = krow; kax[1] = kb; kax[2] = j; Lax[0] = L0; Lax[1] = L1; Lax[2] = L2; nax = 3;
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

hipError_t gen_operands(double* ops, float* /*opsf*/, const int32_t* axis_len, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_operands, dim3(3), dim3(256), 0, s, ops, axis_len[0], axis_len[1], axis_len[2]);
  return hipGetLastError();
}

hipError_t gen_spectrum(double* H, int kind, const int32_t* axis_len, const double* d_taps, int32_t K, int32_t dims,
                        int32_t convention, int32_t radius, int64_t t, hipStream_t s) {
  const int len = kind == KIND_3D16 ? 4096 : 256;
  hipLaunchKernelGGL(k_gen_spectrum, dim3((len + 255) / 256), dim3(256), 0, s, H, kind, axis_len[0], axis_len[1],
                     axis_len[2], d_taps, K, dims, convention, radius, t);
  return hipGetLastError();
}

hipError_t launch_block(const TileArgs& a, int kind, hipStream_t s, int num_cus) {
  if (a.pairs <= 0) return hipSuccess;
  const int64_t waves_needed = a.pairs;
  int64_t blocks = (waves_needed + 3) / 4;
  const int64_t cap = (int64_t)num_cus * 8;  // 8 workgroups (32 waves) per CU at most, grid-stride beyond
  if (blocks > cap) blocks = cap;
  if (kind == KIND_2D16) {
    hipLaunchKernelGGL(k_tile16_f64<KIND_2D16>, dim3((unsigned)blocks), dim3(256), 0, s, a);
  } else if (kind == KIND_1DPFA) {
    hipLaunchKernelGGL(k_tile16_f64<KIND_1DPFA>, dim3((unsigned)blocks), dim3(256), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ffsk
