// ffs_direct.hip -- direct (non-FFT) FP64 stencil kernels with temporal blocking, for gfx950.
//
// The same periodic step as the tile kernels (src/2D/check_correct.hpp:64-100 for the ref convention,
// src/1D/check_correct.hpp:15-38 for the centred one), applied t times per launch without a transform:
// the fallback and the sanity baseline of the FFT-stencil path (the reference has no such kernel; its
// CPU check is the semantics).  FP64 VALU FMAs, taps in kernel arguments (scalar registers).
//
// Window form.  With lo = 2r (ref) or r (centred) and g[e] the taps in window order (ref: g[e] =
// w[2r - e] per axis; centred: g[e] = w[e]), one step on a window u is u'[i] = sum_{e in [0,2r]^d}
// g[e] u[i + e]: the valid part shrinks by 2r per axis and step, from the top.  A window whose first
// point is grid index x0 - t*lo therefore yields S^t(in)[x0 + i] for i < window - 2rt.
//
// 2D and 3D stream along axis 0.  Level s (1..t) keeps, per owned column, 2r+1 partial sums of the
// rows (planes) it is building; every arriving row of level s-1 is read once from LDS (with its 2r
// neighbours along the other axes) and added into all 2r+1 partial sums with the matching row of taps;
// the completed row goes to LDS for level s+1 (or to HBM from level t).  Level s consumes in iteration
// m what level s-1 produced in iteration m-1, so one barrier per iteration orders everything (two LDS
// slots per level).  The unroll by 2r+1 makes every partial-sum index static.
//
// 3D windows are 32 x 64 points per plane, 512 threads.
//
// 1D keeps a window of kDir1Window points in registers (kDir1Pts per thread plus a redundant run of the
// right neighbour's), exchanging through LDS once every kDir1Pts / 2r steps.
#include <hip/hip_runtime.h>

#include <utility>

#include "ffs_internal.h"

namespace ffsk {
namespace {

__device__ __forceinline__ int64_t dwrap(int64_t i, int64_t n) {
  const int64_t r = i % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------------------------------
// 1D: 256 threads x kDir1Pts points in registers.  Each thread also holds the first H = 2r*K values of its
// right neighbour, so K steps run between two exchanges through LDS: the neighbour's values are
// recomputed redundantly (the held run shrinks by 2r per step) in exchange for K times fewer barriers.
template <int R>
__global__ __launch_bounds__(256) void k_direct1d(DirectArgs a) {
  constexpr int E = kDir1Pts;
  constexpr int W = kDir1Window;
  constexpr int K = R > 0 ? E / (2 * R) : E;  // steps per exchange
  constexpr int H = 2 * R * K;                 // neighbour values held, <= E
  constexpr int XS = H + 1;                    // exchange row stride (odd: no LDS bank conflicts)
  static_assert(257 * XS <= W + W / E + 32, "exchange rows must fit the staging buffer");
  __shared__ double win[W + W / E + 32];  // window staging (one pad per thread's run), then the exchange rows
  const int tid = threadIdx.x;
  const int64_t n = a.n0;
  const int64_t x0 = (int64_t)blockIdx.x * a.valid1;
  const int64_t base = x0 - (int64_t)a.steps * a.lo;
  if (base >= 0 && base + W <= n) {  // interior window: no wrap arithmetic
    for (int i = tid; i < W; i += 256) win[i + i / E] = a.src[base + i];
  } else {
    for (int i = tid; i < W; i += 256) win[i + i / E] = a.src[dwrap(base + i, n)];
  }
  for (int i = W + W / E + tid; i < W + W / E + 32; i += 256) win[i] = 0.0;
  __syncthreads();
  double u[E + (H > 0 ? H : 1)];
#pragma unroll
  for (int k = 0; k < E; k++) u[k] = win[tid * (E + 1) + k];
  double g[2 * R + 1];
#pragma unroll
  for (int e = 0; e <= 2 * R; e++) g[e] = a.g[e];
  __syncthreads();  // the staging buffer becomes the exchange buffer
  for (int s = 0; s < a.steps; s += K) {
    if constexpr (H > 0) {
#pragma unroll
      for (int k = 0; k < H; k++) win[tid * XS + k] = u[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < H; k++) u[E + k] = win[(tid + 1) * XS + k];  // row 256: stale, beyond the valid part
      __syncthreads();
    }
    const int ks = min(K, a.steps - s);
#pragma unroll
    for (int j = 1; j <= K; j++) {
      if (j > ks) break;
#pragma unroll
      for (int i = 0; i < E + H - 2 * R * j; i++) {  // ascending, in place: u[i + e] is still the old value
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e <= 2 * R; e++) acc = __builtin_fma(g[e], u[i + e], acc);
        u[i] = acc;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < E; k++) win[tid * (E + 1) + k] = u[k];
  __syncthreads();
  for (int i = tid; i < a.valid1; i += 256)
    if (x0 + i < n) a.dst[x0 + i] = win[i + i / E];
}

// ---------------------------------------------------------------------------------------------------
// 2D: window of kDir2Cols columns (2 per thread) streamed down a segment of rows.
template <int R, int T, bool STAR>
__global__ __launch_bounds__(256) void k_direct2d(DirectArgs a) {
  constexpr int U = 2 * R + 1;
  constexpr int LS = kDir2Cols + 8;  // LDS row: reads run up to 2r + 1 columns past the window
  __shared__ __attribute__((aligned(16))) double buf[T][2][LS];
  const int tid = threadIdx.x;
  const int c = 2 * tid;
  const int sx = blockIdx.x % a.strips1;
  const int sy = blockIdx.x / a.strips1;
  const int64_t n0 = a.n0, n1 = a.n1;
  const int64_t x0 = (int64_t)sx * a.valid1;
  const int64_t y0 = (int64_t)sy * a.seg;
  const int64_t rows = min(a.seg, n0 - y0);
  const int64_t col0 = dwrap(x0 - (int64_t)T * a.lo + c, n1);
  const int64_t col1 = dwrap(x0 - (int64_t)T * a.lo + c + 1, n1);
  for (int i = tid; i < T * 2 * LS; i += 256) (&buf[0][0][0])[i] = 0.0;
  __syncthreads();
  double acc[T][U][2];
#pragma unroll
  for (int s = 0; s < T; s++)
#pragma unroll
    for (int k = 0; k < U; k++) acc[s][k][0] = acc[s][k][1] = 0.0;
  // input rows are loaded two iterations ahead of their LDS write (one row ahead leaves HBM latency exposed)
  int64_t yr = dwrap(y0 - (int64_t)T * a.lo, n0);
  double p0 = a.src[yr * n1 + col0], p1 = a.src[yr * n1 + col1];
  yr = yr + 1 == n0 ? 0 : yr + 1;
  double q0 = a.src[yr * n1 + col0], q1 = a.src[yr * n1 + col1];
  constexpr int64_t kShift = 2 * R * T + T - 1;  // level t emits row m - kShift in iteration m
  const int64_t M = rows + kShift;
  const bool st0 = c < a.valid1 && x0 + c < n1;
  const bool st1 = c + 1 < a.valid1 && x0 + c + 1 < n1;
  for (int64_t m0 = 0; m0 < M; m0 += U) {
#pragma unroll
    for (int p = 0; p < U; p++) {
      const int64_t m = m0 + p;
      const int sl = (int)(m & 1);
      *reinterpret_cast<double2*>(&buf[0][sl][c]) = make_double2(p0, p1);
      p0 = q0;
      p1 = q1;
      yr = yr + 1 == n0 ? 0 : yr + 1;
      q0 = a.src[yr * n1 + col0];
      q1 = a.src[yr * n1 + col1];
      __syncthreads();
#pragma unroll
      for (int s = 0; s < T; s++) {  // level s + 1
        const double* in = s == 0 ? &buf[0][sl][c] : &buf[s][sl ^ 1][c];
        double v[2 * R + 2];
#pragma unroll
        for (int k = 0; k <= R; k++) {
          const double2 q = *reinterpret_cast<const double2*>(in + 2 * k);
          v[2 * k] = q.x;
          v[2 * k + 1] = q.y;
        }
#pragma unroll
        for (int e0 = 0; e0 < U; e0++) {
          const int k = (p - e0 + U) % U;
#pragma unroll
          for (int e1 = 0; e1 < U; e1++) {
            if (STAR && e0 != R && e1 != R) continue;
            const double gw = a.g[e0 * U + e1];
            acc[s][k][0] = __builtin_fma(gw, v[e1], acc[s][k][0]);
            acc[s][k][1] = __builtin_fma(gw, v[e1 + 1], acc[s][k][1]);
          }
        }
        const int ke = (p + 1) % U;  // row m - 2r of this level is complete
        const double o0 = acc[s][ke][0], o1 = acc[s][ke][1];
        acc[s][ke][0] = 0.0;
        acc[s][ke][1] = 0.0;
        if (s + 1 < T) {
          *reinterpret_cast<double2*>(&buf[s + 1][sl][c]) = make_double2(o0, o1);
        } else {
          const int64_t j = m - kShift;
          if (j >= 0 && j < rows) {
            double* row = a.dst + (y0 + j) * n1 + x0 + c;
            if (st0) row[0] = o0;
            if (st1) row[1] = o1;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------------
// 3D: plane window kDir3Rows x kDir3Cols (axes 1 x 2; kDir3PerThread x 2 points per thread) streamed
// along axis 0.  The window is widest along the contiguous axis: 64 columns (512 B rows) load at
// ~1.25x the rate of 32-column rows.
template <int R, int T, bool STAR>
__global__ __launch_bounds__(kDir3Threads) void k_direct3d(DirectArgs a) {
  constexpr int U = 2 * R + 1;
  constexpr int MR = kDir3PerThread;
  constexpr int TX = kDir3Cols / 2;
  constexpr int RS = kDir3Cols + 4;  // LDS row stride (doubles)
  constexpr int PR = kDir3Rows + 4;  // LDS rows per plane (reads run 2r rows past the window)
  __shared__ __attribute__((aligned(16))) double buf[T][2][PR * RS];
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int c = 2 * tx, r0 = MR * ty;
  const int64_t n0 = a.n0, n1 = a.n1, n2 = a.n2;
  const int sx = blockIdx.x % a.strips2;
  const int sy = (blockIdx.x / a.strips2) % a.strips1;
  const int sz = blockIdx.x / (a.strips2 * a.strips1);
  const int64_t x0 = (int64_t)sx * a.valid2;  // axis 2
  const int64_t y0 = (int64_t)sy * a.valid1;  // axis 1
  const int64_t z0 = (int64_t)sz * a.seg;     // axis 0
  const int64_t planes = min(a.seg, n0 - z0);
  int64_t gcol[2], grow[MR];
#pragma unroll
  for (int b = 0; b < 2; b++) gcol[b] = dwrap(x0 - (int64_t)T * a.lo + c + b, n2);
#pragma unroll
  for (int i = 0; i < MR; i++) grow[i] = dwrap(y0 - (int64_t)T * a.lo + r0 + i, n1) * n2;
  for (int i = tid; i < T * 2 * PR * RS; i += kDir3Threads) (&buf[0][0][0])[i] = 0.0;
  __syncthreads();
  double acc[T][U][MR][2];
#pragma unroll
  for (int s = 0; s < T; s++)
#pragma unroll
    for (int k = 0; k < U; k++)
#pragma unroll
      for (int i = 0; i < MR; i++) acc[s][k][i][0] = acc[s][k][i][1] = 0.0;
  // input planes are loaded two iterations ahead of their LDS write
  int64_t zr = dwrap(z0 - (int64_t)T * a.lo, n0);
  double pf[MR][2], pg[MR][2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const double* pl = a.src + zr * n1 * n2;
#pragma unroll
    for (int i = 0; i < MR; i++)
#pragma unroll
      for (int b = 0; b < 2; b++) (h == 0 ? pf : pg)[i][b] = pl[grow[i] + gcol[b]];
    if (h == 0) zr = zr + 1 == n0 ? 0 : zr + 1;
  }
  constexpr int64_t kShift = 2 * R * T + T - 1;
  const int64_t M = planes + kShift;
  bool stc[2];
#pragma unroll
  for (int b = 0; b < 2; b++) stc[b] = c + b < a.valid2 && x0 + c + b < n2;
  bool str[MR];
#pragma unroll
  for (int i = 0; i < MR; i++) str[i] = r0 + i < a.valid1 && y0 + r0 + i < n1;
  for (int64_t m0 = 0; m0 < M; m0 += U) {
#pragma unroll
    for (int p = 0; p < U; p++) {
      const int64_t m = m0 + p;
      const int sl = (int)(m & 1);
#pragma unroll
      for (int i = 0; i < MR; i++) {
        *reinterpret_cast<double2*>(&buf[0][sl][(r0 + i) * RS + c]) = make_double2(pf[i][0], pf[i][1]);
        pf[i][0] = pg[i][0];
        pf[i][1] = pg[i][1];
      }
      zr = zr + 1 == n0 ? 0 : zr + 1;
      {
        const double* pl = a.src + zr * n1 * n2;
#pragma unroll
        for (int i = 0; i < MR; i++)
#pragma unroll
          for (int b = 0; b < 2; b++) pg[i][b] = pl[grow[i] + gcol[b]];
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < T; s++) {
        const double* in = s == 0 ? &buf[0][sl][0] : &buf[s][sl ^ 1][0];
        double v[MR + 2 * R][2 + 2 * R];
#pragma unroll
        for (int i = 0; i < MR + 2 * R; i++)
#pragma unroll
          for (int k = 0; k <= R; k++) {
            const double2 q = *reinterpret_cast<const double2*>(in + (r0 + i) * RS + c + 2 * k);
            v[i][2 * k] = q.x;
            v[i][2 * k + 1] = q.y;
          }
#pragma unroll
        for (int e0 = 0; e0 < U; e0++) {
          const int k = (p - e0 + U) % U;
#pragma unroll
          for (int e1 = 0; e1 < U; e1++)
#pragma unroll
            for (int e2 = 0; e2 < U; e2++) {
              if (STAR && (int(e0 != R) + int(e1 != R) + int(e2 != R)) > 1) continue;
              const double gw = a.g[(e0 * U + e1) * U + e2];
#pragma unroll
              for (int i = 0; i < MR; i++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[s][k][i][b] = __builtin_fma(gw, v[i + e1][b + e2], acc[s][k][i][b]);
            }
        }
        const int ke = (p + 1) % U;
        double o[MR][2];
#pragma unroll
        for (int i = 0; i < MR; i++)
#pragma unroll
          for (int b = 0; b < 2; b++) {
            o[i][b] = acc[s][ke][i][b];
            acc[s][ke][i][b] = 0.0;
          }
        if (s + 1 < T) {
#pragma unroll
          for (int i = 0; i < MR; i++)
            *reinterpret_cast<double2*>(&buf[s + 1][sl][(r0 + i) * RS + c]) = make_double2(o[i][0], o[i][1]);
        } else {
          const int64_t j = m - kShift;
          if (j >= 0 && j < planes) {
            double* pl = a.dst + (z0 + j) * n1 * n2;
#pragma unroll
            for (int i = 0; i < MR; i++)
#pragma unroll
              for (int b = 0; b < 2; b++)
                if (str[i] && stc[b]) pl[(y0 + r0 + i) * n2 + x0 + c + b] = o[i][b];
          }
        }
      }
    }
  }
}

template <int R, int T, bool STAR>
void go2d(const DirectArgs& a, dim3 grid, hipStream_t s) {
  k_direct2d<R, T, STAR><<<grid, 256, 0, s>>>(a);
}

template <int R, int T, bool STAR>
void go3d(const DirectArgs& a, dim3 grid, hipStream_t s) {
  k_direct3d<R, T, STAR><<<grid, kDir3Threads, 0, s>>>(a);
}

template <int R, bool STAR, int... Ts>
hipError_t launch2d(int T, const DirectArgs& a, dim3 grid, hipStream_t s, std::integer_sequence<int, Ts...>) {
  bool done = false;
  (void)((T == Ts + 1 ? (go2d<R, Ts + 1, STAR>(a, grid, s), done = true) : false) || ...);
  return done ? hipGetLastError() : hipErrorInvalidValue;
}

template <int R, bool STAR, int... Ts>
hipError_t launch3d(int T, const DirectArgs& a, dim3 grid, hipStream_t s, std::integer_sequence<int, Ts...>) {
  bool done = false;
  (void)((T == Ts + 1 ? (go3d<R, Ts + 1, STAR>(a, grid, s), done = true) : false) || ...);
  return done ? hipGetLastError() : hipErrorInvalidValue;
}

}  // namespace

// Deepest instantiated depth per (dims, radius, star support): the partial sums of t levels must stay in
// registers (beyond these the accumulators spill).
int direct_max_depth(int dims, int radius, bool star) {
  if (dims == 1) return radius <= 3 ? kDir1MaxDepth : 0;
  if (dims == 2) {
    if (radius <= 1) return 8;
    if (radius == 2) return star ? 6 : 4;
    return radius == 3 ? 2 : 0;
  }
  if (radius == 0) return 3;
  return radius == 1 ? (star ? 3 : 2) : 0;
}

hipError_t launch_direct(const DirectArgs& a, int dims, int radius, bool star, hipStream_t s) {
  const int T = a.steps;
  if (T < 1 || T > direct_max_depth(dims, radius, star)) return hipErrorInvalidValue;
  if (dims == 1) {
    const dim3 grid((unsigned)a.nseg);
    switch (radius) {
      case 0: k_direct1d<0><<<grid, 256, 0, s>>>(a); break;
      case 1: k_direct1d<1><<<grid, 256, 0, s>>>(a); break;
      case 2: k_direct1d<2><<<grid, 256, 0, s>>>(a); break;
      case 3: k_direct1d<3><<<grid, 256, 0, s>>>(a); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (dims == 2) {
    const dim3 grid((unsigned)((int64_t)a.strips1 * a.nseg));
    const auto seq8 = std::make_integer_sequence<int, 8>{};
    switch (radius) {
      case 0: return launch2d<0, false>(T, a, grid, s, seq8);
      case 1: return star ? launch2d<1, true>(T, a, grid, s, seq8) : launch2d<1, false>(T, a, grid, s, seq8);
      case 2:
        return star ? launch2d<2, true>(T, a, grid, s, std::make_integer_sequence<int, 6>{})
                    : launch2d<2, false>(T, a, grid, s, std::make_integer_sequence<int, 4>{});
      case 3: return launch2d<3, false>(T, a, grid, s, std::make_integer_sequence<int, 2>{});
      default: return hipErrorInvalidValue;
    }
  }
  const dim3 grid((unsigned)((int64_t)a.strips1 * a.strips2 * a.nseg));
  switch (radius) {
    case 0: return launch3d<0, false>(T, a, grid, s, std::make_integer_sequence<int, 3>{});
    case 1:
      return star ? launch3d<1, true>(T, a, grid, s, std::make_integer_sequence<int, 3>{})
                  : launch3d<1, false>(T, a, grid, s, std::make_integer_sequence<int, 2>{});
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ffsk
