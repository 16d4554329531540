// Streaming kernel family of the plain GPU form: FP64 (reference-exact) and
// FP32 (the soft-min passes, for the low-precision log-domain solves).
//
// Matrices are column-major. Every kernel that belongs to one iteration of an
// iterative solve takes the solve's LoopState and returns at once when the
// solve is no longer active, so the host can enqueue iterations ahead of
// knowing whether the loop has ended (the reference's per-iteration host
// round trip, src/cuda_sinkhorn.cu:32-64, is gone).
//
// Reductions use fixed trees and fixed chunk orders: a result does not depend
// on timing, only on the problem shape.
#pragma once

#include "common.hpp"
#include "exp2_f64.hpp"

namespace rwig {

constexpr int kTB = 256;          // threads per block of the streaming kernels
constexpr int kWavesPerBlock = kTB / kWave;

// ---------------------------------------------------------------- launch shape

// Column chunking of a row-oriented pass over an M x N matrix: enough blocks
// of 256 rows x cw columns to cover the 256 CUs several times over, with at
// least 16 columns per chunk.
struct Chunking {
  int nchunk, cw;
};
inline Chunking chunking(int M, int N) {
  const int rowblocks = (M + kTB - 1) / kTB;
  int nchunk = 2048 / (rowblocks > 0 ? rowblocks : 1);
  const int maxc = (N + 15) / 16;
  if (nchunk > maxc) nchunk = maxc;
  if (nchunk < 1) nchunk = 1;
  const int cw = (N + nchunk - 1) / nchunk;
  nchunk = (N + cw - 1) / cw;
  return {nchunk, cw};
}

// ---------------------------------------------------------------- elementwise

template <class F>
__global__ void __launch_bounds__(kTB) k_map(long n, F f, const LoopState *ls) {
  if (ls && !ls->active) return;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    f(k);
}

template <class F>
void launch_map(hipStream_t st, long n, F f, const LoopState *ls = nullptr) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_map<F>, dim3(grid_for(n, kTB, 8192)), dim3(kTB), 0, st, n, f, ls);
  check_launch();
}

// ---------------------------------------------------------------- K * X

// part[(c*S + s)*M + i] = sum_{j in chunk c} K(i,j) X(j,s), s in [s0, s0+SB)
template <int SB>
__global__ void __launch_bounds__(kTB)
k_gemv_n_part(const double *__restrict__ K, int M, int N, long ldk,
              const double *__restrict__ X, long ldx, int S, int cw,
              double *__restrict__ part, const LoopState *ls) {
  if (ls && !ls->active) return;
  const int i = blockIdx.x * kTB + threadIdx.x;
  const int c = blockIdx.y, s0 = blockIdx.z * SB;
  if (i >= M) return;
  const int j0 = c * cw, j1 = min(N, j0 + cw);
  const int ns = min(SB, S - s0);
  double acc[SB];
#pragma unroll
  for (int t = 0; t < SB; ++t) acc[t] = 0.0;
  for (int j = j0; j < j1; ++j) {
    const double kij = K[i + j * ldk];
#pragma unroll
    for (int t = 0; t < SB; ++t)
      if (t < ns) acc[t] += kij * X[j + (long)(s0 + t) * ldx];
  }
#pragma unroll
  for (int t = 0; t < SB; ++t)
    if (t < ns) part[((long)c * S + s0 + t) * M + i] = acc[t];
}

// Chunk merges. The partial results of the row passes (nchunk chunks per
// output element, element k of chunk c at c * len + k) are merged by
// workgroups of 8 wavefronts: wavefront w takes chunks w, w + 8, ... of 64
// consecutive elements (coalesced loads, 8 times the parallelism of one
// thread per element), then wavefront 0 combines the 8 results through LDS.
// The order is fixed: a result depends on the shape only.
constexpr int kMergeWaves = 8, kMergeTB = kWave * kMergeWaves;

inline unsigned merge_grid(long len) { return (unsigned)((len + kWave - 1) / kWave); }

// op(k, sum over the chunks of part) for every output element k
template <typename T, class Op>
__global__ void __launch_bounds__(kMergeTB)
k_reduce_part(const T *__restrict__ part, int nchunk, long len, Op op,
              const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T sh[kMergeWaves][kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const long k = blockIdx.x * (long)kWave + lane;
  T v = T(0);
  if (k < len)
    for (int c = w; c < nchunk; c += kMergeWaves) v += part[(long)c * len + k];
  sh[w][lane] = v;
  __syncthreads();
  if (w == 0 && k < len) {
    T r = sh[0][lane];
    for (int u = 1; u < kMergeWaves; ++u) r += sh[u][lane];
    op(k, r);
  }
}

// Y = K X with an epilogue: op(k = s*M + i, (K X)(i,s)).
template <class Op>
void gemv_n(hipStream_t st, const double *K, int M, int N, long ldk,
            const double *X, long ldx, int S, double *part, Op op,
            const LoopState *ls) {
  const Chunking ch = chunking(M, N);
  constexpr int SB = 8;
  dim3 grid((M + kTB - 1) / kTB, ch.nchunk, (S + SB - 1) / SB);
  if (S == 1) {
    hipLaunchKernelGGL(k_gemv_n_part<1>, grid, dim3(kTB), 0, st, K, M, N, ldk, X, ldx, S,
                       ch.cw, part, ls);
  } else {
    hipLaunchKernelGGL(k_gemv_n_part<SB>, grid, dim3(kTB), 0, st, K, M, N, ldk, X, ldx, S,
                       ch.cw, part, ls);
  }
  check_launch();
  const long len = (long)S * M;
  hipLaunchKernelGGL((k_reduce_part<double, Op>), dim3(merge_grid(len)), dim3(kMergeTB), 0, st,
                     part, ch.nchunk, len, op, ls);
  check_launch();
}

inline size_t gemv_n_part_size(int M, int N, int S) {
  return (size_t)chunking(M, N).nchunk * S * M;
}

// ---------------------------------------------------------------- K^T * U

// One wavefront per column j: dots[s] = sum_i K(i,j) U(i,s); op(j, dots) on
// lane 0.
template <int SM, class Op>
__global__ void __launch_bounds__(kTB)
k_gemv_t(const double *__restrict__ K, int M, int N, long ldk,
         const double *__restrict__ U, long ldu, int S, Op op,
         const LoopState *ls) {
  if (ls && !ls->active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (j >= N) return;
  const double *col = K + (long)j * ldk;
  double acc[SM];
#pragma unroll
  for (int t = 0; t < SM; ++t) acc[t] = 0.0;
  for (int i = lane; i < M; i += kWave) {
    const double kij = col[i];
#pragma unroll
    for (int t = 0; t < SM; ++t)
      if (t < S) acc[t] += kij * U[i + (long)t * ldu];
  }
  // the op indexes the dots at run time: staged in LDS, so that acc stays in
  // registers (a private array indexed at run time would live in scratch)
  __shared__ double dots[kWavesPerBlock][SM];
  double *dw = dots[threadIdx.x / kWave];
#pragma unroll
  for (int t = 0; t < SM; ++t)
    if (t < S) {
      const double v = wave_sum(acc[t]);
      if (lane == 0) dw[t] = v;
    }
  if (lane == 0) op(j, (const double *)dw);
}

#define RWIG_DISPATCH_S(S, SM, ...)                                            \
  do {                                                                         \
    if ((S) <= 1) { constexpr int SM = 1; __VA_ARGS__; }                       \
    else if ((S) <= 4) { constexpr int SM = 4; __VA_ARGS__; }                  \
    else if ((S) <= 8) { constexpr int SM = 8; __VA_ARGS__; }                  \
    else if ((S) <= 16) { constexpr int SM = 16; __VA_ARGS__; }                \
    else if ((S) <= 32) { constexpr int SM = 32; __VA_ARGS__; }                \
    else throw Error(RWIG_ERR_INVALID, "at most 32 columns are supported");   \
  } while (0)

template <class Op>
void gemv_t(hipStream_t st, const double *K, int M, int N, long ldk,
            const double *U, long ldu, int S, Op op, const LoopState *ls) {
  const dim3 grid((N + kWavesPerBlock - 1) / kWavesPerBlock);
  RWIG_DISPATCH_S(S, SM, hipLaunchKernelGGL((k_gemv_t<SM, Op>), grid, dim3(kTB), 0, st, K, M,
                                            N, ldk, U, ldu, S, op, ls));
  check_launch();
}

// ---------------------------------------------------------------- soft-min passes
//
// The soft-min of R_s(i,j) = (C(i,j) - F(i,s)) - G(j,s) over i (columns) or
// over j (rows), for all S slices in one sweep over C: each entry of C is read
// once per pass whatever S is, and kept in a register while the S slices use
// it. The log-sum-exp is accumulated online: a running minimum m and
// sum t = sum exp(-(r - m)/reg), rescaled when the minimum drops, one exp per
// entry:
//   d = (r - m)/reg, e = exp(-|d|); d < 0: t = t e + 1, m = r; else t += e.
// Two partial states merge as m = min(m1, m2),
// t = t1 exp(-(m1 - m)/reg) + t2 exp(-(m2 - m)/reg). The result
// m - reg log t is the two-pass form of src/sinkhorn_impl_log.cpp:48-50 up to
// rounding.

// The passes work in base-2 units of reg: every operand of R is multiplied by
// sc = log2(e) / reg once, where it is loaded (an entry of C once for all S
// slices, F, G and the shifts once per row or column), so that
// exp(-r / reg) = exp2(-r * sc) needs no multiply per (entry, slice). A
// result m2 - log2(t) converts back with 1 / sc = reg ln 2.
// exp2: FP64 uses exp2_f64 (range reduction and a degree-11 polynomial,
// about 1 ulp, straight-line); FP32 uses the hardware v_exp_f32 (about 1 ulp;
// the arguments are <= 0 and terms below 2^-126 vanish either way).
constexpr double kLog2e = 1.4426950408889634074;
constexpr double kLn2 = 0.69314718055994530942;
__device__ __forceinline__ double sm_exp2(double x) { return exp2_f64(x); }
__device__ __forceinline__ float sm_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
// exp2 of an argument held in T, evaluated in X: X = float with T = double is
// the mixed form (FP64 arguments, the hardware FP32 exponential). The
// argument is formed in FP64 and only then rounded, so the weight's relative
// error is about 2^-24 times the argument, which is <= 0 and small for the
// weights that count; an FP32 argument would carry the rounding of C and of
// the shifts instead (2^-24 max|C| / reg relative: 1.5e-5 at max C 1 and
// reg 0.004).
template <typename X, typename T> __device__ __forceinline__ T ex2(T x) {
  return (T)sm_exp2((X)x);
}
// The same over T's whole exponent range: in the mixed form x = k + r with
// k = floor(x), 2^r in FP32 on [0, 1) and 2^k applied exactly in FP64. For
// the weighted sums, whose fixed shifts can leave every term of a row far
// below FP32's range (2^-126) while FP64 holds them.
template <typename X, typename T> __device__ __forceinline__ T ex2_wide(T x) {
  if constexpr (sizeof(X) < sizeof(T)) {
    const T k = fmin(fmax(floor(x), T(-1100)), T(1100));
    return ldexp((T)sm_exp2((X)(x - k)), (int)k);
  } else {
    return (T)sm_exp2((X)x);
  }
}

// Online log-sum-exp state in base-2 units: running minimum m and
// t = sum exp2(-(r - m)), rescaled when the minimum drops. X is the type the
// exponentials are evaluated in.
template <typename T, typename X = T> struct SoftMin {
  T m, t;
  __device__ void init() {
    m = T(INFINITY);
    t = T(0);
  }
  __device__ void add(T r) {
    const T d = r - m;
    const T e = ex2<X>(-fabs(d));
    const bool lt = d < T(0);
    t = lt ? fma(t, e, T(1)) : t + e;
    m = lt ? r : m;
  }
  __device__ void merge(T m2, T t2) {
    const T mn = fmin(m, m2);
    const T a = (m == mn) ? t : t * ex2<X>(mn - m);
    const T b = (m2 == mn) ? t2 : t2 * ex2<X>(mn - m2);
    m = mn;
    t = a + b;
  }
  // a group of values whose minimum mg is known: one rescale of the state to
  // the new minimum, then one exponential per value (no compare or select)
  template <int G> __device__ void add_group(T mg, const T (&r)[G]) {
    const T mn = fmin(m, mg);
    if (m != mn) t *= ex2<X>(mn - m);
    m = mn;
    if constexpr (sizeof(X) < sizeof(T)) {
      // mixed form: the group's G terms (each <= 1) are summed in X and
      // converted once, which saves a conversion per term; the group sum's
      // rounding (a few X ulps) is of the order of the X exponentials' own
      X tg = X(0);
#pragma unroll
      for (int u = 0; u < G; ++u) tg += sm_exp2((X)(m - r[u]));
      t += (T)tg;
    } else {
#pragma unroll
      for (int u = 0; u < G; ++u) t += ex2<X>(m - r[u]);
    }
  }
  // the soft-min in the caller's units: (m - log2 t) / sc
  __device__ T value(T back) const { return (m - log2(t)) * back; }
};

template <typename T> __device__ __forceinline__ T shfl_xor_t(T v, int o) {
  return __shfl_xor(v, o, kWave);
}

// Columns per workgroup of the column passes: each thread keeps CB x SM
// partial states.
template <typename T, int SM> struct ColBlock {
  static constexpr int LIM = sizeof(T) == 4 ? 32 : 16;  // CB x SM states per thread
  static constexpr int CB = (LIM / SM >= 8) ? 8 : (LIM / SM >= 1 ? LIM / SM : 1);
};

// Column soft-min: one workgroup per CB columns, threads along the rows.
// op(j, res) with res[s] the soft-min of column j, slice s.
template <typename T, int SM, int CB, class Op>
__global__ void __launch_bounds__(kTB)
k_lse_col(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
          const T *__restrict__ G, long ldg, int S, T reg, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T shm[kWavesPerBlock][CB * SM], sht[kWavesPerBlock][CB * SM];
  const int j0 = blockIdx.x * CB;
  const T sc = T(kLog2e / (double)reg), back = T((double)reg * kLn2);
  T g[CB][SM];
  SoftMin<T> st[CB][SM];
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      st[c][s].init();
      g[c][s] = (s < S && j0 + c < N) ? G[(j0 + c) + (long)s * ldg] * sc : T(0);
    }
  for (int i = threadIdx.x; i < M; i += kTB) {
    if constexpr (CB == 1) {
      // one column: each F value is used once, so it is loaded where it is
      // used rather than held for all slices (no scratch at S = 32 in FP64)
      if (j0 >= N) continue;
      const T x = C[i + (long)j0 * ldc] * sc;
#pragma unroll
      for (int s = 0; s < SM; ++s)
        if (s < S) st[0][s].add((x - F[i + (long)s * ldf] * sc) - g[0][s]);
    } else {
      T f[SM];
#pragma unroll
      for (int s = 0; s < SM; ++s) f[s] = (s < S) ? F[i + (long)s * ldf] * sc : T(0);
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        if (j0 + c >= N) continue;
        const T x = C[i + (long)(j0 + c) * ldc] * sc;
#pragma unroll
        for (int s = 0; s < SM; ++s)
          if (s < S) st[c][s].add((x - f[s]) - g[c][s]);
      }
    }
  }
  // wavefront, then workgroup merge
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      if (s >= S) continue;
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1)
        st[c][s].merge(shfl_xor_t(st[c][s].m, o), shfl_xor_t(st[c][s].t, o));
      if (lane == 0) {
        shm[w][c * SM + s] = st[c][s].m;
        sht[w][c * SM + s] = st[c][s].t;
      }
    }
  __syncthreads();
  if (threadIdx.x < CB) {
    const int c = threadIdx.x;
    if (j0 + c >= N) return;
    // the results go back into this column's own LDS entries (read by this
    // thread only), which the op indexes at run time: a private array would
    // live in scratch
    T *res = &shm[0][c * SM];
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      if (s >= S) continue;
      SoftMin<T> a;
      a.m = shm[0][c * SM + s];
      a.t = sht[0][c * SM + s];
      for (int v = 1; v < kWavesPerBlock; ++v) a.merge(shm[v][c * SM + s], sht[v][c * SM + s]);
      res[s] = a.value(back);
    }
    op(j0 + c, (const T *)res);
  }
}

// Few columns (fewer than 512 workgroups at CB per workgroup): one column per
// workgroup, so that the pass still spreads over the CUs.
template <typename T, int SM> inline bool col_block_one(int N) {
  return ColBlock<T, SM>::CB > 1 && (N + ColBlock<T, SM>::CB - 1) / ColBlock<T, SM>::CB < 512;
}

template <typename T, class Op>
void lse_col(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf,
             const T *G, long ldg, int S, double reg, Op op, const LoopState *ls) {
  RWIG_DISPATCH_S(S, SM, {
    constexpr int CB = ColBlock<T, SM>::CB;
    if (col_block_one<T, SM>(N))
      hipLaunchKernelGGL((k_lse_col<T, SM, 1, Op>), dim3(N), dim3(kTB), 0, st, C, M, N, ldc, F,
                         ldf, G, ldg, S, (T)reg, op, ls);
    else
      hipLaunchKernelGGL((k_lse_col<T, SM, CB, Op>), dim3((N + CB - 1) / CB), dim3(kTB), 0, st,
                         C, M, N, ldc, F, ldf, G, ldg, S, (T)reg, op, ls);
  });
  check_launch();
}

// Row passes: a workgroup holds 64 rows (one per lane) and a chunk of the
// columns, split between its 4 wavefronts; the wavefronts' partial states are
// merged through LDS, one partial per (chunk, slice, row).
constexpr int kRowsPerBlock = kWave;

inline Chunking row_chunking(int M, int N) {
  const int rowblocks = (M + kRowsPerBlock - 1) / kRowsPerBlock;
  int nchunk = 2048 / (rowblocks > 0 ? rowblocks : 1);
  const int maxc = (N + 4 * kWavesPerBlock - 1) / (4 * kWavesPerBlock);  // >= 4 columns per wave
  if (nchunk > maxc) nchunk = maxc;
  if (nchunk < 1) nchunk = 1;
  const int cw = (N + nchunk - 1) / nchunk;
  nchunk = (N + cw - 1) / cw;
  return {nchunk, cw};
}

// Column tile of the row passes: the per-column values (G and, for the
// adjoint, cshift and y) of kColTile columns are staged in LDS as
// [column][slice], so that the wavefronts read them as broadcasts.
constexpr int kColTile = 64;
constexpr int kColUnroll = 4;  // columns per wavefront whose loads are in flight together

// per-column values of columns t0 .. t0 + nc into LDS as [column][slice],
// times sc; with src2, (src + src2) * sc
template <typename T, int SM>
__device__ __forceinline__ void stage_cols(const T *__restrict__ src, long ld, int t0, int nc,
                                           int S, T *dst, T sc,
                                           const T *__restrict__ src2 = nullptr, long ld2 = 0) {
  for (int k = threadIdx.x; k < nc * SM; k += kTB) {
    const int jj = k / SM, s = k % SM;
    T v = T(0);
    if (s < S) {
      v = src[(t0 + jj) + (long)s * ld];
      if (src2) v += src2[(t0 + jj) + (long)s * ld2];
    }
    dst[k] = v * sc;
  }
}

// Row soft-min partials over the column chunk blockIdx.y: pm = m, ps = t of
// the chunk, at [(c * S + s) * M + i]. Slices s >= S (up to SM) run on zeros
// and are not written, which keeps the inner loop free of per-slice branches.
// The row's own shift f is constant along the row: the state runs on
// x * sc - g and f is taken off the partial's minimum at the end. In FP32 the
// columns go in groups of RowGroup: first the group's minimum, then one
// exponential per value against it (SoftMin::add_group), about 4 VALU
// operations and one exponential per value and slice instead of 7 and one for
// the online update (row pass at the C3 shape 47 -> 37 us). FP64 keeps the
// online update: its exponential is a long instruction sequence, which the
// group's extra rescale and unrolled code make slower, not faster.
template <typename T, int SM, typename X = T> struct RowGroup {
  // columns per group: their values and one slice's terms stay in registers;
  // the rescale costs one exponential per group and slice. The mixed form
  // (FP64 values, FP32 exponentials) groups as FP32 does, with fewer columns
  // for its 8-byte registers (no scratch at any S).
  static constexpr int G =
      sizeof(X) == 4 ? (sizeof(T) == 4 ? (SM <= 8 ? 16 : 8) : (SM <= 4 ? 8 : (SM <= 8 ? 4 : 2)))
                     : kColUnroll;
  static constexpr bool two_step = sizeof(X) == 4;
};

template <typename T, int SM, typename X = T>
__global__ void __launch_bounds__(kTB)
k_lse_row_part(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F,
               long ldf, const T *__restrict__ G, long ldg, int S, T reg, int cw,
               T *__restrict__ pm, T *__restrict__ ps, const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T shm[kWavesPerBlock][kRowsPerBlock], sht[kWavesPerBlock][kRowsPerBlock];
  const int lane = threadIdx.x & (kWave - 1), w = wave_id();
  const int i = blockIdx.x * kRowsPerBlock + lane;
  const int c = blockIdx.y;
  const int j0 = c * cw, j1 = min(N, j0 + cw);
  const T sc = T(kLog2e / (double)reg);
  const bool row = i < M;
  SoftMin<T, X> st[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) st[s].init();
  constexpr int GR = RowGroup<T, SM, X>::G;
  __shared__ T gs[kColTile * SM];
  for (int t0 = j0; t0 < j1; t0 += kColTile) {
    const int nc = min(kColTile, j1 - t0);
    __syncthreads();
    stage_cols<T, SM>(G, ldg, t0, nc, S, gs, sc);
    __syncthreads();
    if (row) {
      int jj = w;
      // GR columns' loads in flight together, then their updates
      for (; jj + (GR - 1) * kWavesPerBlock < nc; jj += GR * kWavesPerBlock) {
        T x[GR];
#pragma unroll
        for (int u = 0; u < GR; ++u) x[u] = (C + (long)(t0 + jj + u * kWavesPerBlock) * ldc)[i];
        if constexpr (RowGroup<T, SM, X>::two_step) {
#pragma unroll
          for (int s = 0; s < SM; ++s) {
            T r[GR];
            T mg = T(INFINITY);
#pragma unroll
            for (int u = 0; u < GR; ++u) {
              r[u] = x[u] * sc - gs[(jj + u * kWavesPerBlock) * SM + s];
              mg = fmin(mg, r[u]);
            }
            st[s].add_group(mg, r);
          }
        } else {
#pragma unroll
          for (int u = 0; u < GR; ++u) {
            const T *g = gs + (jj + u * kWavesPerBlock) * SM;
            const T xs = x[u] * sc;
#pragma unroll
            for (int s = 0; s < SM; ++s) st[s].add(xs - g[s]);
          }
        }
      }
      // the last columns of the tile, one at a time
      for (; jj < nc; jj += kWavesPerBlock) {
        const T xs = (C + (long)(t0 + jj) * ldc)[i] * sc;
        const T *g = gs + jj * SM;
#pragma unroll
        for (int s = 0; s < SM; ++s) st[s].add(xs - g[s]);
      }
    }
  }
  // merge the wavefronts' states, one slice at a time through LDS (S is
  // uniform over the workgroup; a guard rather than a loop exit keeps the
  // loop fully unrolled, so st[] stays in registers)
#pragma unroll
  for (int s = 0; s < SM; ++s) {
    if (s < S) {
      __syncthreads();
      shm[w][lane] = st[s].m;
      sht[w][lane] = st[s].t;
      __syncthreads();
      if (w == 0 && row) {
        SoftMin<T, X> a;
        a.m = shm[0][lane];
        a.t = sht[0][lane];
        for (int v = 1; v < kWavesPerBlock; ++v) a.merge(shm[v][lane], sht[v][lane]);
        const long o = ((long)c * S + s) * M + i;
        pm[o] = a.m - F[i + (long)s * ldf] * sc;
        ps[o] = a.t;
      }
    }
  }
}

template <typename T, class Op>
__global__ void __launch_bounds__(kMergeTB)
k_lse_row_merge(const T *__restrict__ pm, const T *__restrict__ ps, int nchunk, long len, T reg,
                Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T shm[kMergeWaves][kWave], sht[kMergeWaves][kWave];
  const T back = T((double)reg * kLn2);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const long k = blockIdx.x * (long)kWave + lane;
  SoftMin<T> a;
  a.init();
  if (k < len)
    for (int c = w; c < nchunk; c += kMergeWaves) a.merge(pm[(long)c * len + k], ps[(long)c * len + k]);
  shm[w][lane] = a.m;
  sht[w][lane] = a.t;
  __syncthreads();
  if (w == 0 && k < len) {
    SoftMin<T> b;
    b.m = shm[0][lane];
    b.t = sht[0][lane];
    for (int u = 1; u < kMergeWaves; ++u) b.merge(shm[u][lane], sht[u][lane]);
    op(k, b.value(back));
  }
}

// Soft-min over columns j of R_s(i,j) for every row i and slice s:
// op(k = s*M + i, value).
template <typename T, typename X = T, class Op>
void lse_row(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf,
             const T *G, long ldg, int S, double reg, T *pm, T *ps, Op op,
             const LoopState *ls) {
  const Chunking ch = row_chunking(M, N);
  const dim3 grid((M + kRowsPerBlock - 1) / kRowsPerBlock, ch.nchunk);
  RWIG_DISPATCH_S(S, SM,
                  hipLaunchKernelGGL((k_lse_row_part<T, SM, X>), grid, dim3(kTB), 0, st, C, M, N,
                                     ldc, F, ldf, G, ldg, S, (T)reg, ch.cw, pm, ps, ls));
  check_launch();
  const long len = (long)S * M;
  hipLaunchKernelGGL((k_lse_row_merge<T, Op>), dim3(merge_grid(len)), dim3(kMergeTB), 0, st, pm,
                     ps, ch.nchunk, len, (T)reg, op, ls);
  check_launch();
}

inline size_t lse_row_part_size(int M, int N, int S) {
  return (size_t)row_chunking(M, N).nchunk * S * M;
}

// ---------------------------------------------------------------- soft-max weighted sums

// Column pass with row soft-max weights (the adjoint X of the row soft-min):
//   res[s] = sum_i exp(-(R_s(i,j) - rshift(i,s)) / reg) * x(i,s)
// where rshift is the row soft-min of R_s, so the weights of each row sum to 1.
// One workgroup per CB columns, threads along the rows, C read once.
template <typename T, int SM, int CB, class Op>
__global__ void __launch_bounds__(kTB)
k_wcol(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
       const T *__restrict__ G, long ldg, int S, T reg, const T *__restrict__ rshift,
       const T *__restrict__ x, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T sha[kWavesPerBlock][CB * SM];
  const int j0 = blockIdx.x * CB;
  const T sc = T(kLog2e / (double)reg);
  T g[CB][SM], acc[CB][SM];
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      acc[c][s] = T(0);
      g[c][s] = (s < S && j0 + c < N) ? G[(j0 + c) + (long)s * ldg] * sc : T(0);
    }
  for (int i = threadIdx.x; i < M; i += kTB) {
    if constexpr (CB == 1) {
      // one column: the row's values are loaded where they are used (no
      // scratch at S = 32 in FP64)
      if (j0 >= N) continue;
      const T cij = C[i + (long)j0 * ldc] * sc;
#pragma unroll
      for (int s = 0; s < SM; ++s)
        if (s < S) {
          const T fs = (F[i + (long)s * ldf] + rshift[i + (long)s * M]) * sc;
          acc[0][s] += sm_exp2(-((cij - fs) - g[0][s])) * x[i + (long)s * M];
        }
    } else {
      T fs[SM], xv[SM];  // fs = (F + rshift) * sc: the row's two shifts in one
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        fs[s] = (s < S) ? (F[i + (long)s * ldf] + rshift[i + (long)s * M]) * sc : T(0);
        xv[s] = (s < S) ? x[i + (long)s * M] : T(0);
      }
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        if (j0 + c >= N) continue;
        const T cij = C[i + (long)(j0 + c) * ldc] * sc;
#pragma unroll
        for (int s = 0; s < SM; ++s)
          if (s < S) acc[c][s] += sm_exp2(-((cij - fs[s]) - g[c][s])) * xv[s];
      }
    }
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      if (s >= S) continue;
      T v = acc[c][s];
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) v += shfl_xor_t(v, o);
      if (lane == 0) sha[w][c * SM + s] = v;
    }
  __syncthreads();
  if (threadIdx.x < CB) {
    const int c = threadIdx.x;
    if (j0 + c >= N) return;
    // results into this column's own LDS entries (read by this thread only),
    // which the op indexes at run time
    T *res = &sha[0][c * SM];
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      if (s >= S) continue;
      T v = sha[0][c * SM + s];
      for (int u = 1; u < kWavesPerBlock; ++u) v += sha[u][c * SM + s];
      res[s] = v;
    }
    op(j0 + c, (const T *)res);
  }
}

template <typename T, class Op>
void wcol(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf, const T *G,
          long ldg, int S, double reg, const T *rshift, const T *x, Op op,
          const LoopState *ls) {
  RWIG_DISPATCH_S(S, SM, {
    constexpr int CB = ColBlock<T, SM>::CB;
    if (col_block_one<T, SM>(N))
      hipLaunchKernelGGL((k_wcol<T, SM, 1, Op>), dim3(N), dim3(kTB), 0, st, C, M, N, ldc, F, ldf,
                         G, ldg, S, (T)reg, rshift, x, op, ls);
    else
      hipLaunchKernelGGL((k_wcol<T, SM, CB, Op>), dim3((N + CB - 1) / CB), dim3(kTB), 0, st, C,
                         M, N, ldc, F, ldf, G, ldg, S, (T)reg, rshift, x, op, ls);
  });
  check_launch();
}

// Row pass with column soft-max weights (the adjoint W of the column soft-min):
//   part = sum_{j in chunk} exp(-(R_s(i,j) - cshift(j,s) + rowsh(i,s)) / reg) * y(j,s)
// rowsh (optional, M x S) scales row i by exp(-rowsh/reg) inside the
// exponent, so rows whose result is tiny can be computed relative to their
// own scale without underflow. Layout as k_lse_row_part.
template <typename T, int SM, typename X = T>
__global__ void __launch_bounds__(kTB)
k_wrow_part(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
            const T *__restrict__ G, long ldg, int S, T reg, const T *__restrict__ cshift,
            const T *__restrict__ y, const T *__restrict__ rowsh, int cw, T *__restrict__ part,
            const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ T sha[kWavesPerBlock][kRowsPerBlock];
  const int lane = threadIdx.x & (kWave - 1), w = wave_id();
  const int i = blockIdx.x * kRowsPerBlock + lane;
  const int c = blockIdx.y;
  const int j0 = c * cw, j1 = min(N, j0 + cw);
  const T sc = T(kLog2e / (double)reg);
  const bool row = i < M;
  T f[SM], acc[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) {
    acc[s] = T(0);
    f[s] = T(0);
    if (row && s < S) {
      f[s] = F[i + (long)s * ldf];
      if (rowsh) f[s] -= rowsh[i + (long)s * M];
      f[s] *= sc;
    }
  }
  // gh = (G + cshift) * sc: the column's two shifts in one
  __shared__ T gh[kColTile * SM], ys[kColTile * SM];
  for (int t0 = j0; t0 < j1; t0 += kColTile) {
    const int nc = min(kColTile, j1 - t0);
    __syncthreads();
    stage_cols<T, SM>(G, ldg, t0, nc, S, gh, sc, cshift, N);
    stage_cols<T, SM>(y, N, t0, nc, S, ys, T(1));
    __syncthreads();
    if (row) {
      int jj = w;
      for (; jj + (kColUnroll - 1) * kWavesPerBlock < nc; jj += kColUnroll * kWavesPerBlock) {
        T x[kColUnroll];
#pragma unroll
        for (int u = 0; u < kColUnroll; ++u)
          x[u] = (C + (long)(t0 + jj + u * kWavesPerBlock) * ldc)[i];
#pragma unroll
        for (int u = 0; u < kColUnroll; ++u) {
          const int q = (jj + u * kWavesPerBlock) * SM;
          const T xs = x[u] * sc;
#pragma unroll
          for (int s = 0; s < SM; ++s)
            acc[s] += ex2_wide<X>(-((xs - f[s]) - gh[q + s])) * ys[q + s];
        }
      }
      for (; jj < nc; jj += kWavesPerBlock) {
        const T xs = (C + (long)(t0 + jj) * ldc)[i] * sc;
        const T *g = gh + jj * SM, *yv = ys + jj * SM;
#pragma unroll
        for (int s = 0; s < SM; ++s) acc[s] += ex2_wide<X>(-((xs - f[s]) - g[s])) * yv[s];
      }
    }
  }
#pragma unroll
  for (int s = 0; s < SM; ++s) {
    if (s < S) {
      __syncthreads();
      sha[w][lane] = acc[s];
      __syncthreads();
      if (w == 0 && row) {
        T v = sha[0][lane];
        for (int u = 1; u < kWavesPerBlock; ++u) v += sha[u][lane];
        part[((long)c * S + s) * M + i] = v;
      }
    }
  }
}

inline size_t wrow_part_size(int M, int N, int S) {
  return (size_t)row_chunking(M, N).nchunk * S * M;
}

template <typename T, typename X = T, class Op>
void wrow(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf, const T *G,
          long ldg, int S, double reg, const T *cshift, const T *y, T *part, Op op,
          const LoopState *ls, const T *rowsh = nullptr) {
  const Chunking ch = row_chunking(M, N);
  const dim3 grid((M + kRowsPerBlock - 1) / kRowsPerBlock, ch.nchunk);
  RWIG_DISPATCH_S(S, SM,
                  hipLaunchKernelGGL((k_wrow_part<T, SM, X>), grid, dim3(kTB), 0, st, C, M, N, ldc,
                                     F, ldf, G, ldg, S, (T)reg, cshift, y, rowsh, ch.cw, part,
                                     ls));
  check_launch();
  const long len = (long)S * M;
  hipLaunchKernelGGL((k_reduce_part<T, Op>), dim3(merge_grid(len)), dim3(kMergeTB), 0, st, part,
                     ch.nchunk, len, op, ls);
  check_launch();
}

// ---------------------------------------------------------------- transpose

// dst (cols x rows, leading dimension ldd) = src^T (src rows x cols, ld lds),
// through 32 x 32 LDS tiles (reads and writes coalesced).
template <typename T>
__global__ void __launch_bounds__(256)
k_transpose(const T *__restrict__ src, int rows, int cols, long lds, T *__restrict__ dst,
            long ldd) {
  __shared__ T tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + tx, c = c0 + k;
    if (r < rows && c < cols) tile[k][tx] = src[r + (long)c * lds];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + tx, r = r0 + k;
    if (r < rows && c < cols) dst[c + (long)r * ldd] = tile[tx][k];
  }
}

template <typename T>
void transpose(hipStream_t st, const T *src, int rows, int cols, long lds, T *dst, long ldd) {
  const dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, st, src, rows, cols, lds, dst, ldd);
  check_launch();
}

// ---------------------------------------------------------------- column passes over C^T
//
// The column passes as row passes over CT = C^T (N x M, leading dimension
// ldct), for callers that keep the transpose: the row kernels read CT
// coalesced along its rows with the per-column values staged in LDS, where the
// column kernels above reduce across all the threads of a column.
// R_s(i,j) = (C(i,j) - G(j,s)) - F(i,s) here (the row kernels' order of the
// two subtractions), equal to the column kernels' up to rounding.

// Column form of a row pass over CT: v holds the merged value of slice s,
// column j at [s * N + j] (k_lse_row_merge writes it, with coalesced loads of
// the chunk partials), and op(j, res) gets column j's S values.
template <typename T, int SM, class Op>
__global__ void __launch_bounds__(kWave)
k_cols_apply(const T *__restrict__ v, int N, int S, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  const int j = blockIdx.x * kWave + threadIdx.x;
  if (j >= N) return;
  // the op indexes the values at run time: staged in LDS rather than in a
  // private array (which would live in scratch at SM = 32)
  __shared__ T res[kWave][SM];
  T *r = res[threadIdx.x];
#pragma unroll
  for (int s = 0; s < SM; ++s) r[s] = s < S ? v[(long)s * N + j] : T(0);
  op(j, (const T *)r);
}

template <typename T, int SM, class Op>
__global__ void __launch_bounds__(kMergeTB)
k_sum_cols(const T *__restrict__ part, int nchunk, int N, int S, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  constexpr int JB = kWave / SM;
  __shared__ T sh[kMergeWaves][kWave];
  __shared__ T r[JB][SM];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int jj = lane / SM, s = lane % SM;
  const int j = blockIdx.x * JB + jj;
  const bool valid = j < N && s < S;
  const long k = (long)s * N + j, len = (long)S * N;
  T v = T(0);
  if (valid)
    for (int c = w; c < nchunk; c += kMergeWaves) v += part[c * len + k];
  sh[w][lane] = v;
  __syncthreads();
  if (w == 0 && valid) {
    T t = sh[0][lane];
    for (int u = 1; u < kMergeWaves; ++u) t += sh[u][lane];
    r[jj][s] = t;
  }
  __syncthreads();
  if (threadIdx.x < JB) {
    const int jc = blockIdx.x * JB + threadIdx.x;
    if (jc < N) op(jc, r[threadIdx.x]);
  }
}

// the chunk partials of a row pass over CT, then S x N for the merged values
inline size_t col_t_part_size(int M, int N, int S) {
  return ((size_t)row_chunking(N, M).nchunk + 1) * S * N;
}

// column soft-min of C (as lse_col) from CT; pm, ps: col_t_part_size each
template <typename T, typename X = T, class Op>
void lse_col_t(hipStream_t st, const T *CT, int M, int N, long ldct, const T *F, long ldf,
               const T *G, long ldg, int S, double reg, T *pm, T *ps, Op op,
               const LoopState *ls) {
  const Chunking ch = row_chunking(N, M);
  const dim3 grid((N + kRowsPerBlock - 1) / kRowsPerBlock, ch.nchunk);
  RWIG_DISPATCH_S(S, SM, {
    hipLaunchKernelGGL((k_lse_row_part<T, SM, X>), grid, dim3(kTB), 0, st, CT, N, M, ldct, G,
                       ldg, F, ldf, S, (T)reg, ch.cw, pm, ps, ls);
    check_launch();
    // merged values into the room after the partials, then the column form
    T *mv = pm + (size_t)ch.nchunk * S * N;
    const long len = (long)S * N;
    auto put = [=] __device__(long k, T r) { mv[k] = r; };
    hipLaunchKernelGGL((k_lse_row_merge<T, decltype(put)>), dim3(merge_grid(len)), dim3(kMergeTB),
                       0, st, pm, ps, ch.nchunk, len, (T)reg, put, ls);
    check_launch();
    // one wavefront per workgroup: N / 64 workgroups spread over the CUs
    hipLaunchKernelGGL((k_cols_apply<T, SM, Op>), dim3((N + kWave - 1) / kWave), dim3(kWave), 0,
                       st, mv, N, S, op, ls);
  });
  check_launch();
}

// column weighted sums of C (as wcol) from CT; part: col_t_part_size
template <typename T, typename X = T, class Op>
void wcol_t(hipStream_t st, const T *CT, int M, int N, long ldct, const T *F, long ldf,
            const T *G, long ldg, int S, double reg, const T *rshift, const T *x, T *part, Op op,
            const LoopState *ls) {
  const Chunking ch = row_chunking(N, M);
  const dim3 grid((N + kRowsPerBlock - 1) / kRowsPerBlock, ch.nchunk);
  RWIG_DISPATCH_S(S, SM, {
    hipLaunchKernelGGL((k_wrow_part<T, SM, X>), grid, dim3(kTB), 0, st, CT, N, M, ldct, G, ldg,
                       F, ldf, S, (T)reg, rshift, x, (const T *)nullptr, ch.cw, part, ls);
    check_launch();
    hipLaunchKernelGGL((k_sum_cols<T, SM, Op>), dim3((N + kWave / SM - 1) / (kWave / SM)),
                       dim3(kMergeTB), 0, st, part, ch.nchunk, N, S, op, ls);
  });
  check_launch();
}

// ---------------------------------------------------------------- single-block reductions

// Two deterministic sums in one block of 1024 threads; fin(a, b) runs on
// thread 0 and may update the LoopState.
template <class F1, class F2, class Fin>
__global__ void __launch_bounds__(1024)
k_reduce2(long n1, F1 f1, long n2, F2 f2, Fin fin, const LoopState *gate) {
  __shared__ double sh[1024 / kWave];
  if (gate && !gate->active) return;
  double a = 0.0;
  for (long k = threadIdx.x; k < n1; k += blockDim.x) a += f1(k);
  a = block_sum(a, sh);
  double b = 0.0;
  for (long k = threadIdx.x; k < n2; k += blockDim.x) b += f2(k);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) fin(a, b);
}

template <class F1, class F2, class Fin>
void reduce2(hipStream_t st, long n1, F1 f1, long n2, F2 f2, Fin fin,
             const LoopState *gate = nullptr) {
  hipLaunchKernelGGL((k_reduce2<F1, F2, Fin>), dim3(1), dim3(1024), 0, st, n1, f1, n2, f2, fin,
                     gate);
  check_launch();
}

// Deterministic sum of f(k) over a large range: per-block partials in a fixed
// grid, then one block adds them in order. out receives the sum.
template <class F>
__global__ void __launch_bounds__(kTB) k_sum_part(long n, F f, double *part) {
  __shared__ double sh[kTB / kWave];
  double a = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    a += f(k);
  a = block_sum(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

constexpr int kSumBlocks = 1024;

template <class F>
void big_sum(hipStream_t st, long n, F f, double *part /* kSumBlocks */, double *out) {
  hipLaunchKernelGGL(k_sum_part<F>, dim3(kSumBlocks), dim3(kTB), 0, st, n, f, part);
  check_launch();
  reduce2(
      st, (long)kSumBlocks, [=] __device__(long k) { return part[k]; }, 0L,
      [=] __device__(long) { return 0.0; }, [=] __device__(double a, double) { *out = a; });
}

// Jacobi eigenvalue sweeps on the s x s symmetric matrix G (in place); returns
// the largest eigenvalue. Same rotation sequence as the oracle's
// spectral_norm(); the sweeps stop once the off-diagonal mass is below 1e-32 of
// the diagonal's (the eigenvalues are then exact to rounding, the error being
// quadratic in the off-diagonal part), where the oracle sweeps on to an exact
// zero -- typically 100 sweeps that change nothing but the time.
__device__ inline double jacobi_lmax(double *G, int s) {
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < s; ++p) {
      dia += G[p + p * s] * G[p + p * s];
      for (int q = p + 1; q < s; ++q) off += G[p + q * s] * G[p + q * s];
    }
    if (off <= 1e-32 * dia) break;
    for (int p = 0; p < s; ++p)
      for (int q = p + 1; q < s; ++q) {
        const double apq = G[p + q * s];
        if (apq == 0.0) continue;
        const double app = G[p + p * s], aqq = G[q + q * s];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < s; ++k) {
          const double gkp = G[k + p * s], gkq = G[k + q * s];
          G[k + p * s] = c * gkp - sn * gkq;
          G[k + q * s] = sn * gkp + c * gkq;
        }
        for (int k = 0; k < s; ++k) {
          const double gpk = G[p + k * s], gqk = G[q + k * s];
          G[p + k * s] = c * gpk - sn * gqk;
          G[q + k * s] = sn * gpk + c * gqk;
        }
      }
  }
  double lmax = 0.0;
  for (int p = 0; p < s; ++p) lmax = fmax(lmax, G[p + p * s]);
  return lmax;
}

// Spectral norm of the M x S matrix e(i, s) (Armadillo norm(X, 2),
// src/barycenter_impl_log.cpp:252): the Gram matrix X^T X by partial sums over
// row blocks (k_gram_part, kGramBlocks blocks, rows staged through LDS), then
// one block adds the partials in block order and runs Jacobi on thread 0;
// fin(norm) runs there.
constexpr int kGramBlocks = 64;
constexpr int kGramRows = 256;

inline size_t gram_part_size(int S) { return (size_t)kGramBlocks * S * (S + 1) / 2; }

__device__ inline void gram_pair(int k, int S, int &p, int &q) {
  p = 0;
  while (k >= S - p) {
    k -= S - p;
    ++p;
  }
  q = p + k;
}

template <class E>
__global__ void __launch_bounds__(256)
k_gram_part(long M, int S, E e, double *__restrict__ part, const LoopState *gate) {
  if (gate && !gate->active) return;
  __shared__ double tile[kGramRows * 32];
  const int np = S * (S + 1) / 2;
  constexpr int PR = (32 * 33 / 2 + 255) / 256;  // pairs per thread, S <= 32
  double acc[PR];
  int pp[PR], qq[PR];
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    acc[r] = 0.0;
    pp[r] = qq[r] = 0;
    const int k = threadIdx.x + 256 * r;
    if (k < np) gram_pair(k, S, pp[r], qq[r]);
  }
  const long rows = (M + gridDim.x - 1) / gridDim.x;
  const long i0 = blockIdx.x * rows, i1 = min(M, i0 + rows);
  for (long t0 = i0; t0 < i1; t0 += kGramRows) {
    const int nr = (int)min((long)kGramRows, i1 - t0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nr * S; idx += blockDim.x) {
      const int rr = idx % nr, c = idx / nr;
      tile[c * kGramRows + rr] = e(t0 + rr, c);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PR; ++r) {
      if (threadIdx.x + 256 * r >= np) continue;
      const double *xp = tile + pp[r] * kGramRows, *xq = tile + qq[r] * kGramRows;
      double a = acc[r];
      for (int rr = 0; rr < nr; ++rr) a += xp[rr] * xq[rr];
      acc[r] = a;
    }
  }
#pragma unroll
  for (int r = 0; r < PR; ++r) {
    const int k = threadIdx.x + 256 * r;
    if (k < np) part[(long)blockIdx.x * np + k] = acc[r];
  }
}

// jacobi_lmax on a matrix of at most 8 x 8 held in registers (the loops are
// unrolled; padding rows and columns are zero and skipped): the same rotations
// without the LDS round trip of every element access.
__device__ inline double jacobi_lmax8(const double *Gs, int s) {
  constexpr int SM = 8;
  double G[SM][SM];
#pragma unroll
  for (int p = 0; p < SM; ++p)
#pragma unroll
    for (int q = 0; q < SM; ++q) G[p][q] = (p < s && q < s) ? Gs[p + q * s] : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
#pragma unroll
    for (int p = 0; p < SM; ++p) {
      dia += G[p][p] * G[p][p];
#pragma unroll
      for (int q = p + 1; q < SM; ++q) off += G[p][q] * G[p][q];
    }
    if (off <= 1e-32 * dia) break;
#pragma unroll
    for (int p = 0; p < SM; ++p)
#pragma unroll
      for (int q = p + 1; q < SM; ++q) {
        const double apq = G[p][q];
        if (apq == 0.0) continue;
        const double app = G[p][p], aqq = G[q][q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
#pragma unroll
        for (int k = 0; k < SM; ++k) {
          const double gkp = G[k][p], gkq = G[k][q];
          G[k][p] = c * gkp - sn * gkq;
          G[k][q] = sn * gkp + c * gkq;
        }
#pragma unroll
        for (int k = 0; k < SM; ++k) {
          const double gpk = G[p][k], gqk = G[q][k];
          G[p][k] = c * gpk - sn * gqk;
          G[q][k] = sn * gpk + c * gqk;
        }
      }
  }
  double lmax = 0.0;
#pragma unroll
  for (int p = 0; p < SM; ++p) lmax = fmax(lmax, G[p][p]);
  return lmax;
}

template <class Fin>
__global__ void __launch_bounds__(256)
k_gram_fin(int S, const double *__restrict__ part, Fin fin, const LoopState *gate) {
  if (gate && !gate->active) return;
  __shared__ double G[32 * 32];
  const int np = S * (S + 1) / 2;
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    double a = 0.0;
    for (int b = 0; b < kGramBlocks; ++b) a += part[(long)b * np + k];
    int p, q;
    gram_pair(k, S, p, q);
    G[p + q * S] = a;
    G[q + p * S] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // With a non-positive tolerance the stopping test err >= tol of a
    // non-final iteration depends only on err being a number: the eigenvalue
    // sweeps run for the last iteration (whose err is reported) and whenever
    // the tolerance can stop the loop; otherwise sqrt(trace) (an upper bound
    // of the norm, NaN exactly when the matrix has one) decides the test.
    const bool exact =
        !gate || gate->tol > 0.0 || gate->iter + 1 >= gate->maxiter;
    double v;
    if (exact) {
      v = sqrt(S <= 8 ? jacobi_lmax8(G, S) : jacobi_lmax(G, S));
    } else {
      double tr = 0.0;
      for (int p = 0; p < S; ++p) tr += G[p + p * S];
      v = sqrt(tr);
    }
    fin(v);
  }
}

// part: gram_part_size(S) doubles of device memory
template <class E, class Fin>
void spectral(hipStream_t st, long M, int S, E e, Fin fin, const LoopState *gate, double *part) {
  RWIG_REQUIRE(S >= 1 && S <= 32, "spectral norm of at most 32 columns");
  hipLaunchKernelGGL((k_gram_part<E>), dim3(kGramBlocks), dim3(256), 0, st, M, S, e, part, gate);
  check_launch();
  hipLaunchKernelGGL((k_gram_fin<Fin>), dim3(1), dim3(256), 0, st, S, part, fin, gate);
  check_launch();
}

}  // namespace rwig
