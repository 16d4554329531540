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

// out index k = s*M + i: op(k, sum over chunks of part)
template <typename T, class Op>
__global__ void __launch_bounds__(kTB)
k_reduce_part(const T *__restrict__ part, int nchunk, long len, Op op,
              const LoopState *ls) {
  if (ls && !ls->active) return;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < len;
       k += (long)gridDim.x * blockDim.x) {
    T s = T(0);
    for (int c = 0; c < nchunk; ++c) s += part[(long)c * len + k];
    op(k, s);
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
  hipLaunchKernelGGL((k_reduce_part<double, Op>), dim3(grid_for(len, kTB, 4096)), dim3(kTB), 0,
                     st, part, ch.nchunk, len, op, ls);
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
#pragma unroll
  for (int t = 0; t < SM; ++t)
    if (t < S) acc[t] = wave_sum(acc[t]);
  if (lane == 0) op(j, acc);
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

// Soft-min over rows i of R_s(i,j) = (C(i,j) - F(i,s)) - G(j,s), one
// wavefront per column j, for every slice s:
//   res[s] = min_i R - reg * log(sum_i exp(-(R - min_i R) / reg))
// (the two-pass form of src/sinkhorn_impl_log.cpp:48-50).
template <typename T, int SM, class Op>
__global__ void __launch_bounds__(kTB)
k_lse_col(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
          const T *__restrict__ G, long ldg, int S, T reg, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (j >= N) return;
  const T *col = C + (long)j * ldc;
  T res[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) {
    if (s < S) {
      const T *f = F + (long)s * ldf;
      const T gj = G[j + (long)s * ldg];
      T mn = T(INFINITY);
      for (int i = lane; i < M; i += kWave) mn = fmin(mn, (col[i] - f[i]) - gj);
      mn = wave_min(mn);
      T sm = T(0);
      for (int i = lane; i < M; i += kWave) sm += exp(-(((col[i] - f[i]) - gj) - mn) / reg);
      sm = wave_sum(sm);
      res[s] = mn - reg * log(sm);
    }
  }
  if (lane == 0) op(j, res);
}

template <typename T, class Op>
void lse_col(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf,
             const T *G, long ldg, int S, double reg, Op op, const LoopState *ls) {
  const dim3 grid((N + kWavesPerBlock - 1) / kWavesPerBlock);
  RWIG_DISPATCH_S(S, SM, hipLaunchKernelGGL((k_lse_col<T, SM, Op>), grid, dim3(kTB), 0, st, C,
                                            M, N, ldc, F, ldf, G, ldg, S, (T)reg, op, ls));
  check_launch();
}

// Row soft-min partials over the column chunk c: min and the shifted sum.
template <typename T>
__global__ void __launch_bounds__(kTB)
k_lse_row_part(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F,
               long ldf, const T *__restrict__ G, long ldg, int S, T reg, int cw,
               T *__restrict__ pm, T *__restrict__ ps, const LoopState *ls) {
  if (ls && !ls->active) return;
  const int i = blockIdx.x * kTB + threadIdx.x;
  const int c = blockIdx.y, s = blockIdx.z;
  if (i >= M) return;
  const int j0 = c * cw, j1 = min(N, j0 + cw);
  const T fi = F[i + (long)s * ldf];
  const T *g = G + (long)s * ldg;
  T mn = T(INFINITY);
  for (int j = j0; j < j1; ++j) mn = fmin(mn, (C[i + j * ldc] - fi) - g[j]);
  T sm = T(0);
  for (int j = j0; j < j1; ++j) sm += exp(-(((C[i + j * ldc] - fi) - g[j]) - mn) / reg);
  const long o = ((long)c * S + s) * M + i;
  pm[o] = mn;
  ps[o] = sm;
}

template <typename T, class Op>
__global__ void __launch_bounds__(kTB)
k_lse_row_merge(const T *__restrict__ pm, const T *__restrict__ ps, int nchunk, long len, T reg,
                Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < len;
       k += (long)gridDim.x * blockDim.x) {
    T m = pm[k];
    for (int c = 1; c < nchunk; ++c) m = fmin(m, pm[(long)c * len + k]);
    T t = T(0);
    for (int c = 0; c < nchunk; ++c) {
      const T mc = pm[(long)c * len + k];
      t += (mc == m) ? ps[(long)c * len + k] : ps[(long)c * len + k] * exp(-(mc - m) / reg);
    }
    op(k, m - reg * log(t));
  }
}

// Soft-min over columns j of R_s(i,j) for every row i and slice s:
// op(k = s*M + i, value).
template <typename T, class Op>
void lse_row(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf,
             const T *G, long ldg, int S, double reg, T *pm, T *ps, Op op,
             const LoopState *ls) {
  const Chunking ch = chunking(M, N);
  dim3 grid((M + kTB - 1) / kTB, ch.nchunk, S);
  hipLaunchKernelGGL(k_lse_row_part<T>, grid, dim3(kTB), 0, st, C, M, N, ldc, F, ldf, G, ldg, S,
                     (T)reg, ch.cw, pm, ps, ls);
  check_launch();
  const long len = (long)S * M;
  hipLaunchKernelGGL((k_lse_row_merge<T, Op>), dim3(grid_for(len, kTB, 4096)), dim3(kTB), 0, st,
                     pm, ps, ch.nchunk, len, (T)reg, op, ls);
  check_launch();
}

inline size_t lse_row_part_size(int M, int N, int S) {
  return (size_t)chunking(M, N).nchunk * S * M;
}

// ---------------------------------------------------------------- soft-max weighted sums

// Column pass with row soft-max weights (the adjoint X of the row soft-min):
//   res[s] = sum_i exp(-(R_s(i,j) - rshift(i,s)) / reg) * x(i,s)
// where rshift is the row soft-min of R_s, so the weights of each row sum to 1.
template <typename T, int SM, class Op>
__global__ void __launch_bounds__(kTB)
k_wcol(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
       const T *__restrict__ G, long ldg, int S, T reg, const T *__restrict__ rshift,
       const T *__restrict__ x, Op op, const LoopState *ls) {
  if (ls && !ls->active) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int j = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (j >= N) return;
  const T *col = C + (long)j * ldc;
  T res[SM];
#pragma unroll
  for (int s = 0; s < SM; ++s) {
    if (s < S) {
      const T *f = F + (long)s * ldf;
      const T *sh = rshift + (long)s * M;
      const T *xs = x + (long)s * M;
      const T gj = G[j + (long)s * ldg];
      T acc = T(0);
      for (int i = lane; i < M; i += kWave)
        acc += exp(-(((col[i] - f[i]) - gj) - sh[i]) / reg) * xs[i];
      res[s] = wave_sum(acc);
    }
  }
  if (lane == 0) op(j, res);
}

template <typename T, class Op>
void wcol(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf, const T *G,
          long ldg, int S, double reg, const T *rshift, const T *x, Op op,
          const LoopState *ls) {
  const dim3 grid((N + kWavesPerBlock - 1) / kWavesPerBlock);
  RWIG_DISPATCH_S(S, SM, hipLaunchKernelGGL((k_wcol<T, SM, Op>), grid, dim3(kTB), 0, st, C, M,
                                            N, ldc, F, ldf, G, ldg, S, (T)reg, rshift, x, op,
                                            ls));
  check_launch();
}

// Row pass with column soft-max weights (the adjoint W of the column soft-min):
//   part = sum_{j in chunk} exp(-(R_s(i,j) - cshift(j,s) + rowsh(i,s)) / reg) * y(j,s)
// rowsh (optional, M x S) scales row i by exp(-rowsh/reg) inside the
// exponent, so rows whose result is tiny can be computed relative to their
// own scale without underflow.
template <typename T>
__global__ void __launch_bounds__(kTB)
k_wrow_part(const T *__restrict__ C, int M, int N, long ldc, const T *__restrict__ F, long ldf,
            const T *__restrict__ G, long ldg, int S, T reg, const T *__restrict__ cshift,
            const T *__restrict__ y, const T *__restrict__ rowsh, int cw, T *__restrict__ part,
            const LoopState *ls) {
  if (ls && !ls->active) return;
  const int i = blockIdx.x * kTB + threadIdx.x;
  const int c = blockIdx.y, s = blockIdx.z;
  if (i >= M) return;
  const int j0 = c * cw, j1 = min(N, j0 + cw);
  const T fi = rowsh ? F[i + (long)s * ldf] - rowsh[i + (long)s * M] : F[i + (long)s * ldf];
  const T *g = G + (long)s * ldg;
  const T *sh = cshift + (long)s * N;
  const T *ys = y + (long)s * N;
  T acc = T(0);
  for (int j = j0; j < j1; ++j)
    acc += exp(-(((C[i + j * ldc] - fi) - g[j]) - sh[j]) / reg) * ys[j];
  part[((long)c * S + s) * M + i] = acc;
}

template <typename T, class Op>
void wrow(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf, const T *G,
          long ldg, int S, double reg, const T *cshift, const T *y, T *part, Op op,
          const LoopState *ls, const T *rowsh = nullptr) {
  const Chunking ch = chunking(M, N);
  dim3 grid((M + kTB - 1) / kTB, ch.nchunk, S);
  hipLaunchKernelGGL(k_wrow_part<T>, grid, dim3(kTB), 0, st, C, M, N, ldc, F, ldf, G, ldg, S,
                     (T)reg, cshift, y, rowsh, ch.cw, part, ls);
  check_launch();
  const long len = (long)S * M;
  hipLaunchKernelGGL((k_reduce_part<T, Op>), dim3(grid_for(len, kTB, 4096)), dim3(kTB), 0, st,
                     part, ch.nchunk, len, op, ls);
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
// spectral_norm().
__device__ inline double jacobi_lmax(double *G, int s) {
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < s; ++p)
      for (int q = p + 1; q < s; ++q) off += G[p + q * s] * G[p + q * s];
    if (off == 0.0) break;
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
// src/barycenter_impl_log.cpp:252): Gram matrix by block sums, then Jacobi
// on thread 0. fin(norm) runs on thread 0.
template <class E, class Fin>
__global__ void __launch_bounds__(1024)
k_spectral(long M, int S, E e, Fin fin, const LoopState *gate) {
  __shared__ double G[32 * 32];
  __shared__ double sh[1024 / kWave];
  if (gate && !gate->active) return;
  for (int p = 0; p < S; ++p)
    for (int q = p; q < S; ++q) {
      double acc = 0.0;
      for (long i = threadIdx.x; i < M; i += blockDim.x) acc += e(i, p) * e(i, q);
      acc = block_sum(acc, sh);
      if (threadIdx.x == 0) {
        G[p + q * S] = acc;
        G[q + p * S] = acc;
      }
    }
  __syncthreads();
  if (threadIdx.x == 0) fin(sqrt(jacobi_lmax(G, S)));
}

template <class E, class Fin>
void spectral(hipStream_t st, long M, int S, E e, Fin fin, const LoopState *gate) {
  hipLaunchKernelGGL((k_spectral<E, Fin>), dim3(1), dim3(1024), 0, st, M, S, e, fin, gate);
  check_launch();
}

}  // namespace rwig
