// Log-stabilised barycenter with its unrolled adjoint, in FP64 or FP32, with
// checkpointed recomputation of the iterates.
//
// Forward and adjoint follow Barycenter::_fwd_log / _bwd_log
// (src/barycenter_impl_log.cpp:189-382). The iterates are the log-domain form
// (F = reg log U, G = reg log V) of the parallel barycenter
// (src/barycenter_impl_parallel.cpp:60-151), which is what lets the FP32 form
// run at regularisations where exp(-C/reg) underflows FP32.
//
// Precision: FP64; FP32 (C, iterates and passes in FP32); or the mixed form,
// FP64 C, iterates and exponent arguments with the exponentials of the passes
// in FP32. FP32 storage of C alone perturbs every exponent by up to
// 2^-24 max|C| / reg (1.5e-5 at the 64 x 64 grid's reg 0.004), which bounds
// the FP32 gradient's accuracy near 1e-5; the mixed form forms each argument
// in FP64 and rounds only that, at about the FP32 form's cost (FP64 add and
// FMA run at the FP32 rate on CDNA4; the FP64 exponential is the slow part).
//
// Memory: the reference keeps all L+1 iterates for the adjoint
// (src/barycenter_impl_log.cpp:197-209). Here the forward stores
// (F, G, row soft-min) only at iterations 0, ck, 2 ck, ...; the adjoint walks
// back one segment at a time, first recomputing the segment's iterates from
// its checkpoint with the forward's own kernels (so bit-identical to the
// forward), then stepping back through it. Memory falls from L+1 to
// L/ck + ck + 2 slices of (M + N) x S.
//
// Seed of the adjoint at l = L (the gradient of the loss with respect to
// logb_L):
//   seed 0: 2 (b - b_ext) .* b with the normalised b, the reference's
//           (src/barycenter_impl_log.cpp:323-325);
//   seed 1: 2 (b_L - b_ext) .* b_L with the unnormalised b_L = exp(logb_L):
//           the exact gradient of ||b_L - b_ext||^2 for fixed L;
//   seed 2: b .* (g - <g, b>) with g = 2 (b - b_ext), b normalised: the exact
//           gradient of ||b - b_ext||^2 through the normalisation.
// The loss reported is the one the seed differentiates (seeds 0 and 2 report
// the reference's ||b - b_ext||^2).
#include <algorithm>

#include <type_traits>

#include "gemm.hpp"
#include "kernels.hpp"

namespace rwig {

namespace {

// ---------------------------------------------------------------- FP64 soft-min as a product
//
// With FP64 iterates and a cost whose range (max C - min C) / reg is at most
// kScaledRange, the row soft-min of R_s(i,j) = C(i,j) - F(i,s) - G(j,s) is
//   -F(i,s) - reg (log(sum_j Kt(i,j) E(j,s)) + m_s) + cmin,
// Kt = exp(-(C - cmin) / reg), E = exp(G / reg - m_s), m_s = max_j G(j,s) / reg
// (the column soft-min likewise, with F and G exchanged): one FP64 MFMA
// product of the M x N scaled kernel with the S slices per pass, in place of
// an exponential per entry and slice. Every factor is at most 1, and in every
// row the term of j = argmax_j G(j,s) is at least exp(-kScaledRange), so no
// sum underflows and the terms lost below FP64's range (exp(-744)) are under
// exp(-144) of it; the value is the log-sum-exp form's up to rounding.
constexpr double kScaledRange = 600.0;

// part[2 b], part[2 b + 1] = min and max of block b's share of x[0 .. n);
// with part == out and one block, the min and max of part's 2 nb entries
__global__ void __launch_bounds__(kTB) k_minmax(const double *x, long n, double *out) {
  __shared__ double smin[kTB / kWave], smax[kTB / kWave];
  double lo = INFINITY, hi = -INFINITY;
  for (long k = blockIdx.x * (long)kTB + threadIdx.x; k < n; k += (long)gridDim.x * kTB) {
    lo = fmin(lo, x[k]);
    hi = fmax(hi, x[k]);
  }
  for (int o = kWave / 2; o > 0; o /= 2) {
    lo = fmin(lo, __shfl_xor(lo, o));
    hi = fmax(hi, __shfl_xor(hi, o));
  }
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (lane == 0) {
    smin[w] = lo;
    smax[w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kTB / kWave; ++i) {
      lo = fmin(lo, smin[i]);
      hi = fmax(hi, smax[i]);
    }
    out[2 * blockIdx.x] = lo;
    out[2 * blockIdx.x + 1] = hi;
  }
}

// min and max of C (M x N entries) into mm[0], mm[1]; part: 2 * kMinmaxBlocks
constexpr int kMinmaxBlocks = 1024;
void minmax(hipStream_t st, const double *C, long n, double *part, double *mm) {
  hipLaunchKernelGGL(k_minmax, dim3(kMinmaxBlocks), dim3(kTB), 0, st, C, n, part);
  check_launch();
  // the 1024 partial minima and maxima, read as 2048 values: the min of all
  // is the min of the minima (a maximum is never below its block's minimum)
  // and likewise for the max
  hipLaunchKernelGGL(k_minmax, dim3(1), dim3(kTB), 0, st, part, 2L * kMinmaxBlocks, mm);
  check_launch();
}

// *flag = 1 if some entry of x differs from the one of y (x, y: n entries;
// *flag starts at 0)
__global__ void __launch_bounds__(kTB) k_differ(const double *x, const double *y, long n,
                                                  int *flag) {
  bool d = false;
  for (long k = blockIdx.x * (long)kTB + threadIdx.x; k < n; k += (long)gridDim.x * kTB)
    d = d || x[k] != y[k];
  if (__any(d) && (threadIdx.x % kWave) == 0) atomicOr(flag, 1);
}

constexpr int kMaxTB = 1024;  // threads of the per-slice operand kernel

// The slice operand of a scaled pass, one block of 1024 threads per slice s:
// x(i) = arg(i + s n) for i < n, m[s] = max_i x(i) (0 where no x is finite),
// E[i + s n] = exp(x(i) - m[s]) * wt(i + s n); the arguments are staged in E
// between the two sweeps.
template <class Arg, class Wt>
__global__ void __launch_bounds__(kMaxTB)
k_slice_operand(int n, Arg arg, Wt wt, double *E, double *m, const LoopState *ls) {
  if (ls && !ls->active) return;
  __shared__ double sh[kMaxTB / kWave];
  __shared__ double msh;
  const int s = blockIdx.x;
  double *e = E + (long)s * n;
  double hi = -INFINITY;
  for (int i = threadIdx.x; i < n; i += kMaxTB) {
    const double x = arg((long)i + (long)s * n);
    e[i] = x;
    hi = fmax(hi, x);
  }
  for (int o = kWave / 2; o > 0; o /= 2) hi = fmax(hi, __shfl_xor(hi, o));
  if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = hi;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kMaxTB / kWave; ++i) hi = fmax(hi, sh[i]);
    msh = isfinite(hi) ? hi : 0.0;
    m[s] = msh;
  }
  __syncthreads();
  const double ms = msh;
  for (int i = threadIdx.x; i < n; i += kMaxTB)
    e[i] = exp(e[i] - ms) * wt((long)i + (long)s * n);
}

template <class Arg, class Wt>
void slice_operand(hipStream_t st, int n, int S, Arg arg, Wt wt, double *E, double *m,
                   const LoopState *ls) {
  hipLaunchKernelGGL((k_slice_operand<Arg, Wt>), dim3(S), dim3(kMaxTB), 0, st, n, arg, wt, E, m,
                     ls);
  check_launch();
}

// The scaled kernel in both layouts and the passes' scratch.
struct ScaledSoftmin {
  const double *Kr = nullptr;  // Kt(i, j) at Kr[j + i * N] (C^T's layout): rows
  const double *Kc = nullptr;  // Kt(i, j) at Kc[i + j * M] (C's layout): columns
  double cmin = 0.0;
  double *m = nullptr, *E = nullptr, *mv = nullptr, *wr = nullptr, *wc = nullptr;
  int ksr = 1, ksc = 1;
};

// k splits of a scaled pass: twice the split the product would take alone
// (the passes have S <= 32 columns and stream the kernel once).
inline int scaled_splits(int rows, int k) {
  return std::min(8, 2 * splitk_for<double, F64Small>(rows, k));
}

// row soft-min (as lse_row): op(i + s M, value)
template <class Op>
void scaled_row(hipStream_t st, const ScaledSoftmin &sp, int M, int N, int S, double reg,
                const double *F, const double *G, Op op, const LoopState *ls) {
  double *m = sp.m, *E = sp.E;
  slice_operand(
      st, N, S, [=] __device__(long k) { return G[k] / reg; },
      [] __device__(long) { return 1.0; }, E, m, ls);
  SplitK<double> sk;
  sk.ks = sp.ksr;
  sk.work = sk.ks > 1 ? sp.wr : nullptr;
  const double cm = sp.cmin;
  gemm_tn<double, F64Small>(st, M, S, N, sp.Kr, N, E, N,
                            [=] __device__(int i, int s, double y) {
                              if (ls && !ls->active) return;
                              const long k = i + (long)s * M;
                              op(k, -F[k] - reg * (log(y) + m[s]) + cm);
                            },
                            nullptr, nullptr, sk);
}

// column soft-min (as lse_col_t): op(j, values of the S slices)
template <class Op>
void scaled_col(hipStream_t st, const ScaledSoftmin &sp, int M, int N, int S, double reg,
                const double *F, const double *G, Op op, const LoopState *ls) {
  double *m = sp.m, *E = sp.E, *mv = sp.mv;
  slice_operand(
      st, M, S, [=] __device__(long k) { return F[k] / reg; },
      [] __device__(long) { return 1.0; }, E, m, ls);
  SplitK<double> sk;
  sk.ks = sp.ksc;
  sk.work = sk.ks > 1 ? sp.wc : nullptr;
  const double cm = sp.cmin;
  gemm_tn<double, F64Small>(st, N, S, M, sp.Kc, M, E, M,
                            [=] __device__(int j, int s, double z) {
                              if (ls && !ls->active) return;
                              const long k = j + (long)s * N;
                              mv[k] = -G[k] - reg * (log(z) + m[s]) + cm;
                            },
                            nullptr, nullptr, sk);
  RWIG_DISPATCH_S(S, SM,
                  hipLaunchKernelGGL((k_cols_apply<double, SM, Op>),
                                     dim3((N + kWave - 1) / kWave), dim3(kWave), 0, st, mv, N, S,
                                     op, ls));
  check_launch();
}

// z * exp(x) for a sum z of the scaled form and its exponent x in natural
// units, as ldexp(z 2^f, n) with x log2(e) = n + f: no overflow of exp(x)
// on the way to a representable result
__device__ __forceinline__ double scale_exp(double z, double x) {
  const double t = x * kLog2e, n = floor(t);
  return ldexp(z * exp2(t - n), (int)fmax(fmin(n, 4000.0), -4000.0));
}

// The adjoint's weighted sums as scaled products (the same bound: the largest
// weight of every row or column is at least exp(-kScaledRange) before its
// factor). Column sums with row soft-max weights (as wcol_t):
//   res(j,s) = sum_i exp(-(C(i,j) - F(i,s) - G(j,s) - rsh(i,s)) / reg) x(i,s)
//            = exp(G(j,s)/reg + m_s - cmin/reg) sum_i Kt(i,j) E(i,s),
//   E = exp(a - m_s) x, a = (F + rsh) / reg, m_s = max_i a(i,s);
// op(j, values of the S slices).
template <class Op>
void scaled_wcol(hipStream_t st, const ScaledSoftmin &sp, int M, int N, int S, double reg,
                 const double *F, const double *G, const double *rsh, const double *x, Op op) {
  double *m = sp.m, *E = sp.E, *mv = sp.mv;
  slice_operand(
      st, M, S, [=] __device__(long k) { return (F[k] + rsh[k]) / reg; },
      [=] __device__(long k) { return x[k]; }, E, m, nullptr);
  SplitK<double> sk;
  sk.ks = sp.ksc;
  sk.work = sk.ks > 1 ? sp.wc : nullptr;
  const double cm = sp.cmin;
  gemm_tn<double, F64Small>(st, N, S, M, sp.Kc, M, E, M,
                            [=] __device__(int j, int s, double z) {
                              const long k = j + (long)s * N;
                              mv[k] = scale_exp(z, G[k] / reg + m[s] - cm / reg);
                            },
                            nullptr, nullptr, sk);
  RWIG_DISPATCH_S(S, SM,
                  hipLaunchKernelGGL((k_cols_apply<double, SM, Op>),
                                     dim3((N + kWave - 1) / kWave), dim3(kWave), 0, st, mv, N, S,
                                     op, (const LoopState *)nullptr));
  check_launch();
}

// Row sums with column soft-max weights (as wrow with a row shift):
//   part(i,s) = sum_j exp(-(C(i,j) - F(i,s) - G(j,s) - csh(j,s) + rowsh(i,s)) / reg) y(j,s)
//             = exp((F - rowsh)(i,s)/reg + m_s - cmin/reg) sum_j Kt(i,j) E(j,s),
//   E = exp(v - m_s) y, v = (G + csh) / reg, m_s = max_j v(j,s); op(i + s M, part).
template <class Op>
void scaled_wrow(hipStream_t st, const ScaledSoftmin &sp, int M, int N, int S, double reg,
                 const double *F, const double *G, const double *csh, const double *y,
                 const double *rowsh, Op op) {
  double *m = sp.m, *E = sp.E;
  slice_operand(
      st, N, S, [=] __device__(long k) { return (G[k] + csh[k]) / reg; },
      [=] __device__(long k) { return y[k]; }, E, m, nullptr);
  SplitK<double> sk;
  sk.ks = sp.ksr;
  sk.work = sk.ks > 1 ? sp.wr : nullptr;
  const double cm = sp.cmin;
  gemm_tn<double, F64Small>(st, M, S, N, sp.Kr, N, E, N,
                            [=] __device__(int i, int s, double z) {
                              const long k = i + (long)s * M;
                              op(k, scale_exp(z, (F[k] - rowsh[k]) / reg + m[s] - cm / reg));
                            },
                            nullptr, nullptr, sk);
}

__device__ __forceinline__ void finish_iteration_ck(LoopState *ls, double err) {
  ls->err = err;
  ls->iter = ls->iter + 1;
  ls->active = (ls->iter < ls->maxiter && err >= ls->tol) ? 1 : 0;
}

// out(s) = scale * sum_j acc(j, s), one block per s.
__global__ void __launch_bounds__(kTB)
k_colsum_ck(const double *acc, int N, double scale, double *out) {
  __shared__ double sh[kTB / kWave];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) a += acc[j + (long)s * N];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) out[s] = scale * a;
}

// Device buffers owned for the duration of one call.
struct Arena {
  std::vector<DBuf<char>> bufs;
  template <typename T> T *make(size_t n) {
    bufs.emplace_back(n * sizeof(T));
    return reinterpret_cast<T *>(bufs.back().get());
  }
};

int status_of(double err, double tol, int iter, int maxiter) {
  if (err <= tol) return 0;
  if (iter == maxiter) return 1;
  return 2;
}

struct BaryCkResult {
  int iter = 0;
  double err = 0.0;
};

// Device pointers of type T: C (M x N), w (S), bext (N) in; b (N), F (M x S),
// G (N x S) out. FP64: A (M x S) in (its log and the division of the
// gradient by A stay in FP64, so FP32 runs take entries of A below FP32's
// normal range); gA (M x S), gw (S), loss out. X is the type the soft-min and
// weighted-sum passes evaluate their exponentials in (X = float, T = double:
// FP64 arguments, FP32 exponentials).
template <typename T, typename X>
BaryCkResult barycenter_log_ck(hipStream_t st, const double *A, const T *C, const T *w,
                               const T *bext, int M, int N, int S, double reg, bool wg,
                               int maxiter, double tol, int ck, int seed, T *bout, T *Fout,
                               T *Gout, double *gA, double *gw, double *loss) {
  const long MS = (long)M * S, NS = (long)N * S;
  Arena ar;
  T *F = ar.make<T>(MS), *G = ar.make<T>(NS);
  T *logA = ar.make<T>(MS), *KV = ar.make<T>(MS), *logb = ar.make<T>(N);
  T *pm = ar.make<T>(lse_row_part_size(M, N, S)), *ps = ar.make<T>(lse_row_part_size(M, N, S));
  // C^T for the column passes (row passes over C^T), and their partials
  T *CT = ar.make<T>((size_t)M * N);
  transpose<T>(st, C, M, N, M, CT, N);
  T *cpm = ar.make<T>(col_t_part_size(M, N, S)), *cps = ar.make<T>(col_t_part_size(M, N, S));
  LoopState *ls = ar.make<LoopState>(1);
  double *gpart = ar.make<double>(gram_part_size(S));
  const T treg = (T)reg;
  // FP64: the soft-min passes as products with the scaled kernel when the
  // cost's range allows it (the same values up to rounding)
  constexpr bool kScalable = std::is_same<T, double>::value && std::is_same<X, double>::value;
  bool scaled = false;
  ScaledSoftmin sp;
  if constexpr (kScalable) {
    double *mm = ar.make<double>(2), *mpart = ar.make<double>(2 * kMinmaxBlocks), hmm[2];
    minmax(st, C, (long)M * N, mpart, mm);
    d2h(hmm, mm, 2, st);
    // a symmetric cost (C = C^T, as on a grid) keeps one scaled kernel for
    // both passes: half the memory, and the passes then stream the same
    // 8 M N bytes
    int *dflag = ar.make<int>(1), hdiff = 1;
    if (M == N) {
      launch_map(st, 1L, [=] __device__(long) { *dflag = 0; });
      hipLaunchKernelGGL(k_differ, dim3(1024), dim3(kTB), 0, st, (const double *)C,
                         (const double *)CT, (long)M * N, dflag);
      check_launch();
      d2h(&hdiff, dflag, 1, st);
    }
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
    if (std::isfinite(hmm[0]) && std::isfinite(hmm[1]) && (hmm[1] - hmm[0]) / reg <= kScaledRange) {
      scaled = true;
      const double cmin = hmm[0];
      double *Kc = ar.make<double>((size_t)M * N);
      double *Kr = hdiff ? ar.make<double>((size_t)M * N) : Kc;
      const double *Cd = C, *CTd = CT;
      launch_map(st, (long)M * N, [=] __device__(long k) {
        const double kc = exp(-(Cd[k] - cmin) / reg);
        Kc[k] = kc;
        if (Kr != Kc) Kr[k] = exp(-(CTd[k] - cmin) / reg);
      });
      sp.Kr = Kr;
      sp.Kc = Kc;
      sp.cmin = cmin;
      sp.m = ar.make<double>(S);
      sp.E = ar.make<double>((size_t)std::max(M, N) * S);
      sp.mv = ar.make<double>((size_t)N * S);
      sp.ksr = scaled_splits(M, N);
      sp.ksc = scaled_splits(N, M);
      sp.wr = ar.make<double>(splitk_work_elems<double, F64Small>(M, S, sp.ksr));
      sp.wc = ar.make<double>(splitk_work_elems<double, F64Small>(N, S, sp.ksc));
    }
  }
  // the forward's two soft-min passes over the working F, G
  auto row_pass = [&](auto op, const LoopState *gate) {
    if constexpr (kScalable) {
      if (scaled) {
        scaled_row(st, sp, M, N, S, reg, F, G, op, gate);
        return;
      }
    }
    lse_row<T, X>(st, C, M, N, M, F, M, G, N, S, reg, pm, ps, op, gate);
  };
  auto col_pass = [&](auto op, const LoopState *gate) {
    if constexpr (kScalable) {
      if (scaled) {
        scaled_col(st, sp, M, N, S, reg, F, G, op, gate);
        return;
      }
    }
    lse_col_t<T, X>(st, CT, M, N, N, F, M, G, N, S, reg, cpm, cps, op, gate);
  };
  const int ncp = wg ? maxiter / ck + 1 : 0;
  T *Fc = wg ? ar.make<T>((size_t)ncp * MS) : nullptr;
  T *Gc = wg ? ar.make<T>((size_t)ncp * NS) : nullptr;
  T *KVc = wg ? ar.make<T>((size_t)ncp * MS) : nullptr;

  launch_map(st, (long)N, [=] __device__(long j) { logb[j] = T(0); });
  launch_map(st, MS, [=] __device__(long k) {
    F[k] = T(0);
    logA[k] = (T)log(A[k]);
    if (Fc) Fc[k] = T(0);
  });
  launch_map(st, NS, [=] __device__(long k) {
    G[k] = T(0);
    if (Gc) Gc[k] = T(0);
  });
  {
    LoopState h{};
    h.err = 1000.0;  // inst/include/sinkhorn_impl.hpp:85-88
    h.maxiter = maxiter;
    h.tol = tol;
    h.active = (0 < maxiter && h.err >= tol) ? 1 : 0;
    h2d(ls, &h, 1, st);
  }
  row_pass(
      [=] __device__(long k, T r) {
        KV[k] = r;
        if (KVc) KVc[k] = r;
      },
      nullptr);

  // One iteration (src/barycenter_impl_log.cpp:226-253). In the forward
  // (gate set) the iteration number l = iter + 1 is read on the device and the
  // checkpoint slots are written when l % ck == 0; in the recomputation
  // (gate null) hF, hG, hKV, hKTU are the iteration's own history slices.
  auto iteration = [&](LoopState *gate, T *hF, T *hG, T *hKV, T *hKTU) {
    const int ckk = ck;
    // F += reg log A + KV
    launch_map(
        st, MS,
        [=] __device__(long k) {
          const T fk = F[k] + (treg * logA[k] + KV[k]);
          F[k] = fk;
          if (gate) {
            const int l = gate->iter + 1;
            if (Fc && l % ckk == 0) Fc[(long)(l / ckk) * MS + k] = fk;
          } else {
            hF[k] = fk;
          }
        },
        gate);
    // KTU = column soft-min of R(F, G); logb = -(G + KTU) w / reg;
    // G += reg logb 1^T + KTU
    col_pass(
        [=] __device__(int j, const T *r) {
          T acc = T(0);
          for (int s = 0; s < S; ++s) acc += -(G[j + (long)s * N] + r[s]) * w[s];
          const T lb = acc / treg;
          logb[j] = lb;
          const int l = gate ? gate->iter + 1 : 0;
          for (int s = 0; s < S; ++s) {
            const long k = j + (long)s * N;
            const T gk = G[k] + (treg * lb + r[s]);
            G[k] = gk;
            if (gate) {
              if (Gc && l % ckk == 0) Gc[(long)(l / ckk) * NS + k] = gk;
            } else {
              hG[k] = gk;
              hKTU[k] = r[s];
            }
          }
        },
        gate);
    // KV = row soft-min of R(F, G)
    row_pass(
        [=] __device__(long k, T r) {
          KV[k] = r;
          if (gate) {
            const int l = gate->iter + 1;
            if (KVc && l % ckk == 0) KVc[(long)(l / ckk) * MS + k] = r;
          } else {
            hKV[k] = r;
          }
        },
        gate);
    // err = norm(-KV / reg - log A, 2), the spectral norm
    if (gate)
      spectral(
          st, (long)M, S,
          [=] __device__(long i, int s) {
            const long k = i + (long)s * M;
            return -(double)KV[k] / reg - (double)logA[k];
          },
          [=] __device__(double nrm) { finish_iteration_ck(gate, nrm); }, gate, gpart);
  };

  LoopState hs{};
  d2h(&hs, ls, 1, st);
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  for (int batch = 4; hs.active; batch = std::min(batch * 2, 64)) {
    const int n = std::min(batch, maxiter - hs.iter);
    for (int k = 0; k < n; ++k) iteration(ls, nullptr, nullptr, nullptr, nullptr);
    d2h(&hs, ls, 1, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
  }
  const int L = hs.iter;
  d2d(Fout, F, MS, st);
  d2d(Gout, G, NS, st);

  // b = exp(logb) / sum  (src/barycenter_impl_log.cpp:268-271)
  double *tot = ar.make<double>(2);
  reduce2(
      st, (long)N, [=] __device__(long j) { return exp((double)logb[j]); }, 0L,
      [=] __device__(long) { return 0.0; }, [=] __device__(double s1, double) { tot[0] = s1; });
  launch_map(st, (long)N,
             [=] __device__(long j) { bout[j] = (T)(exp((double)logb[j]) / tot[0]); });

  if (wg) {
    // loss and, for seed 2, <g, b> with g = 2 (b - b_ext)
    const int sd = seed;
    reduce2(
        st, (long)N,
        [=] __device__(long j) {
          const double e = exp((double)logb[j]);
          const double bj = sd == 1 ? e : (sd == 0 ? (double)bout[j] : e / tot[0]);
          const double r = bj - (double)bext[j];
          return r * r;
        },
        (long)N,
        [=] __device__(long j) {
          const double bj = exp((double)logb[j]) / tot[0];
          return 2.0 * (bj - (double)bext[j]) * bj;
        },
        [=] __device__(double s1, double s2) {
          *loss = s1;
          tot[1] = s2;
        });
    // adjoint seed, computed in FP64
    double *seedv = ar.make<double>(N);
    launch_map(st, (long)N, [=] __device__(long j) {
      const double e = exp((double)logb[j]);
      const double bn = e / tot[0];
      double v;
      if (sd == 1)
        v = 2.0 * (e - (double)bext[j]) * e;
      else if (sd == 2)
        v = bn * (2.0 * (bn - (double)bext[j]) - tot[1]);
      else
        v = 2.0 * ((double)bout[j] - (double)bext[j]) * (double)bout[j];
      seedv[j] = v;
    });

    // history of one segment: slices 0 .. ck of F, G, KV, KTU
    T *hF = ar.make<T>((size_t)(ck + 1) * MS), *hG = ar.make<T>((size_t)(ck + 1) * NS);
    T *hKV = ar.make<T>((size_t)(ck + 1) * MS), *hKTU = ar.make<T>((size_t)(ck + 1) * NS);
    T *Fbar = ar.make<T>(MS), *Gbar = ar.make<T>(NS), *lbb = ar.make<T>(N), *Y = ar.make<T>(NS);
    double *gwa = ar.make<double>(NS);
    T *part = ar.make<T>(wrow_part_size(M, N, S)), *cpart = ar.make<T>(col_t_part_size(M, N, S));
    T *rlogA = ar.make<T>(MS);
    launch_map(st, MS, [=] __device__(long k) { rlogA[k] = (T)(reg * log(A[k])); });
    launch_map(st, MS, [=] __device__(long k) {
      gA[k] = 0.0;
      Fbar[k] = T(0);
    });
    launch_map(st, NS, [=] __device__(long k) {
      gwa[k] = 0.0;
      Gbar[k] = T(0);
    });

    for (int l1 = L; l1 > 0;) {
      const int l0 = ((l1 - 1) / ck) * ck;
      const long cp = l0 / ck;
      // restore checkpoint l0 into the working state and history slice 0
      {
        const T *fc = Fc + cp * MS, *gc = Gc + cp * NS, *kvc = KVc + cp * MS;
        launch_map(st, MS, [=] __device__(long k) {
          F[k] = fc[k];
          KV[k] = kvc[k];
          hF[k] = fc[k];
          hKV[k] = kvc[k];
        });
        launch_map(st, NS, [=] __device__(long k) {
          G[k] = gc[k];
          hG[k] = gc[k];
        });
      }
      for (int l = l0 + 1; l <= l1; ++l) {
        const long q = l - l0;
        iteration(nullptr, hF + q * MS, hG + q * NS, hKV + q * MS, hKTU + q * NS);
      }
      // adjoint steps l = l1 .. l0 + 1  (src/barycenter_impl_log.cpp:301-373)
      for (int l = l1; l > l0; --l) {
        const long q = l - l0;
        const T *Fl = hF + q * MS, *Gl = hG + q * NS, *Glm1 = hG + (q - 1) * NS;
        const T *KVl = hKV + q * MS, *KTUl = hKTU + q * NS;
        if (l != L) {
          // Gbar(:,s) = -X_s^T Fbar(:,s); logbbar = reg * rowsum(Gbar)
          auto gop = [=] __device__(int j, const T *r) {
            T acc = T(0);
            for (int s = 0; s < S; ++s) {
              Gbar[j + (long)s * N] = -r[s];
              acc += -r[s];
            }
            lbb[j] = treg * acc;
          };
          bool done = false;
          if constexpr (kScalable) {
            if (scaled) {
              scaled_wcol(st, sp, M, N, S, reg, Fl, Gl, KVl, Fbar, gop);
              done = true;
            }
          }
          if (!done) wcol_t<T, X>(st, CT, M, N, N, Fl, M, Gl, N, S, reg, KVl, Fbar, cpart, gop,
                                  nullptr);
        } else {
          launch_map(st, (long)N, [=] __device__(long j) { lbb[j] = (T)seedv[j]; });
        }
        // Fbar(:,s) = W_s ((w_s / reg) logbbar - Gbar(:,s)), W_s the column
        // soft-max of -R_s(F_l, G_{l-1}) / reg
        const bool last = (l == L);
        launch_map(st, NS, [=] __device__(long k) {
          const long j = k % N, s = k / N;
          const T t = (w[s] / treg) * lbb[j];
          Y[k] = last ? t : t - Gbar[k];
          gwa[k] += (double)(Glm1[k] + KTUl[k]) * (double)lbb[j];
        });
        // computed relative to each row's A (Q = Fbar / A, rows shifted by
        // reg log A): rows of A far below the FP32 range keep their gradient
        auto fop = [=] __device__(long k, T q) {
          Fbar[k] = (T)((double)q * A[k]);
          gA[k] += (double)q;
        };
        bool done = false;
        if constexpr (kScalable) {
          if (scaled) {
            scaled_wrow(st, sp, M, N, S, reg, Fl, Glm1, KTUl, Y, rlogA, fop);
            done = true;
          }
        }
        if (!done) wrow<T, X>(st, C, M, N, M, Fl, M, Glm1, N, S, reg, KTUl, Y, part, fop, nullptr,
                              rlogA);
      }
      l1 = l0;
    }
    launch_map(st, MS, [=] __device__(long k) { gA[k] *= reg; });
    hipLaunchKernelGGL(k_colsum_ck, dim3(S), dim3(kTB), 0, st, gwa, N, -1.0 / reg, gw);
    check_launch();
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return {L, hs.err};
}

// WDL inference in the log domain (src/wdl_impl.cpp:354-377): D barycenters
// sharing A (M x S) and the cost, document d weighted by W(:, d), each the
// log barycenter of src/barycenter_impl_log.cpp:189-272 without gradient,
// stopping on its own spectral-norm error. Documents run in groups of
// G = 32 / S (at least 1): the soft-min passes of a group sweep C once for
// all its G S slices. A document that has stopped keeps its state while the
// others go on (its updates are masked; a slice's arithmetic does not depend
// on the other slices of a pass), so its result is the one it would have
// alone. Device pointers: A (M x S), W (S x D) FP64; C (M x N) and CT (N x M)
// of type T; B (N x D, normalised barycenters), iters, errs (D) out.
template <typename T, typename X>
void barycenter_log_batch(hipStream_t st, const double *A, const T *C, const T *CT,
                          const double *W, int M, int N, int S, int D, double reg, int maxiter,
                          double tol, double *B, int *iters, double *errs) {
  const int G = std::max(1, 32 / S), QM = G * S;
  const T treg = (T)reg;
  Arena ar;
  T *F = ar.make<T>((size_t)M * QM), *Gp = ar.make<T>((size_t)N * QM);
  T *KV = ar.make<T>((size_t)M * QM), *logb = ar.make<T>((size_t)N * G);
  T *wq = ar.make<T>(QM), *logA = ar.make<T>((size_t)M * S);
  T *pm = ar.make<T>(lse_row_part_size(M, N, QM)), *ps = ar.make<T>(lse_row_part_size(M, N, QM));
  T *cpm = ar.make<T>(col_t_part_size(M, N, QM)), *cps = ar.make<T>(col_t_part_size(M, N, QM));
  const size_t gps = gram_part_size(S);
  double *gpart = ar.make<double>(gps * G), *tot = ar.make<double>(G);
  LoopState *ls = ar.make<LoopState>(G + 1);  // G documents, then the group's gate
  LoopState *gate = ls + G;
  launch_map(st, (long)M * S, [=] __device__(long k) { logA[k] = (T)log(A[k]); });
  for (int d0 = 0; d0 < D; d0 += G) {
    const int g = std::min(G, D - d0), Q = g * S;
    launch_map(st, (long)M * Q, [=] __device__(long k) { F[k] = T(0); });
    launch_map(st, (long)N * Q, [=] __device__(long k) { Gp[k] = T(0); });
    launch_map(st, (long)N * g, [=] __device__(long k) { logb[k] = T(0); });
    launch_map(st, (long)Q, [=] __device__(long q) {
      wq[q] = (T)W[(q % S) + (long)(d0 + q / S) * S];
    });
    {
      std::vector<LoopState> h(g + 1);
      for (auto &x : h) {
        x.iter = 0;
        x.err = 1000.0;  // inst/include/sinkhorn_impl.hpp:85-88
        x.maxiter = maxiter;
        x.tol = tol;
        x.active = (0 < maxiter && x.err >= tol) ? 1 : 0;
      }
      h2d(ls, h.data(), (size_t)g, st);
      h2d(gate, &h[g], 1, st);
    }
    lse_row<T, X>(st, C, M, N, M, F, M, Gp, N, Q, reg, pm, ps,
                  [=] __device__(long k, T r) { KV[k] = r; }, nullptr);
    auto iteration = [&]() {
      // F += reg log A + KV, for the slices of running documents
      launch_map(
          st, (long)M * Q,
          [=] __device__(long k) {
            const long q = k / M;
            if (!ls[q / S].active) return;
            F[k] = F[k] + (treg * logA[k % M + (q % S) * (long)M] + KV[k]);
          },
          gate);
      // KTU = column soft-min; logb = -(G + KTU) w / reg; G += reg logb + KTU
      lse_col_t<T, X>(
          st, CT, M, N, N, F, M, Gp, N, Q, reg, cpm, cps,
          [=] __device__(int j, const T *r) {
            for (int d = 0; d < g; ++d) {
              if (!ls[d].active) continue;
              T acc = T(0);
              for (int s = 0; s < S; ++s)
                acc += -(Gp[j + (long)(d * S + s) * N] + r[d * S + s]) * wq[d * S + s];
              const T lb = acc / treg;
              logb[j + (long)d * N] = lb;
              for (int s = 0; s < S; ++s) {
                const long k = j + (long)(d * S + s) * N;
                Gp[k] = Gp[k] + (treg * lb + r[d * S + s]);
              }
            }
          },
          gate);
      // KV = row soft-min
      lse_row<T, X>(
          st, C, M, N, M, F, M, Gp, N, Q, reg, pm, ps,
          [=] __device__(long k, T r) {
            if (ls[(k / M) / S].active) KV[k] = r;
          },
          gate);
      // each document's error: the spectral norm of -KV / reg - log A
      for (int d = 0; d < g; ++d)
        spectral(
            st, (long)M, S,
            [=] __device__(long i, int s) {
              return -(double)KV[i + (long)(d * S + s) * M] / reg - (double)logA[i + (long)s * M];
            },
            [=] __device__(double nrm) { finish_iteration_ck(ls + d, nrm); }, ls + d,
            gpart + d * gps);
      // the group runs while one of its documents does
      launch_map(
          st, 1L,
          [=] __device__(long) {
            int a = 0;
            for (int d = 0; d < g; ++d) a |= ls[d].active;
            gate->iter = gate->iter + 1;
            gate->active = a;
          },
          gate);
    };
    LoopState hs{};
    d2h(&hs, gate, 1, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
    for (int batch = 4; hs.active; batch = std::min(batch * 2, 64)) {
      const int n = std::min(batch, maxiter - hs.iter);
      for (int k = 0; k < n; ++k) iteration();
      d2h(&hs, gate, 1, st);
      RWIG_HIP_CHECK(hipStreamSynchronize(st));
    }
    // b = exp(logb) / sum  (src/barycenter_impl_log.cpp:268-271)
    for (int d = 0; d < g; ++d) {
      const T *lb = logb + (long)d * N;
      double *t = tot + d;
      reduce2(
          st, (long)N, [=] __device__(long j) { return exp((double)lb[j]); }, 0L,
          [=] __device__(long) { return 0.0; }, [=] __device__(double s1, double) { *t = s1; });
    }
    launch_map(st, (long)N * g, [=] __device__(long k) {
      const long d = k / N;
      B[k + (long)d0 * N] = exp((double)logb[k]) / tot[d];
    });
    launch_map(st, (long)g, [=] __device__(long d) {
      if (iters) iters[d0 + d] = ls[d].iter;
      if (errs) errs[d0 + d] = ls[d].err;
    });
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
}

template <typename T, typename X>
BaryCkResult run_ck(hipStream_t st, double *Fh, double *Gh, double *bh, double *gAh, double *gwh,
                    double *lossh, const double *A, const double *w, const double *C,
                    const double *bext, int M, int N, int S, double reg, bool wg, int maxiter,
                    double tol, int ck, int seed) {
  const long MN = (long)M * N, MS = (long)M * S, NS = (long)N * S;
  Arena ar;
  // inputs and outputs of type T, staged through an FP64 device buffer
  double *stage = ar.make<double>(std::max({MN, MS, NS}));
  auto upload = [&](const double *h, long n) {
    T *d = ar.make<T>(n);
    h2d(stage, h, n, st);
    launch_map(st, n, [=] __device__(long k) { d[k] = (T)stage[k]; });
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
    return d;
  };
  double *dA = ar.make<double>(MS);
  h2d(dA, A, MS, st);
  const T *dC = upload(C, MN), *dw = upload(w, S);
  const T *dbext = wg ? upload(bext, N) : nullptr;
  T *db = ar.make<T>(N), *dF = ar.make<T>(MS), *dG = ar.make<T>(NS);
  double *dgA = ar.make<double>(MS), *dgw = ar.make<double>(S), *dloss = ar.make<double>(1);
  const BaryCkResult res = barycenter_log_ck<T, X>(st, dA, dC, dw, dbext, M, N, S, reg, wg, maxiter,
                                                tol, ck, seed, db, dF, dG, dgA, dgw, dloss);
  auto download = [&](double *h, const T *d, long n) {
    if (!h) return;
    launch_map(st, n, [=] __device__(long k) { stage[k] = (double)d[k]; });
    d2h(h, stage, n, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
  };
  download(bh, db, N);
  download(Fh, dF, MS);
  download(Gh, dG, NS);
  if (wg) {
    if (gAh) d2h(gAh, dgA, MS, st);
    if (gwh) d2h(gwh, dgw, S, st);
    if (lossh) d2h(lossh, dloss, 1, st);
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return res;
}

// The same solve on device buffers: C in T (column-major, leading dimension
// M), A, w and b_ext in FP64; F, G and b are written in FP64 (converted from
// T), the gradient and loss as the solve leaves them. Null F, G or b: not
// returned.
template <typename T, typename X>
BaryCkResult run_ck_dev(hipStream_t st, double *F, double *G, double *b, double *gA, double *gw,
                        double *loss, const double *A, const double *w, const T *C,
                        const double *bext, int M, int N, int S, double reg, bool wg, int maxiter,
                        double tol, int ck, int seed) {
  const long MS = (long)M * S, NS = (long)N * S;
  Arena ar;
  T *dw = ar.make<T>(S);
  launch_map(st, (long)S, [=] __device__(long k) { dw[k] = (T)w[k]; });
  T *dbext = nullptr;
  if (wg) {
    dbext = ar.make<T>(N);
    launch_map(st, (long)N, [=] __device__(long k) { dbext[k] = (T)bext[k]; });
  }
  constexpr bool same = std::is_same<T, double>::value;
  auto out = [&](double *h, long n) -> T * {
    if constexpr (same) {
      if (h) return h;
    }
    return ar.make<T>(n);
  };
  T *db = out(b, N), *dF = out(F, MS), *dG = out(G, NS);
  double *dgA = wg && gA ? gA : ar.make<double>(MS);
  double *dgw = wg && gw ? gw : ar.make<double>(S);
  double *dloss = wg && loss ? loss : ar.make<double>(1);
  const BaryCkResult res = barycenter_log_ck<T, X>(st, A, C, dw, dbext, M, N, S, reg, wg, maxiter,
                                                tol, ck, seed, db, dF, dG, dgA, dgw, dloss);
  if constexpr (!same) {
    auto widen = [&](double *h, const T *d, long n) {
      if (h) launch_map(st, n, [=] __device__(long k) { h[k] = (double)d[k]; });
    };
    widen(b, db, N);
    widen(F, dF, MS);
    widen(G, dG, NS);
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return res;
}

}  // namespace

}  // namespace rwig

using namespace rwig;

extern "C" int rwig_barycenter_log_ck_dev(double *F, double *G, double *b, double *grad_A,
                                          double *grad_w, double *loss, int *iter_out,
                                          double *err_out, const double *A, const double *w,
                                          const void *C, const double *b_ext, int M, int N,
                                          int S, double reg, int withgrad, int max_iter,
                                          double zero_tol, int checkpoint_every, int seed_mode,
                                          int precision, int *return_status_out, void *stream) {
  return guarded("rwig_barycenter_log_ck_dev", [&] {
    RWIG_REQUIRE(M > 0 && N > 0, "matrix dimensions must be positive");
    RWIG_REQUIRE(max_iter >= 0, "max_iter must be non-negative");
    RWIG_REQUIRE(reg > 0.0, "reg must be positive");
    RWIG_REQUIRE(S > 0 && S <= 32, "S must be in [1, 32]");
    RWIG_REQUIRE(A && w && C, "A, w and C are required");
    RWIG_REQUIRE(!withgrad || b_ext, "b_ext is required with withgrad");
    RWIG_REQUIRE(checkpoint_every >= 1, "checkpoint_every must be positive");
    RWIG_REQUIRE(seed_mode >= 0 && seed_mode <= 2, "seed_mode must be 0, 1 or 2");
    RWIG_REQUIRE(precision >= 0 && precision <= 2,
                 "precision must be 0 (FP64), 1 (FP32) or 2 (FP64 arguments, FP32 exponentials)");
    hipStream_t st = (hipStream_t)stream;
    const bool wg = withgrad != 0;
    const BaryCkResult res =
        precision == 0
            ? run_ck_dev<double, double>(st, F, G, b, grad_A, grad_w, loss, A, w,
                                         (const double *)C, b_ext, M, N, S, reg, wg, max_iter,
                                         zero_tol, checkpoint_every, seed_mode)
        : precision == 1
            ? run_ck_dev<float, float>(st, F, G, b, grad_A, grad_w, loss, A, w, (const float *)C,
                                       b_ext, M, N, S, reg, wg, max_iter, zero_tol,
                                       checkpoint_every, seed_mode)
            : run_ck_dev<double, float>(st, F, G, b, grad_A, grad_w, loss, A, w,
                                        (const double *)C, b_ext, M, N, S, reg, wg, max_iter,
                                        zero_tol, checkpoint_every, seed_mode);
    if (iter_out) *iter_out = res.iter;
    if (err_out) *err_out = res.err;
    if (return_status_out) *return_status_out = status_of(res.err, zero_tol, res.iter, max_iter);
  });
}

extern "C" int rwig_barycenter_log_batched(double *B, int *iters_out, double *errs_out,
                                           const double *A, const double *W, const double *C,
                                           int M, int N, int S, int D, double reg,
                                           int max_iter, double zero_tol) {
  return guarded("rwig_barycenter_log_batched", [&] {
    RWIG_REQUIRE(M > 0 && N > 0 && D >= 0, "matrix dimensions must be positive");
    RWIG_REQUIRE(S > 0 && S <= 32, "S must be in [1, 32]");
    RWIG_REQUIRE(max_iter >= 0, "max_iter must be non-negative");
    RWIG_REQUIRE(reg > 0.0, "reg must be positive");
    RWIG_REQUIRE(B && A && W && C, "B, A, W and C are required");
    if (D == 0) return;
    ensure_device();
    Stream s;
    hipStream_t st = s.get();
    Arena ar;
    const long MN = (long)M * N;
    double *dA = ar.make<double>((size_t)M * S), *dW = ar.make<double>((size_t)S * D);
    double *dC = ar.make<double>(MN), *dCT = ar.make<double>(MN);
    double *dB = ar.make<double>((size_t)N * D), *dE = ar.make<double>(D);
    int *dI = ar.make<int>(D);
    h2d(dA, A, (size_t)M * S, st);
    h2d(dW, W, (size_t)S * D, st);
    h2d(dC, C, (size_t)MN, st);
    transpose<double>(st, dC, M, N, M, dCT, N);
    barycenter_log_batch<double, double>(st, dA, dC, dCT, dW, M, N, S, D, reg, max_iter,
                                         zero_tol, dB, dI, dE);
    d2h(B, dB, (size_t)N * D, st);
    if (iters_out) d2h(iters_out, dI, (size_t)D, st);
    if (errs_out) d2h(errs_out, dE, (size_t)D, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
  });
}

extern "C" int rwig_barycenter_log_ck(double *F, double *G, double *b, double *grad_A,
                                      double *grad_w, double *loss, int *iter_out,
                                      double *err_out, const double *A, const double *w,
                                      const double *C, const double *b_ext, int M, int N, int S,
                                      double reg, int withgrad, int max_iter, double zero_tol,
                                      int checkpoint_every, int seed_mode, int precision,
                                      int *return_status_out) {
  return guarded("rwig_barycenter_log_ck", [&] {
    RWIG_REQUIRE(M > 0 && N > 0, "matrix dimensions must be positive");
    RWIG_REQUIRE(max_iter >= 0, "max_iter must be non-negative");
    RWIG_REQUIRE(reg > 0.0, "reg must be positive");
    RWIG_REQUIRE(S > 0 && S <= 32, "S must be in [1, 32]");
    RWIG_REQUIRE(A && w && C, "A, w and C are required");
    RWIG_REQUIRE(!withgrad || b_ext, "b_ext is required with withgrad");
    RWIG_REQUIRE(checkpoint_every >= 1, "checkpoint_every must be positive");
    RWIG_REQUIRE(seed_mode >= 0 && seed_mode <= 2, "seed_mode must be 0, 1 or 2");
    RWIG_REQUIRE(precision >= 0 && precision <= 2,
                 "precision must be 0 (FP64), 1 (FP32) or 2 (FP64 arguments, FP32 exponentials)");
    ensure_device();
    Stream s;
    auto run = [&](auto tag_t, auto tag_x) {
      using T = decltype(tag_t);
      using X = decltype(tag_x);
      return run_ck<T, X>(s.get(), F, G, b, grad_A, grad_w, loss, A, w, C, b_ext, M, N, S, reg,
                          withgrad != 0, max_iter, zero_tol, checkpoint_every, seed_mode);
    };
    const BaryCkResult res = precision == 0   ? run(double(), double())
                             : precision == 1 ? run(float(), float())
                                              : run(double(), float());
    if (iter_out) *iter_out = res.iter;
    if (err_out) *err_out = res.err;
    if (return_status_out) *return_status_out = status_of(res.err, zero_tol, res.iter, max_iter);
  });
}
