// Wasserstein Dictionary Learning on the GPU: the batched training step of
// WassersteinDictionaryLearning::_train_batch_batched (src/wdl_impl.cpp:87-247)
// with its unrolled adjoint, the optimizers (inst/include/optimizer.hpp:72-99),
// the column softmaxes (inst/include/wdl_impl.hpp:80-96), the epoch x batch
// loop of _compute_serial (src/wdl_impl.cpp:254-347) and the inference of
// Yhat (:354-377), the last one batched over documents.
//
// The D documents of a batch are stacked as S*D columns (column c = d*S + s),
// so every Gibbs product of the L unrolled iterations -- forward and adjoint
// -- is one MFMA GEMM of K (N x N) with an N x S*D matrix (FP64:
// v_mfma_f64_16x16x4, FP32: v_mfma_f32_16x16x4), with the elementwise
// recurrences fused into the GEMM epilogues. The forward keeps the iterates
// of all L steps (as the reference does) for the adjoint, or only every
// ck-th VB, the adjoint then recomputing one segment at a time.
#include <algorithm>
#include <random>

#include "gemm.hpp"
#include "kernels.hpp"

namespace rwig {

// tiles of the WDL products (S * D columns: the wide FP64 tiles)
template <typename T> using WdlCfg = typename WideCfg<T>::type;

// prod_s x[s * ld]^w[s] (the barycenter of src/wdl_impl.cpp:142-150) as
// exp2(sum_s w_s log2 x_s): one exponential in place of S powers, the same
// value up to rounding and without the partial products' underflow. A zero
// weight contributes x^0 = 1, also at x = 0.
__device__ __forceinline__ float log2_t(float x) { return log2f(x); }
__device__ __forceinline__ double log2_t(double x) { return log2(x); }
__device__ __forceinline__ float exp2_t(float x) { return exp2f(x); }
__device__ __forceinline__ double exp2_t(double x) { return exp2(x); }
template <typename T>
__device__ __forceinline__ T weighted_geomean(const T *x, long ld, const T *w, int S) {
  T l = T(0);
  for (int s = 0; s < S; ++s)
    if (w[s] != T(0)) l += w[s] * log2_t(x[s * ld]);
  return exp2_t(l);
}

namespace {

template <typename T> struct WdlBuf {
  std::vector<DBuf<T>> bufs;
  T *make(size_t n) {
    bufs.emplace_back(n);
    return bufs.back().get();
  }
};

// out(c) = sum_i X(i, c) over rows, one block per column (fixed order)
template <typename T, typename O>
__global__ void __launch_bounds__(256)
k_colsum_to(const T *X, long ldx, int N, O *out) {
  __shared__ double sh[256 / kWave];
  const int c = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a += (double)X[i + (long)c * ldx];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) out[c] = (O)a;
}

// column dots dot(c) = sum_i A(i, c % S) X(i, c)
template <typename T>
__global__ void __launch_bounds__(256)
k_coldot(const T *A, const T *X, long ldx, int N, int S, double *out) {
  __shared__ double sh[256 / kWave];
  const int c = blockIdx.x, s = c % S;
  double a = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    a += (double)A[i + (long)s * N] * (double)X[i + (long)c * ldx];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) out[c] = a;
}

// Inference stopping test of document d = blockIdx.x (one block per
// document, fixed order): err = ||U .* KV - A||_F over its S columns; the
// document stops when err < tol or at the last iteration.
template <typename T>
__global__ void __launch_bounds__(256)
k_infer_err(const T *U, const T *KV, long ldn, const T *A, int N, int S, int itp1,
            int maxiter, double tol, int *done, double *errv) {
  __shared__ double sh[256 / kWave];
  const int d = blockIdx.x;
  if (done[d]) return;  // the same for every thread of the block
  double acc = 0.0;
  for (int s = 0; s < S; ++s)
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const long k = i + ((long)d * S + s) * ldn;
      const double r = (double)U[k] * (double)KV[k] - (double)A[i + (long)s * N];
      acc += r * r;
    }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    errv[d] = sqrt(acc);
    if (!(itp1 < maxiter && errv[d] >= tol)) done[d] = 1;
  }
}

// Yhat(:, d) = b(:, d) / sum b(:, d), one block per document
template <typename T>
__global__ void __launch_bounds__(256)
k_infer_normalise(const T *bb, int N, double *Yhat) {
  __shared__ double sh[256 / kWave];
  const int d = blockIdx.x;
  double a = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a += (double)bb[i + (long)d * N];
  const double tot = block_sum(a, sh);
  for (int i = threadIdx.x; i < N; i += blockDim.x)
    Yhat[i + (long)d * N] = (double)bb[i + (long)d * N] / tot;
}

// glambda(s) = mean over d of w(s,d) (wbar(s,d) - sum_q w(q,d) wbar(q,d)), the
// softmax Jacobian of the weights; one block per atom s, the documents split
// over its threads (fixed order)
template <typename T>
__global__ void __launch_bounds__(256)
k_lambda_grad(const T *Wb, const double *wB, int S, int D, double *glambda) {
  __shared__ double sh[256 / kWave];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int d = threadIdx.x; d < D; d += blockDim.x) {
    double dw = 0.0;
    for (int q = 0; q < S; ++q) dw += (double)Wb[q + (long)d * S] * wB[(long)d * S + q];
    const double ws = (double)Wb[s + (long)d * S];
    a += ws * wB[(long)d * S + s] - ws * dw;
  }
  a = block_sum(a, sh);
  if (threadIdx.x == 0) glambda[s] = a / D;
}

}  // namespace

// One batched WDL gradient (src/wdl_impl.cpp:87-247).
//   K, KT: the Gibbs kernel exp(-C/reg) and its transpose, column-major with
//          leading dimension ldk (a multiple of 16 bytes)
//   A (N x S), Yb (N x D), Wb (S x D): column-major, leading dims N, N, S
// Outputs (FP64): gAlpha (N x S), glambda (S), loss = sum_d ||b_d - y_d||^2
// of the unnormalised barycenters of the L-th iteration (may be null).
// ck: 0 keeps the iterates of all L steps for the adjoint, as the reference
// does (src/wdl_impl.cpp:111-121); 0 < ck < L keeps VB every ck iterations
// and recomputes each segment's iterates (UB, KVB, KTUB, bB, VB) with the
// forward's own kernels before stepping back through it: the same gradient
// bit for bit, with ck + L / ck + 1 slices of history instead of L + 1, for
// L more forward steps (2L more GEMMs).
template <typename T>
void wdl_batch_grad(hipStream_t st, const T *K, const T *KT, long ldk, int N, int S, int D,
                    int L, const T *A, const T *Yb, const T *Wb, double *gAlpha,
                    double *glambda, double *loss, int ck) {
  RWIG_REQUIRE(L >= 1, "the WDL step needs at least one iteration");
  RWIG_REQUIRE(ck >= 0, "checkpoint interval must be non-negative");
  const int SD = S * D;
  const long ldn = (N + 7) / 8 * 8;  // leading dimension of the N x SD work matrices
  const long MSD = ldn * SD;
  const bool ckm = ck > 0 && ck < L;
  const int SL = ckm ? ck : L;  // steps of history held at once
  WdlBuf<T> wb;
  // history slots: iterate l sits in slot l - base (base 0 without
  // checkpoints; the first step of the current segment with them)
  T *UBh = wb.make((size_t)(SL + 1) * MSD), *VBh = wb.make((size_t)(SL + 1) * MSD);
  T *KVBh = wb.make((size_t)SL * MSD), *KTUBh = wb.make((size_t)(SL + 1) * MSD);
  T *bBh = wb.make((size_t)(SL + 1) * N * D);
  T *ckV = ckm ? wb.make((size_t)((L - 1) / ck + 1) * MSD) : nullptr;  // VB at l = 0, ck, ...
  int base = 0;
  auto UB = [&](int l) { return UBh + (long)(l - base) * MSD; };
  auto VB = [&](int l) { return VBh + (long)(l - base) * MSD; };
  auto KVB = [&](int l) { return KVBh + (long)(l - base) * MSD; };
  auto KTUB = [&](int l) { return KTUBh + (long)(l - base) * MSD; };
  auto bB = [&](int l) { return bBh + (long)(l - base) * N * D; };

  // VB_0 = UB_0 = 1 (padding rows 0), bB_0 = 0
  {
    T *u0 = UB(0), *v0 = VB(0);
    launch_map(st, MSD, [=] __device__(long k) {
      const T one = (k % ldn < N) ? T(1) : T(0);
      u0[k] = one;
      v0[k] = one;
    });
    T *b0 = bB(0);
    launch_map(st, (long)N * D, [=] __device__(long k) { b0[k] = T(0); });
  }
  // forward step l: VB_l -> KVB_l, UB_{l+1}, KTUB_{l+1}, bB_{l+1}, VB_{l+1}
  auto fwd = [&](int l) {
    T *kvb = KVB(l), *ub1 = UB(l + 1), *ktub1 = KTUB(l + 1), *vb1 = VB(l + 1);
    T *bb1 = bB(l + 1);
    // KVB = K VB_l ; UB_{l+1} = A / KVB. The product only stores its result,
    // and the division runs in its own streaming kernel: at the end of a
    // product tile its loads and divisions held every workgroup's tail (the
    // same arithmetic, so the same values)
    gemm_tn<T, WdlCfg<T>>(st, N, SD, N, KT, ldk, VB(l), ldn,
                          [=] __device__(int i, int c, T v) { kvb[i + (long)c * ldn] = v; });
    launch_map(st, (long)N * SD, [=] __device__(long k) {
      const long i = k % N, c = k / N, kk = i + c * ldn;
      ub1[kk] = A[i + (c % S) * N] / kvb[kk];
    });
    // KTUB = K^T UB_{l+1}
    gemm_tn<T, WdlCfg<T>>(st, N, SD, N, K, ldk, ub1, ldn,
               [=] __device__(int j, int c, T v) { ktub1[j + (long)c * ldn] = v; });
    // bB = prod_s KTUB^w ; VB = bB / KTUB
    launch_map(st, (long)N * D, [=] __device__(long k) {
      const long i = k % N, d = k / N;
      const T p = weighted_geomean(ktub1 + i + d * S * ldn, ldn, Wb + d * S, S);
      bb1[k] = p;
      for (int s = 0; s < S; ++s) {
        const long kk = i + (d * S + s) * ldn;
        vb1[kk] = p / ktub1[kk];
      }
    });
  };
  // ------------------------------------------------------------ forward
  for (int l = 0; l < L; ++l) {
    if (ckm && l - base == SL) {  // the slots are full: VB_l starts a segment
      d2d(VBh, VB(l), (size_t)MSD, st);
      base = l;
    }
    if (ckm && l % ck == 0) d2d(ckV + (long)(l / ck) * MSD, VB(l), (size_t)MSD, st);
    fwd(l);
  }
  if (loss) {
    const T *bbL = bB(L);
    reduce2(
        st, (long)N * D,
        [=] __device__(long k) {
          const double r = (double)bbL[k] - (double)Yb[k];
          return r * r;
        },
        0L, [=] __device__(long) { return 0.0; },
        [=] __device__(double s1, double) { *loss = s1; });
  }
  // ------------------------------------------------------------ adjoint
  T *ABbar = wb.make(MSD), *Tw = wb.make(MSD), *Xb = wb.make(MSD), *tmp = wb.make(MSD);
  T *VBbar = wb.make(MSD), *UBbar = wb.make(MSD);
  launch_map(st, MSD, [=] __device__(long k) {
    ABbar[k] = T(0);
    Tw[k] = T(0);
    tmp[k] = T(0);
    Xb[k] = T(0);
  });
  for (int l = L; l > 0; --l) {
    if (ckm && l == base) {
      // step back into the previous segment: its iterates from its checkpoint
      base = l - ck;
      d2d(VB(base), ckV + (long)(base / ck) * MSD, (size_t)MSD, st);
      for (int q = base; q < l; ++q) fwd(q);
    }
    const T *ktubl = KTUB(l), *vbl = VB(l), *bbl = bB(l), *ubl = UB(l);
    const T *kvblm1 = KVB(l - 1);
    if (l == L) {
      // bBbar = 2 (bB_L - y); tmp = (bBbar w^T) .* VB_L;
      // wBbar += log(KTUB_L)^T (bBbar .* bB_L)
      launch_map(st, (long)N * SD, [=] __device__(long k) {
        const long i = k % N, c = k / N, d = c / S, s = c % S;
        const long kk = i + c * ldn;
        const T bbar = T(2) * (bbl[i + d * N] - Yb[i + d * N]);
        tmp[kk] = (bbar * Wb[s + d * S]) * vbl[kk];
        Tw[kk] += log(ktubl[kk]) * (bbar * bbl[i + d * N]);
      });
    } else {
      // VBbar = -K^T Xb
      gemm_tn<T, WdlCfg<T>>(st, N, SD, N, K, ldk, Xb, ldn,
                 [=] __device__(int j, int c, T v) { VBbar[j + (long)c * ldn] = -v; });
      // bBbar = rowsum(VBbar / KTUB); tmp = (bBbar w^T - VBbar / KTUB) .* VB_l
      launch_map(st, (long)N * D, [=] __device__(long k) {
        const long i = k % N, d = k / N;
        T bb = T(0);
        for (int s = 0; s < S; ++s) {
          const long kk = i + (d * S + s) * ldn;
          bb += VBbar[kk] / ktubl[kk];
        }
        for (int s = 0; s < S; ++s) {
          const long kk = i + (d * S + s) * ldn;
          tmp[kk] = (bb * Wb[s + d * S] - VBbar[kk] / ktubl[kk]) * vbl[kk];
          Tw[kk] += log(ktubl[kk]) * (bb * bbl[i + d * N]);
        }
      });
    }
    // UBbar = K tmp ; ABbar += UBbar / KVB_{l-1} ; Xb = (UBbar .* UB_l) / KVB_{l-1}
    // (the product stores UBbar; the update streams over it afterwards)
    gemm_tn<T, WdlCfg<T>>(st, N, SD, N, KT, ldk, tmp, ldn,
                          [=] __device__(int i, int c, T v) { UBbar[i + (long)c * ldn] = v; });
    launch_map(st, (long)N * SD, [=] __device__(long k) {
      const long i = k % N, c = k / N, kk = i + c * ldn;
      const T v = UBbar[kk];
      ABbar[kk] += v / kvblm1[kk];
      Xb[kk] = (v * ubl[kk]) / kvblm1[kk];
    });
  }
  // ------------------------------------------------------------ softmax Jacobians
  DBuf<double> dotb(SD), wbar(SD);
  double *dot = dotb.get(), *wB = wbar.get();
  hipLaunchKernelGGL(k_coldot<T>, dim3(SD), dim3(256), 0, st, A, ABbar, ldn, N, S, dot);
  check_launch();
  launch_map(st, (long)N * S, [=] __device__(long k) {
    const long i = k % N, s = k / N;
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
      const long c = d * S + s;
      const double as = (double)A[i + s * N];
      acc += as * (double)ABbar[i + c * ldn] - as * dot[c];
    }
    gAlpha[k] = acc / D;
  });
  hipLaunchKernelGGL((k_colsum_to<T, double>), dim3(SD), dim3(256), 0, st, Tw, ldn, N, wB);
  check_launch();
  hipLaunchKernelGGL(k_lambda_grad<T>, dim3(S), dim3(256), 0, st, Wb, wB, S, D, glambda);
  check_launch();
}

template void wdl_batch_grad<double>(hipStream_t, const double *, const double *, long, int, int,
                                     int, int, const double *, const double *, const double *,
                                     double *, double *, double *, int);
template void wdl_batch_grad<float>(hipStream_t, const float *, const float *, long, int, int,
                                    int, int, const float *, const float *, const float *,
                                    double *, double *, double *, int);

namespace {

// column softmax with max shift (inst/include/wdl_impl.hpp:80-96), one
// thread per column (columns are short: N or S entries)
__global__ void k_softmax_cols(const double *X, double *Y, int m, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double *x = X + (long)j * m;
  double *y = Y + (long)j * m;
  double mx = -INFINITY;
  for (int i = 0; i < m; ++i) mx = fmax(mx, x[i]);
  double tot = 0.0;
  for (int i = 0; i < m; ++i) y[i] = exp(x[i] - mx);
  for (int i = 0; i < m; ++i) tot += y[i];
  for (int i = 0; i < m; ++i) y[i] = y[i] / tot;
}

// the same for long columns (Alpha's N entries): one block per column
__global__ void __launch_bounds__(256) k_softmax_cols_block(const double *X, double *Y, int m) {
  __shared__ double sh[256 / kWave];
  const double *x = X + (long)blockIdx.x * m;
  double *y = Y + (long)blockIdx.x * m;
  double mx = -INFINITY;
  for (int i = threadIdx.x; i < m; i += blockDim.x) mx = fmax(mx, x[i]);
  mx = -block_min(-mx, sh);
  double tot = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double e = exp(x[i] - mx);
    y[i] = e;
    tot += e;
  }
  tot = block_sum(tot, sh);
  for (int i = threadIdx.x; i < m; i += blockDim.x) y[i] = y[i] / tot;
}

void softmax_cols(hipStream_t st, const double *X, double *Y, int m, int n) {
  if (m >= 256)
    hipLaunchKernelGGL(k_softmax_cols_block, dim3(n), dim3(256), 0, st, X, Y, m);
  else
    hipLaunchKernelGGL(k_softmax_cols, dim3((n + 255) / 256), dim3(256), 0, st, X, Y, m, n);
  check_launch();
}

// inst/include/optimizer.hpp:72-99 (opt 0 sgd, 1 adam, 2 adamw)
void optimizer_step(hipStream_t st, int opt, double *theta, const double *g, double *m,
                    double *v, long n, int step, double eta, double gamma, double beta1,
                    double beta2, double eps) {
  const double c1 = 1.0 - std::pow(beta1, step), c2 = 1.0 - std::pow(beta2, step);
  launch_map(st, n, [=] __device__(long k) {
    if (opt == 0) {
      theta[k] = theta[k] - eta * g[k];
      return;
    }
    const double mk = beta1 * m[k] + (1.0 - beta1) * g[k];
    const double vk = beta2 * v[k] + (1.0 - beta2) * pow(g[k], 2.0);
    m[k] = mk;
    v[k] = vk;
    const double upd = eta * ((mk / c1) / (sqrt(vk / c2) + eps));
    theta[k] = (opt == 1) ? theta[k] - upd : (1.0 - eta * gamma) * theta[k] - upd;
  });
}

template <typename T> void cast_to(hipStream_t st, const double *src, T *dst, long n) {
  launch_map(st, n, [=] __device__(long k) { dst[k] = (T)src[k]; });
}

}  // namespace

// Batched barycenter forward (Barycenter::_fwd_parallel without gradient,
// src/barycenter_impl_parallel.cpp:60-151) for the columns of W (S x Dn):
// every document is a barycenter of the topics A with its weights; each
// document stops at its own iteration (err < tol), as the reference's
// per-document solve does. Writes the normalised barycenters to Yhat (N x Dn).
template <typename T>
void wdl_infer(hipStream_t st, const T *K, const T *KT, long ldk, int N, int S, int Dn,
               const T *A, const T *Wn, int maxiter, double tol, double *Yhat) {
  const int SD = S * Dn;
  const long ldn = (N + 7) / 8 * 8;
  const long MSD = ldn * SD;
  WdlBuf<T> wb;
  T *U = wb.make(MSD), *V = wb.make(MSD), *KV = wb.make(MSD), *KTU = wb.make(MSD);
  T *bb = wb.make((size_t)N * Dn);
  DBuf<int> doneb(Dn);
  DBuf<double> errb(Dn);
  int *done = doneb.get();
  double *errv = errb.get();
  launch_map(st, MSD, [=] __device__(long k) { V[k] = (k % ldn < N) ? T(1) : T(0); });
  launch_map(st, (long)N * Dn, [=] __device__(long k) { bb[k] = T(0); });
  launch_map(st, (long)Dn, [=] __device__(long d) { done[d] = 0; });
  gemm_tn<T, WdlCfg<T>>(st, N, SD, N, KT, ldk, V, ldn,
             [=] __device__(int i, int c, T v) { KV[i + (long)c * ldn] = v; });
  for (int it = 0; it < maxiter; ++it) {
    // U = A / KV (running documents)
    launch_map(st, (long)N * SD, [=] __device__(long k) {
      const long i = k % N, c = k / N;
      if (done[c / S]) return;
      U[i + c * ldn] = A[i + (c % S) * N] / KV[i + c * ldn];
    });
    gemm_tn<T, WdlCfg<T>>(st, N, SD, N, K, ldk, U, ldn,
               [=] __device__(int j, int c, T v) { KTU[j + (long)c * ldn] = v; });
    launch_map(st, (long)N * Dn, [=] __device__(long k) {
      const long i = k % N, d = k / N;
      if (done[d]) return;
      const T p = weighted_geomean(KTU + i + d * S * ldn, ldn, Wn + d * S, S);
      bb[k] = p;
      for (int s = 0; s < S; ++s) V[i + (d * S + s) * ldn] = p / KTU[i + (d * S + s) * ldn];
    });
    gemm_tn<T, WdlCfg<T>>(st, N, SD, N, KT, ldk, V, ldn, [=] __device__(int i, int c, T v) {
      if (!done[c / S]) KV[i + (long)c * ldn] = v;
    });
    // per-document err = ||U .* KV - A||_F ; stop when err < tol
    hipLaunchKernelGGL(k_infer_err<T>, dim3(Dn), dim3(256), 0, st, U, KV, ldn, A, N, S, it + 1,
                       maxiter, tol, done, errv);
    check_launch();
  }
  // b / sum(b)
  hipLaunchKernelGGL(k_infer_normalise<T>, dim3(Dn), dim3(256), 0, st, bb, N, Yhat);
  check_launch();
}

// Buffers of one WDL step: the softmax images of Alpha (N x S) and Lambda
// (S x M) in FP64 and in T, the gradients, and g_Lambda (S x M).
template <typename T> struct WdlStepBufs {
  DBuf<double> Ad, Wd, gA, gl, gL;
  DBuf<T> Ab, Wt;
  WdlStepBufs(int N, int S, int M)
      : Ad((size_t)N * S), Wd((size_t)S * M), gA((size_t)N * S), gl(S), gL((size_t)S * M),
        Ab((size_t)N * S), Wt((size_t)S * M) {}
};

// One step of the batch loop (src/wdl_impl.cpp:316-337): A = softmax(Alpha),
// W = softmax(Lambda); the batched gradient of documents [first, first + D)
// of Y (N x M in T); g_Lambda = g_lambda 1^T over all M documents
// (src/wdl_impl.cpp:246); the optimizer step `step` (1-based) on Alpha and
// Lambda with their moments (mA, vA: N x S; mL, vL: S x M).
template <typename T>
void wdl_step(hipStream_t st, const T *K, const T *KT, long ldk, const T *Y, int N, int M,
              int S, int L, int first, int D, double *Alpha, double *Lambda, double *mA,
              double *vA, double *mL, double *vL, int step, int opt, double eta, double gamma,
              double beta1, double beta2, double eps, double *loss, int ck,
              WdlStepBufs<T> &b) {
  softmax_cols(st, Alpha, b.Ad.get(), N, S);
  softmax_cols(st, Lambda, b.Wd.get(), S, M);
  cast_to<T>(st, b.Ad.get(), b.Ab.get(), (long)N * S);
  cast_to<T>(st, b.Wd.get() + (long)first * S, b.Wt.get(), (long)S * D);
  wdl_batch_grad<T>(st, K, KT, ldk, N, S, D, L, b.Ab.get(), Y + (long)first * N, b.Wt.get(),
                    b.gA.get(), b.gl.get(), loss, ck);
  {
    double *gLp = b.gL.get();
    const double *glp = b.gl.get();
    launch_map(st, (long)S * M, [=] __device__(long k) { gLp[k] = glp[k % S]; });
  }
  optimizer_step(st, opt, Alpha, b.gA.get(), mA, vA, (long)N * S, step, eta, gamma, beta1,
                 beta2, eps);
  optimizer_step(st, opt, Lambda, b.gL.get(), mL, vL, (long)S * M, step, eta, gamma, beta1,
                 beta2, eps);
}

// The whole WDL fit (src/wdl_impl.cpp:254-378) from explicit Alpha0, Lambda0.
template <typename T>
void wdl_fit(hipStream_t st, const double *Y, const double *C, int N, int M, int S, double reg,
             int max_iter, double zero_tol, int Bsz, int E, int opt, double eta, double gamma,
             double beta1, double beta2, double eps, const double *Alpha0,
             const double *Lambda0, double *A_out, double *W_out, double *Yhat_out, int ck) {
  const long ldk = (N + 7) / 8 * 8;
  DBuf<T> Kb(ldk * N), KTb(ldk * N), Yt((size_t)N * M);
  DBuf<double> dC((size_t)N * N), dY((size_t)N * M), Alpha((size_t)N * S),
      Lambda((size_t)S * M), mA((size_t)N * S), vA((size_t)N * S), mL((size_t)S * M),
      vL((size_t)S * M), Yh((size_t)N * M);
  WdlStepBufs<T> b(N, S, M);
  h2d(dC.get(), C, (size_t)N * N, st);
  h2d(dY.get(), Y, (size_t)N * M, st);
  h2d(Alpha.get(), Alpha0, (size_t)N * S, st);
  h2d(Lambda.get(), Lambda0, (size_t)S * M, st);
  {
    T *K = Kb.get(), *KT = KTb.get();
    const double *Cp = dC.get();
    launch_map(st, ldk * N, [=] __device__(long k) {
      const long i = k % ldk, j = k / ldk;
      K[k] = (i < N) ? (T)exp(-Cp[i + j * N] / reg) : T(0);   // K(i,j)
      KT[k] = (i < N) ? (T)exp(-Cp[j + i * N] / reg) : T(0);  // K(j,i)
    });
  }
  cast_to<T>(st, dY.get(), Yt.get(), (long)N * M);
  double *pm[4] = {mA.get(), vA.get(), mL.get(), vL.get()};
  const long pn[4] = {(long)N * S, (long)N * S, (long)S * M, (long)S * M};
  for (int q = 0; q < 4; ++q) {
    double *p = pm[q];
    launch_map(st, pn[q], [=] __device__(long k) { p[k] = 0.0; });
  }
  const int batches = (M % Bsz) ? (M / Bsz + 1) : (M / Bsz);
  int step = 0;
  for (int e = 0; e < E; ++e) {
    for (int bid = 0; bid < batches; ++bid) {
      // the reference's batch size rule (src/wdl_impl.cpp:91)
      const int D = (bid == (M / Bsz)) ? (M % Bsz) : Bsz;
      ++step;
      wdl_step<T>(st, Kb.get(), KTb.get(), ldk, Yt.get(), N, M, S, max_iter, bid * Bsz, D,
                  Alpha.get(), Lambda.get(), mA.get(), vA.get(), mL.get(), vL.get(), step, opt,
                  eta, gamma, beta1, beta2, eps, nullptr, ck, b);
    }
  }
  // A and W of the final parameters (src/wdl_impl.cpp:336)
  DBuf<double> &Ad = b.Ad, &Wd = b.Wd;
  DBuf<T> &Ab = b.Ab, &Wt = b.Wt;
  softmax_cols(st, Alpha.get(), Ad.get(), N, S);
  softmax_cols(st, Lambda.get(), Wd.get(), S, M);
  if (A_out) d2h(A_out, Ad.get(), (size_t)N * S, st);
  if (W_out) d2h(W_out, Wd.get(), (size_t)S * M, st);
  if (Yhat_out) {
    cast_to<T>(st, Ad.get(), Ab.get(), (long)N * S);
    cast_to<T>(st, Wd.get(), Wt.get(), (long)S * M);
    for (int m0 = 0; m0 < M; m0 += Bsz) {
      const int Dn = std::min(Bsz, M - m0);
      wdl_infer<T>(st, Kb.get(), KTb.get(), ldk, N, S, Dn, Ab.get(), Wt.get() + (long)m0 * S,
                   max_iter, zero_tol, Yh.get() + (long)m0 * N);
    }
    d2h(Yhat_out, Yh.get(), (size_t)N * M, st);
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace rwig

using namespace rwig;

namespace {

void wdl_check(int N, int M, int S, double reg, int max_iter, int B, int E, int opt) {
  RWIG_REQUIRE(N > 0 && M > 0 && S > 0, "dimensions must be positive");
  RWIG_REQUIRE(N > S, "Number of topics S must be smaller than the vocab size N!");
  RWIG_REQUIRE(opt >= 0 && opt <= 2, "optimizer must be: 0, 1, 2!");
  RWIG_REQUIRE(reg > 0.0 && max_iter >= 1 && B >= 1 && E >= 0, "invalid argument");
}

}  // namespace

extern "C" {

// Replaces cuda_wdl (inst/include/cuda_interface.cuh:31-39,
// src/cuda_wdl.cu:539-818), same argument order; Alpha and Lambda are
// initialised from std::mt19937(seed) standard normals (Alpha first, column
// major, then Lambda), as the reference's CUDA path does
// (src/cuda_wdl.cu:592-593, :677-689). FP64 throughout.
int rwig_wdl(double *A, double *W, double *Yhat, const double *Y, const double *C, int N, int M,
             int S, double reg, int max_iter, double zero_tol, int B, int E, int opt, double eta,
             double gamma, double beta1, double beta2, double eps, int seed, int verbose) {
  return guarded("rwig_wdl", [&] {
    wdl_check(N, M, S, reg, max_iter, B, E, opt);
    RWIG_REQUIRE(Y && C, "Y and C are required");
    ensure_device();
    std::mt19937 rng(seed);
    std::normal_distribution<double> randn(0.0, 1.0);
    std::vector<double> Alpha0((size_t)N * S), Lambda0((size_t)S * M);
    for (auto &x : Alpha0) x = randn(rng);
    for (auto &x : Lambda0) x = randn(rng);
    Stream s;
    wdl_fit<double>(s.get(), Y, C, N, M, S, reg, max_iter, zero_tol, B, E, opt, eta, gamma,
                    beta1, beta2, eps, Alpha0.data(), Lambda0.data(), A, W, Yhat, 0);
    (void)verbose;
  });
}

// WDL fit from explicit initial Alpha0 (N x S) and Lambda0 (S x M), the
// parameters the reference draws with R's RNG (src/wdl_impl.cpp:277-278).
// precision 0: FP64, 1: FP32 GEMMs and recurrences (optimizer and softmax in
// FP64). checkpoint_every: 0 keeps every iterate for the adjoint, k > 0 every
// k-th (same result, less memory, more GEMMs).
int rwig_wdl_init(double *A, double *W, double *Yhat, const double *Y, const double *C, int N,
                  int M, int S, double reg, int max_iter, double zero_tol, int B, int E, int opt,
                  double eta, double gamma, double beta1, double beta2, double eps,
                  const double *Alpha0, const double *Lambda0, int precision,
                  int checkpoint_every) {
  return guarded("rwig_wdl_init", [&] {
    wdl_check(N, M, S, reg, max_iter, B, E, opt);
    RWIG_REQUIRE(Y && C && Alpha0 && Lambda0, "Y, C, Alpha0 and Lambda0 are required");
    RWIG_REQUIRE(checkpoint_every >= 0, "checkpoint_every must be non-negative");
    ensure_device();
    Stream s;
    if (precision == 0)
      wdl_fit<double>(s.get(), Y, C, N, M, S, reg, max_iter, zero_tol, B, E, opt, eta, gamma,
                      beta1, beta2, eps, Alpha0, Lambda0, A, W, Yhat, checkpoint_every);
    else if (precision == 1)
      wdl_fit<float>(s.get(), Y, C, N, M, S, reg, max_iter, zero_tol, B, E, opt, eta, gamma,
                     beta1, beta2, eps, Alpha0, Lambda0, A, W, Yhat, checkpoint_every);
    else
      throw Error(RWIG_ERR_INVALID, "precision must be 0 (fp64) or 1 (fp32)");
  });
}

// One batched WDL gradient on device pointers (src/wdl_impl.cpp:87-247):
// K = exp(-C/reg) and KT its transpose (column-major, leading dimension ldk,
// a multiple of 8 elements), A (N x S), Yb (N x D), Wb (S x D) in the
// precision given (0: double, 1: float); gAlpha (N x S), glambda (S) and
// loss (sum of squared residuals of the unnormalised b_L, may be null) FP64.
// checkpoint_every as in rwig_wdl_init.
int rwig_wdl_batch_grad_dev(const void *K, const void *KT, long ldk, int N, int S, int D, int L,
                            const void *A, const void *Yb, const void *Wb, double *gAlpha,
                            double *glambda, double *loss, int precision, int checkpoint_every,
                            void *stream) {
  return guarded("rwig_wdl_batch_grad_dev", [&] {
    RWIG_REQUIRE(K && KT && A && Yb && Wb && gAlpha && glambda, "null argument");
    RWIG_REQUIRE(N > 0 && S > 0 && D > 0 && L > 0 && ldk >= N && ldk % 8 == 0,
                 "invalid dimensions");
    RWIG_REQUIRE(checkpoint_every >= 0, "checkpoint_every must be non-negative");
    hipStream_t st = (hipStream_t)stream;
    if (precision == 0)
      wdl_batch_grad<double>(st, (const double *)K, (const double *)KT, ldk, N, S, D, L,
                             (const double *)A, (const double *)Yb, (const double *)Wb, gAlpha,
                             glambda, loss, checkpoint_every);
    else if (precision == 1)
      wdl_batch_grad<float>(st, (const float *)K, (const float *)KT, ldk, N, S, D, L,
                            (const float *)A, (const float *)Yb, (const float *)Wb, gAlpha,
                            glambda, loss, checkpoint_every);
    else
      throw Error(RWIG_ERR_INVALID, "precision must be 0 (fp64) or 1 (fp32)");
  });
}

// One training step on device buffers, the body of the reference's batch loop
// (src/wdl_impl.cpp:316-337): A = softmax(Alpha) and W = softmax(Lambda) by
// columns; the batched gradient of documents [first, first + D) of Y;
// g_Lambda = g_lambda 1^T over all M documents; optimizer step `step` (1 for
// the first, as the reference counts) on Alpha (N x S) and Lambda (S x M),
// with their moments mA, vA (N x S) and mL, vL (S x M) updated in place
// (zero before the first step). K, KT (leading dimension ldk) and Y (N x M)
// in the precision given (0: double, 1: float); Alpha, Lambda, the moments
// and loss (the batch's sum of squared residuals, may be null) FP64.
// A fit of E epochs is the sequence of these steps over the batches; it
// gives rwig_wdl_init's A and W bit for bit.
int rwig_wdl_step_dev(const void *K, const void *KT, long ldk, const void *Y, int N, int M,
                      int S, int L, int first, int D, double *Alpha, double *Lambda, double *mA,
                      double *vA, double *mL, double *vL, int step, int optimizer, double eta,
                      double gamma, double beta1, double beta2, double eps, double *loss,
                      int precision, int checkpoint_every, void *stream) {
  return guarded("rwig_wdl_step_dev", [&] {
    RWIG_REQUIRE(K && KT && Y && Alpha && Lambda && mA && vA && mL && vL, "null argument");
    wdl_check(N, M, S, 1.0, L, std::max(D, 1), 1, optimizer);
    RWIG_REQUIRE(ldk >= N && ldk % 8 == 0, "ldk must be >= N and a multiple of 8");
    RWIG_REQUIRE(D > 0 && first >= 0 && first + D <= M, "the batch must lie within the M documents");
    RWIG_REQUIRE(step >= 1, "step counts from 1");
    RWIG_REQUIRE(checkpoint_every >= 0, "checkpoint_every must be non-negative");
    hipStream_t st = (hipStream_t)stream;
    if (precision == 0) {
      WdlStepBufs<double> b(N, S, M);
      wdl_step<double>(st, (const double *)K, (const double *)KT, ldk, (const double *)Y, N, M,
                       S, L, first, D, Alpha, Lambda, mA, vA, mL, vL, step, optimizer, eta,
                       gamma, beta1, beta2, eps, loss, checkpoint_every, b);
    } else if (precision == 1) {
      WdlStepBufs<float> b(N, S, M);
      wdl_step<float>(st, (const float *)K, (const float *)KT, ldk, (const float *)Y, N, M, S,
                      L, first, D, Alpha, Lambda, mA, vA, mL, vL, step, optimizer, eta, gamma,
                      beta1, beta2, eps, loss, checkpoint_every, b);
    } else {
      throw Error(RWIG_ERR_INVALID, "precision must be 0 (fp64) or 1 (fp32)");
    }
  });
}

}  // extern "C"
