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

#include "kernels.hpp"

namespace rwig {

namespace {

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
  lse_row<T, X>(
      st, C, M, N, M, F, M, G, N, S, reg, pm, ps,
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
    lse_col_t<T, X>(
        st, CT, M, N, N, F, M, G, N, S, reg, cpm, cps,
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
    lse_row<T, X>(
        st, C, M, N, M, F, M, G, N, S, reg, pm, ps,
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
          wcol_t<T, X>(
              st, CT, M, N, N, Fl, M, Gl, N, S, reg, KVl, Fbar, cpart,
              [=] __device__(int j, const T *r) {
                T acc = T(0);
                for (int s = 0; s < S; ++s) {
                  Gbar[j + (long)s * N] = -r[s];
                  acc += -r[s];
                }
                lbb[j] = treg * acc;
              },
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
        wrow<T, X>(
            st, C, M, N, M, Fl, M, Glm1, N, S, reg, KTUl, Y, part,
            [=] __device__(long k, T q) {
              Fbar[k] = (T)((double)q * A[k]);
              gA[k] += (double)q;
            },
            nullptr, rlogA);
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

}  // namespace

}  // namespace rwig

using namespace rwig;

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
