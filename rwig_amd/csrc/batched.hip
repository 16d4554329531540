// Batched log-stabilised Sinkhorn, one source measure against many target
// measures (the C2 workload: one a against 1024 sparse b's), resident form:
// one workgroup owns one target's whole solve and runs its iteration loop to
// the target's own stopping point inside a single kernel launch.
//
// Per iteration the reference (Sinkhorn::_fwd_log,
// src/sinkhorn_impl_log.cpp:327-375) makes three soft-min passes and two R
// writes over the reduced cost. Here one sweep over the target's cost
// columns does all of it: for each column the column soft-min of
// R(f, g) gives the new g_j (src/sinkhorn_impl_log.cpp:348-349), and the same
// column values, still in registers, feed an online row soft-min of
// R(f, g_new) (the minrowcol pass, :359) that is the next f update's input.
// The column half of the reference's error, computed from R(f, g_new), is the
// column soft-min of R(f, g) shifted by g_new - g; it is evaluated in that
// form.
//
// Layout: C is column-major with leading dimension ldc; the rows are the
// source points (all of them carry mass), target t owns the columns
// col_idx[col_ptr[t] .. col_ptr[t+1]) with masses b[] at the same positions.
// Thread k of a workgroup owns rows k, k + TB, k + 2 TB, ... so every column
// read is one contiguous, coalesced stretch of the column.
#include "kernels.hpp"

namespace rwig {

namespace {

constexpr int kBTB = 512;  // threads per workgroup of the resident kernels
constexpr int kBW = kBTB / kWave;

// err_mode 0: the reference's ||-Rminrow/reg - log a|| + ||-Rmincol/reg - log b||
// err_mode 1: marginal L1 error sum|P 1 - a| + sum|P^T 1 - b| (P1 = exp(-Rminrow/reg))
template <int RPT, int CB>
__global__ void __launch_bounds__(kBTB)
k_sinkhorn_log_resident(const double *__restrict__ C, long ldc, int M,
                        const double *__restrict__ a, const int *__restrict__ col_ptr,
                        const int *__restrict__ col_idx, const double *__restrict__ b,
                        double reg, int maxiter, double tol, int err_mode,
                        double *__restrict__ Fout, double *__restrict__ Gout,
                        int *__restrict__ iters, double *__restrict__ errs,
                        double *__restrict__ loss) {
  extern __shared__ double smem[];
  double *gs = smem;                          // target's g, nt entries
  double *red = smem + (col_ptr[blockIdx.x + 1] - col_ptr[blockIdx.x]);  // kBW * CB * 2
  __shared__ double redsum[kBW];

  const int t = blockIdx.x;
  const int c0 = col_ptr[t], nt = col_ptr[t + 1] - c0;
  // exp(-x / reg) of the per-entry terms as exp2(-x * sc): one multiply in
  // place of an FP64 division per term
  const double sc = kLog2e / reg;
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;

  double f[RPT], rmin[RPT];
  bool own[RPT];
#pragma unroll
  for (int k = 0; k < RPT; ++k) {
    const int i = threadIdx.x + k * kBTB;
    own[k] = i < M;
    f[k] = 0.0;
  }
  for (int q = threadIdx.x; q < nt; q += kBTB) gs[q] = 0.0;
  __syncthreads();

  // one sweep over the columns. With update_g, the column soft-min of R(f, g)
  // updates g; every sweep ends with the row soft-min of R(f, g) in rmin.
  // Returns (thread-uniform) the column part of the error.
  auto sweep = [&](bool update_g) -> double {
    double rmn[RPT], rsm[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      rmn[k] = INFINITY;
      rsm[k] = 0.0;
    }
    double colacc = 0.0;  // lanes 0..CB-1: error terms of their columns
    for (int jb = 0; jb < nt; jb += CB) {
      double c[RPT][CB], gq[CB], lbq[CB], bq[CB];
#pragma unroll
      for (int q = 0; q < CB; ++q) {
        const int jj = jb + q;
        const bool ok = jj < nt;
        const long col = ok ? (long)col_idx[c0 + jj] : 0;
        gq[q] = ok ? gs[jj] : 0.0;
        bq[q] = ok ? b[c0 + jj] : 1.0;
        lbq[q] = log(bq[q]);
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const int i = threadIdx.x + k * kBTB;
          c[k][q] = (ok && own[k]) ? C[i + col * ldc] : 0.0;
        }
      }
      if (update_g) {
        // column minima of R(f, g): per-wave minima, then lanes 0..CB-1 of
        // the workgroup finish one column each
        double *mS = red + 2 * kBW * CB;  // CB column minima
        double *gS2 = mS + CB;            // CB new g values
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          double m = INFINITY;
#pragma unroll
          for (int k = 0; k < RPT; ++k)
            if (own[k]) m = fmin(m, (c[k][q] - f[k]) - gq[q]);
          m = wave_min(m);
          if (lane == 0) red[wv * CB + q] = m;
        }
        __syncthreads();
        if (threadIdx.x < CB) {
          const int q = threadIdx.x;
          double m = red[q];
          for (int w = 1; w < kBW; ++w) m = fmin(m, red[w * CB + q]);
          mS[q] = m;
        }
        __syncthreads();
        // column sums of exp(-(R - min) / reg)
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          const double mn = mS[q];
          double sacc = 0.0;
#pragma unroll
          for (int k = 0; k < RPT; ++k)
            if (own[k]) sacc += exp2_f64(-(((c[k][q] - f[k]) - gq[q]) - mn) * sc);
          sacc = wave_sum(sacc);
          if (lane == 0) red[kBW * CB + wv * CB + q] = sacc;
        }
        __syncthreads();
        if (threadIdx.x < CB) {
          const int q = threadIdx.x;
          double gn = gq[q];
          if (jb + q < nt) {
            double sacc = 0.0;
            for (int w = 0; w < kBW; ++w) sacc += red[kBW * CB + w * CB + q];
            const double rmc = mS[q] - reg * log(sacc);     // Rmincol of R(f, g)
            gn = gq[q] + (reg * lbq[q] + rmc);              // g update
            const double rmc2 = rmc - (gn - gq[q]);         // Rmincol of R(f, g_new)
            double e;
            if (err_mode == 0) {
              const double r = -rmc2 / reg - lbq[q];
              e = r * r;
            } else {
              e = fabs(exp(-rmc2 / reg) - bq[q]);
            }
            colacc += e;
            gs[jb + q] = gn;
          }
          gS2[q] = gn;
        }
        __syncthreads();  // also orders the reuse of red by the next chunk
#pragma unroll
        for (int q = 0; q < CB; ++q) gq[q] = gS2[q];
      }
      // online row soft-min of R(f, g) over this chunk
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        if (!own[k]) continue;
        double m = rmn[k];
#pragma unroll
        for (int q = 0; q < CB; ++q)
          if (jb + q < nt) m = fmin(m, (c[k][q] - f[k]) - gq[q]);
        double s = (m == rmn[k]) ? rsm[k] : rsm[k] * exp2_f64(-(rmn[k] - m) * sc);
#pragma unroll
        for (int q = 0; q < CB; ++q)
          if (jb + q < nt) s += exp2_f64(-(((c[k][q] - f[k]) - gq[q]) - m) * sc);
        rmn[k] = m;
        rsm[k] = s;
      }
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) rmin[k] = own[k] ? rmn[k] - reg * log(rsm[k]) : 0.0;
    __syncthreads();
    return colacc;
  };

  auto block_total = [&](double v) -> double {
    v = wave_sum(v);
    if (lane == 0) redsum[wv] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kBW; ++w) s += redsum[w];
    __syncthreads();
    return s;
  };

  sweep(false);  // Rminrow of R(0, 0) (src/sinkhorn_impl_log.cpp:325)
  int iter = 0;
  double err = 1000.0;
  while (iter < maxiter && err >= tol) {
    ++iter;
#pragma unroll
    for (int k = 0; k < RPT; ++k)
      if (own[k]) f[k] = f[k] + (reg * log(a[threadIdx.x + k * kBTB]) + rmin[k]);
    const double colpart = sweep(true);
    double rowerr = 0.0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      if (!own[k]) continue;
      if (err_mode == 0) {
        const double r = -rmin[k] / reg - log(a[threadIdx.x + k * kBTB]);
        rowerr += r * r;
      } else {
        rowerr += fabs(exp(-rmin[k] / reg) - a[threadIdx.x + k * kBTB]);
      }
    }
    const double rt = block_total(rowerr);
    const double ct = block_total(colpart);
    err = (err_mode == 0) ? sqrt(rt) + sqrt(ct) : rt + ct;
  }
#pragma unroll
  for (int k = 0; k < RPT; ++k)
    if (own[k]) Fout[(long)t * M + threadIdx.x + k * kBTB] = f[k];
  for (int q = threadIdx.x; q < nt; q += kBTB) Gout[c0 + q] = gs[q];
  double cost = 0.0, regloss = 0.0;
  if (loss) {
    // transport cost sum C.*P and regularised loss sum C.*P + reg sum P.*(log P - 1)
    // of P = exp(-R(f, g)/reg) (src/sinkhorn_impl_log.cpp:266, :277)
    double cc = 0.0, ce = 0.0;
    for (int j = 0; j < nt; ++j) {
      const long col = col_idx[c0 + j];
      const double gq = gs[j];
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        if (!own[k]) continue;
        const double cij = C[threadIdx.x + k * kBTB + col * ldc];
        const double rr = (cij - f[k]) - gq;
        const double p = exp(-rr / reg);
        cc += cij * p;
        ce += p * (-rr / reg - 1.0);
      }
    }
    cost = block_total(cc);
    regloss = cost + reg * block_total(ce);
  }
  if (threadIdx.x == 0) {
    iters[t] = iter;
    errs[t] = err;
    if (loss) {
      loss[2 * t] = cost;
      loss[2 * t + 1] = regloss;
    }
  }
}

}  // namespace

void sinkhorn_log_batched_fp64(hipStream_t st, const double *C, long ldc, int M,
                               const double *a, int B, const int *col_ptr, const int *col_idx,
                               const double *b, int max_nt, double reg, int maxiter, double tol,
                               int err_mode, double *F, double *G, int *iters, double *errs,
                               double *loss) {
  RWIG_REQUIRE(M >= 1 && M <= 16 * kBTB, "batched solve supports 1..8192 source points");
  const size_t shmem = sizeof(double) * ((size_t)max_nt + 2 * kBW * 8 + 2 * 8);
  RWIG_REQUIRE(max_nt >= 1 && shmem <= 65536, "targets must have 1..8000 support points");
  const dim3 grid(B), block(kBTB);
#define RWIG_LAUNCH_LOG(RPT_, CB_)                                                            \
  hipLaunchKernelGGL((k_sinkhorn_log_resident<RPT_, CB_>), grid, block, shmem, st, C, ldc, M, \
                     a, col_ptr, col_idx, b, reg, maxiter, tol, err_mode, F, G, iters, errs,  \
                     loss)
  if (M <= kBTB) RWIG_LAUNCH_LOG(1, 8);
  else if (M <= 2 * kBTB) RWIG_LAUNCH_LOG(2, 4);
  else if (M <= 4 * kBTB) RWIG_LAUNCH_LOG(4, 4);
  else if (M <= 8 * kBTB) RWIG_LAUNCH_LOG(8, 2);
  else RWIG_LAUNCH_LOG(16, 1);
#undef RWIG_LAUNCH_LOG
  check_launch();
}

void gibbs_lowp(hipStream_t st, const double *C, long ldc, int M, int N, double reg,
                int precision, void *K, long ldk);
void sinkhorn_batched_fast(hipStream_t st, const double *C, long ldc, const void *K, long ldk,
                           int precision, int M, const double *a, int B, const int *col_ptr,
                           const int *col_idx, const double *b, int max_nt, double reg,
                           int maxiter, double tol, int err_mode, int absorb_every, double *F,
                           double *G, int *iters, double *errs, double *loss);

}  // namespace rwig

using namespace rwig;

extern "C" {

int rwig_gibbs_lowp_dev(const double *C, long ldc, int M, int N, double reg, int precision,
                        void *K, long ldk, void *stream) {
  return guarded("rwig_gibbs_lowp_dev", [&] {
    RWIG_REQUIRE(C && K && M > 0 && N > 0 && reg > 0.0, "invalid argument");
    gibbs_lowp((hipStream_t)stream, C, ldc, M, N, reg, precision, K, ldk);
  });
}

int rwig_sinkhorn_batched_dev(const double *C, long ldc, int M, const double *a, int B,
                              const int *col_ptr, const int *col_idx, const double *b,
                              int max_nt, double reg, int max_iter, double zero_tol,
                              int err_mode, int precision, const void *Klo, long ldk,
                              int absorb_every, double *f, double *g, int *iters, double *errs,
                              double *loss, void *stream) {
  return guarded("rwig_sinkhorn_batched_dev", [&] {
    RWIG_REQUIRE(B >= 1, "B must be positive");
    RWIG_REQUIRE(reg > 0.0, "reg must be positive");
    RWIG_REQUIRE(max_iter >= 0, "max_iter must be non-negative");
    RWIG_REQUIRE(err_mode == 0 || err_mode == 1, "err_mode must be 0 or 1");
    RWIG_REQUIRE(C && a && col_ptr && col_idx && b && f && g && iters && errs,
                 "null argument");
    if (precision == 0) {
      sinkhorn_log_batched_fp64((hipStream_t)stream, C, ldc, M, a, B, col_ptr, col_idx, b,
                                max_nt, reg, max_iter, zero_tol, err_mode, f, g, iters, errs,
                                loss);
    } else {
      RWIG_REQUIRE(Klo != nullptr, "precision 1/2 needs the low-precision Gibbs kernel");
      sinkhorn_batched_fast((hipStream_t)stream, C, ldc, Klo, ldk, precision, M, a, B, col_ptr,
                            col_idx, b, max_nt, reg, max_iter, zero_tol, err_mode,
                            absorb_every, f, g, iters, errs, loss);
    }
  });
}

}  // extern "C"
