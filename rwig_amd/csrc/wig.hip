// Truncated SVD and the WIG index: the last steps of wig() (R/wig.R:53-170)
// after WDL has produced topics A, weights W and reconstructions Yhat.
//   tsvd:       the leading k singular triplets with the reference's three
//               sign rules (tsvd_cpp, src/tsvd.cpp:7-86);
//   wig scores: one score per document from a rank-1 TSVD of the topics or
//               of the reconstructed documents (R/wig.R:96-106);
//   wig index:  scores summed per calendar period, optionally standardised
//               (R/wig.R:2-6, :139-159).
//
// The reference computes a full SVD (LAPACK) and keeps k columns. Here only
// the k leading triplets are computed: subspace iteration on G = T^T T for
// the tall orientation T of the matrix (T = X when rows >= columns, else
// X^T), where each step is two FP64 MFMA GEMMs streaming T from HBM (Y = T Q,
// Z = T^T Y) and the Rayleigh-Ritz step on the small block runs on the host.
// Msvd = U_k diag(s_k) is formed as X v (no division by s).
#include <algorithm>
#include <cstring>
#include <map>

#include "gemm.hpp"
#include "kernels.hpp"
#include "tsvd_host.hpp"

namespace rwig {

namespace {

inline long ld8(long n) { return (n + 7) & ~7L; }

// out(c) = sum_r w(r) * (S(r, c) - sum_l Pa(r, l) Pb(c, l)) for the columns c
// of S (rows x cols, column-major): the projections on w of the columns of the
// rank-kk residual S - Pa Pb^T, without forming it. One wave per column,
// lanes over rows (coalesced along the column).
__global__ void __launch_bounds__(256)
k_resid_proj(const double *__restrict__ S, int rows, int cols, long lds,
             const double *__restrict__ w, const double *__restrict__ Pa, long lda,
             const double *__restrict__ Pb, long ldb, int kk, double *__restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int c = blockIdx.x * (256 / kWave) + threadIdx.x / kWave;
  if (c >= cols) return;  // the whole wave leaves together
  double acc = 0.0;
  for (int r = lane; r < rows; r += kWave) {
    double y = S[r + (long)c * lds];
    for (int l = 0; l < kk; ++l) y -= Pa[r + (long)l * lda] * Pb[c + (long)l * ldb];
    acc += w[r] * y;
  }
  acc = wave_sum(acc);
  if (lane == 0) out[c] = acc;
}

void h2d_2d(double *dst, long ldd, const double *src, long lds, long rows, long cols,
            hipStream_t st) {
  if (rows > 0 && cols > 0)
    RWIG_HIP_CHECK(hipMemcpy2DAsync(dst, ldd * sizeof(double), src, lds * sizeof(double),
                                    rows * sizeof(double), cols, hipMemcpyHostToDevice, st));
}
void d2h_2d(double *dst, long ldd, const double *src, long lds, long rows, long cols,
            hipStream_t st) {
  if (rows > 0 && cols > 0)
    RWIG_HIP_CHECK(hipMemcpy2DAsync(dst, ldd * sizeof(double), src, lds * sizeof(double),
                                    rows * sizeof(double), cols, hipMemcpyDeviceToHost, st));
}

// sign(x) with sign(0) = 0 (arma::sign)
inline double sgn(double x) { return (x > 0.0) - (x < 0.0); }

struct TsvdOut {
  std::vector<double> Msvd;  // m x kk
  int kk = 0;
  int iters = 0;
  double residual = 0.0;
};

// tsvd_cpp (src/tsvd.cpp:7-86) for the m x n column-major host matrix X.
TsvdOut tsvd(const double *X, int m, int n, int k, int flip_sign) {
  RWIG_REQUIRE(X && m > 0 && n > 0, "invalid argument");
  RWIG_REQUIRE(k >= 1, "k must be at least 1");
  RWIG_REQUIRE(flip_sign >= 0 && flip_sign <= 2,
               "flip_sign must be 0 (auto), 1 (sklearn) or 2 (none)");
  TsvdOut res;
  const int kk = std::min(k, std::min(m, n));  // src/tsvd.cpp:37
  res.kk = kk;
  ensure_device();
  Stream s;
  const hipStream_t st = s.get();
  // X (m x n) and X^T (n x m) on the device; T is the tall one (R x d, R >= d)
  const long ldx = ld8(m), ldxt = ld8(n);
  DBuf<double> dX((size_t)ldx * n), dXT((size_t)ldxt * m);
  h2d_2d(dX.get(), ldx, X, m, m, n, st);
  transpose(st, dX.get(), m, n, ldx, dXT.get(), ldxt);
  const bool tall = m >= n;
  const int R = tall ? m : n, d = tall ? n : m;
  const double *T = tall ? dX.get() : dXT.get();
  const double *TT = tall ? dXT.get() : dX.get();
  const long ldT = tall ? ldx : ldxt, ldTT = tall ? ldxt : ldx;
  const int bmax = std::min(d, std::max(kk + 16, 32));
  const long ldd = ld8(d), ldr = ld8(R);
  DBuf<double> dQ((size_t)ldd * bmax), dY((size_t)ldr * bmax), dZ((size_t)ldd * bmax);
  const int ks1 = splitk_for<double>(R, d), ks2 = splitk_for<double>(d, R);
  const size_t wn = std::max(ks1 > 1 ? splitk_work_elems<double>(R, bmax, ks1) : 0,
                             ks2 > 1 ? splitk_work_elems<double>(d, bmax, ks2) : 0);
  DBuf<double> work(std::max<size_t>(wn, 1));
  SplitK<double> sk1, sk2;
  if (ks1 > 1) { sk1.ks = ks1; sk1.work = work.get(); }
  if (ks2 > 1) { sk2.ks = ks2; sk2.work = work.get(); }
  // out(R x cols) = T * B (B: d x cols on the device, leading dimension ldd)
  auto times_T = [&](const double *B, int cols, double *out) {
    gemm_tn<double>(st, R, cols, d, TT, ldTT, B, ldd,
                    [=] __device__(int r, int j, double v) { out[r + (long)j * ldr] = v; },
                    nullptr, nullptr, sk1);
  };
  auto apply = [&](const std::vector<double> &Q, std::vector<double> &Z) {
    const int b = (int)(Q.size() / d);
    h2d_2d(dQ.get(), ldd, Q.data(), d, d, b, st);
    times_T(dQ.get(), b, dY.get());
    double *z = dZ.get();
    gemm_tn<double>(st, d, b, R, T, ldT, dY.get(), ldr,
                    [=] __device__(int c, int j, double v) { z[c + (long)j * ldd] = v; },
                    nullptr, nullptr, sk2);
    d2h_2d(Z.data(), d, dZ.get(), ldd, d, b, st);
    s.sync();
  };
  tsvd_host::EigResult eig = tsvd_host::leading_eigvecs(d, kk, apply);
  res.iters = eig.iters;
  res.residual = eig.residual;
  // P = T v_i (R x kk): |P_i| = s_i
  std::vector<double> P((size_t)R * kk);
  h2d_2d(dQ.get(), ldd, eig.V.data(), d, d, kk, st);
  times_T(dQ.get(), kk, dY.get());
  d2h_2d(P.data(), R, dY.get(), ldr, R, kk, st);
  s.sync();
  std::vector<double> sv(kk);
  for (int i = 0; i < kk; ++i)
    sv[i] = std::sqrt(tsvd_host::dot(&P[(size_t)i * R], &P[(size_t)i * R], R));
  // Msvd = U_k diag(s_k), U_k (m x kk), V_k (n x kk)
  std::vector<double> U((size_t)m * kk), V((size_t)n * kk);
  res.Msvd.assign((size_t)m * kk, 0.0);
  for (int i = 0; i < kk; ++i) {
    const double inv = sv[i] > 0.0 ? 1.0 / sv[i] : 0.0;  // s_i = 0: the column is 0 anyway
    for (int r = 0; r < R; ++r) {
      const double p = P[r + (size_t)i * R];
      if (tall) {
        res.Msvd[r + (size_t)i * m] = p;
        U[r + (size_t)i * m] = p * inv;
      } else {
        V[r + (size_t)i * n] = p * inv;
      }
    }
    for (int c = 0; c < d; ++c) {
      const double v = eig.V[c + (size_t)i * d];
      if (tall) {
        V[c + (size_t)i * n] = v;
      } else {
        U[c + (size_t)i * m] = v;
        res.Msvd[c + (size_t)i * m] = v * sv[i];
      }
    }
  }
  if (flip_sign == 0) {
    // Bro, Acar & Kolda (2008) as src/tsvd.cpp:39-71 states it: Y = X - U_k
    // S_k V_k^T with all kk components removed (:41), sk_left = sum_j
    // sign(u_K^T y_j) (u_K^T y_j)^2 over the columns (:48-52), sk_right over
    // the rows (:54-58), the smaller one flipped when they disagree (:60-66).
    const long ldm = ld8(m), ldn = ld8(n);
    DBuf<double> dU((size_t)ldm * kk), dV((size_t)ldn * kk), dM((size_t)ldm * kk),
        dt((size_t)n), dr((size_t)m);
    h2d_2d(dU.get(), ldm, U.data(), m, m, kk, st);
    h2d_2d(dV.get(), ldn, V.data(), n, n, kk, st);
    h2d_2d(dM.get(), ldm, res.Msvd.data(), m, m, kk, st);
    std::vector<double> t(n), r(m), sign_u(kk);
    for (int K = 0; K < kk; ++K) {
      // t_j = u_K^T (X(:, j) - Msvd V(j, :)^T); r_i = v_K^T (X(i, :)^T - V Msvd(i, :)^T)
      hipLaunchKernelGGL(k_resid_proj, dim3((n + 3) / 4), dim3(256), 0, st, dX.get(), m, n, ldx,
                         dU.get() + (long)K * ldm, dM.get(), ldm, dV.get(), ldn, kk, dt.get());
      check_launch();
      hipLaunchKernelGGL(k_resid_proj, dim3((m + 3) / 4), dim3(256), 0, st, dXT.get(), n, m,
                         ldxt, dV.get() + (long)K * ldn, dV.get(), ldn, dM.get(), ldm, kk,
                         dr.get());
      check_launch();
      d2h(t.data(), dt.get(), (size_t)n, st);
      d2h(r.data(), dr.get(), (size_t)m, st);
      s.sync();
      double sk_left = 0.0, sk_right = 0.0;
      for (int j = 0; j < n; ++j) sk_left += sgn(t[j]) * t[j] * t[j];
      for (int i = 0; i < m; ++i) sk_right += sgn(r[i]) * r[i] * r[i];
      if (sk_left * sk_right < 0) {
        if (std::fabs(sk_left) < std::fabs(sk_right))
          sk_left = -sk_left;
        else
          sk_right = -sk_right;
      }
      sign_u[K] = sgn(sk_left);  // a zero statistic zeroes the column, as arma::sign does
    }
    for (int K = 0; K < kk; ++K)
      for (int i = 0; i < m; ++i) res.Msvd[i + (size_t)K * m] *= sign_u[K];
  } else if (flip_sign == 1) {
    // the entry of largest magnitude of each column made positive (:76-80;
    // index_max: the first of equal magnitudes)
    for (int K = 0; K < kk; ++K) {
      double *col = &res.Msvd[(size_t)K * m];
      int imax = 0;
      for (int i = 1; i < m; ++i)
        if (std::fabs(col[i]) > std::fabs(col[imax])) imax = i;
      if (col[imax] < 0)
        for (int i = 0; i < m; ++i) col[i] = -col[i];
    }
  }
  return res;
}

// ---------------------------------------------------------------- calendar

// Proleptic Gregorian calendar <-> days since 1970-01-01 (H. Hinnant,
// "chrono-Compatible Low-Level Date Algorithms": days_from_civil and
// civil_from_days).
long days_from_civil(long y, int m, int d) {
  y -= m <= 2;
  const long era = (y >= 0 ? y : y - 399) / 400;
  const long yoe = y - era * 400;
  const long doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
  const long doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}
void civil_from_days(long z, long &y, int &m, int &d) {
  z += 719468;
  const long era = (z >= 0 ? z : z - 146096) / 146097;
  const long doe = z - era * 146097;
  const long yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const long doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const long mp = (5 * doy + 2) / 153;
  d = (int)(doy - (153 * mp + 2) / 5 + 1);
  m = (int)(mp < 10 ? mp + 3 : mp - 9);
  y = yoe + era * 400 + (m <= 2);
}

// First day of the period holding `day` (lubridate::floor_date on a Date,
// R/wig.R:3): 0 day, 1 week (starting on Sunday, lubridate's default
// week_start = 7), 2 month, 3 bimonth, 4 quarter, 5 halfyear, 6 year,
// 7 season (starting in December, March, June and September).
long floor_period(long day, int unit) {
  if (unit == 0) return day;
  if (unit == 1) {
    const long wd = ((day + 4) % 7 + 7) % 7;  // 1970-01-01 was a Thursday; Sunday = 0
    return day - wd;
  }
  long y;
  int m, dd;
  civil_from_days(day, y, m, dd);
  if (unit == 7) {
    const int s = (m % 12) / 3;
    return s == 0 ? days_from_civil(m == 12 ? y : y - 1, 12, 1) : days_from_civil(y, 3 * s, 1);
  }
  static const int months[] = {0, 0, 1, 2, 3, 6, 12};
  const int span = months[unit];
  return days_from_civil(y, ((m - 1) / span) * span + 1, 1);
}

}  // namespace

}  // namespace rwig

using namespace rwig;

extern "C" {

int rwig_tsvd(double *out, int *k_out, double *residual, const double *M, int m, int n, int k,
              int flip_sign) {
  return guarded("rwig_tsvd", [&] {
    RWIG_REQUIRE(out, "null output");
    TsvdOut r = tsvd(M, m, n, k, flip_sign);
    std::memcpy(out, r.Msvd.data(), r.Msvd.size() * sizeof(double));
    if (k_out) *k_out = r.kk;
    if (residual) *residual = r.residual;
  });
}

int rwig_wig_scores(double *scores, const double *A, const double *W, const double *Yhat,
                    int N, int S, int M, int svd_method) {
  return guarded("rwig_wig_scores", [&] {
    RWIG_REQUIRE(scores && N >= 2 && S >= 1 && M >= 1, "invalid argument");
    RWIG_REQUIRE(svd_method == 0 || svd_method == 1,
                 "svd_method must be 0 (\"topics\") or 1 (\"docs\")");
    if (svd_method == 0) {
      // c(t(tsvd(t(A), k = 1)) %*% W) (R/wig.R:99-100): the S leading-vector
      // entries of A^T (S x N) weight each document's topic weights
      RWIG_REQUIRE(A && W, "null A or W");
      std::vector<double> At((size_t)S * N);
      for (int s = 0; s < S; ++s)
        for (int i = 0; i < N; ++i) At[s + (size_t)i * S] = A[i + (size_t)s * N];
      TsvdOut r = tsvd(At.data(), S, N, 1, 0);
      for (int j = 0; j < M; ++j) {
        double acc = 0.0;
        for (int s = 0; s < S; ++s) acc += r.Msvd[s] * W[s + (size_t)j * S];
        scores[j] = acc;
      }
    } else {
      // c(tsvd(t(Yhat), k = 1)) (R/wig.R:103)
      RWIG_REQUIRE(Yhat, "null Yhat");
      std::vector<double> Yt((size_t)M * N);
      for (int j = 0; j < M; ++j)
        for (int i = 0; i < N; ++i) Yt[j + (size_t)i * M] = Yhat[i + (size_t)j * N];
      TsvdOut r = tsvd(Yt.data(), M, N, 1, 0);
      std::memcpy(scores, r.Msvd.data(), (size_t)M * sizeof(double));
    }
  });
}

int rwig_wig_index(int *period_days, double *index, int *n_periods, const double *scores,
                   const int *days, int M, int group_unit, int standardize) {
  return guarded("rwig_wig_index", [&] {
    RWIG_REQUIRE(period_days && index && n_periods && scores && days && M >= 1,
                 "invalid argument");
    RWIG_REQUIRE(group_unit >= 0 && group_unit <= 7, "unknown group_unit");
    // aggregate(values, by = floor_date(dates), FUN = sum) (R/wig.R:2-6):
    // groups in ascending order of period start, each summed in document
    // order with an extended-precision accumulator (R's sum())
    std::map<long, long double> sums;
    for (int j = 0; j < M; ++j) sums[floor_period(days[j], group_unit)] += scores[j];
    const int P = (int)sums.size();
    std::vector<double> x;
    x.reserve(P);
    int p = 0;
    for (const auto &kv : sums) {
      period_days[p++] = (int)kv.first;
      x.push_back((double)kv.second);
    }
    if (standardize) {
      // (x - mean(x)) / sd(x) + 100 (R/wig.R:155-159); mean as R's mean()
      // (extended-precision sum, then one correction pass), sd as R's
      // sd() = sqrt(var()) with n - 1; a single period has sd NA (NaN here)
      long double s = 0.0L;
      for (double v : x) s += v;
      s /= P;
      if (std::isfinite((double)s)) {
        long double t = 0.0L;
        for (double v : x) t += (v - s);
        s += t / P;
      }
      const double mean = (double)s;
      double sd = NAN;
      if (P > 1) {
        long double q = 0.0L;  // double products, extended sum (R's cov.c)
        for (double v : x) {
          const double dv = v - mean;
          q += dv * dv;
        }
        sd = std::sqrt((double)(q / (P - 1)));
      }
      for (int i = 0; i < P; ++i) index[i] = (x[i] - mean) / sd + 100.0;
    } else {
      for (int i = 0; i < P; ++i) index[i] = x[i];
    }
    *n_periods = P;
  });
}

}  // extern "C"
