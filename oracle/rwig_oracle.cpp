// rwig oracle: FP64 CPU restatement of rwig's entropic-OT algorithms.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
// path (rwig_amd/librwig_hip.so) and the CPU baseline timed beside it. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// Nothing in the product path links, loads or falls back to it.
//
// It restates, operation for operation and in plain C++ (no Armadillo), the
// reference's CPU core:
//   src/sinkhorn_impl_vanilla.cpp   (Sinkhorn::compute_vanilla, _fwd_vanilla,
//                                    _bwd_vanilla)
//   src/sinkhorn_impl_log.cpp       (compute_log, _fwd_log, _bwd_log, the
//                                    serial and std::thread soft-min passes)
//   src/barycenter_impl_parallel.cpp (compute_parallel, _fwd_parallel,
//                                    _bwd_parallel)
//   src/barycenter_impl_log.cpp     (compute_log, _fwd_log, _bwd_log,
//                                    serial and threaded soft-min passes)
//   src/wdl_impl.cpp                (_train_batch_batched, _compute_serial)
//   inst/include/wdl_impl.hpp       (_softmax, _optimize)
//   inst/include/optimizer.hpp      (sgd, adam, adamw)
//   src/cuda_wdl.cu:592-593, :677-689 (the CUDA path's seeded initial Alpha
//                                    and Lambda)
//   src/sinkhorn.cpp, src/barycenter.cpp (output assembly, return codes)
//
// Armadillo semantics reproduced here:
//   norm(vec, 2)        Euclidean norm          -> nrm2()
//   norm(mat, 2)        largest singular value  -> spectral_norm()
//   find(x != 0)        support indices         -> support()
//   min(X, 1), min(X,0) row / column minima
//   prod(pow(KTU, 1 w^T), 1)  row-wise product of powers
// symmatu(K) after is_symmetric(C) is not reproduced: when C is exactly
// symmetric, exp(-C/reg) is exactly symmetric elementwise already, so
// symmatu is the identity on it.
//
// All matrices are column-major (R / Armadillo layout): X(i, j) = X[i + j*ld].

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

namespace {

using Vec = std::vector<double>;

struct Mat {
  int m = 0, n = 0;
  Vec d;
  Mat() {}
  Mat(int m_, int n_, double fill = 0.0) : m(m_), n(n_), d((size_t)m_ * n_, fill) {}
  double &operator()(int i, int j) { return d[(size_t)i + (size_t)j * m]; }
  double operator()(int i, int j) const { return d[(size_t)i + (size_t)j * m]; }
  double *col(int j) { return d.data() + (size_t)j * m; }
  const double *col(int j) const { return d.data() + (size_t)j * m; }
};

const double kInf = std::numeric_limits<double>::infinity();

double nrm2(const Vec &x) {
  double s = 0.0;
  for (double v : x) s += v * v;
  return std::sqrt(s);
}

// y = K * x   (M x N times N)
Vec matvec(const Mat &K, const Vec &x) {
  Vec y(K.m, 0.0);
  for (int j = 0; j < K.n; ++j) {
    const double xj = x[j];
    const double *c = K.col(j);
    for (int i = 0; i < K.m; ++i) y[i] += c[i] * xj;
  }
  return y;
}

// y = K^T * x   (N x M times M)
Vec matvec_t(const Mat &K, const Vec &x) {
  Vec y(K.n, 0.0);
  for (int j = 0; j < K.n; ++j) {
    const double *c = K.col(j);
    double s = 0.0;
    for (int i = 0; i < K.m; ++i) s += c[i] * x[i];
    y[j] = s;
  }
  return y;
}

// Y = K * X   (M x N times N x S)
Mat matmul(const Mat &K, const Mat &X) {
  Mat Y(K.m, X.n, 0.0);
  for (int s = 0; s < X.n; ++s) {
    double *y = Y.col(s);
    for (int j = 0; j < K.n; ++j) {
      const double xj = X(j, s);
      const double *c = K.col(j);
      for (int i = 0; i < K.m; ++i) y[i] += c[i] * xj;
    }
  }
  return Y;
}

// Y = K^T * X  (N x M times M x S)
Mat matmul_t(const Mat &K, const Mat &X) {
  Mat Y(K.n, X.n, 0.0);
  for (int s = 0; s < X.n; ++s) {
    const double *x = X.col(s);
    for (int j = 0; j < K.n; ++j) {
      const double *c = K.col(j);
      double acc = 0.0;
      for (int i = 0; i < K.m; ++i) acc += c[i] * x[i];
      Y(j, s) = acc;
    }
  }
  return Y;
}

// Largest singular value of X (m x s), Armadillo norm(X, 2)
// (src/barycenter_impl_log.cpp:252). Computed as sqrt of the largest
// eigenvalue of the s x s Gram matrix X^T X by cyclic Jacobi rotations.
double spectral_norm(const Mat &X) {
  const int s = X.n;
  if (s == 1) {
    double acc = 0.0;
    for (int i = 0; i < X.m; ++i) acc += X(i, 0) * X(i, 0);
    return std::sqrt(acc);
  }
  std::vector<double> G((size_t)s * s, 0.0);
  for (int p = 0; p < s; ++p)
    for (int q = p; q < s; ++q) {
      double acc = 0.0;
      for (int i = 0; i < X.m; ++i) acc += X(i, p) * X(i, q);
      G[p + q * s] = acc;
      G[q + p * s] = acc;
    }
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
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
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
  for (int p = 0; p < s; ++p) lmax = std::max(lmax, G[p + p * s]);
  return std::sqrt(lmax);
}

std::vector<int> support(const double *x, int n) {
  std::vector<int> idx;
  for (int i = 0; i < n; ++i)
    if (x[i] != 0.0) idx.push_back(i);
  return idx;
}

// Split [0, n) into n_threads chunks the way the reference does
// (src/sinkhorn_impl_log.cpp:89-94): chunk = n / n_threads, last one takes
// the remainder. Runs fn(start, end) on one std::thread per chunk.
template <class F> void run_chunks(int n, int n_threads, F fn) {
  std::vector<std::thread> threads;
  const int chunk = n / n_threads;
  for (int t = 0; t < n_threads; ++t) {
    const int start = t * chunk;
    const int end = (t == n_threads - 1) ? n : (t + 1) * chunk;
    threads.emplace_back([=, &fn]() { fn(start, end); });
  }
  for (auto &th : threads) th.join();
}

int return_code_of(double err, double tol, int iter, int maxiter) {
  // src/sinkhorn_impl_vanilla.cpp:79-85
  if (err <= tol) return 0;
  if (iter == maxiter) return 1;
  return 2;
}

// ---------------------------------------------------------------------------
// Sinkhorn
// ---------------------------------------------------------------------------

struct SinkhornState {
  int M = 0, N = 0;
  Vec a, b;
  Mat C, K;  // K holds the Gibbs kernel (vanilla) or R (log)
  Vec u, v;  // u, v (vanilla) or f, g (log)
  Vec Rminrow, Rmincol;
  std::vector<Vec> uhist, vhist;
  double reg = 0.1;
  int maxiter = 1000;
  double tol = 1e-6;
  bool withgrad = false;
  int iter = 0;
  double err = 1000.0;
};

void reduce_problem(SinkhornState &st, const double *a, const double *b,
                    const double *C, int M, int N, std::vector<int> &ai,
                    std::vector<int> &bi) {
  // src/sinkhorn_impl_vanilla.cpp:29-40
  ai = support(a, M);
  bi = support(b, N);
  st.M = (int)ai.size();
  st.N = (int)bi.size();
  st.a.resize(st.M);
  st.b.resize(st.N);
  for (int i = 0; i < st.M; ++i) st.a[i] = a[ai[i]];
  for (int j = 0; j < st.N; ++j) st.b[j] = b[bi[j]];
  st.C = Mat(st.M, st.N);
  for (int j = 0; j < st.N; ++j)
    for (int i = 0; i < st.M; ++i) st.C(i, j) = C[ai[i] + (size_t)bi[j] * M];
}

// Sinkhorn::_fwd_vanilla, src/sinkhorn_impl_vanilla.cpp:93-166
void fwd_vanilla(SinkhornState &st) {
  st.u.assign(st.M, 1.0);
  st.v.assign(st.N, 1.0);
  if (st.withgrad) {
    st.uhist.clear();
    st.vhist.clear();
    st.uhist.push_back(st.u);
    st.vhist.push_back(st.v);
  }
  Vec Kv = matvec(st.K, st.v), KTu;
  while (st.iter < st.maxiter && st.err >= st.tol) {
    st.iter++;
    for (int i = 0; i < st.M; ++i) st.u[i] = st.a[i] / Kv[i];
    if (st.withgrad) st.uhist.push_back(st.u);
    KTu = matvec_t(st.K, st.u);
    for (int j = 0; j < st.N; ++j) st.v[j] = st.b[j] / KTu[j];
    if (st.withgrad) st.vhist.push_back(st.v);
    Kv = matvec(st.K, st.v);
    Vec r1(st.M), r2(st.N);
    for (int i = 0; i < st.M; ++i) r1[i] = st.u[i] * Kv[i] - st.a[i];
    for (int j = 0; j < st.N; ++j) r2[j] = st.v[j] * KTu[j] - st.b[j];
    st.err = nrm2(r1) + nrm2(r2);
  }
}

// Sinkhorn::_bwd_vanilla, src/sinkhorn_impl_vanilla.cpp:169-214
Vec bwd_vanilla(SinkhornState &st, const Mat &P) {
  const int M = st.M, N = st.N;
  Mat PbarK(M, N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i)
      PbarK(i, j) = (st.C(i, j) + st.reg * std::log(P(i, j))) * st.K(i, j);
  Vec ubar(M, 0.0), vbar(N, 0.0), grad(M, 0.0);
  for (int l = st.iter; l > 0; --l) {
    if (l == st.iter) {
      vbar = matvec_t(PbarK, st.u);
      Vec t1 = matvec(PbarK, st.v);
      Vec KTu = matvec_t(st.K, st.u);
      Vec x(N);
      for (int j = 0; j < N; ++j) x[j] = (vbar[j] * st.v[j]) / KTu[j];
      Vec t2 = matvec(st.K, x);
      for (int i = 0; i < M; ++i) ubar[i] = t1[i] - t2[i];
    } else {
      Vec Kv = matvec(st.K, st.vhist[l]);
      Vec x(M);
      for (int i = 0; i < M; ++i) x[i] = (ubar[i] * st.uhist[l + 1][i]) / Kv[i];
      Vec t = matvec_t(st.K, x);
      for (int j = 0; j < N; ++j) vbar[j] = -t[j];
      Vec KTu = matvec_t(st.K, st.uhist[l]);
      Vec y(N);
      for (int j = 0; j < N; ++j) y[j] = (vbar[j] * st.vhist[l][j]) / KTu[j];
      Vec t2 = matvec(st.K, y);
      for (int i = 0; i < M; ++i) ubar[i] = -t2[i];
    }
    Vec Kvp = matvec(st.K, st.vhist[l - 1]);
    for (int i = 0; i < M; ++i) grad[i] += ubar[i] / Kvp[i];
  }
  return grad;
}

// R = C - f 1^T - 1 g^T, src/sinkhorn_impl_log.cpp:23-28
void compute_R(SinkhornState &st) {
  st.K = st.C;
  for (int j = 0; j < st.N; ++j)
    for (int i = 0; i < st.M; ++i) {
      double r = st.K(i, j) - st.u[i];
      st.K(i, j) = r - st.v[j];
    }
}

double softmin_row(const Mat &R, int i, double reg) {
  double mn = kInf;
  for (int j = 0; j < R.n; ++j) mn = std::min(mn, R(i, j));
  double s = 0.0;
  for (int j = 0; j < R.n; ++j) s += std::exp(-(R(i, j) - mn) / reg);
  return mn - reg * std::log(s);
}

double softmin_col(const Mat &R, int j, double reg) {
  const double *c = R.col(j);
  double mn = kInf;
  for (int i = 0; i < R.m; ++i) mn = std::min(mn, c[i]);
  double s = 0.0;
  for (int i = 0; i < R.m; ++i) s += std::exp(-(c[i] - mn) / reg);
  return mn - reg * std::log(s);
}

// _minrow_serial / _minrow_thread, src/sinkhorn_impl_log.cpp:31-41, 82-115
void minrow(SinkhornState &st, int n_threads) {
  compute_R(st);
  st.Rminrow.assign(st.M, 0.0);
  if (n_threads) {
    run_chunks(st.M, n_threads, [&](int s, int e) {
      for (int i = s; i < e; ++i) st.Rminrow[i] = softmin_row(st.K, i, st.reg);
    });
  } else {
    for (int i = 0; i < st.M; ++i) st.Rminrow[i] = softmin_row(st.K, i, st.reg);
  }
}

// _mincol_serial / _mincol_thread, src/sinkhorn_impl_log.cpp:44-54, 119-152
void mincol(SinkhornState &st, int n_threads) {
  compute_R(st);
  st.Rmincol.assign(st.N, 0.0);
  if (n_threads) {
    run_chunks(st.N, n_threads, [&](int s, int e) {
      for (int j = s; j < e; ++j) st.Rmincol[j] = softmin_col(st.K, j, st.reg);
    });
  } else {
    for (int j = 0; j < st.N; ++j) st.Rmincol[j] = softmin_col(st.K, j, st.reg);
  }
}

// _minrowcol_serial / _minrowcol_thread, src/sinkhorn_impl_log.cpp:58-75,
// 156-217
void minrowcol(SinkhornState &st, int n_threads) {
  compute_R(st);
  st.Rminrow.assign(st.M, 0.0);
  st.Rmincol.assign(st.N, 0.0);
  if (n_threads) {
    run_chunks(st.M, n_threads, [&](int s, int e) {
      for (int i = s; i < e; ++i) st.Rminrow[i] = softmin_row(st.K, i, st.reg);
    });
    run_chunks(st.N, n_threads, [&](int s, int e) {
      for (int j = s; j < e; ++j) st.Rmincol[j] = softmin_col(st.K, j, st.reg);
    });
  } else {
    for (int i = 0; i < st.M; ++i) st.Rminrow[i] = softmin_row(st.K, i, st.reg);
    for (int j = 0; j < st.N; ++j) st.Rmincol[j] = softmin_col(st.K, j, st.reg);
  }
}

// Sinkhorn::_fwd_log, src/sinkhorn_impl_log.cpp:298-376
void fwd_log(SinkhornState &st, int n_threads) {
  st.u.assign(st.M, 0.0);
  st.v.assign(st.N, 0.0);
  if (st.withgrad) {
    st.uhist.clear();
    st.vhist.clear();
    st.uhist.push_back(st.u);
    st.vhist.push_back(st.v);
  }
  Vec loga(st.M), logb(st.N);
  for (int i = 0; i < st.M; ++i) loga[i] = std::log(st.a[i]);
  for (int j = 0; j < st.N; ++j) logb[j] = std::log(st.b[j]);
  minrow(st, n_threads);
  while (st.iter < st.maxiter && st.err >= st.tol) {
    st.iter++;
    for (int i = 0; i < st.M; ++i) st.u[i] += st.reg * loga[i] + st.Rminrow[i];
    if (st.withgrad) st.uhist.push_back(st.u);
    mincol(st, n_threads);
    for (int j = 0; j < st.N; ++j) st.v[j] += st.reg * logb[j] + st.Rmincol[j];
    if (st.withgrad) st.vhist.push_back(st.v);
    minrowcol(st, n_threads);
    Vec r1(st.M), r2(st.N);
    for (int i = 0; i < st.M; ++i) r1[i] = -st.Rminrow[i] / st.reg - loga[i];
    for (int j = 0; j < st.N; ++j) r2[j] = -st.Rmincol[j] / st.reg - logb[j];
    st.err = nrm2(r1) + nrm2(r2);
  }
}

// Sinkhorn::_bwd_log, src/sinkhorn_impl_log.cpp:379-520
Vec bwd_log(SinkhornState &st, const Mat &P, int n_threads) {
  const int M = st.M, N = st.N;
  const double reg = st.reg;
  Mat PbarP(M, N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i)
      PbarP(i, j) = (st.C(i, j) + reg * std::log(P(i, j))) * P(i, j);
  Vec fbar(M, 0.0), gbar(N, 0.0), grad(M, 0.0);
  Mat W(M, N);
  for (int l = st.iter; l > 0; --l) {
    if (l == st.iter) {
      for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (int i = 0; i < M; ++i) s += PbarP(i, j);
        gbar[j] = s / reg;
      }
    } else {
      st.u = st.uhist[l];
      st.v = st.vhist[l];
      compute_R(st);
      // X = row-softmax(-R/reg)^T, gbar = -X fbar
      Vec g(N, 0.0);
      Mat X(M, N);
      auto rowfn = [&](int s, int e) {
        for (int i = s; i < e; ++i) {
          double mn = kInf;
          for (int j = 0; j < N; ++j) mn = std::min(mn, st.K(i, j));
          double tot = 0.0;
          for (int j = 0; j < N; ++j) {
            X(i, j) = std::exp(-(st.K(i, j) - mn) / reg);
            tot += X(i, j);
          }
          for (int j = 0; j < N; ++j) X(i, j) /= tot;
        }
      };
      if (n_threads) run_chunks(M, n_threads, rowfn); else rowfn(0, M);
      for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (int i = 0; i < M; ++i) s += X(i, j) * fbar[i];
        gbar[j] = -s;
      }
    }
    // W = col-softmax(-R/reg), fbar = -W gbar (+ PbarP 1 / reg at l == L)
    auto colfn = [&](int s, int e) {
      for (int j = s; j < e; ++j) {
        double mn = kInf;
        for (int i = 0; i < M; ++i) mn = std::min(mn, st.K(i, j));
        double tot = 0.0;
        for (int i = 0; i < M; ++i) {
          W(i, j) = std::exp(-(st.K(i, j) - mn) / reg);
          tot += W(i, j);
        }
        for (int i = 0; i < M; ++i) W(i, j) /= tot;
      }
    };
    if (n_threads) run_chunks(N, n_threads, colfn); else colfn(0, N);
    Vec Wg = matvec(W, gbar);
    for (int i = 0; i < M; ++i) fbar[i] = -Wg[i];
    if (l == st.iter) {
      for (int i = 0; i < M; ++i) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += PbarP(i, j);
        fbar[i] += s / reg;
      }
    }
    for (int i = 0; i < M; ++i) grad[i] += fbar[i] / st.a[i];
  }
  for (int i = 0; i < M; ++i) grad[i] *= reg;
  st.u = st.uhist[st.iter];
  st.v = st.vhist[st.iter];
  return grad;
}

double sinkhorn_loss(const Mat &C, const Mat &P, double reg) {
  // src/sinkhorn_impl_vanilla.cpp:69
  double a = 0.0, b = 0.0;
  for (size_t k = 0; k < P.d.size(); ++k) {
    a += C.d[k] * P.d[k];
    b += P.d[k] * (std::log(P.d[k]) - 1.0);
  }
  return a + reg * b;
}

// ---------------------------------------------------------------------------
// Barycenter
// ---------------------------------------------------------------------------

struct BaryState {
  int M = 0, N = 0, S = 0;
  Mat A, C, K, U, V, KV, KTU;
  Vec w, b_ext, b, logb;
  double reg = 0.1;
  int maxiter = 1000;
  double tol = 1e-6;
  bool withgrad = false;
  int iter = 0;
  double err = 1000.0;
  std::vector<Mat> Uhist, Vhist;
  std::vector<Vec> bhist, logbhist;
  Mat grad_A;
  Vec grad_w;
  double loss = 0.0;
};

// b_i = prod_s KTU_is^{w_s}, Armadillo prod(pow(KTU, 1 w^T), 1)
// (src/barycenter_impl_parallel.cpp:120)
Vec rowprod_pow(const Mat &KTU, const Vec &w) {
  Vec b(KTU.m, 1.0);
  for (int i = 0; i < KTU.m; ++i) {
    double p = 1.0;
    for (int s = 0; s < KTU.n; ++s) p *= std::pow(KTU(i, s), w[s]);
    b[i] = p;
  }
  return b;
}

// Barycenter::_fwd_parallel, src/barycenter_impl_parallel.cpp:60-151
void fwd_parallel(BaryState &st) {
  const int M = st.M, N = st.N, S = st.S;
  st.b.assign(N, 0.0);
  st.U = Mat(M, S, 1.0);
  st.V = Mat(N, S, 1.0);
  if (st.withgrad) {
    st.Uhist.clear();
    st.Vhist.clear();
    st.bhist.clear();
    st.Uhist.push_back(st.U);
    st.Vhist.push_back(st.V);
    st.bhist.push_back(st.b);
  }
  st.K = Mat(M, N);
  for (size_t k = 0; k < st.C.d.size(); ++k) st.K.d[k] = std::exp(-st.C.d[k] / st.reg);
  st.KV = matmul(st.K, st.V);
  while (st.iter < st.maxiter && st.err >= st.tol) {
    st.iter++;
    for (size_t k = 0; k < st.U.d.size(); ++k) st.U.d[k] = st.A.d[k] / st.KV.d[k];
    if (st.withgrad) st.Uhist.push_back(st.U);
    st.KTU = matmul_t(st.K, st.U);
    st.b = rowprod_pow(st.KTU, st.w);
    if (st.withgrad) st.bhist.push_back(st.b);
    for (int s = 0; s < S; ++s)
      for (int j = 0; j < N; ++j) st.V(j, s) = st.b[j] / st.KTU(j, s);
    if (st.withgrad) st.Vhist.push_back(st.V);
    st.KV = matmul(st.K, st.V);
    double acc = 0.0;
    for (size_t k = 0; k < st.U.d.size(); ++k) {
      const double r = st.U.d[k] * st.KV.d[k] - st.A.d[k];
      acc += r * r;
    }
    st.err = std::sqrt(acc);
  }
}

// Barycenter::_bwd_parallel, src/barycenter_impl_parallel.cpp:154-207
void bwd_parallel(BaryState &st) {
  const int M = st.M, N = st.N, S = st.S;
  st.grad_A = Mat(M, S, 0.0);
  st.grad_w.assign(S, 0.0);
  Mat Ubar(M, S, 0.0), Vbar(N, S, 0.0), KTU;
  Vec bbar(N, 0.0);
  for (int l = st.iter; l > 0; --l) {
    KTU = matmul_t(st.K, st.Uhist[l]);
    if (l == st.iter) {
      for (int j = 0; j < N; ++j) bbar[j] = 2.0 * (st.b[j] - st.b_ext[j]);
      Mat X(N, S);
      for (int s = 0; s < S; ++s)
        for (int j = 0; j < N; ++j) X(j, s) = (bbar[j] * st.w[s]) * st.V(j, s);
      Ubar = matmul(st.K, X);
    } else {
      Mat KVl = matmul(st.K, st.Vhist[l]);
      Mat X(M, S);
      for (size_t k = 0; k < X.d.size(); ++k)
        X.d[k] = (Ubar.d[k] * st.Uhist[l + 1].d[k]) / KVl.d[k];
      Mat T = matmul_t(st.K, X);
      for (size_t k = 0; k < T.d.size(); ++k) Vbar.d[k] = -T.d[k];
      for (int j = 0; j < N; ++j) {
        double s_ = 0.0;
        for (int s = 0; s < S; ++s) s_ += Vbar(j, s) / KTU(j, s);
        bbar[j] = s_;
      }
      Mat Y(N, S);
      for (int s = 0; s < S; ++s)
        for (int j = 0; j < N; ++j)
          Y(j, s) = (bbar[j] * st.w[s] - Vbar(j, s) / KTU(j, s)) * st.Vhist[l](j, s);
      Ubar = matmul(st.K, Y);
    }
    Mat KVp = matmul(st.K, st.Vhist[l - 1]);
    for (size_t k = 0; k < st.grad_A.d.size(); ++k) st.grad_A.d[k] += Ubar.d[k] / KVp.d[k];
    for (int s = 0; s < S; ++s) {
      double acc = 0.0;
      for (int j = 0; j < N; ++j) acc += std::log(KTU(j, s)) * (bbar[j] * st.bhist[l][j]);
      st.grad_w[s] += acc;
    }
  }
}

// soft-min of R_s = C - F(:,s) 1^T - 1 G(:,s)^T along rows (into out(:,s))
void bary_minrow_s(const BaryState &st, const Mat &F, const Mat &G, int s,
                   Mat &out, int n_threads) {
  const int M = st.M, N = st.N;
  Mat R(M, N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) {
      const double r = st.C(i, j) - F(i, s);
      R(i, j) = r - G(j, s);
    }
  auto fn = [&](int b0, int e0) {
    for (int i = b0; i < e0; ++i) out(i, s) = softmin_row(R, i, st.reg);
  };
  if (n_threads) run_chunks(M, n_threads, fn); else fn(0, M);
}

// soft-min of R_s along columns. The threaded reference pass shifts column j
// by R.row(j).min() instead of the column minimum
// (src/barycenter_impl_log.cpp:166); that path is reproduced when
// quirk_rowshift is set (valid only for N <= M, as in the reference).
void bary_mincol_s(const BaryState &st, const Mat &F, const Mat &G, int s,
                   Mat &out, int n_threads) {
  const int M = st.M, N = st.N;
  Mat R(M, N);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < M; ++i) {
      const double r = st.C(i, j) - F(i, s);
      R(i, j) = r - G(j, s);
    }
  auto fn_serial = [&](int b0, int e0) {
    for (int j = b0; j < e0; ++j) out(j, s) = softmin_col(R, j, st.reg);
  };
  auto fn_thread = [&](int b0, int e0) {
    for (int j = b0; j < e0; ++j) {
      double mn = kInf;
      for (int k = 0; k < N; ++k) mn = std::min(mn, R(j, k));
      double acc = 0.0;
      for (int i = 0; i < M; ++i) acc += std::exp(-(R(i, j) - mn) / st.reg);
      out(j, s) = mn - st.reg * std::log(acc);
    }
  };
  if (n_threads) run_chunks(N, n_threads, fn_thread); else fn_serial(0, N);
}

// Barycenter::_fwd_log, src/barycenter_impl_log.cpp:189-272
void fwd_log_bary(BaryState &st, int n_threads) {
  const int M = st.M, N = st.N, S = st.S;
  st.logb.assign(N, 0.0);
  st.U = Mat(M, S, 0.0);
  st.V = Mat(N, S, 0.0);
  if (st.withgrad) {
    st.Uhist.clear();
    st.Vhist.clear();
    st.logbhist.clear();
    st.Uhist.push_back(st.U);
    st.Vhist.push_back(st.V);
    st.logbhist.push_back(st.logb);
  }
  Mat logA(M, S);
  for (size_t k = 0; k < logA.d.size(); ++k) logA.d[k] = std::log(st.A.d[k]);
  st.KV = Mat(M, S, 0.0);
  st.KTU = Mat(N, S, 0.0);
  for (int s = 0; s < S; ++s) bary_minrow_s(st, st.U, st.V, s, st.KV, n_threads);
  while (st.iter < st.maxiter && st.err >= st.tol) {
    st.iter++;
    for (size_t k = 0; k < st.U.d.size(); ++k) st.U.d[k] += st.reg * logA.d[k] + st.KV.d[k];
    if (st.withgrad) st.Uhist.push_back(st.U);
    for (int s = 0; s < S; ++s) bary_mincol_s(st, st.U, st.V, s, st.KTU, n_threads);
    for (int j = 0; j < N; ++j) {
      double acc = 0.0;
      for (int s = 0; s < S; ++s) acc += -(st.V(j, s) + st.KTU(j, s)) * st.w[s];
      st.logb[j] = acc / st.reg;
    }
    if (st.withgrad) st.logbhist.push_back(st.logb);
    for (int s = 0; s < S; ++s)
      for (int j = 0; j < N; ++j) st.V(j, s) += st.reg * st.logb[j] + st.KTU(j, s);
    if (st.withgrad) st.Vhist.push_back(st.V);
    for (int s = 0; s < S; ++s) bary_minrow_s(st, st.U, st.V, s, st.KV, n_threads);
    Mat E(M, S);
    for (size_t k = 0; k < E.d.size(); ++k) E.d[k] = -st.KV.d[k] / st.reg - logA.d[k];
    st.err = spectral_norm(E);
  }
  st.b.assign(N, 0.0);
  double tot = 0.0;
  for (int j = 0; j < N; ++j) {
    st.b[j] = std::exp(st.logb[j]);
  }
  for (int j = 0; j < N; ++j) tot += st.b[j];
  for (int j = 0; j < N; ++j) st.b[j] /= tot;
  if (st.withgrad) {
    double acc = 0.0;
    for (int j = 0; j < N; ++j) {
      const double r = st.b[j] - st.b_ext[j];
      acc += r * r;
    }
    st.loss = acc;
  }
}

// Barycenter::_bwd_log, src/barycenter_impl_log.cpp:275-382
void bwd_log_bary(BaryState &st, int n_threads) {
  const int M = st.M, N = st.N, S = st.S;
  const double reg = st.reg;
  st.grad_A = Mat(M, S, 0.0);
  st.grad_w.assign(S, 0.0);
  Mat Fbar(M, S, 0.0), Gbar(N, S, 0.0);
  Vec logbbar(N, 0.0);
  Mat R(M, N), X(M, N), W(M, N);
  for (int l = st.iter; l > 0; --l) {
    if (l != st.iter) {
      for (int s = 0; s < S; ++s) {
        for (int j = 0; j < N; ++j)
          for (int i = 0; i < M; ++i) {
            const double r = st.C(i, j) - st.Uhist[l](i, s);
            R(i, j) = r - st.Vhist[l](j, s);
          }
        for (int i = 0; i < M; ++i) {
          double mn = kInf;
          for (int j = 0; j < N; ++j) mn = std::min(mn, R(i, j));
          double tot = 0.0;
          for (int j = 0; j < N; ++j) {
            X(i, j) = std::exp(-(R(i, j) - mn) / reg);
            tot += X(i, j);
          }
          for (int j = 0; j < N; ++j) X(i, j) /= tot;
        }
        for (int j = 0; j < N; ++j) {
          double acc = 0.0;
          for (int i = 0; i < M; ++i) acc += X(i, j) * Fbar(i, s);
          Gbar(j, s) = -acc;
        }
      }
    }
    if (l == st.iter) {
      for (int j = 0; j < N; ++j) logbbar[j] = 2.0 * (st.b[j] - st.b_ext[j]) * st.b[j];
    } else {
      for (int j = 0; j < N; ++j) {
        double acc = 0.0;
        for (int s = 0; s < S; ++s) acc += Gbar(j, s);
        logbbar[j] = reg * acc;
      }
    }
    for (int s = 0; s < S; ++s) {
      for (int j = 0; j < N; ++j)
        for (int i = 0; i < M; ++i) {
          const double r = st.C(i, j) - st.Uhist[l](i, s);
          R(i, j) = r - st.Vhist[l - 1](j, s);
        }
      for (int j = 0; j < N; ++j) {
        double mn = kInf;
        for (int i = 0; i < M; ++i) mn = std::min(mn, R(i, j));
        double tot = 0.0;
        for (int i = 0; i < M; ++i) {
          W(i, j) = std::exp(-(R(i, j) - mn) / reg);
          tot += W(i, j);
        }
        for (int i = 0; i < M; ++i) W(i, j) /= tot;
      }
      Vec x(N);
      if (l == st.iter) {
        Vec y = matvec(W, logbbar);
        for (int i = 0; i < M; ++i) Fbar(i, s) = (st.w[s] / reg) * y[i];
      } else {
        for (int j = 0; j < N; ++j) x[j] = (st.w[s] / reg) * logbbar[j] - Gbar(j, s);
        Vec y = matvec(W, x);
        for (int i = 0; i < M; ++i) Fbar(i, s) = y[i];
      }
    }
    for (size_t k = 0; k < Fbar.d.size(); ++k) st.grad_A.d[k] += Fbar.d[k] / st.A.d[k];
    Mat KTU(N, S);
    for (int s = 0; s < S; ++s)
      bary_mincol_s(st, st.Uhist[l], st.Vhist[l - 1], s, KTU, n_threads);
    for (int s = 0; s < S; ++s) {
      double acc = 0.0;
      for (int j = 0; j < N; ++j) acc += (st.Vhist[l - 1](j, s) + KTU(j, s)) * logbbar[j];
      st.grad_w[s] -= acc;
    }
  }
  for (auto &x : st.grad_A.d) x *= reg;
  for (auto &x : st.grad_w) x /= reg;
  st.U = st.Uhist[st.iter];
  st.V = st.Vhist[st.iter];
}

// ---------------------------------------------------------------------------
// WDL
// ---------------------------------------------------------------------------

// A = softmax(Alpha) per column with max shift, inst/include/wdl_impl.hpp:80-96
void softmax_cols(const double *X, double *Y, int m, int n) {
  for (int j = 0; j < n; ++j) {
    const double *x = X + (size_t)j * m;
    double *y = Y + (size_t)j * m;
    double mx = -kInf;
    for (int i = 0; i < m; ++i) mx = std::max(mx, x[i]);
    double tot = 0.0;
    for (int i = 0; i < m; ++i) {
      y[i] = std::exp(x[i] - mx);
    }
    for (int i = 0; i < m; ++i) tot += y[i];
    for (int i = 0; i < m; ++i) y[i] /= tot;
  }
}

struct OptState {
  Vec m, v;
  int step = 0;
};

// inst/include/optimizer.hpp:72-99
void opt_step(int opt, OptState &os, double *theta, const double *g, size_t n,
              double eta, double gamma, double beta1, double beta2, double eps) {
  os.step++;
  if (opt == 0) {
    for (size_t k = 0; k < n; ++k) theta[k] = theta[k] - eta * g[k];
    return;
  }
  if (os.m.size() != n) {
    os.m.assign(n, 0.0);
    os.v.assign(n, 0.0);
  }
  const double c1 = 1.0 - std::pow(beta1, os.step);
  const double c2 = 1.0 - std::pow(beta2, os.step);
  for (size_t k = 0; k < n; ++k) {
    os.m[k] = beta1 * os.m[k] + (1.0 - beta1) * g[k];
    os.v[k] = beta2 * os.v[k] + (1.0 - beta2) * std::pow(g[k], 2);
    const double mhat = os.m[k] / c1;
    const double vhat = os.v[k] / c2;
    const double upd = eta * (mhat / (std::sqrt(vhat) + eps));
    if (opt == 1)
      theta[k] = theta[k] - upd;
    else
      theta[k] = (1.0 - eta * gamma) * theta[k] - upd;
  }
}

// WassersteinDictionaryLearning::_train_batch_batched,
// src/wdl_impl.cpp:87-247. Y_b (N x D) are the batch's documents, W_b (S x D)
// the batch's weights. Outputs g_Alpha (N x S) and g_lambda (S).
void wdl_batch_grad(const Mat &K, const Mat &A, const Mat &Yb, const Mat &Wb,
                    int L, Mat &g_Alpha, Vec &g_lambda, Mat *bB_out) {
  const int N = K.m, S = A.n, D = Yb.n, SD = S * D;
  Mat UB(N, SD, 1.0), VB(N, SD, 1.0), bB(N, D, 0.0), KVB, KTUB;
  std::vector<Mat> UBh(L + 1), VBh(L + 1), bBh(L + 1), KVBh(L), KTUBh(L + 1);
  UBh[0] = UB;
  VBh[0] = VB;
  bBh[0] = bB;
  for (int l = 0; l < L; ++l) {
    KVB = matmul(K, VB);
    KVBh[l] = KVB;
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < S; ++s)
        for (int i = 0; i < N; ++i) UB(i, d * S + s) = A(i, s) / KVB(i, d * S + s);
    UBh[l + 1] = UB;
    KTUB = matmul_t(K, UB);
    KTUBh[l + 1] = KTUB;
    for (int d = 0; d < D; ++d)
      for (int i = 0; i < N; ++i) {
        double p = 1.0;
        for (int s = 0; s < S; ++s) p *= std::pow(KTUB(i, d * S + s), Wb(s, d));
        bB(i, d) = p;
      }
    bBh[l + 1] = bB;
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < S; ++s)
        for (int i = 0; i < N; ++i) VB(i, d * S + s) = bB(i, d) / KTUB(i, d * S + s);
    VBh[l + 1] = VB;
  }
  if (bB_out) *bB_out = bB;
  Mat UBbar(N, SD, 0.0), VBbar(N, SD, 0.0), bBbar(N, D, 0.0), ABbar(N, SD, 0.0);
  Mat wBbar(S, D, 0.0);
  for (int l = L; l > 0; --l) {
    Mat tmp(N, SD);
    if (l == L) {
      for (int d = 0; d < D; ++d)
        for (int i = 0; i < N; ++i) bBbar(i, d) = 2.0 * (bBh[l](i, d) - Yb(i, d));
      for (int d = 0; d < D; ++d)
        for (int s = 0; s < S; ++s)
          for (int i = 0; i < N; ++i)
            tmp(i, d * S + s) = (bBbar(i, d) * Wb(s, d)) * VBh[l](i, d * S + s);
      UBbar = matmul(K, tmp);
    } else {
      Mat X(N, SD);
      for (size_t k = 0; k < X.d.size(); ++k)
        X.d[k] = (UBbar.d[k] * UBh[l + 1].d[k]) / KVBh[l].d[k];
      Mat T = matmul_t(K, X);
      for (size_t k = 0; k < T.d.size(); ++k) VBbar.d[k] = -T.d[k];
      for (int d = 0; d < D; ++d)
        for (int i = 0; i < N; ++i) {
          double acc = 0.0;
          for (int s = 0; s < S; ++s) acc += VBbar(i, d * S + s) / KTUBh[l](i, d * S + s);
          bBbar(i, d) = acc;
        }
      for (int d = 0; d < D; ++d)
        for (int s = 0; s < S; ++s)
          for (int i = 0; i < N; ++i) {
            const int c = d * S + s;
            tmp(i, c) = (bBbar(i, d) * Wb(s, d) - VBbar(i, c) / KTUBh[l](i, c)) * VBh[l](i, c);
          }
      UBbar = matmul(K, tmp);
    }
    for (size_t k = 0; k < ABbar.d.size(); ++k) ABbar.d[k] += UBbar.d[k] / KVBh[l - 1].d[k];
    for (int d = 0; d < D; ++d)
      for (int s = 0; s < S; ++s) {
        double acc = 0.0;
        for (int i = 0; i < N; ++i)
          acc += std::log(KTUBh[l](i, d * S + s)) * (bBbar(i, d) * bBh[l](i, d));
        wBbar(s, d) += acc;
      }
  }
  // softmax Jacobians (src/wdl_impl.cpp:222-233)
  for (int j = 0; j < SD; ++j) {
    const int s = j % S;
    double dot = 0.0;
    for (int i = 0; i < N; ++i) dot += A(i, s) * ABbar(i, j);
    Vec out(N);
    for (int i = 0; i < N; ++i) out[i] = A(i, s) * ABbar(i, j) - A(i, s) * dot;
    for (int i = 0; i < N; ++i) ABbar(i, j) = out[i];
  }
  for (int d = 0; d < D; ++d) {
    double dot = 0.0;
    for (int s = 0; s < S; ++s) dot += Wb(s, d) * wBbar(s, d);
    Vec out(S);
    for (int s = 0; s < S; ++s) out[s] = Wb(s, d) * wBbar(s, d) - Wb(s, d) * dot;
    for (int s = 0; s < S; ++s) wBbar(s, d) = out[s];
  }
  g_Alpha = Mat(N, S, 0.0);
  for (int d = 0; d < D; ++d)
    for (int s = 0; s < S; ++s)
      for (int i = 0; i < N; ++i) g_Alpha(i, s) += ABbar(i, d * S + s);
  for (auto &x : g_Alpha.d) x /= D;
  g_lambda.assign(S, 0.0);
  for (int s = 0; s < S; ++s) {
    double acc = 0.0;
    for (int d = 0; d < D; ++d) acc += wBbar(s, d);
    g_lambda[s] = acc / D;
  }
}

void copy_out(const Vec &x, double *out) {
  if (out) std::memcpy(out, x.data(), sizeof(double) * x.size());
}
void copy_out(const Mat &x, double *out) {
  if (out) std::memcpy(out, x.d.data(), sizeof(double) * x.d.size());
}

}  // namespace

extern "C" {

// Sinkhorn::compute_vanilla (src/sinkhorn_impl_vanilla.cpp:18-86) plus the
// output list of sinkhorn_vanilla_cpu (src/sinkhorn.cpp:13-44).
int oracle_sinkhorn_vanilla(const double *a, const double *b, const double *C,
                            int M, int N, double reg, int withgrad, int maxiter,
                            double zerotol, double *P, double *grad_a,
                            double *u, double *v, double *loss, int *iter,
                            double *err, int *return_code) {
  SinkhornState st;
  std::vector<int> ai, bi;
  reduce_problem(st, a, b, C, M, N, ai, bi);
  st.reg = reg;
  st.maxiter = maxiter;
  st.tol = zerotol;
  st.withgrad = withgrad != 0;
  st.K = Mat(st.M, st.N);
  for (size_t k = 0; k < st.C.d.size(); ++k) st.K.d[k] = std::exp(-st.C.d[k] / reg);
  fwd_vanilla(st);
  Mat Pr(st.M, st.N);
  for (int j = 0; j < st.N; ++j)
    for (int i = 0; i < st.M; ++i) Pr(i, j) = (st.u[i] * st.K(i, j)) * st.v[j];
  if (u) {
    for (int i = 0; i < M; ++i) u[i] = 0.0;
    for (int i = 0; i < st.M; ++i) u[ai[i]] = st.u[i];
  }
  if (v) {
    for (int j = 0; j < N; ++j) v[j] = 0.0;
    for (int j = 0; j < st.N; ++j) v[bi[j]] = st.v[j];
  }
  if (P) {
    for (size_t k = 0; k < (size_t)M * N; ++k) P[k] = 0.0;
    for (int j = 0; j < st.N; ++j)
      for (int i = 0; i < st.M; ++i) P[ai[i] + (size_t)bi[j] * M] = Pr(i, j);
  }
  if (loss) *loss = sinkhorn_loss(st.C, Pr, reg);
  if (st.withgrad && grad_a) {
    Vec g = bwd_vanilla(st, Pr);
    for (int i = 0; i < M; ++i) grad_a[i] = 0.0;
    for (int i = 0; i < st.M; ++i) grad_a[ai[i]] = g[i];
  }
  if (iter) *iter = st.iter;
  if (err) *err = st.err;
  if (return_code) *return_code = return_code_of(st.err, zerotol, st.iter, maxiter);
  return 0;
}

// Sinkhorn::compute_log (src/sinkhorn_impl_log.cpp:225-294) plus the output
// list of sinkhorn_log_cpp (src/sinkhorn.cpp:142-170).
int oracle_sinkhorn_log(const double *a, const double *b, const double *C,
                        int M, int N, double reg, int withgrad, int n_threads,
                        int maxiter, double zerotol, double *P, double *grad_a,
                        double *f, double *g, double *loss, int *iter,
                        double *err, int *return_code) {
  if (n_threads < 0) return -1;
  SinkhornState st;
  std::vector<int> ai, bi;
  reduce_problem(st, a, b, C, M, N, ai, bi);
  st.reg = reg;
  st.maxiter = maxiter;
  st.tol = zerotol;
  st.withgrad = withgrad != 0;
  fwd_log(st, n_threads);
  Mat Pr(st.M, st.N);
  for (size_t k = 0; k < Pr.d.size(); ++k) Pr.d[k] = std::exp(-st.K.d[k] / reg);
  if (f) {
    for (int i = 0; i < M; ++i) f[i] = -kInf;
    for (int i = 0; i < st.M; ++i) f[ai[i]] = st.u[i];
  }
  if (g) {
    for (int j = 0; j < N; ++j) g[j] = -kInf;
    for (int j = 0; j < st.N; ++j) g[bi[j]] = st.v[j];
  }
  if (P) {
    for (size_t k = 0; k < (size_t)M * N; ++k) P[k] = 0.0;
    for (int j = 0; j < st.N; ++j)
      for (int i = 0; i < st.M; ++i) P[ai[i] + (size_t)bi[j] * M] = Pr(i, j);
  }
  if (loss) *loss = sinkhorn_loss(st.C, Pr, reg);
  if (st.withgrad && grad_a) {
    Vec gr = bwd_log(st, Pr, n_threads);
    for (int i = 0; i < M; ++i) grad_a[i] = 0.0;
    for (int i = 0; i < st.M; ++i) grad_a[ai[i]] = gr[i];
  }
  if (iter) *iter = st.iter;
  if (err) *err = st.err;
  if (return_code) *return_code = return_code_of(st.err, zerotol, st.iter, maxiter);
  return 0;
}

// Barycenter::compute_parallel (src/barycenter_impl_parallel.cpp:19-53) plus
// the output list of barycenter_parallel_cpu (src/barycenter.cpp:13-60).
// b_bwd receives the unnormalised b_L the backward pass used (may be null).
int oracle_barycenter_parallel(const double *A, const double *C,
                               const double *w, const double *b_ext, int M,
                               int N, int S, double reg, int withgrad,
                               int maxiter, double zerotol, double *b,
                               double *U, double *V, double *grad_A,
                               double *grad_w, double *loss, int *iter,
                               double *err, int *return_code) {
  BaryState st;
  st.M = M;
  st.N = N;
  st.S = S;
  st.A = Mat(M, S);
  std::memcpy(st.A.d.data(), A, sizeof(double) * M * S);
  st.C = Mat(M, N);
  std::memcpy(st.C.d.data(), C, sizeof(double) * (size_t)M * N);
  st.w.assign(w, w + S);
  st.reg = reg;
  st.maxiter = maxiter;
  st.tol = zerotol;
  st.withgrad = withgrad != 0;
  if (st.withgrad) st.b_ext.assign(b_ext, b_ext + N);
  fwd_parallel(st);
  if (st.withgrad) bwd_parallel(st);
  double tot = 0.0;
  for (int j = 0; j < N; ++j) tot += st.b[j];
  for (int j = 0; j < N; ++j) st.b[j] /= tot;
  if (st.withgrad) {
    double acc = 0.0;
    for (int j = 0; j < N; ++j) {
      const double r = st.b[j] - st.b_ext[j];
      acc += r * r;
    }
    st.loss = acc;
  }
  copy_out(st.b, b);
  copy_out(st.U, U);
  copy_out(st.V, V);
  if (st.withgrad) {
    copy_out(st.grad_A, grad_A);
    copy_out(st.grad_w, grad_w);
    if (loss) *loss = st.loss;
  }
  if (iter) *iter = st.iter;
  if (err) *err = st.err;
  if (return_code) *return_code = return_code_of(st.err, zerotol, st.iter, maxiter);
  return 0;
}

// Barycenter::compute_log (src/barycenter_impl_log.cpp:23-48) plus the output
// list of barycenter_log_cpp (src/barycenter.cpp:168-207). With n_threads > 0
// the threaded soft-min passes are used, including their row-minimum column
// shift; that form needs N <= M, as in the reference.
int oracle_barycenter_log(const double *A, const double *C, const double *w,
                          const double *b_ext, int M, int N, int S, double reg,
                          int withgrad, int n_threads, int maxiter,
                          double zerotol, double *b, double *F, double *G,
                          double *grad_A, double *grad_w, double *loss,
                          int *iter, double *err, int *return_code) {
  if (n_threads < 0) return -1;
  if (n_threads > 0 && N > M) return -2;
  BaryState st;
  st.M = M;
  st.N = N;
  st.S = S;
  st.A = Mat(M, S);
  std::memcpy(st.A.d.data(), A, sizeof(double) * M * S);
  st.C = Mat(M, N);
  std::memcpy(st.C.d.data(), C, sizeof(double) * (size_t)M * N);
  st.w.assign(w, w + S);
  st.reg = reg;
  st.maxiter = maxiter;
  st.tol = zerotol;
  st.withgrad = withgrad != 0;
  if (st.withgrad) st.b_ext.assign(b_ext, b_ext + N);
  fwd_log_bary(st, n_threads);
  if (st.withgrad) bwd_log_bary(st, n_threads);
  copy_out(st.b, b);
  copy_out(st.U, F);
  copy_out(st.V, G);
  if (st.withgrad) {
    copy_out(st.grad_A, grad_A);
    copy_out(st.grad_w, grad_w);
    if (loss) *loss = st.loss;
  }
  if (iter) *iter = st.iter;
  if (err) *err = st.err;
  if (return_code) *return_code = return_code_of(st.err, zerotol, st.iter, maxiter);
  return 0;
}

// Column softmax with max shift (inst/include/wdl_impl.hpp:80-96).
void oracle_softmax_cols(const double *X, double *Y, int m, int n) {
  softmax_cols(X, Y, m, n);
}

// One optimizer step on theta (inst/include/optimizer.hpp:72-99). m_state and
// v_state hold the moment estimates (n each) and *step the step counter; all
// are updated in place. opt: 0 sgd, 1 adam, 2 adamw.
void oracle_optimizer_step(int opt, double *theta, const double *g, int n,
                           double *m_state, double *v_state, int *step,
                           double eta, double gamma, double beta1,
                           double beta2, double eps) {
  OptState os;
  os.step = *step;
  if (opt != 0) {
    os.m.assign(m_state, m_state + n);
    os.v.assign(v_state, v_state + n);
  }
  opt_step(opt, os, theta, g, (size_t)n, eta, gamma, beta1, beta2, eps);
  *step = os.step;
  if (opt != 0) {
    std::memcpy(m_state, os.m.data(), sizeof(double) * n);
    std::memcpy(v_state, os.v.data(), sizeof(double) * n);
  }
}

// Gradient of one batched WDL step (src/wdl_impl.cpp:87-247) for a batch of
// D documents: K (N x N Gibbs kernel), A (N x S topics = softmax(Alpha)),
// Yb (N x D documents), Wb (S x D weights = softmax(Lambda) columns), L fixed
// iterations. Writes g_Alpha (N x S), g_lambda (S) and, if bB is non-null,
// the unnormalised barycenters b_L (N x D).
void oracle_wdl_batch_grad(const double *K, const double *A, const double *Yb,
                           const double *Wb, int N, int S, int D, int L,
                           double *g_Alpha, double *g_lambda, double *bB) {
  Mat Km(N, N), Am(N, S), Ym(N, D), Wm(S, D);
  std::memcpy(Km.d.data(), K, sizeof(double) * (size_t)N * N);
  std::memcpy(Am.d.data(), A, sizeof(double) * N * S);
  std::memcpy(Ym.d.data(), Yb, sizeof(double) * (size_t)N * D);
  std::memcpy(Wm.d.data(), Wb, sizeof(double) * S * D);
  Mat gA, bBm;
  Vec gl;
  wdl_batch_grad(Km, Am, Ym, Wm, L, gA, gl, bB ? &bBm : nullptr);
  copy_out(gA, g_Alpha);
  copy_out(gl, g_lambda);
  if (bB) copy_out(bBm, bB);
}

// WassersteinDictionaryLearning::_compute_serial (src/wdl_impl.cpp:254-378)
// with the random initialisation replaced by explicit Alpha0 (N x S) and
// Lambda0 (S x M). Y is N x M, C is N x N. sinkhorn_mode 1 = parallel,
// 2 = log (inference only). Outputs A (N x S), W (S x M), Yhat (N x M).
int oracle_wdl(const double *Y, const double *C, int N, int M, int S,
               double reg, int batch_size, int epochs, int sinkhorn_mode,
               int n_threads, int max_iter, double zero_tol, int optimizer,
               double eta, double gamma, double beta1, double beta2,
               double eps, const double *Alpha0, const double *Lambda0,
               double *A_out, double *W_out, double *Yhat_out) {
  if (sinkhorn_mode != 1 && sinkhorn_mode != 2) return -1;
  if (optimizer < 0 || optimizer > 2) return -2;
  if (N <= S) return -3;
  Mat Alpha(N, S), Lambda(S, M), A(N, S), W(S, M), K(N, N);
  std::memcpy(Alpha.d.data(), Alpha0, sizeof(double) * N * S);
  std::memcpy(Lambda.d.data(), Lambda0, sizeof(double) * S * M);
  softmax_cols(Alpha.d.data(), A.d.data(), N, S);
  softmax_cols(Lambda.d.data(), W.d.data(), S, M);
  for (size_t k = 0; k < K.d.size(); ++k) K.d[k] = std::exp(-C[k] / reg);
  OptState optA, optL;
  const int B = batch_size;
  const int batches = (M % B) ? (M / B + 1) : (M / B);
  Mat g_Lambda(S, M);
  for (int e = 0; e < epochs; ++e) {
    for (int bid = 0; bid < batches; ++bid) {
      const int D = (bid == (M / B)) ? (M % B) : B;
      Mat Yb(N, D), Wb(S, D);
      for (int d = 0; d < D; ++d) {
        for (int i = 0; i < N; ++i) Yb(i, d) = Y[i + (size_t)(bid * B + d) * N];
        for (int s = 0; s < S; ++s) Wb(s, d) = W(s, bid * B + d);
      }
      Mat gA;
      Vec gl;
      wdl_batch_grad(K, A, Yb, Wb, max_iter, gA, gl, nullptr);
      for (int m = 0; m < M; ++m)
        for (int s = 0; s < S; ++s) g_Lambda(s, m) = gl[s];
      opt_step(optimizer, optA, Alpha.d.data(), gA.d.data(), Alpha.d.size(), eta,
               gamma, beta1, beta2, eps);
      opt_step(optimizer, optL, Lambda.d.data(), g_Lambda.d.data(),
               Lambda.d.size(), eta, gamma, beta1, beta2, eps);
      softmax_cols(Alpha.d.data(), A.d.data(), N, S);
      softmax_cols(Lambda.d.data(), W.d.data(), S, M);
    }
  }
  copy_out(A, A_out);
  copy_out(W, W_out);
  if (Yhat_out) {
    for (int m = 0; m < M; ++m) {
      BaryState st;
      st.M = N;
      st.N = N;
      st.S = S;
      st.A = A;
      st.C = Mat(N, N);
      std::memcpy(st.C.d.data(), C, sizeof(double) * (size_t)N * N);
      st.w.assign(S, 0.0);
      for (int s = 0; s < S; ++s) st.w[s] = W(s, m);
      st.reg = reg;
      st.maxiter = max_iter;
      st.tol = zero_tol;
      st.withgrad = false;
      if (sinkhorn_mode == 1) {
        fwd_parallel(st);
        double tot = 0.0;
        for (int j = 0; j < N; ++j) tot += st.b[j];
        for (int j = 0; j < N; ++j) st.b[j] /= tot;
      } else {
        fwd_log_bary(st, n_threads);
      }
      std::memcpy(Yhat_out + (size_t)m * N, st.b.data(), sizeof(double) * N);
    }
  }
  return 0;
}

// Initial Alpha (N x S) and Lambda (S x M) of the reference's CUDA WDL path
// (src/cuda_wdl.cu:592-593, :677-689): std::mt19937(seed) and
// std::normal_distribution<double>(0, 1), N*S draws into Alpha in column-major
// order first, then S*M draws into Lambda.
void oracle_wdl_init_mt19937(int seed, int N, int S, int M, double *Alpha0,
                             double *Lambda0) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> randn(0.0, 1.0);
  for (long k = 0; k < (long)N * S; ++k) Alpha0[k] = randn(rng);
  for (long k = 0; k < (long)S * M; ++k) Lambda0[k] = randn(rng);
}

// Largest singular value of an m x s column-major matrix (Armadillo
// norm(X, 2)); exposed so tests can pin it against numpy.
double oracle_spectral_norm(const double *X, int m, int s) {
  Mat Xm(m, s);
  std::memcpy(Xm.d.data(), X, sizeof(double) * m * s);
  return spectral_norm(Xm);
}

}  // extern "C"
