// hip_backend.cpp -- the Rcpp side of rwig's GPU path on ROCm (declarations
// and the functions they replace: hip_backend.hpp).
//
// Each function allocates its R outputs, calls the C ABI of librwig_hip.so on
// their memory (host pointers, FP64, column-major, as R stores them) and
// returns the list the reference's CUDA shim returns, in the same order. The
// library follows the reference's CPU semantics (support reduction, the
// while (iter < maxiter && err >= tol) loop) and computes return_status as
// the CPU core does (src/sinkhorn_impl_vanilla.cpp:79-85).
#ifdef HAVE_HIP

#include "hip_backend.hpp"

#include <string>

#include "rwig_hip.h"

namespace {

void check(int rc, const char *what) {
  if (rc != RWIG_OK) Rcpp::stop(std::string(what) + ": " + rwig_last_error());
}

}  // namespace

Rcpp::List sinkhorn_vanilla_hip(const SEXP &a, const SEXP &b, const SEXP &C, double reg,
                                bool withgrad, int maxiter, double zerotol, int verbose) {
  (void)verbose;
  const int m = Rf_nrows(C), n = Rf_ncols(C);
  Rcpp::NumericMatrix P(m, n);
  Rcpp::NumericVector u(m), v(n), grad_a(m);
  double loss = 0.0, err = 0.0;
  int iter = 0, status = 2;
  check(rwig_sinkhorn_vanilla(P.begin(), grad_a.begin(), u.begin(), v.begin(), &loss, &iter,
                              &err, REAL(a), REAL(b), REAL(C), m, n, reg, withgrad ? 1 : 0,
                              maxiter, zerotol, &status),
        "sinkhorn (vanilla) on the GPU");
  if (withgrad)
    return Rcpp::List::create(Rcpp::Named("P") = P, Rcpp::Named("grad_a") = grad_a,
                              Rcpp::Named("u") = u, Rcpp::Named("v") = v,
                              Rcpp::Named("loss") = loss, Rcpp::Named("iter") = iter,
                              Rcpp::Named("err") = err,
                              Rcpp::Named("return_status") = status);
  return Rcpp::List::create(Rcpp::Named("P") = P, Rcpp::Named("u") = u,
                            Rcpp::Named("v") = v, Rcpp::Named("loss") = loss,
                            Rcpp::Named("iter") = iter, Rcpp::Named("err") = err,
                            Rcpp::Named("return_status") = status);
}

Rcpp::List barycenter_parallel_hip(const SEXP &A, const SEXP &C, const SEXP &w, double reg,
                                   const SEXP &b_ext, bool withgrad, int maxiter,
                                   double zerotol, int verbose) {
  (void)verbose;
  const int M = Rf_nrows(C), N = Rf_ncols(C), S = Rf_ncols(A);
  Rcpp::NumericVector b(N), grad_w(S);
  Rcpp::NumericMatrix U(M, S), V(N, S), grad_A(M, S);
  double loss = 0.0, err = 0.0;
  int iter = 0, status = 2;
  // b_ext is read only with withgrad: R passes a length-2 dummy without it
  // (R/barycenter.R:159)
  check(rwig_barycenter_parallel(U.begin(), V.begin(), b.begin(), grad_A.begin(),
                                 grad_w.begin(), &loss, &iter, &err, REAL(A), REAL(w), REAL(C),
                                 withgrad ? REAL(b_ext) : nullptr, M, N, S, reg,
                                 withgrad ? 1 : 0, maxiter, zerotol, &status),
        "barycenter (parallel) on the GPU");
  if (withgrad)
    return Rcpp::List::create(Rcpp::Named("b") = b, Rcpp::Named("grad_A") = grad_A,
                              Rcpp::Named("grad_w") = grad_w, Rcpp::Named("loss") = loss,
                              Rcpp::Named("U") = U, Rcpp::Named("V") = V,
                              Rcpp::Named("iter") = iter, Rcpp::Named("err") = err,
                              Rcpp::Named("return_status") = status);
  return Rcpp::List::create(Rcpp::Named("b") = b, Rcpp::Named("U") = U,
                            Rcpp::Named("V") = V, Rcpp::Named("iter") = iter,
                            Rcpp::Named("err") = err,
                            Rcpp::Named("return_status") = status);
}

Rcpp::List wdl_cpp_hip(const SEXP &Y, const SEXP &C, double reg, int S, int batch_size,
                       int epochs, int sinkhorn_mode, int max_iter, double zero_tol,
                       int optimizer, double eta, double gamma, double beta1, double beta2,
                       double eps, bool verbose, int seed) {
  const int N = Rf_nrows(Y), M = Rf_ncols(Y);
  Rcpp::NumericMatrix A(N, S), W(S, M), Yhat(N, M);
  // training is the same in both modes; mode 1 infers Yhat inside rwig_wdl
  check(rwig_wdl(A.begin(), W.begin(), sinkhorn_mode == 1 ? Yhat.begin() : nullptr, REAL(Y),
                 REAL(C), N, M, S, reg, max_iter, zero_tol, batch_size, epochs, optimizer,
                 eta, gamma, beta1, beta2, eps, seed, verbose ? 1 : 0),
        "wdl on the GPU");
  if (sinkhorn_mode == 2)
    check(rwig_barycenter_log_batched(Yhat.begin(), nullptr, nullptr, A.begin(), W.begin(),
                                      REAL(C), N, N, S, M, reg, max_iter, zero_tol),
          "wdl inference (log barycenters) on the GPU");
  return Rcpp::List::create(Rcpp::Named("A") = A, Rcpp::Named("W") = W,
                            Rcpp::Named("Yhat") = Yhat);
}

Rcpp::List sinkhorn_log_hip(const Rcpp::NumericVector &a, const Rcpp::NumericVector &b,
                            const Rcpp::NumericMatrix &C, double reg, bool withgrad,
                            int maxiter, double zerotol) {
  const int m = C.nrow(), n = C.ncol();
  Rcpp::NumericMatrix P(m, n);
  Rcpp::NumericVector f(m), g(n), grad_a(m);
  double loss = 0.0, err = 0.0;
  int iter = 0, status = 2;
  check(rwig_sinkhorn_log(P.begin(), grad_a.begin(), f.begin(), g.begin(), &loss, &iter, &err,
                          a.begin(), b.begin(), C.begin(), m, n, reg, withgrad ? 1 : 0,
                          maxiter, zerotol, &status),
        "sinkhorn (log) on the GPU");
  // element names and order of sinkhorn_log_cpp (src/sinkhorn.cpp:154-169)
  if (withgrad)
    return Rcpp::List::create(Rcpp::Named("P") = P, Rcpp::Named("grad_a") = grad_a,
                              Rcpp::Named("f") = f, Rcpp::Named("g") = g,
                              Rcpp::Named("loss") = loss, Rcpp::Named("iter") = iter,
                              Rcpp::Named("err") = err,
                              Rcpp::Named("return_status") = status);
  return Rcpp::List::create(Rcpp::Named("P") = P, Rcpp::Named("f") = f,
                            Rcpp::Named("g") = g, Rcpp::Named("loss") = loss,
                            Rcpp::Named("iter") = iter, Rcpp::Named("err") = err,
                            Rcpp::Named("return_status") = status);
}

Rcpp::List barycenter_log_hip(const Rcpp::NumericMatrix &A, const Rcpp::NumericMatrix &C,
                              const Rcpp::NumericVector &w, double reg,
                              const Rcpp::NumericVector &b_ext, bool withgrad, int maxiter,
                              double zerotol) {
  const int M = C.nrow(), N = C.ncol(), S = A.ncol();
  Rcpp::NumericVector b(N), grad_w(S);
  Rcpp::NumericMatrix F(M, S), G(N, S), grad_A(M, S);
  double loss = 0.0, err = 0.0;
  int iter = 0, status = 2;
  check(rwig_barycenter_log(F.begin(), G.begin(), b.begin(), grad_A.begin(), grad_w.begin(),
                            &loss, &iter, &err, A.begin(), w.begin(), C.begin(),
                            withgrad ? b_ext.begin() : nullptr, M, N, S, reg,
                            withgrad ? 1 : 0, maxiter, zerotol, &status),
        "barycenter (log) on the GPU");
  // element names and order of barycenter_log_cpp (src/barycenter.cpp:190-206)
  if (withgrad)
    return Rcpp::List::create(Rcpp::Named("b") = b, Rcpp::Named("grad_A") = grad_A,
                              Rcpp::Named("grad_w") = grad_w, Rcpp::Named("loss") = loss,
                              Rcpp::Named("F") = F, Rcpp::Named("G") = G,
                              Rcpp::Named("iter") = iter, Rcpp::Named("err") = err,
                              Rcpp::Named("return_status") = status);
  return Rcpp::List::create(Rcpp::Named("b") = b, Rcpp::Named("F") = F,
                            Rcpp::Named("G") = G, Rcpp::Named("iter") = iter,
                            Rcpp::Named("err") = err,
                            Rcpp::Named("return_status") = status);
}

bool hip_available() { return rwig_available() == 1; }

#endif  // HAVE_HIP
