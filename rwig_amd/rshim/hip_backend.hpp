// hip_backend.hpp -- the Rcpp side of rwig's GPU path on ROCm.
//
// Counterparts of the functions rwig compiles under
// #if defined(HAVE_CUBLAS) && defined(HAVE_CUDA_RUNTIME):
//   sinkhorn_vanilla_cuda     src/sinkhorn.cpp:51-109
//   barycenter_parallel_cuda  src/barycenter.cpp:67-136
//   wdl_cpp_cuda              src/wdl.cpp:57-99
//   cuda_available_cpp        src/check_cuda.cpp:6-12
// with the same arguments and the same output lists (element names and
// order), calling librwig_hip.so through include/rwig_hip.h. Compiled when
// configure defines HAVE_HIP (configure.hip.ac in this directory). A failing
// call raises an R error (Rcpp::stop) with the library's message; the
// reference's CUDA path returned silently.
#pragma once

#ifdef HAVE_HIP

#include <Rcpp.h>

Rcpp::List sinkhorn_vanilla_hip(const SEXP &a, const SEXP &b, const SEXP &C, double reg,
                                bool withgrad = false, int maxiter = 1000,
                                double zerotol = 1e-6, int verbose = 0);

Rcpp::List barycenter_parallel_hip(const SEXP &A, const SEXP &C, const SEXP &w, double reg,
                                   const SEXP &b_ext, bool withgrad = false,
                                   int maxiter = 1000, double zerotol = 1e-6,
                                   int verbose = 0);

// sinkhorn_mode 1 infers Yhat with the parallel barycenter, 2 with the log
// barycenter (src/wdl_impl.cpp:354-377), both batched over the documents.
Rcpp::List wdl_cpp_hip(const SEXP &Y, const SEXP &C, double reg, int S, int batch_size,
                       int epochs, int sinkhorn_mode = 1, int max_iter = 1000,
                       double zero_tol = 1e-6, int optimizer = 2, double eta = .001,
                       double gamma = .01, double beta1 = .9, double beta2 = .999,
                       double eps = 1e-8, bool verbose = false, int seed = 42);

// The GPU forms of the reference's CPU-only log solvers (src/sinkhorn.cpp:142,
// src/barycenter.cpp:168), for an R-side option that routes them to the GPU.
Rcpp::List sinkhorn_log_hip(const Rcpp::NumericVector &a, const Rcpp::NumericVector &b,
                            const Rcpp::NumericMatrix &C, double reg, bool withgrad = false,
                            int maxiter = 1000, double zerotol = 1e-6);

Rcpp::List barycenter_log_hip(const Rcpp::NumericMatrix &A, const Rcpp::NumericMatrix &C,
                              const Rcpp::NumericVector &w, double reg,
                              const Rcpp::NumericVector &b_ext, bool withgrad = false,
                              int maxiter = 1000, double zerotol = 1e-6);

// cuda_available_cpp() returns this under HAVE_HIP.
bool hip_available();

#endif  // HAVE_HIP
