// C ABI of the MFMA GEMM (gemm.hpp), used by tests and by host code that
// drives the batched products directly.
#include "gemm.hpp"

using namespace rwig;

extern "C" {

int rwig_gemm_tn_bf16_dev(int M, int N, int K, const void *A, long lda, const void *B,
                          long ldb, float *D, long ldd, void *stream) {
  return guarded("rwig_gemm_tn_bf16_dev", [&] {
    RWIG_REQUIRE(A && B && D, "null argument");
    gemm_tn<uint16_t>((hipStream_t)stream, M, N, K, (const uint16_t *)A, lda,
                      (const uint16_t *)B, ldb,
                      [=] __device__(int m, int n, float v) { D[m + (long)n * ldd] = v; });
  });
}

int rwig_gemm_tn_f64_dev(int M, int N, int K, const double *A, long lda, const double *B,
                         long ldb, double *D, long ldd, void *stream) {
  return guarded("rwig_gemm_tn_f64_dev", [&] {
    RWIG_REQUIRE(A && B && D, "null argument");
    gemm_tn<double>((hipStream_t)stream, M, N, K, A, lda, B, ldb,
                    [=] __device__(int m, int n, double v) { D[m + (long)n * ldd] = v; });
  });
}

}  // extern "C"
