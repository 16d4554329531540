// C ABI of the MFMA GEMM (gemm.hpp), used by tests and by host code that
// drives the batched products directly.
#include "gemm.hpp"

using namespace rwig;

extern "C" {

int rwig_gemm_workspace_bytes(int dtype, int M, int N, int ks, long *bytes) {
  return guarded("rwig_gemm_workspace_bytes", [&] {
    RWIG_REQUIRE(bytes && M > 0 && N > 0 && ks >= 1, "invalid argument");
    RWIG_REQUIRE(dtype == 0 || dtype == 1, "dtype must be 0 (bf16) or 1 (f64)");
    *bytes = dtype == 0 ? (long)(4 * tiles_work_elems<uint16_t>(M, N, ks))
                        : (long)(8 * tiles_work_elems<double>(M, N, ks));
  });
}

int rwig_gemm_tn_bf16_dev(int M, int N, int K, const void *A, long lda, const void *B,
                          long ldb, float *D, long ldd, int ks, void *work, void *stream) {
  return guarded("rwig_gemm_tn_bf16_dev", [&] {
    RWIG_REQUIRE(A && B && (D || work), "null argument");
    SplitK<uint16_t> sk;
    sk.ks = ks;
    sk.work = (float *)work;
    with_tiles<uint16_t>(N, [&](auto cfg) {
      gemm_tn<uint16_t, decltype(cfg)>(
          (hipStream_t)stream, M, N, K, (const uint16_t *)A, lda, (const uint16_t *)B, ldb,
          [=] __device__(int m, int n, float v) { D[m + (long)n * ldd] = v; }, nullptr, nullptr,
          sk, D != nullptr);
    });
  });
}

int rwig_gemm_tn_f64_dev(int M, int N, int K, const double *A, long lda, const double *B,
                         long ldb, double *D, long ldd, int ks, void *work, void *stream) {
  return guarded("rwig_gemm_tn_f64_dev", [&] {
    RWIG_REQUIRE(A && B && (D || work), "null argument");
    SplitK<double> sk;
    sk.ks = ks;
    sk.work = (double *)work;
    with_tiles<double>(N, [&](auto cfg) {
      gemm_tn<double, decltype(cfg)>(
          (hipStream_t)stream, M, N, K, A, lda, B, ldb,
          [=] __device__(int m, int n, double v) { D[m + (long)n * ldd] = v; }, nullptr, nullptr,
          sk, D != nullptr);
    });
  });
}

}  // extern "C"
