// C ABI of one fused soft-min pass (the log-sum-exp reductions of the
// log-domain solvers, kernels.hpp) on device pointers: for callers that build
// their own log-domain iterations, and for measuring the pass alone. The
// calls enqueue on the stream and return; the caller provides the workspace
// of the chunk partials.
#include "kernels.hpp"

using namespace rwig;

namespace {

template <typename T> size_t part_elems(int which, int M, int N, int S) {
  return which == 0 ? lse_row_part_size(M, N, S) : col_t_part_size(M, N, S);
}

void check_args(int M, int N, int S, double reg, int precision) {
  RWIG_REQUIRE(M > 0 && N > 0, "dimensions must be positive");
  RWIG_REQUIRE(S >= 1 && S <= 32, "S must be in [1, 32]");
  RWIG_REQUIRE(reg > 0.0, "reg must be positive");
  RWIG_REQUIRE(precision == 0 || precision == 1, "precision must be 0 (FP64) or 1 (FP32)");
}

template <typename T>
void rows(hipStream_t st, const T *C, int M, int N, long ldc, const T *F, long ldf, const T *G,
          long ldg, int S, double reg, T *out, T *work) {
  const size_t n = part_elems<T>(0, M, N, S);
  lse_row<T>(st, C, M, N, ldc, F, ldf, G, ldg, S, reg, work, work + n,
             [=] __device__(long k, T r) { out[k] = r; }, nullptr);
}

template <typename T>
void cols(hipStream_t st, const T *CT, int M, int N, long ldct, const T *F, long ldf,
          const T *G, long ldg, int S, double reg, T *out, T *work) {
  const size_t n = part_elems<T>(1, M, N, S);
  lse_col_t<T>(
      st, CT, M, N, ldct, F, ldf, G, ldg, S, reg, work, work + n,
      [=] __device__(int j, const T *r) {
        for (int s = 0; s < S; ++s) out[j + (long)s * N] = r[s];
      },
      nullptr);
}

}  // namespace

extern "C" {

int rwig_softmin_workspace_bytes(int which, int M, int N, int S, int precision, long *bytes) {
  return guarded("rwig_softmin_workspace_bytes", [&] {
    RWIG_REQUIRE(bytes && (which == 0 || which == 1), "invalid argument");
    check_args(M, N, S, 1.0, precision);
    const size_t n = precision == 0 ? part_elems<double>(which, M, N, S)
                                    : part_elems<float>(which, M, N, S);
    *bytes = (long)(2 * n * (precision == 0 ? 8 : 4));
  });
}

int rwig_softmin_rows_dev(const void *C, int M, int N, long ldc, const void *F, long ldf,
                          const void *G, long ldg, int S, double reg, void *out, int precision,
                          void *work, void *stream) {
  return guarded("rwig_softmin_rows_dev", [&] {
    check_args(M, N, S, reg, precision);
    RWIG_REQUIRE(C && F && G && out && work && ldc >= M && ldf >= M && ldg >= N,
                 "invalid argument");
    if (precision == 0)
      rows<double>((hipStream_t)stream, (const double *)C, M, N, ldc, (const double *)F, ldf,
                   (const double *)G, ldg, S, reg, (double *)out, (double *)work);
    else
      rows<float>((hipStream_t)stream, (const float *)C, M, N, ldc, (const float *)F, ldf,
                  (const float *)G, ldg, S, reg, (float *)out, (float *)work);
  });
}

int rwig_softmin_cols_dev(const void *CT, int M, int N, long ldct, const void *F, long ldf,
                          const void *G, long ldg, int S, double reg, void *out, int precision,
                          void *work, void *stream) {
  return guarded("rwig_softmin_cols_dev", [&] {
    check_args(M, N, S, reg, precision);
    RWIG_REQUIRE(CT && F && G && out && work && ldct >= N && ldf >= M && ldg >= N,
                 "invalid argument");
    if (precision == 0)
      cols<double>((hipStream_t)stream, (const double *)CT, M, N, ldct, (const double *)F, ldf,
                   (const double *)G, ldg, S, reg, (double *)out, (double *)work);
    else
      cols<float>((hipStream_t)stream, (const float *)CT, M, N, ldct, (const float *)F, ldf,
                  (const float *)G, ldg, S, reg, (float *)out, (float *)work);
  });
}

}  // extern "C"
