// C ABI of one fused soft-min pass (the log-sum-exp reductions of the
// log-domain solvers, kernels.hpp) on device pointers: for callers that build
// their own log-domain iterations, and for measuring the pass alone. The
// calls enqueue on the stream and return; the caller provides the workspace
// of the chunk partials, except in rwig_softmin_rows_noalloc_dev, whose
// partials live in the file-scope device array below.
#include <mutex>

#include "kernels.hpp"

using namespace rwig;

// Chunk partials (running minimum and scaled sum per chunk, row and slice) of
// rwig_softmin_rows_noalloc_dev: 32 MiB per device, enough for FP64 at
// M = N = 4096 with S = 16 (2 x 32 chunks x 16 x 4096 x 8 B). Every call on a
// device shares its array; calls that overlap on one device race. Nothing in
// it needs to start at zero: the pass writes every partial it reads.
static constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device with that device current (a fixed table: no host allocation either).
// The array's name is a host-side stand-in in host code: kernels must be
// given this address instead.
constexpr int kMaxDevices = 64;
void *scratch_address() {
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  int dev = 0;
  RWIG_HIP_CHECK(hipGetDevice(&dev));
  RWIG_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (addr[dev]) return addr[dev];
  void *p = nullptr;
  RWIG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  RWIG_REQUIRE(p != nullptr, "no device address for the soft-min scratch array");
  addr[dev] = p;
  return p;
}

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

int rwig_softmin_rows_noalloc_dev(const void *C, int M, int N, long ldc, const void *F,
                                  long ldf, const void *G, long ldg, int S, double reg,
                                  void *out, int precision, void *stream) {
  return guarded("rwig_softmin_rows_noalloc_dev", [&] {
    check_args(M, N, S, reg, precision);
    RWIG_REQUIRE(C && F && G && out && ldc >= M && ldf >= M && ldg >= N, "invalid argument");
    const size_t need = 2 * (precision == 0 ? part_elems<double>(0, M, N, S) * sizeof(double)
                                            : part_elems<float>(0, M, N, S) * sizeof(float));
    RWIG_REQUIRE(need <= kScratchBytes,
                 "the pass needs " + std::to_string(need) + " bytes of chunk partials, more than "
                 "the " + std::to_string(kScratchBytes) + "-byte static scratch array holds; "
                 "use rwig_softmin_rows_dev with a workspace");
    void *work = scratch_address();
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
