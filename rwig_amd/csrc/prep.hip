// Cost and histogram construction on the device: the inputs of wdl() that
// the reference builds on the CPU (src/utils.cpp:17-68), so that WDL and the
// serving engine can be fed without a host round trip.
//   euclidean: the V x V distance matrix between the rows of an embedding
//              (euclidean_cpp, src/utils.cpp:17-37), or its square;
//   doc2dist:  documents as normalised bags of words over a dictionary
//              (doc2dist_cpp, src/utils.cpp:39-68).
#include <string>
#include <unordered_map>

#include "kernels.hpp"

namespace rwig {

namespace {

constexpr int kTile = 16;

// out(i, j) = sqrt(sum_k (E(i,k) - E(j,k))^2), k in order (the reference's
// accu(square(row_i - row_j))); 16 x 16 output tiles, embedding columns
// staged through LDS. (i, j) and (j, i) use the same arithmetic, so the
// result is exactly symmetric with an exactly zero diagonal.
__global__ void __launch_bounds__(kTile *kTile)
k_euclidean(const double *__restrict__ E, int V, int d, long lde, double *__restrict__ out,
            long ldo, int squared) {
  __shared__ double Ei[kTile][kTile + 1], Ej[kTile][kTile + 1];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int i = blockIdx.x * kTile + tx, j = blockIdx.y * kTile + ty;
  double acc = 0.0;
  for (int k0 = 0; k0 < d; k0 += kTile) {
    // stage E(rows of this tile, k0 .. k0 + 16)
    {
      const int r = blockIdx.x * kTile + tx, k = k0 + ty;
      Ei[ty][tx] = (r < V && k < d) ? E[r + (long)k * lde] : 0.0;
      const int r2 = blockIdx.y * kTile + tx;
      Ej[ty][tx] = (r2 < V && k < d) ? E[r2 + (long)k * lde] : 0.0;
    }
    __syncthreads();
    const int kn = min(kTile, d - k0);
    for (int kk = 0; kk < kn; ++kk) {
      const double diff = Ei[kk][tx] - Ej[kk][ty];
      acc += diff * diff;
    }
    __syncthreads();
  }
  if (i < V && j < V) out[i + (long)j * ldo] = squared ? acc : sqrt(acc);
}

// counts of word w in document c from the (document, word) pairs, then each
// column divided by its sum (src/utils.cpp:61-63); a document without tokens
// divides 0 by 0, as in the reference.
__global__ void __launch_bounds__(256)
k_bow(const int *__restrict__ words, const long *__restrict__ offsets, int D, int V,
      double *__restrict__ out) {
  const int c = blockIdx.x;
  if (c >= D) return;
  double *col = out + (long)c * V;
  for (int w = threadIdx.x; w < V; w += blockDim.x) col[w] = 0.0;
  __syncthreads();
  const long t0 = offsets[c], t1 = offsets[c + 1];
  // one thread adds the document's tokens in order: exact integer counts with
  // no atomics (documents are short; the columns run in parallel)
  if (threadIdx.x == 0)
    for (long t = t0; t < t1; ++t) col[words[t]] += 1.0;
  __syncthreads();
  __shared__ double sh[256 / kWave];
  double s = 0.0;
  for (int w = threadIdx.x; w < V; w += blockDim.x) s += col[w];
  s = block_sum(s, sh);
  for (int w = threadIdx.x; w < V; w += blockDim.x) col[w] /= s;
}

}  // namespace

}  // namespace rwig

using namespace rwig;

extern "C" {

int rwig_euclidean_dev(const double *E, int V, int d, long lde, double *out, long ldo,
                       int squared, void *stream) {
  return guarded("rwig_euclidean_dev", [&] {
    RWIG_REQUIRE(E && out && V > 0 && d > 0 && lde >= V && ldo >= V, "invalid argument");
    const dim3 grid((V + kTile - 1) / kTile, (V + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_euclidean, grid, dim3(kTile * kTile), 0, (hipStream_t)stream, E, V, d,
                       lde, out, ldo, squared);
    check_launch();
  });
}

int rwig_euclidean(const double *E, int V, int d, double *out, int squared) {
  return guarded("rwig_euclidean", [&] {
    RWIG_REQUIRE(E && out && V > 0 && d > 0, "invalid argument");
    ensure_device();
    Stream s;
    DBuf<double> dE((size_t)V * d), dO((size_t)V * V);
    h2d(dE.get(), E, (size_t)V * d, s.get());
    const dim3 grid((V + kTile - 1) / kTile, (V + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_euclidean, grid, dim3(kTile * kTile), 0, s.get(), dE.get(), V, d,
                       (long)V, dO.get(), (long)V, squared);
    check_launch();
    d2h(out, dO.get(), (size_t)V * V, s.get());
    s.sync();
  });
}

int rwig_doc2dist(const char *const *tokens, const long *offsets, int D,
                  const char *const *dict, int V, double *out) {
  return guarded("rwig_doc2dist", [&] {
    RWIG_REQUIRE(offsets && dict && out && D >= 0 && V > 0, "invalid argument");
    RWIG_REQUIRE(offsets[0] == 0, "offsets must start at 0");
    for (int c = 0; c < D; ++c)
      RWIG_REQUIRE(offsets[c + 1] >= offsets[c], "offsets must be non-decreasing");
    const long T = offsets[D];
    RWIG_REQUIRE(T == 0 || tokens, "null tokens");
    // token -> dictionary index: the first equal entry, or the last index for
    // a token not in the dictionary (std::find and the clamp of
    // src/utils.cpp:53-57)
    std::unordered_map<std::string, int> index;
    index.reserve((size_t)V * 2);
    for (int w = 0; w < V; ++w) index.emplace(dict[w], w);
    std::vector<int> words((size_t)T);
    for (long t = 0; t < T; ++t) {
      auto it = index.find(tokens[t]);
      words[t] = (it == index.end()) ? V - 1 : it->second;
    }
    if (D == 0) return;
    ensure_device();
    Stream s;
    DBuf<int> dw(T > 0 ? (size_t)T : 1);
    DBuf<long> doff((size_t)D + 1);
    DBuf<double> dout((size_t)V * D);
    h2d(dw.get(), words.data(), (size_t)T, s.get());
    h2d(doff.get(), offsets, (size_t)D + 1, s.get());
    hipLaunchKernelGGL(k_bow, dim3(D), dim3(256), 0, s.get(), dw.get(), doff.get(), D, V,
                       dout.get());
    check_launch();
    d2h(out, dout.get(), (size_t)V * D, s.get());
    s.sync();
  });
}

}  // extern "C"
