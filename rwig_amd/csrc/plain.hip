// Plain FP64 GPU form of rwig's four solvers and their unrolled adjoints:
// the drop-in for sinkhorn_vanilla_cpp / sinkhorn_log_cpp /
// barycenter_parallel_cpp / barycenter_log_cpp (src/sinkhorn.cpp,
// src/barycenter.cpp), with the semantics of the reference's CPU core
// (src/sinkhorn_impl_vanilla.cpp, src/sinkhorn_impl_log.cpp,
// src/barycenter_impl_parallel.cpp, src/barycenter_impl_log.cpp).
//
// Each iteration is a short chain of streaming kernels over the cost or
// Gibbs matrix; the loop condition lives on the device (LoopState) and the
// host polls it once per batch of enqueued iterations. The forward pass
// records, beside the iterates, the products the adjoint needs (K v_l,
// K^T u_l, the soft-min vectors of every iteration), so each backward step is
// two passes over the matrix instead of the reference's four or five.

#include <algorithm>

#include "kernels.hpp"

namespace rwig {

namespace {

void init_loop(hipStream_t st, LoopState *dls, int maxiter, double tol) {
  LoopState h{};
  h.iter = 0;
  h.err = 1000.0;  // inst/include/sinkhorn_impl.hpp:85-88
  h.maxiter = maxiter;
  h.tol = tol;
  h.active = (h.iter < maxiter && h.err >= tol) ? 1 : 0;
  h2d(dls, &h, 1, st);
}

// Enqueue iterations in growing batches until the device loop reports done.
template <class Iter>
LoopState run_loop(hipStream_t st, LoopState *dls, int maxiter, Iter &&iteration) {
  LoopState hs{};
  d2h(&hs, dls, 1, st);
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  int batch = 4;
  while (hs.active) {
    const int n = std::min(batch, maxiter - hs.iter);
    for (int k = 0; k < n; ++k) iteration();
    d2h(&hs, dls, 1, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
    batch = std::min(batch * 2, 64);
  }
  return hs;
}

// Device-side end of one iteration: iter++, err, and the loop condition
// while (iter < maxiter && err >= tol) of the reference.
__device__ __forceinline__ void finish_iteration(LoopState *ls, double err) {
  ls->err = err;
  ls->iter = ls->iter + 1;
  ls->active = (ls->iter < ls->maxiter && err >= ls->tol) ? 1 : 0;
}

template <class T> void fill(hipStream_t st, T *p, long n, T v) {
  launch_map(st, n, [=] __device__(long k) { p[k] = v; });
}

// out(s) = scale * sum_j acc(j, s), one block per s.
__global__ void __launch_bounds__(kTB)
k_colsum(const double *acc, int N, double scale, double *out) {
  __shared__ double sh[kTB / kWave];
  const int s = blockIdx.x;
  double a = 0.0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) a += acc[j + (long)s * N];
  a = block_sum(a, sh);
  if (threadIdx.x == 0) out[s] = scale * a;
}

// Partial sums of the regularised transport cost
// sum C .* P + reg * sum P .* (log P - 1)  (src/sinkhorn_impl_vanilla.cpp:69)
__global__ void __launch_bounds__(kTB)
k_loss_part(long n, const double *__restrict__ C, const double *__restrict__ P,
            double *part) {
  __shared__ double sh[kTB / kWave];
  double a = 0.0, b = 0.0;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const double p = P[k];
    a += C[k] * p;
    b += p * (log(p) - 1.0);
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kSumBlocks + blockIdx.x] = b;
  }
}

void sinkhorn_loss(hipStream_t st, long n, const double *C, const double *P, double reg,
                   double *part /* 2*kSumBlocks */, double *loss) {
  hipLaunchKernelGGL(k_loss_part, dim3(kSumBlocks), dim3(kTB), 0, st, n, C, P, part);
  check_launch();
  reduce2(
      st, (long)kSumBlocks, [=] __device__(long k) { return part[k]; }, (long)kSumBlocks,
      [=] __device__(long k) { return part[kSumBlocks + k]; },
      [=] __device__(double a, double b) { *loss = a + reg * b; });
}

// ---------------------------------------------------------------- support reduction

struct Reduced {
  std::vector<int> ai, bi;
  std::vector<double> a, b, C;
  int M = 0, N = 0;
};

// find(a != 0), find(b != 0), C.submat(a_ind, b_ind)
// (src/sinkhorn_impl_vanilla.cpp:29-35)
Reduced reduce_support(const double *a, const double *b, const double *C, int m, int n) {
  Reduced r;
  for (int i = 0; i < m; ++i)
    if (a[i] != 0.0) r.ai.push_back(i);
  for (int j = 0; j < n; ++j)
    if (b[j] != 0.0) r.bi.push_back(j);
  r.M = (int)r.ai.size();
  r.N = (int)r.bi.size();
  r.a.resize(r.M);
  r.b.resize(r.N);
  for (int i = 0; i < r.M; ++i) r.a[i] = a[r.ai[i]];
  for (int j = 0; j < r.N; ++j) r.b[j] = b[r.bi[j]];
  r.C.resize((size_t)r.M * r.N);
  for (int j = 0; j < r.N; ++j)
    for (int i = 0; i < r.M; ++i) r.C[i + (size_t)j * r.M] = C[r.ai[i] + (size_t)r.bi[j] * m];
  return r;
}

int return_status(double err, double tol, int iter, int maxiter) {
  if (err <= tol) return 0;  // src/sinkhorn.cpp:88-95 with the CPU's == test
  if (iter == maxiter) return 1;
  return 2;
}

}  // namespace

// ============================================================================
// Vanilla Sinkhorn (src/sinkhorn_impl_vanilla.cpp:18-214)
// ============================================================================

struct SinkhornResult {
  int iter = 0;
  double err = 0.0;
};

// ---------------------------------------------------------------- small problems
//
// A problem whose Gibbs kernel fits one XCD's L2 with room to spare runs its
// whole forward loop in ONE workgroup of 1024 threads: K and K^T, built
// beforehand over the whole GPU, are read from that CU's L2 after the first
// iteration, the vectors u, v, K v, K^T u live in LDS, and the loop condition of the reference,
// while (iter < maxiter && err >= tol), is decided by thread 0 and read by
// every thread after a workgroup barrier, so the loop ends for all waves
// together and within maxiter iterations. The multi-kernel form pays five
// launches and a host poll per batch of iterations instead; at 256 x 256
// those cost ~24 us per iteration. Reductions run in a fixed order.
constexpr int kTinyTB = 1024;
constexpr int kTinyWaves = kTinyTB / kWave;
constexpr long kTinyMaxMN = 1L << 17;  // 1 MiB of FP64 K
constexpr int kTinyMaxMplusN = 2048;   // 4 vectors and the pass partials in LDS: < 48 KiB

inline bool tiny_vanilla(int M, int N) {
  return (long)M * N <= kTinyMaxMN && M + N <= kTinyMaxMplusN;
}

// Workgroup sum in a fixed order; every thread gets the result.
__device__ __forceinline__ double tiny_sum(double x, double *red) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  x = wave_sum(x);
  __syncthreads();
  if (lane == 0) red[w] = x;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kTinyWaves; ++k) t += red[k];
  return t;
}

// out[r] = sum_c A[r + c * R] x[c] for r < R (A column-major R x Cn, in
// global memory and L2-resident), by the whole workgroup: G = kTinyTB / R
// column groups when R <= kTinyTB (their partials summed in group order),
// else rows looped. Each thread keeps four accumulators and unrolls its loads
// so that many are in flight: the pass is L2-latency-bound, not
// bandwidth-bound, at these sizes.
// sum_{c0 <= c < c1} a(c) x[c] with four accumulators: element c goes to
// s[(c - c0) % 4] in order, whatever block of loads it falls in (16 in
// flight where the range allows, then 8, then one at a time).
template <class Acc>
__device__ __forceinline__ double tiny_dot(Acc a, const double *x, int c0, int c1) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  int c = c0;
  for (; c + 15 < c1; c += 16) {
    double k[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) k[q] = a(c + q);
#pragma unroll
    for (int q = 0; q < 16; ++q) s[q & 3] = fma(k[q], x[c + q], s[q & 3]);
  }
  for (; c + 7 < c1; c += 8) {
    double k[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) k[q] = a(c + q);
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q & 3] = fma(k[q], x[c + q], s[q & 3]);
  }
  for (; c < c1; ++c) s[0] = fma(a(c), x[c], s[0]);
  return (s[0] + s[1]) + (s[2] + s[3]);
}

__device__ __forceinline__ void tiny_row_pass(const double *__restrict__ A, int R, int Cn,
                                              const double *x, double *out, double *part) {
  const int tid = threadIdx.x;
  auto dot = [&](int r, int c0, int c1) {
    return tiny_dot([&](int c) { return A[r + (long)c * R]; }, x, c0, c1);
  };
  __syncthreads();  // A and x complete
  if (R <= kTinyTB) {
    const int G = kTinyTB / R, cw = (Cn + G - 1) / G;
    const int r = tid % R, g = tid / R;
    if (g < G) part[g * R + r] = dot(r, min(Cn, g * cw), min(Cn, (g + 1) * cw));
    __syncthreads();
    if (tid < R) {
      double acc = 0.0;
      for (int q = 0; q < G; ++q) acc += part[q * R + tid];
      out[tid] = acc;
    }
  } else {
    for (int r = tid; r < R; r += kTinyTB) out[r] = dot(r, 0, Cn);
  }
  __syncthreads();
}

// K = exp(-C / reg) and K^T, 32 x 32 tiles through LDS (reads of C and
// writes of both coalesced), over the whole GPU: the one-workgroup loop then
// reads both from L2.
__global__ void __launch_bounds__(1024)
k_gibbs_both(const double *__restrict__ C, int M, int N, double reg, double *__restrict__ K,
             double *__restrict__ KT) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int i = i0 + tx, j = j0 + ty;
  double kv = 0.0;
  if (i < M && j < N) {
    kv = exp(-C[i + (long)j * M] / reg);
    K[i + (long)j * M] = kv;
  }
  tile[ty][tx] = kv;
  __syncthreads();
  const int jt = j0 + tx, it = i0 + ty;
  if (jt < N && it < M) KT[jt + (long)it * N] = tile[tx][ty];
}

// Forward of Sinkhorn::_fwd_vanilla (src/sinkhorn_impl_vanilla.cpp:93-166):
// K = exp(-C/reg); Kv = K 1; loop { u = a / Kv; z = K^T u; v = b / z;
// Kv = K v; err = ||u .* Kv - a||_2 + ||v .* z - b||_2 } over K and its
// transpose KT (k_gibbs_both; both read from L2 by the row passes). Writes
// u, v and the final (iter, err) to ls; with uh non-null also the history slices u_l, v_l,
// K v_l (yh) and K^T u_l (zh) the adjoint reads.
__global__ void __launch_bounds__(kTinyTB)
k_sinkhorn_vanilla_tiny(const double *__restrict__ a, const double *__restrict__ b,
                        const double *__restrict__ C, int M, int N, double reg, int maxiter,
                        double tol, const double *__restrict__ K,
                        const double *__restrict__ KT,
                        double *__restrict__ u, double *__restrict__ v, LoopState *ls,
                        double *uh, double *vh, double *yh, double *zh) {
  extern __shared__ double lds[];
  double *su = lds, *sKv = su + M, *sv = sKv + M, *sz = sv + N;
  double *part = sz + N;  // max(kTinyTB, M, N) doubles: the column groups of a pass
  __shared__ double red[kTinyWaves];
  __shared__ int active;
  const int tid = threadIdx.x;
  for (int j = tid; j < N; j += kTinyTB) sv[j] = 1.0;
  tiny_row_pass(K, M, N, sv, sKv, part);
  if (uh) {
    for (int i = tid; i < M; i += kTinyTB) {
      uh[i] = 1.0;
      yh[i] = sKv[i];
    }
    for (int j = tid; j < N; j += kTinyTB) vh[j] = 1.0;
  }
  int iter = 0;
  double err = 1000.0;  // inst/include/sinkhorn_impl.hpp:85-88
  if (tid == 0) active = (iter < maxiter && err >= tol) ? 1 : 0;
  __syncthreads();
  while (active) {
    const long l = iter + 1;
    for (int i = tid; i < M; i += kTinyTB) {
      const double ui = a[i] / sKv[i];
      su[i] = ui;
      if (uh) uh[l * M + i] = ui;
    }
    tiny_row_pass(KT, N, M, su, sz, part);  // z = K^T u
    for (int j = tid; j < N; j += kTinyTB) {
      const double vj = b[j] / sz[j];
      sv[j] = vj;
      if (uh) {
        zh[l * N + j] = sz[j];
        vh[l * N + j] = vj;
      }
    }
    tiny_row_pass(K, M, N, sv, sKv, part);  // Kv = K v
    double e1 = 0.0, e2 = 0.0;
    for (int i = tid; i < M; i += kTinyTB) {
      const double r = su[i] * sKv[i] - a[i];
      e1 += r * r;
      if (uh) yh[l * M + i] = sKv[i];
    }
    for (int j = tid; j < N; j += kTinyTB) {
      const double r = sv[j] * sz[j] - b[j];
      e2 += r * r;
    }
    e1 = tiny_sum(e1, red);
    e2 = tiny_sum(e2, red);
    ++iter;
    err = sqrt(e1) + sqrt(e2);
    if (tid == 0) active = (iter < maxiter && err >= tol) ? 1 : 0;
    __syncthreads();
  }
  for (int i = tid; i < M; i += kTinyTB) u[i] = iter > 0 ? su[i] : 1.0;
  for (int j = tid; j < N; j += kTinyTB) v[j] = sv[j];
  if (tid == 0) {
    ls->iter = iter;
    ls->err = err;
    ls->active = 0;
    ls->maxiter = maxiter;
    ls->tol = tol;
  }
}

// The same loop over G workgroups (G in 4 .. 64, one CU each):
// workgroup w keeps rows w pm .. of K and w pn .. of K^T in LDS, and u, v
// and the error terms cross between workgroups through global memory at
// grid-wide barriers, two per iteration: the next u is computed before the
// barrier that publishes the error terms, and kept only if the loop goes on.
// One CU reads its rows of K from L2 at about 60 GB/s; G CUs hold theirs in
// LDS. The launch is an ordinary one with its own barrier (grid_barrier):
// after one cooperative launch, the concurrent streams of a serving session
// later in the same process ran a fifth slower. Every value is computed as in
// k_sinkhorn_vanilla_tiny: the same column groups and accumulators per row
// (tiny_dot), u = a / K v and v = b / K^T u, and the two error sums over the
// same 64-lane trees, so both forms give the same bits.
constexpr int kCoopTB = 256;
constexpr int kCoopWaves = kCoopTB / kWave;

struct CoopBufs {
  double *u, *v, *t1, *t2;  // u (M), v (N), error terms (2 M, 2 N: by iteration parity)
  unsigned long long *bar;  // barrier arrivals, barrier failure flag (zero at launch)
};

// Polls of one barrier wait before it gives up (about a second)
constexpr long kBarrierPolls = 1L << 24;

// Grid-wide barrier of the G workgroups: the k-th barrier of the launch
// completes when bar[0] reaches k G. Every thread releases its global writes
// at agent scope before the workgroup barrier, one thread adds the
// workgroup's arrival and polls, and every thread acquires after it. The
// G workgroups need G CUs at once: they come up together on an idle device,
// and otherwise as the kernels of other streams (which do not wait on this
// one) leave their CUs. A wait is bounded: past kBarrierPolls polls it sets
// bar[1], every later wait returns at once, the loop runs out, and the host
// reports the failure.
__device__ inline void grid_barrier(unsigned long long *bar, unsigned long long target) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&bar[0], 1ull);
    long polls = 0;
    while (__hip_atomic_load(&bar[0], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      // the failure flag every 256 polls, so that a poll is one load
      if ((polls & 255) == 255 &&
          __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        break;
      if (++polls > kBarrierPolls) {
        atomicExch(&bar[1], 1ull);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __threadfence();
}

// LDS of the multi-workgroup loop, in doubles
inline size_t coop_lds(int M, int N, int G) {
  const int pm = (M + G - 1) / G, pn = (N + G - 1) / G;
  const int GM = kTinyTB / M, GN = kTinyTB / N;
  return (size_t)pm * N + (size_t)pn * M + std::max(M, N) + 3 * (size_t)pm + 2 * (size_t)pn +
         std::max((size_t)pm * GM, (size_t)pn * GN);
}

__global__ void __launch_bounds__(kCoopTB)
k_sinkhorn_vanilla_coop(const double *__restrict__ a, const double *__restrict__ b, int M, int N,
                        int maxiter, double tol, const double *__restrict__ K,
                        const double *__restrict__ KT, double *__restrict__ u,
                        double *__restrict__ v, LoopState *ls, double *uh, double *vh,
                        double *yh, double *zh, CoopBufs xb, int pm, int pn) {
  extern __shared__ double lds[];
  unsigned long long nbar = 0;  // barriers passed
  auto grid_sync = [&] { grid_barrier(xb.bar, ++nbar * gridDim.x); };
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m0 = w * pm, n0 = w * pn;
  const int mr = max(0, min(M, m0 + pm) - m0), nr = max(0, min(N, n0 + pn) - n0);
  double *Ks = lds;                    // Ks[rl + c pm] = K(m0 + rl, c)
  double *KTs = Ks + (size_t)pm * N;   // KTs[rl + c pn] = K^T(n0 + rl, c)
  double *xs = KTs + (size_t)pn * M;   // the vector a pass multiplies
  double *kv = xs + max(M, N), *uo = kv + pm, *un = uo + pm, *zo = un + pm, *vo = zo + pn;
  double *part = vo + pn;              // column-group partials of a pass
  __shared__ double red[kTinyWaves];
  for (long k = tid; k < (long)pm * N; k += kCoopTB) {
    const int rl = (int)(k % pm), c = (int)(k / pm);
    Ks[k] = rl < mr ? K[m0 + rl + (long)c * M] : 0.0;
  }
  for (long k = tid; k < (long)pn * M; k += kCoopTB) {
    const int rl = (int)(k % pn), c = (int)(k / pn);
    KTs[k] = rl < nr ? KT[n0 + rl + (long)c * N] : 0.0;
  }
  // own rows of a pass over an R x Cn matrix held as S[rl + c p]: the column
  // groups of tiny_row_pass (G = kTinyTB / R of width ceil(Cn / G)), their
  // partials added in group order
  auto pass = [&](const double *S, int p, int rows, int R, int Cn, double *out) {
    const int G = kTinyTB / R, cw = (Cn + G - 1) / G;
    __syncthreads();  // xs complete
    for (int t = tid; t < rows * G; t += kCoopTB) {
      const int rl = t % rows, g = t / rows;
      part[g * p + rl] = tiny_dot([&](int c) { return S[rl + c * p]; }, xs, min(Cn, g * cw),
                                  min(Cn, (g + 1) * cw));
    }
    __syncthreads();
    for (int rl = tid; rl < rows; rl += kCoopTB) {
      double acc = 0.0;
      for (int q = 0; q < G; ++q) acc += part[q * p + rl];
      out[rl] = acc;
    }
    __syncthreads();
  };
  // tiny_sum over a block of kTinyTB threads whose thread i holds terms[i]
  // (0 past n): the same 64-lane trees, then the waves in order
  auto tree_sum = [&](const double *terms, int n) {
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    for (int k = wv; k < kTinyWaves; k += kCoopWaves) {
      const int i = k * kWave + lane;
      const double x = wave_sum(i < n ? terms[i] : 0.0);
      if (lane == 0) red[k] = x;
    }
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < kTinyWaves; ++k) t += red[k];
    __syncthreads();
    return t;
  };
  for (int j = tid; j < N; j += kCoopTB) xs[j] = 1.0;
  pass(Ks, pm, mr, M, N, kv);  // K v, v = 1
  for (int rl = tid; rl < mr; rl += kCoopTB) uo[rl] = 1.0;
  for (int rl = tid; rl < nr; rl += kCoopTB) vo[rl] = 1.0;
  if (uh) {
    for (int rl = tid; rl < mr; rl += kCoopTB) {
      uh[m0 + rl] = 1.0;
      yh[m0 + rl] = kv[rl];
    }
    for (int rl = tid; rl < nr; rl += kCoopTB) vh[n0 + rl] = 1.0;
  }
  int iter = 0;
  double err = 1000.0;  // inst/include/sinkhorn_impl.hpp:85-88
  bool active = iter < maxiter && err >= tol;
  if (active) {  // u of the first iteration
    for (int rl = tid; rl < mr; rl += kCoopTB) {
      const int i = m0 + rl;
      const double ui = a[i] / kv[rl];
      uo[rl] = ui;
      xb.u[i] = ui;
      if (uh) uh[M + i] = ui;
    }
  }
  grid_sync();
  while (active) {
    const long l = iter + 1;
    double *t1 = xb.t1 + (l & 1) * M, *t2 = xb.t2 + (l & 1) * N;
    for (int i = tid; i < M; i += kCoopTB) xs[i] = xb.u[i];
    pass(KTs, pn, nr, N, M, zo);  // z = K^T u
    for (int rl = tid; rl < nr; rl += kCoopTB) {
      const int j = n0 + rl;
      const double z = zo[rl], vj = b[j] / z;
      vo[rl] = vj;
      xb.v[j] = vj;
      const double r = vj * z - b[j];
      t2[j] = r * r;
      if (uh) {
        zh[l * N + j] = z;
        vh[l * N + j] = vj;
      }
    }
    grid_sync();
    for (int j = tid; j < N; j += kCoopTB) xs[j] = xb.v[j];
    pass(Ks, pm, mr, M, N, kv);  // K v
    for (int rl = tid; rl < mr; rl += kCoopTB) {
      const int i = m0 + rl;
      const double r = uo[rl] * kv[rl] - a[i];
      t1[i] = r * r;
      if (uh) yh[l * M + i] = kv[rl];
      const double ui = a[i] / kv[rl];  // the next iteration's u, if there is one
      un[rl] = ui;
      xb.u[i] = ui;
    }
    grid_sync();
    const double e1 = tree_sum(t1, M), e2 = tree_sum(t2, N);
    ++iter;
    err = sqrt(e1) + sqrt(e2);
    active = iter < maxiter && err >= tol;
    if (active)
      for (int rl = tid; rl < mr; rl += kCoopTB) {
        uo[rl] = un[rl];
        if (uh) uh[(iter + 1) * (long)M + m0 + rl] = un[rl];
      }
  }
  for (int rl = tid; rl < mr; rl += kCoopTB) u[m0 + rl] = iter > 0 ? uo[rl] : 1.0;
  for (int rl = tid; rl < nr; rl += kCoopTB) v[n0 + rl] = vo[rl];
  if (w == 0 && tid == 0) {
    ls->iter = iter;
    ls->err = err;
    ls->active = 0;
    ls->maxiter = maxiter;
    ls->tol = tol;
  }
}

// Steps l = L - 1 .. 1 of the vanilla adjoint (src/sinkhorn_impl_vanilla.cpp:
// 169-214) over the same G workgroups and LDS slices as
// k_sinkhorn_vanilla_coop: y = (-(K^T x) .* v_l) / (K^T u_l) over the
// workgroup's columns, a grid barrier, then ubar = -K y, grad += ubar ./ (K
// v_{l-1}) and x = ubar .* u_l ./ (K v_{l-1}) over its rows, a grid barrier.
// x (M) and grad (M) hold step L's values on entry; y (N) is scratch. Each
// workgroup writes only its own rows of x and grad and its own columns of y,
// and reads x and y whole only between the barriers that publish them.
__global__ void __launch_bounds__(kCoopTB)
k_vanilla_adjoint_coop(int M, int N, int L, const double *__restrict__ K,
                       const double *__restrict__ KT, const double *__restrict__ uh,
                       const double *__restrict__ vh, const double *__restrict__ yh,
                       const double *__restrict__ zh, double *x, double *y,
                       double *__restrict__ grad, unsigned long long *bar, int pm, int pn) {
  extern __shared__ double lds[];
  unsigned long long nbar = 0;
  auto grid_sync = [&] { grid_barrier(bar, ++nbar * gridDim.x); };
  const int w = blockIdx.x, tid = threadIdx.x;
  const int m0 = w * pm, n0 = w * pn;
  const int mr = max(0, min(M, m0 + pm) - m0), nr = max(0, min(N, n0 + pn) - n0);
  double *Ks = lds;
  double *KTs = Ks + (size_t)pm * N;
  double *xs = KTs + (size_t)pn * M;
  double *acc = xs + max(M, N);        // max(pm, pn) row results
  double *part = acc + max(pm, pn);
  for (long k = tid; k < (long)pm * N; k += kCoopTB) {
    const int rl = (int)(k % pm), c = (int)(k / pm);
    Ks[k] = rl < mr ? K[m0 + rl + (long)c * M] : 0.0;
  }
  for (long k = tid; k < (long)pn * M; k += kCoopTB) {
    const int rl = (int)(k % pn), c = (int)(k / pn);
    KTs[k] = rl < nr ? KT[n0 + rl + (long)c * N] : 0.0;
  }
  auto pass = [&](const double *S, int p, int rows, int R, int Cn) {
    const int G = kTinyTB / R, cw = (Cn + G - 1) / G;
    __syncthreads();
    for (int t = tid; t < rows * G; t += kCoopTB) {
      const int rl = t % rows, g = t / rows;
      part[g * p + rl] = tiny_dot([&](int c) { return S[rl + c * p]; }, xs, min(Cn, g * cw),
                                  min(Cn, (g + 1) * cw));
    }
    __syncthreads();
    for (int rl = tid; rl < rows; rl += kCoopTB) {
      double a = 0.0;
      for (int q = 0; q < G; ++q) a += part[q * p + rl];
      acc[rl] = a;
    }
    __syncthreads();
  };
  for (int l = L - 1; l > 0; --l) {
    const double *ylm1 = yh + (long)(l - 1) * M, *ul = uh + (long)l * M;
    const double *zl = zh + (long)l * N, *vl = vh + (long)l * N;
    for (int i = tid; i < M; i += kCoopTB) xs[i] = x[i];
    pass(KTs, pn, nr, N, M);  // K^T x
    for (int rl = tid; rl < nr; rl += kCoopTB) {
      const int j = n0 + rl;
      const double vb = -acc[rl];
      y[j] = (vb * vl[j]) / zl[j];
    }
    grid_sync();
    for (int j = tid; j < N; j += kCoopTB) xs[j] = y[j];
    pass(Ks, pm, mr, M, N);  // K y
    for (int rl = tid; rl < mr; rl += kCoopTB) {
      const int i = m0 + rl;
      const double ub = -acc[rl];
      grad[i] += ub / ylm1[i];
      x[i] = (ub * ul[i]) / ylm1[i];
    }
    grid_sync();
  }
}

// Workgroups of the multi-workgroup loop for an M x N problem (0: not this form):
// both sides at most kTinyTB (the tiny loop's column groups), and the
// smallest G >= 4 whose LDS fits.
int coop_groups(int M, int N) {
  if (M > kTinyTB || N > kTinyTB) return 0;
  for (int G = 4; G <= 64; G *= 2)
    if (coop_lds(M, N, G) * sizeof(double) <= 150 * 1024) return G;
  return 0;
}

// Which form runs a small vanilla problem: RWIG_VANILLA_FORM = "tiny" keeps
// the single-workgroup loop (for comparisons); anything else lets the
// multi-workgroup loop run where it applies.
bool coop_allowed() {
  const char *f = std::getenv("RWIG_VANILLA_FORM");
  return !(f && std::string(f) == "tiny");
}

// All pointers are device pointers of the reduced (support) problem.
// u (M), v (N), P (M*N), loss (1), grad (M, withgrad only) are outputs.
SinkhornResult sinkhorn_vanilla_dev(hipStream_t st, const double *a, const double *b,
                                    const double *C, int M, int N, double reg, bool wg,
                                    int maxiter, double tol, double *u, double *v, double *P,
                                    double *loss, double *grad) {
  const long MN = (long)M * N;
  DBuf<double> Kb(MN), Kvb(M), KTub(N), partb(gemv_n_part_size(M, N, 1));
  DBuf<double> sump(2 * kSumBlocks);
  DBuf<LoopState> lsb(1);
  double *K = Kb.get(), *Kv = Kvb.get(), *KTu = KTub.get(), *part = partb.get();
  LoopState *ls = lsb.get();
  const int L = maxiter;
  DBuf<double> uhb, vhb, yhb, zhb;
  if (wg) {
    uhb.reset((size_t)(L + 1) * M);
    vhb.reset((size_t)(L + 1) * N);
    yhb.reset((size_t)(L + 1) * M);
    zhb.reset((size_t)(L + 1) * N);
  }
  double *uh = uhb.get(), *vh = vhb.get(), *yh = yhb.get(), *zh = zhb.get();

  LoopState fin{};
  DBuf<double> KTb;  // K^T, kept for the adjoint of the multi-workgroup form
  int Gv = 0;        // workgroups of that form (0: not run)
  if (tiny_vanilla(M, N)) {
    KTb.reset(MN);
    const size_t lds = sizeof(double) * (2 * (size_t)M + 2 * (size_t)N +
                                         std::max({kTinyTB, M, N}));
    hipLaunchKernelGGL(k_gibbs_both, dim3((M + 31) / 32, (N + 31) / 32), dim3(1024), 0, st, C,
                       M, N, reg, K, KTb.get());
    check_launch();
    const int G = coop_allowed() ? coop_groups(M, N) : 0;
    Gv = G;
    if (G > 0) {
      const int pm = (M + G - 1) / G, pn = (N + G - 1) / G;
      const size_t clds = coop_lds(M, N, G) * sizeof(double);
      DBuf<double> xbuf(3 * (size_t)M + 3 * (size_t)N);
      DBuf<unsigned long long> barb(2);
      fill(st, barb.get(), 2, 0ull);
      CoopBufs xb{xbuf.get(), xbuf.get() + M, xbuf.get() + M + N, xbuf.get() + 3 * M + N,
                  barb.get()};
      RWIG_HIP_CHECK(hipFuncSetAttribute((const void *)k_sinkhorn_vanilla_coop,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)clds));
      hipLaunchKernelGGL(k_sinkhorn_vanilla_coop, dim3(G), dim3(kCoopTB), clds, st, a, b, M, N,
                         maxiter, tol, (const double *)K, (const double *)KTb.get(), u, v, ls,
                         uh, vh, yh, zh, xb, pm, pn);
      check_launch();
      unsigned long long failed = 0;
      d2h(&failed, barb.get() + 1, 1, st);
      RWIG_HIP_CHECK(hipStreamSynchronize(st));  // xbuf is released after the loop
      if (failed)
        throw Error(RWIG_ERR_HIP, "vanilla loop: a grid barrier timed out (the " +
                                      std::to_string(G) + " workgroups were not resident together)");
    } else {
      hipLaunchKernelGGL(k_sinkhorn_vanilla_tiny, dim3(1), dim3(kTinyTB), lds, st, a, b, C, M,
                         N, reg, maxiter, tol, K, KTb.get(), u, v, ls, uh, vh, yh, zh);
      check_launch();
    }
    d2h(&fin, ls, 1, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
  } else {
  launch_map(st, MN, [=] __device__(long k) { K[k] = exp(-C[k] / reg); });
  fill(st, u, M, 1.0);
  fill(st, v, N, 1.0);
  if (wg) {
    fill(st, uh, M, 1.0);
    fill(st, vh, N, 1.0);
  }
  init_loop(st, ls, maxiter, tol);
  // Kv = K v  (v = 1)
  gemv_n(st, K, M, N, M, v, N, 1, part,
         [=] __device__(long k, double s) {
           Kv[k] = s;
           if (wg) yh[k] = s;
         },
         nullptr);

  auto iteration = [&]() {
    // u = a / Kv
    launch_map(
        st, M,
        [=] __device__(long i) {
          const double ui = a[i] / Kv[i];
          u[i] = ui;
          if (wg) uh[(long)(ls->iter + 1) * M + i] = ui;
        },
        ls);
    // KTu = K^T u, v = b / KTu
    gemv_t(
        st, K, M, N, M, u, M, 1,
        [=] __device__(int j, const double *acc) {
          const double z = acc[0];
          KTu[j] = z;
          const double vj = b[j] / z;
          v[j] = vj;
          if (wg) {
            const long l = ls->iter + 1;
            vh[l * N + j] = vj;
            zh[l * N + j] = z;
          }
        },
        ls);
    // Kv = K v
    gemv_n(
        st, K, M, N, M, v, N, 1, part,
        [=] __device__(long i, double s) {
          Kv[i] = s;
          if (wg) yh[(long)(ls->iter + 1) * M + i] = s;
        },
        ls);
    // err = ||u .* Kv - a|| + ||v .* KTu - b||
    reduce2(
        st, (long)M,
        [=] __device__(long i) {
          const double r = u[i] * Kv[i] - a[i];
          return r * r;
        },
        (long)N,
        [=] __device__(long j) {
          const double r = v[j] * KTu[j] - b[j];
          return r * r;
        },
        [=] __device__(double s1, double s2) { finish_iteration(ls, sqrt(s1) + sqrt(s2)); }, ls);
  };
  fin = run_loop(st, ls, maxiter, iteration);
  }

  // P = diag(u) K diag(v)
  launch_map(st, MN, [=] __device__(long k) {
    const long i = k % M, j = k / M;
    P[k] = (u[i] * K[k]) * v[j];
  });
  sinkhorn_loss(st, MN, C, P, reg, sump.get(), loss);

  if (wg) {
    // adjoint: src/sinkhorn_impl_vanilla.cpp:169-214
    const int Lf = fin.iter;
    DBuf<double> PbKb(MN), ubarb(M), xb(M), yb(N);
    double *PbK = PbKb.get(), *ubar = ubarb.get(), *x = xb.get(), *y = yb.get();
    launch_map(st, MN, [=] __device__(long k) { PbK[k] = (C[k] + reg * log(P[k])) * K[k]; });
    fill(st, grad, M, 0.0);
    for (int l = Lf; l > 0; --l) {
      const double *ylm1 = yh + (long)(l - 1) * M;  // K v_{l-1}
      const double *zl = zh + (long)l * N;        // K^T u_l
      const double *ul = uh + (long)l * M;
      const double *vl = vh + (long)l * N;
      if (l == Lf) {
        // vbar = PbarK^T u; y = (vbar .* v) / (K^T u)
        gemv_t(
            st, PbK, M, N, M, u, M, 1,
            [=] __device__(int j, const double *acc) { y[j] = (acc[0] * v[j]) / zl[j]; },
            nullptr);
        // ubar = PbarK v - K y
        gemv_n(st, PbK, M, N, M, v, N, 1, part,
               [=] __device__(long i, double s) { ubar[i] = s; }, nullptr);
        gemv_n(
            st, K, M, N, M, y, N, 1, part,
            [=] __device__(long i, double s) {
              const double ub = ubar[i] - s;
              ubar[i] = ub;
              grad[i] += ub / ylm1[i];
              x[i] = (ub * ul[i]) / ylm1[i];
            },
            nullptr);
      } else if (Gv > 0) {
        // steps l .. 1 in one launch over the forward's workgroups
        const int pm = (M + Gv - 1) / Gv, pn = (N + Gv - 1) / Gv;
        const size_t clds = coop_lds(M, N, Gv) * sizeof(double);
        DBuf<unsigned long long> barb(2);
        fill(st, barb.get(), 2, 0ull);
        RWIG_HIP_CHECK(hipFuncSetAttribute((const void *)k_vanilla_adjoint_coop,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)clds));
        hipLaunchKernelGGL(k_vanilla_adjoint_coop, dim3(Gv), dim3(kCoopTB), clds, st, M, N,
                           l + 1, (const double *)K, (const double *)KTb.get(),
                           (const double *)uh, (const double *)vh, (const double *)yh,
                           (const double *)zh, x, y, grad, barb.get(), pm, pn);
        check_launch();
        unsigned long long failed = 0;
        d2h(&failed, barb.get() + 1, 1, st);
        RWIG_HIP_CHECK(hipStreamSynchronize(st));
        if (failed)
          throw Error(RWIG_ERR_HIP, "vanilla adjoint: a grid barrier timed out (the " +
                                        std::to_string(Gv) +
                                        " workgroups were not resident together)");
        break;
      } else {
        // vbar = -K^T x; y = (vbar .* v_l) / (K^T u_l)
        gemv_t(
            st, K, M, N, M, x, M, 1,
            [=] __device__(int j, const double *acc) {
              const double vb = -acc[0];
              y[j] = (vb * vl[j]) / zl[j];
            },
            nullptr);
        // ubar = -K y; grad += ubar / (K v_{l-1}); x for step l-1
        gemv_n(
            st, K, M, N, M, y, N, 1, part,
            [=] __device__(long i, double s) {
              const double ub = -s;
              ubar[i] = ub;
              grad[i] += ub / ylm1[i];
              x[i] = (ub * ul[i]) / ylm1[i];
            },
            nullptr);
      }
    }
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return {fin.iter, fin.err};
}

// ============================================================================
// Log-stabilised Sinkhorn (src/sinkhorn_impl_log.cpp:225-520)
// ============================================================================

SinkhornResult sinkhorn_log_dev(hipStream_t st, const double *a, const double *b,
                                const double *C, int M, int N, double reg, bool wg,
                                int maxiter, double tol, double *f, double *g, double *P,
                                double *loss, double *grad) {
  const long MN = (long)M * N;
  DBuf<double> logab(M), logbb(N), rminb(M), cminb(N), cmin2b(N);
  DBuf<double> pmb(lse_row_part_size(M, N, 1)), psb(lse_row_part_size(M, N, 1));
  DBuf<double> sump(2 * kSumBlocks);
  DBuf<LoopState> lsb(1);
  double *loga = logab.get(), *logb = logbb.get(), *rmin = rminb.get(), *cmin = cminb.get(),
         *cmin2 = cmin2b.get(), *pm = pmb.get(), *ps = psb.get();
  LoopState *ls = lsb.get();
  const int L = maxiter;
  DBuf<double> fhb, ghb, rmhb, cmhb;
  if (wg) {
    fhb.reset((size_t)(L + 1) * M);
    ghb.reset((size_t)(L + 1) * N);
    rmhb.reset((size_t)(L + 1) * M);
    cmhb.reset((size_t)(L + 1) * N);
  }
  double *fh = fhb.get(), *gh = ghb.get(), *rmh = rmhb.get(), *cmh = cmhb.get();

  fill(st, f, M, 0.0);
  fill(st, g, N, 0.0);
  if (wg) {
    fill(st, fh, M, 0.0);
    fill(st, gh, N, 0.0);
  }
  launch_map(st, M, [=] __device__(long i) { loga[i] = log(a[i]); });
  launch_map(st, N, [=] __device__(long j) { logb[j] = log(b[j]); });
  init_loop(st, ls, maxiter, tol);
  // Rminrow of R(0, 0)  (src/sinkhorn_impl_log.cpp:325)
  lse_row(st, C, M, N, M, f, M, g, N, 1, reg, pm, ps,
          [=] __device__(long i, double r) { rmin[i] = r; }, nullptr);

  auto iteration = [&]() {
    // f += reg * log a + Rminrow
    launch_map(
        st, M,
        [=] __device__(long i) {
          const double fi = f[i] + (reg * loga[i] + rmin[i]);
          f[i] = fi;
          if (wg) fh[(long)(ls->iter + 1) * M + i] = fi;
        },
        ls);
    // g += reg * log b + Rmincol(f, g)
    lse_col(
        st, C, M, N, M, f, M, g, N, 1, reg,
        [=] __device__(int j, const double *r) {
          cmin[j] = r[0];
          const double gj = g[j] + (reg * logb[j] + r[0]);
          g[j] = gj;
          if (wg) gh[(long)(ls->iter + 1) * N + j] = gj;
        },
        ls);
    // Rminrow, Rmincol of R(f, g) for the error (and the next f update)
    lse_row(
        st, C, M, N, M, f, M, g, N, 1, reg, pm, ps,
        [=] __device__(long i, double r) {
          rmin[i] = r;
          if (wg) rmh[(long)(ls->iter + 1) * M + i] = r;
        },
        ls);
    lse_col(
        st, C, M, N, M, f, M, g, N, 1, reg,
        [=] __device__(int j, const double *r) {
          cmin2[j] = r[0];
          if (wg) cmh[(long)(ls->iter + 1) * N + j] = r[0];
        },
        ls);
    // err = ||-Rminrow/reg - log a|| + ||-Rmincol/reg - log b||
    reduce2(
        st, (long)M,
        [=] __device__(long i) {
          const double r = -rmin[i] / reg - loga[i];
          return r * r;
        },
        (long)N,
        [=] __device__(long j) {
          const double r = -cmin2[j] / reg - logb[j];
          return r * r;
        },
        [=] __device__(double s1, double s2) { finish_iteration(ls, sqrt(s1) + sqrt(s2)); }, ls);
  };
  const LoopState fin = run_loop(st, ls, maxiter, iteration);

  // P = exp(-R(f, g) / reg)
  launch_map(st, MN, [=] __device__(long k) {
    const long i = k % M, j = k / M;
    const double r = (C[k] - f[i]) - g[j];
    P[k] = exp(-r / reg);
  });
  sinkhorn_loss(st, MN, C, P, reg, sump.get(), loss);

  if (wg) {
    const int Lf = fin.iter;
    DBuf<double> PbPb(MN), fbarb(M), gbarb(N), rsb(M), onesb(M > N ? M : N),
        partb(std::max(gemv_n_part_size(M, N, 1), wrow_part_size(M, N, 1)));
    double *PbP = PbPb.get(), *fbar = fbarb.get(), *gbar = gbarb.get(), *rs = rsb.get(),
           *ones = onesb.get(), *part = partb.get();
    launch_map(st, MN, [=] __device__(long k) { PbP[k] = (C[k] + reg * log(P[k])) * P[k]; });
    fill(st, ones, (long)(M > N ? M : N), 1.0);
    fill(st, grad, M, 0.0);
    for (int l = Lf; l > 0; --l) {
      const double *fl = fh + (long)l * M, *gl = gh + (long)l * N;
      const double *rml = rmh + (long)l * M, *cml = cmh + (long)l * N;
      if (l == Lf) {
        // gbar = PbarP^T 1 / reg ; fbar = -W gbar + PbarP 1 / reg
        gemv_t(st, PbP, M, N, M, ones, M, 1,
               [=] __device__(int j, const double *acc) { gbar[j] = acc[0] / reg; }, nullptr);
        gemv_n(st, PbP, M, N, M, ones, N, 1, part,
               [=] __device__(long i, double s) { rs[i] = s / reg; }, nullptr);
        wrow(
            st, C, M, N, M, fl, M, gl, N, 1, reg, cml, gbar, part,
            [=] __device__(long i, double s) {
              const double fb = -s + rs[i];
              fbar[i] = fb;
              grad[i] += fb / a[i];
            },
            nullptr);
      } else {
        // gbar = -X fbar, X the row soft-max of -R(f_l, g_l)/reg
        wcol(
            st, C, M, N, M, fl, M, gl, N, 1, reg, rml, fbar,
            [=] __device__(int j, const double *r) { gbar[j] = -r[0]; }, nullptr);
        // fbar = -W gbar, W the column soft-max of -R(f_l, g_l)/reg
        wrow(
            st, C, M, N, M, fl, M, gl, N, 1, reg, cml, gbar, part,
            [=] __device__(long i, double s) {
              const double fb = -s;
              fbar[i] = fb;
              grad[i] += fb / a[i];
            },
            nullptr);
      }
    }
    launch_map(st, M, [=] __device__(long i) { grad[i] *= reg; });
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return {fin.iter, fin.err};
}

// ============================================================================
// Parallel barycenter (src/barycenter_impl_parallel.cpp:19-207)
// ============================================================================

// Device pointers. A (M x S), C (M x N), w (S), bext (N, withgrad) inputs;
// b (N, normalised on return), U (M x S), V (N x S), gA (M x S), gw (S),
// loss (1) outputs.
SinkhornResult barycenter_parallel_dev(hipStream_t st, const double *A, const double *C,
                                       const double *w, const double *bext, int M, int N, int S,
                                       double reg, bool wg, int maxiter, double tol, double *b,
                                       double *U, double *V, double *gA, double *gw,
                                       double *loss) {
  const long MN = (long)M * N, MS = (long)M * S, NS = (long)N * S;
  DBuf<double> Kb(MN), KVb(MS), KTUb(NS), partb(gemv_n_part_size(M, N, S));
  DBuf<LoopState> lsb(1);
  double *K = Kb.get(), *KV = KVb.get(), *KTU = KTUb.get(), *part = partb.get();
  LoopState *ls = lsb.get();
  const int L = maxiter;
  DBuf<double> Uhb, Vhb, bhb, KVhb, KTUhb;
  if (wg) {
    Uhb.reset((size_t)(L + 1) * MS);
    Vhb.reset((size_t)(L + 1) * NS);
    bhb.reset((size_t)(L + 1) * N);
    KVhb.reset((size_t)(L + 1) * MS);
    KTUhb.reset((size_t)(L + 1) * NS);
  }
  double *Uh = Uhb.get(), *Vh = Vhb.get(), *bh = bhb.get(), *KVh = KVhb.get(),
         *KTUh = KTUhb.get();

  launch_map(st, MN, [=] __device__(long k) { K[k] = exp(-C[k] / reg); });
  fill(st, b, N, 0.0);
  fill(st, U, MS, 1.0);
  fill(st, V, NS, 1.0);
  if (wg) {
    fill(st, Uh, MS, 1.0);
    fill(st, Vh, NS, 1.0);
    fill(st, bh, N, 0.0);
  }
  init_loop(st, ls, maxiter, tol);
  gemv_n(st, K, M, N, M, V, N, S, part,
         [=] __device__(long k, double s) {
           KV[k] = s;
           if (wg) KVh[k] = s;
         },
         nullptr);

  auto iteration = [&]() {
    // U = A / KV
    launch_map(
        st, MS,
        [=] __device__(long k) {
          const double uk = A[k] / KV[k];
          U[k] = uk;
          if (wg) Uh[(long)(ls->iter + 1) * MS + k] = uk;
        },
        ls);
    // KTU = K^T U; b = prod_s KTU^w_s; V = b / KTU
    gemv_t(
        st, K, M, N, M, U, M, S,
        [=] __device__(int j, const double *acc) {
          double p = 1.0;
          for (int s = 0; s < S; ++s) p *= pow(acc[s], w[s]);
          b[j] = p;
          const long l = ls->iter + 1;
          if (wg) bh[l * N + j] = p;
          for (int s = 0; s < S; ++s) {
            const double vj = p / acc[s];
            KTU[j + (long)s * N] = acc[s];
            V[j + (long)s * N] = vj;
            if (wg) {
              Vh[l * NS + j + (long)s * N] = vj;
              KTUh[l * NS + j + (long)s * N] = acc[s];
            }
          }
        },
        ls);
    // KV = K V
    gemv_n(
        st, K, M, N, M, V, N, S, part,
        [=] __device__(long k, double s) {
          KV[k] = s;
          if (wg) KVh[(long)(ls->iter + 1) * MS + k] = s;
        },
        ls);
    // err = ||U .* KV - A||_F
    reduce2(
        st, MS,
        [=] __device__(long k) {
          const double r = U[k] * KV[k] - A[k];
          return r * r;
        },
        0L, [=] __device__(long) { return 0.0; },
        [=] __device__(double s1, double) { finish_iteration(ls, sqrt(s1)); }, ls);
  };
  const LoopState fin = run_loop(st, ls, maxiter, iteration);

  if (wg) {
    const int Lf = fin.iter;
    DBuf<double> Xb(NS), Ymb(MS), gwab(NS);
    double *X = Xb.get(), *Xm = Ymb.get(), *gwa = gwab.get();
    fill(st, gA, MS, 0.0);
    fill(st, gwa, NS, 0.0);
    for (int l = Lf; l > 0; --l) {
      const double *KTUl = KTUh + (long)l * NS, *Vl = Vh + (long)l * NS, *bl = bh + (long)l * N;
      const double *Ul = Uh + (long)l * MS, *KVlm1 = KVh + (long)(l - 1) * MS;
      if (l == Lf) {
        // bbar = 2 (b_L - b_ext); X = (bbar w^T) .* V_L; grad_w += log(KTU)^T (bbar .* b_L)
        launch_map(st, NS, [=] __device__(long k) {
          const long j = k % N, s = k / N;
          const double bb = 2.0 * (b[j] - bext[j]);
          X[k] = (bb * w[s]) * Vl[k];
          gwa[k] += log(KTUl[k]) * (bb * bl[j]);
        });
      } else {
        // Vbar = -K^T Xm; bbar = rowsum(Vbar / KTU);
        // X = (bbar w^T - Vbar / KTU) .* V_l; grad_w += log(KTU)^T (bbar .* b_l)
        gemv_t(
            st, K, M, N, M, Xm, M, S,
            [=] __device__(int j, const double *acc) {
              double bb = 0.0;
              for (int s = 0; s < S; ++s) bb += (-acc[s]) / KTUl[j + (long)s * N];
              for (int s = 0; s < S; ++s) {
                const long k = j + (long)s * N;
                X[k] = (bb * w[s] - (-acc[s]) / KTUl[k]) * Vl[k];
                gwa[k] += log(KTUl[k]) * (bb * bl[j]);
              }
            },
            nullptr);
      }
      // Ubar = K X; grad_A += Ubar / (K V_{l-1}); Xm = (Ubar .* U_l) / (K V_{l-1})
      gemv_n(
          st, K, M, N, M, X, N, S, part,
          [=] __device__(long k, double s) {
            gA[k] += s / KVlm1[k];
            Xm[k] = (s * Ul[k]) / KVlm1[k];
          },
          nullptr);
    }
    hipLaunchKernelGGL(k_colsum, dim3(S), dim3(kTB), 0, st, gwa, N, 1.0, gw);
    check_launch();
  }
  // b = b / sum(b); loss = ||b - b_ext||^2  (src/barycenter_impl_parallel.cpp:39-43)
  DBuf<double> totb(1);
  double *tot = totb.get();
  reduce2(
      st, (long)N, [=] __device__(long j) { return b[j]; }, 0L,
      [=] __device__(long) { return 0.0; }, [=] __device__(double s, double) { *tot = s; });
  launch_map(st, N, [=] __device__(long j) { b[j] = b[j] / *tot; });
  if (wg) {
    reduce2(
        st, (long)N,
        [=] __device__(long j) {
          const double r = b[j] - bext[j];
          return r * r;
        },
        0L, [=] __device__(long) { return 0.0; },
        [=] __device__(double s, double) { *loss = s; });
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return {fin.iter, fin.err};
}

// ============================================================================
// Log barycenter (src/barycenter_impl_log.cpp:23-382)
// ============================================================================

SinkhornResult barycenter_log_dev(hipStream_t st, const double *A, const double *C,
                                  const double *w, const double *bext, int M, int N, int S,
                                  double reg, bool wg, int maxiter, double tol, double *b,
                                  double *F, double *G, double *gA, double *gw,
                                  double *loss) {
  RWIG_REQUIRE(S <= 32, "log barycenter supports at most 32 measures");
  const long MS = (long)M * S, NS = (long)N * S;
  DBuf<double> logAb(MS), KVb(MS), KTUb(NS), logbb(N);
  DBuf<double> pmb(lse_row_part_size(M, N, S)), psb(lse_row_part_size(M, N, S));
  DBuf<LoopState> lsb(1);
  DBuf<double> gpartb(gram_part_size(S));
  double *logA = logAb.get(), *KV = KVb.get(), *KTU = KTUb.get(), *logb = logbb.get(),
         *pm = pmb.get(), *ps = psb.get(), *gpart = gpartb.get();
  LoopState *ls = lsb.get();
  const int L = maxiter;
  DBuf<double> Fhb, Ghb, KVhb, KTUhb;
  if (wg) {
    Fhb.reset((size_t)(L + 1) * MS);
    Ghb.reset((size_t)(L + 1) * NS);
    KVhb.reset((size_t)(L + 1) * MS);
    KTUhb.reset((size_t)(L + 1) * NS);
  }
  double *Fh = Fhb.get(), *Gh = Ghb.get(), *KVh = KVhb.get(), *KTUh = KTUhb.get();

  fill(st, logb, N, 0.0);
  fill(st, F, MS, 0.0);
  fill(st, G, NS, 0.0);
  if (wg) {
    fill(st, Fh, MS, 0.0);
    fill(st, Gh, NS, 0.0);
  }
  launch_map(st, MS, [=] __device__(long k) { logA[k] = log(A[k]); });
  init_loop(st, ls, maxiter, tol);
  lse_row(st, C, M, N, M, F, M, G, N, S, reg, pm, ps,
          [=] __device__(long k, double r) { KV[k] = r; }, nullptr);

  auto iteration = [&]() {
    // F += reg * log A + KV
    launch_map(
        st, MS,
        [=] __device__(long k) {
          const double fk = F[k] + (reg * logA[k] + KV[k]);
          F[k] = fk;
          if (wg) Fh[(long)(ls->iter + 1) * MS + k] = fk;
        },
        ls);
    // KTU = mincol(R_s(F, G)); logb = -(G + KTU) w / reg; G += reg logb 1^T + KTU
    lse_col(
        st, C, M, N, M, F, M, G, N, S, reg,
        [=] __device__(int j, const double *r) {
          double acc = 0.0;
          for (int s = 0; s < S; ++s) acc += -(G[j + (long)s * N] + r[s]) * w[s];
          const double lb = acc / reg;
          logb[j] = lb;
          const long l = ls->iter + 1;
          for (int s = 0; s < S; ++s) {
            const long k = j + (long)s * N;
            KTU[k] = r[s];
            const double gk = G[k] + (reg * lb + r[s]);
            G[k] = gk;
            if (wg) {
              Gh[l * NS + k] = gk;
              KTUh[l * NS + k] = r[s];
            }
          }
        },
        ls);
    // KV = minrow(R_s(F, G))
    lse_row(
        st, C, M, N, M, F, M, G, N, S, reg, pm, ps,
        [=] __device__(long k, double r) {
          KV[k] = r;
          if (wg) KVh[(long)(ls->iter + 1) * MS + k] = r;
        },
        ls);
    // err = norm(-KV/reg - log A, 2): spectral norm
    spectral(
        st, (long)M, S,
        [=] __device__(long i, int s) {
          const long k = i + (long)s * M;
          return -KV[k] / reg - logA[k];
        },
        [=] __device__(double nrm) { finish_iteration(ls, nrm); }, ls, gpart);
  };
  const LoopState fin = run_loop(st, ls, maxiter, iteration);

  // b = exp(logb) / sum  (src/barycenter_impl_log.cpp:268-271)
  DBuf<double> totb(1);
  double *tot = totb.get();
  launch_map(st, N, [=] __device__(long j) { b[j] = exp(logb[j]); });
  reduce2(
      st, (long)N, [=] __device__(long j) { return b[j]; }, 0L,
      [=] __device__(long) { return 0.0; }, [=] __device__(double s, double) { *tot = s; });
  launch_map(st, N, [=] __device__(long j) { b[j] = b[j] / *tot; });
  if (wg) {
    reduce2(
        st, (long)N,
        [=] __device__(long j) {
          const double r = b[j] - bext[j];
          return r * r;
        },
        0L, [=] __device__(long) { return 0.0; },
        [=] __device__(double s, double) { *loss = s; });

    const int Lf = fin.iter;
    DBuf<double> Fbarb(MS), Gbarb(NS), lbbb(N), Yb(NS), gwab(NS),
        partb(std::max(gemv_n_part_size(M, N, S), wrow_part_size(M, N, S)));
    double *Fbar = Fbarb.get(), *Gbar = Gbarb.get(), *lbb = lbbb.get(), *Y = Yb.get(),
           *gwa = gwab.get(), *part = partb.get();
    fill(st, gA, MS, 0.0);
    fill(st, gwa, NS, 0.0);
    fill(st, Fbar, MS, 0.0);
    fill(st, Gbar, NS, 0.0);
    for (int l = Lf; l > 0; --l) {
      const double *Fl = Fh + (long)l * MS, *Gl = Gh + (long)l * NS;
      const double *Glm1 = Gh + (long)(l - 1) * NS;
      const double *KVl = KVh + (long)l * MS, *KTUl = KTUh + (long)l * NS;
      if (l != Lf) {
        // Gbar(:,s) = -X_s^T Fbar(:,s); logbbar = reg * rowsum(Gbar)
        wcol(
            st, C, M, N, M, Fl, M, Gl, N, S, reg, KVl, Fbar,
            [=] __device__(int j, const double *r) {
              double acc = 0.0;
              for (int s = 0; s < S; ++s) {
                Gbar[j + (long)s * N] = -r[s];
                acc += -r[s];
              }
              lbb[j] = reg * acc;
            },
            nullptr);
      } else {
        // logbbar = 2 (b - b_ext) .* b with the normalised b
        launch_map(st, N, [=] __device__(long j) { lbb[j] = 2.0 * (b[j] - bext[j]) * b[j]; });
      }
      // Fbar(:,s) = W_s ((w_s / reg) logbbar - Gbar(:,s)); W_s the column
      // soft-max of -R_s(F_l, G_{l-1}) / reg
      const bool last = (l == Lf);
      launch_map(st, NS, [=] __device__(long k) {
        const long j = k % N, s = k / N;
        const double t = (w[s] / reg) * lbb[j];
        Y[k] = last ? t : t - Gbar[k];
        gwa[k] += (Glm1[k] + KTUl[k]) * lbb[j];
      });
      wrow(
          st, C, M, N, M, Fl, M, Glm1, N, S, reg, KTUl, Y, part,
          [=] __device__(long k, double s) {
            Fbar[k] = s;
            gA[k] += s / A[k];
          },
          nullptr);
    }
    launch_map(st, MS, [=] __device__(long k) { gA[k] *= reg; });
    hipLaunchKernelGGL(k_colsum, dim3(S), dim3(kTB), 0, st, gwa, N, -1.0 / reg, gw);
    check_launch();
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
  return {fin.iter, fin.err};
}

}  // namespace rwig

// ============================================================================
// C ABI (include/rwig_hip.h)
// ============================================================================

using namespace rwig;

namespace {

void check_counts(int m, int n, int maxiter, double reg) {
  RWIG_REQUIRE(m > 0 && n > 0, "matrix dimensions must be positive");
  RWIG_REQUIRE(maxiter >= 0, "max_iter must be non-negative");
  RWIG_REQUIRE(reg > 0.0, "reg must be positive");
}

template <class Core>
int sinkhorn_abi(const char *name, Core core, double *P, double *grad_a, double *u, double *v,
                 double *loss, int *iter_out, double *err_out, const double *a, const double *b,
                 const double *C, int m, int n, double reg, int withgrad, int max_iter,
                 double zero_tol, int *return_status_out, double off_support) {
  return guarded(name, [&] {
    check_counts(m, n, max_iter, reg);
    RWIG_REQUIRE(a && b && C, "a, b and C are required");
    ensure_device();
    Reduced r = reduce_support(a, b, C, m, n);
    RWIG_REQUIRE(r.M > 0 && r.N > 0, "a and b must each have a non-zero entry");
    Stream s;
    const long MN = (long)r.M * r.N;
    DBuf<double> da(r.M), db(r.N), dC(MN), du(r.M), dv(r.N), dP(MN), dloss(1), dg(r.M);
    h2d(da.get(), r.a.data(), r.M, s.get());
    h2d(db.get(), r.b.data(), r.N, s.get());
    h2d(dC.get(), r.C.data(), MN, s.get());
    const SinkhornResult res =
        core(s.get(), da.get(), db.get(), dC.get(), r.M, r.N, reg, withgrad != 0, max_iter,
             zero_tol, du.get(), dv.get(), dP.get(), dloss.get(), dg.get());
    std::vector<double> hu(r.M), hv(r.N), hP(P ? MN : 0), hg(withgrad ? r.M : 0);
    double hl = 0.0;
    d2h(hu.data(), du.get(), r.M, s.get());
    d2h(hv.data(), dv.get(), r.N, s.get());
    if (P) d2h(hP.data(), dP.get(), MN, s.get());
    d2h(&hl, dloss.get(), 1, s.get());
    if (withgrad) d2h(hg.data(), dg.get(), r.M, s.get());
    s.sync();
    if (u) {
      for (int i = 0; i < m; ++i) u[i] = off_support;
      for (int i = 0; i < r.M; ++i) u[r.ai[i]] = hu[i];
    }
    if (v) {
      for (int j = 0; j < n; ++j) v[j] = off_support;
      for (int j = 0; j < r.N; ++j) v[r.bi[j]] = hv[j];
    }
    if (P) {
      std::fill(P, P + (size_t)m * n, 0.0);
      for (int j = 0; j < r.N; ++j)
        for (int i = 0; i < r.M; ++i) P[r.ai[i] + (size_t)r.bi[j] * m] = hP[i + (size_t)j * r.M];
    }
    if (withgrad && grad_a) {
      std::fill(grad_a, grad_a + m, 0.0);
      for (int i = 0; i < r.M; ++i) grad_a[r.ai[i]] = hg[i];
    }
    if (loss) *loss = hl;
    if (iter_out) *iter_out = res.iter;
    if (err_out) *err_out = res.err;
    if (return_status_out)
      *return_status_out = return_status(res.err, zero_tol, res.iter, max_iter);
  });
}

template <class Core>
int barycenter_abi(const char *name, Core core, double *U, double *V, double *b, double *grad_A,
                   double *grad_w, double *loss, int *iter_out, double *err_out,
                   const double *A, const double *w, const double *C, const double *b_ext, int M,
                   int N, int S, double reg, int withgrad, int max_iter, double zero_tol,
                   int *return_status_out) {
  return guarded(name, [&] {
    check_counts(M, N, max_iter, reg);
    RWIG_REQUIRE(S > 0 && S <= 32, "S must be in [1, 32]");
    RWIG_REQUIRE(A && w && C, "A, w and C are required");
    RWIG_REQUIRE(!withgrad || b_ext, "b_ext is required with withgrad");
    ensure_device();
    Stream s;
    const long MN = (long)M * N, MS = (long)M * S, NS = (long)N * S;
    DBuf<double> dA(MS), dC(MN), dw(S), dbext(N), db(N), dU(MS), dV(NS), dgA(MS), dgw(S),
        dloss(1);
    h2d(dA.get(), A, MS, s.get());
    h2d(dC.get(), C, MN, s.get());
    h2d(dw.get(), w, S, s.get());
    if (withgrad) h2d(dbext.get(), b_ext, N, s.get());
    const SinkhornResult res =
        core(s.get(), dA.get(), dC.get(), dw.get(), dbext.get(), M, N, S, reg, withgrad != 0,
             max_iter, zero_tol, db.get(), dU.get(), dV.get(), dgA.get(), dgw.get(), dloss.get());
    if (b) d2h(b, db.get(), N, s.get());
    if (U) d2h(U, dU.get(), MS, s.get());
    if (V) d2h(V, dV.get(), NS, s.get());
    if (withgrad) {
      if (grad_A) d2h(grad_A, dgA.get(), MS, s.get());
      if (grad_w) d2h(grad_w, dgw.get(), S, s.get());
      if (loss) d2h(loss, dloss.get(), 1, s.get());
    }
    s.sync();
    if (iter_out) *iter_out = res.iter;
    if (err_out) *err_out = res.err;
    if (return_status_out)
      *return_status_out = return_status(res.err, zero_tol, res.iter, max_iter);
  });
}

}  // namespace

extern "C" {

int rwig_sinkhorn_vanilla(double *P, double *grad_a, double *u, double *v, double *loss,
                          int *iter_out, double *err_out, const double *a, const double *b,
                          const double *C, int m, int n, double reg, int withgrad, int max_iter,
                          double zero_tol, int *return_status_out) {
  return sinkhorn_abi("rwig_sinkhorn_vanilla", sinkhorn_vanilla_dev, P, grad_a, u, v, loss,
                      iter_out, err_out, a, b, C, m, n, reg, withgrad, max_iter, zero_tol,
                      return_status_out, 0.0);
}

int rwig_sinkhorn_log(double *P, double *grad_a, double *f, double *g, double *loss,
                      int *iter_out, double *err_out, const double *a, const double *b,
                      const double *C, int m, int n, double reg, int withgrad, int max_iter,
                      double zero_tol, int *return_status_out) {
  return sinkhorn_abi("rwig_sinkhorn_log", sinkhorn_log_dev, P, grad_a, f, g, loss, iter_out,
                      err_out, a, b, C, m, n, reg, withgrad, max_iter, zero_tol,
                      return_status_out, -INFINITY);
}

int rwig_barycenter_parallel(double *U, double *V, double *b, double *grad_A, double *grad_w,
                             double *loss, int *iter_out, double *err_out, const double *A,
                             const double *w, const double *C, const double *b_ext, int M, int N,
                             int S, double reg, int withgrad, int max_iter, double zero_tol,
                             int *return_status_out) {
  return barycenter_abi("rwig_barycenter_parallel", barycenter_parallel_dev, U, V, b, grad_A,
                        grad_w, loss, iter_out, err_out, A, w, C, b_ext, M, N, S, reg, withgrad,
                        max_iter, zero_tol, return_status_out);
}

int rwig_barycenter_log(double *F, double *G, double *b, double *grad_A, double *grad_w,
                        double *loss, int *iter_out, double *err_out, const double *A,
                        const double *w, const double *C, const double *b_ext, int M, int N,
                        int S, double reg, int withgrad, int max_iter, double zero_tol,
                        int *return_status_out) {
  return barycenter_abi("rwig_barycenter_log", barycenter_log_dev, F, G, b, grad_A, grad_w,
                        loss, iter_out, err_out, A, w, C, b_ext, M, N, S, reg, withgrad,
                        max_iter, zero_tol, return_status_out);
}

int rwig_sinkhorn_log_dev(const double *a, const double *b, const double *C, int m, int n,
                          double reg, int withgrad, int max_iter, double zero_tol, double *P,
                          double *grad_a, double *f, double *g, double *loss, int *iter_out,
                          double *err_out, void *stream) {
  return guarded("rwig_sinkhorn_log_dev", [&] {
    check_counts(m, n, max_iter, reg);
    RWIG_REQUIRE(a && b && C && P && f && g && loss, "null argument");
    RWIG_REQUIRE(!withgrad || grad_a, "grad_a is required with withgrad");
    const SinkhornResult res =
        sinkhorn_log_dev((hipStream_t)stream, a, b, C, m, n, reg, withgrad != 0, max_iter,
                         zero_tol, f, g, P, loss, grad_a);
    if (iter_out) *iter_out = res.iter;
    if (err_out) *err_out = res.err;
  });
}

}  // extern "C"
