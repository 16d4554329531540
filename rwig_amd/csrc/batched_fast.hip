// Mixed-precision batched Sinkhorn (one source against many targets) with
// FP64 log-domain absorption and defect-corrected low-precision iterations.
//
// State of target t: dual potentials f (rows), g (its support columns) in
// FP64, and scalings u, v (FP32) on top of them. The transport plan is
// P_ij = u_i exp((f_i + g_j - C_ij)/reg) v_j.
//
// Absorption (FP64, every absorb_every iterations and at the end):
//   f += reg log u, g += reg log v, u = v = 1;
//   one sweep over the target's cost columns evaluates the exact plan
//   E_ij = exp(-((C_ij - f_i) - g_j)/reg), its row and column sums r0, c0,
//   the marginal error (checked against the tolerance), the transport cost
//   and the regularised loss. The same sweep evaluates, with exactly the
//   arithmetic of the iterations below, the low-precision row and column
//   sums rl = alpha .* (K beta), cl = beta .* (K^T alpha) of the same plan
//   and keeps the ratios rho = r0 / rl, sigma = c0 / cl.
// Iteration (FP32 arithmetic on a shared low-precision Gibbs kernel
// K = exp(-C/reg), stored as FP32 or BF16; alpha = exp(f/reg),
// beta = exp(g/reg)):
//   v = b / (sigma .* beta .* K^T (alpha .* u))
//   u = a / (rho .* alpha .* K (beta .* v))
// This is the Sinkhorn iteration on the absorbed plan with the rounding of
// the low-precision kernel divided out multiplicatively: at u = v = 1 the
// corrected products equal the exact FP64 marginals bit for bit, so the fixed
// point is the FP64 one (defect correction), and every product sums positive
// terms only. Both products of one iteration come from a single sweep over
// K: each column's K^T product gives v_j, and the same column values, still
// in registers, accumulate K (beta .* v).
//
// One workgroup owns one target; thread k owns rows 4(k + m TB) .. +3 for
// m < RG so a wavefront's loads of a column are contiguous 8- or 16-byte
// vectors.
#include "kernels.hpp"

namespace rwig {

namespace {

constexpr int kFTB = 512;
constexpr int kFW = kFTB / kWave;

template <typename KT> struct KLoad;
template <> struct KLoad<float> {
  __device__ static void load4(const float *p, float o[4]) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
};
template <> struct KLoad<uint16_t> {
  __device__ static void load4(const uint16_t *p, float o[4]) {
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    o[0] = __uint_as_float(v.x << 16);
    o[1] = __uint_as_float(v.x & 0xffff0000u);
    o[2] = __uint_as_float(v.y << 16);
    o[3] = __uint_as_float(v.y & 0xffff0000u);
  }
};

__device__ __forceinline__ int row_of(int k) {
  return 4 * (threadIdx.x + (k / 4) * kFTB) + (k % 4);
}

// Loads the CB columns jb.. of K for this thread's rows (zeros past nt / M).
template <int RG, int CB, typename KT>
__device__ __forceinline__ void load_chunk(const KT *__restrict__ K, long ldk, int M,
                                           const int *__restrict__ cols, int jb, int nt,
                                           float kv[CB][RG * 4]) {
#pragma unroll
  for (int q = 0; q < CB; ++q) {
    const int jj = jb + q;
    const bool ok = jj < nt;
    const long col = ok ? (long)cols[jj] : 0;
#pragma unroll
    for (int m = 0; m < RG; ++m) {
      float o[4] = {0.f, 0.f, 0.f, 0.f};
      const int i0 = 4 * (threadIdx.x + m * kFTB);
      if (ok && i0 < M) KLoad<KT>::load4(K + i0 + col * ldk, o);
#pragma unroll
      for (int e = 0; e < 4; ++e) kv[q][m * 4 + e] = o[e];
    }
  }
}

template <int RG, int CB, typename KT>
__global__ void __launch_bounds__(kFTB)
k_sinkhorn_fast_resident(const double *__restrict__ C, long ldc, const KT *__restrict__ K,
                         long ldk, int M, const double *__restrict__ a,
                         const int *__restrict__ col_ptr, const int *__restrict__ col_idx,
                         const double *__restrict__ b, double reg, int maxiter, double tol,
                         int err_mode, int absorb_every, double *__restrict__ Fout,
                         double *__restrict__ Gout, int *__restrict__ iters,
                         double *__restrict__ errs, double *__restrict__ loss) {
  extern __shared__ double smem[];
  const int t = blockIdx.x;
  const int c0i = col_ptr[t], nt = col_ptr[t + 1] - c0i;
  const int *cols = col_idx + c0i;
  const double *bt = b + c0i;
  double *gS = smem;                                           // nt: g
  double *red = gS + nt;                                       // 4 * kFW * CB
  float *beS = reinterpret_cast<float *>(red + 4 * kFW * CB);  // nt: beta
  float *sbS = beS + nt;                                       // nt: sigma * beta
  float *vS = sbS + nt;                                        // nt: v
  __shared__ double redsum[kFW];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  constexpr int R = RG * 4;
  // f lives in this target's row of Fout between absorptions
  double *fT = Fout + (long)t * M;

  float al[R], ra[R], u[R];
  bool own[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    own[k] = row_of(k) < M;
    if (own[k]) fT[row_of(k)] = 0.0;
    u[k] = 1.0f;
    al[k] = 0.0f;
    ra[k] = 0.0f;
  }
  for (int q = threadIdx.x; q < nt; q += kFTB) {
    gS[q] = 0.0;
    vS[q] = 1.0f;
  }
  __syncthreads();

  auto block_total = [&](double x) -> double {
    x = wave_sum(x);
    if (lane == 0) redsum[wv] = x;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < kFW; ++w) s += redsum[w];
    __syncthreads();
    return s;
  };

  int iter = 0;
  double err = 1000.0, cost = 0.0, regloss = 0.0;
  while (true) {
    // ------------------------------------------------ absorption (FP64)
    double fr[R], r0[R];
    float y[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      fr[k] = 0.0;
      r0[k] = 0.0;
      y[k] = 0.0f;
      if (!own[k]) continue;
      const double fk = fT[row_of(k)] + reg * log((double)u[k]);
      fT[row_of(k)] = fk;
      fr[k] = fk;
      al[k] = (float)exp(fk / reg);
      u[k] = 1.0f;
    }
    for (int q = threadIdx.x; q < nt; q += kFTB) {
      const double gq = gS[q] + reg * log((double)vS[q]);
      gS[q] = gq;
      beS[q] = (float)exp(gq / reg);
      vS[q] = 1.0f;
    }
    __syncthreads();
    // the absorption sweep carries FP64 and FP32 state: half-width column chunks
    constexpr int CA = (CB >= 4) ? CB / 2 : CB;
    double ccost = 0.0, cent = 0.0, colerr = 0.0;
    for (int jb = 0; jb < nt; jb += CA) {
      float kv[CA][R];
      load_chunk<RG, CA, KT>(K, ldk, M, cols, jb, nt, kv);
      double cs[CA];
      float zs[CA];
#pragma unroll
      for (int q = 0; q < CA; ++q) {
        const int jj = jb + q;
        const bool ok = jj < nt;
        const long col = ok ? (long)cols[jj] : 0;
        const double gq = ok ? gS[jj] : 0.0;
        double s = 0.0;
        float z = 0.0f;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if (ok && own[k]) {
            const double cij = C[row_of(k) + col * ldc];
            const double rr = (cij - fr[k]) - gq;
            const double p = exp(-rr / reg);
            r0[k] += p;
            s += p;
            ccost += cij * p;
            cent += p * (-rr / reg - 1.0);
          }
          z += kv[q][k] * (al[k] * u[k]);
        }
        cs[q] = wave_sum(s);
        zs[q] = wave_sum(z);
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < CA; ++q) {
          red[wv * CA + q] = cs[q];
          red[kFW * CA + wv * CA + q] = (double)zs[q];
        }
      }
      __syncthreads();
      float X[CA];
#pragma unroll
      for (int q = 0; q < CA; ++q) {
        const int jj = jb + q;
        X[q] = 0.0f;
        if (jj < nt) {
          double c0 = 0.0, z = 0.0;
          for (int w = 0; w < kFW; ++w) {
            c0 += red[w * CA + q];
            z += red[kFW * CA + w * CA + q];
          }
          const float be = beS[jj];
          const double cl = (double)be * z;  // low-precision column sum at u = 1
          if (threadIdx.x == 0) {
            sbS[jj] = (float)((c0 / cl) * (double)be);
            const double bq = bt[jj];
            if (err_mode == 0) {
              const double r = log(c0) - log(bq);
              colerr += r * r;
            } else {
              colerr += fabs(c0 - bq);
            }
          }
          X[q] = be * vS[jj];
        }
      }
#pragma unroll
      for (int q = 0; q < CA; ++q)
#pragma unroll
        for (int k = 0; k < R; ++k) y[k] += kv[q][k] * X[q];
      __syncthreads();  // red is reused by the next chunk
    }
    double rowerr = 0.0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (!own[k]) continue;
      const double ak = a[row_of(k)];
      if (err_mode == 0) {
        const double r = log(r0[k]) - log(ak);
        rowerr += r * r;
      } else {
        rowerr += fabs(r0[k] - ak);
      }
      // rho * alpha with rho = r0 / (alpha * (K beta))
      ra[k] = (float)(r0[k] / (double)y[k]);
    }
    const double rt = block_total(rowerr);
    const double ct = block_total(colerr);  // only thread 0 contributed
    err = (err_mode == 0) ? sqrt(rt) + sqrt(ct) : rt + ct;
    cost = block_total(ccost);
    regloss = cost + reg * block_total(cent);
    if (!(iter < maxiter && err >= tol)) break;

    // ------------------------------------------------ low-precision iterations
    // u = a / (rho alpha K (beta v)) with v = 1: y still holds K beta
#pragma unroll
    for (int k = 0; k < R; ++k)
      u[k] = own[k] ? (float)(a[row_of(k)] / ((double)ra[k] * (double)y[k])) : 0.0f;
    const int nin = min(absorb_every, maxiter - iter);
    for (int it = 0; it < nin; ++it) {
      float W[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        W[k] = al[k] * u[k];
        y[k] = 0.0f;
      }
      for (int jb = 0; jb < nt; jb += CB) {
        float kv[CB][R];
        load_chunk<RG, CB, KT>(K, ldk, M, cols, jb, nt, kv);
        float zq[CB];
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          float z = 0.0f;
#pragma unroll
          for (int k = 0; k < R; ++k) z += kv[q][k] * W[k];
          zq[q] = wave_sum(z);
        }
        if (lane == 0) {
#pragma unroll
          for (int q = 0; q < CB; ++q) red[wv * CB + q] = (double)zq[q];
        }
        __syncthreads();
        // lanes 0..CB-1 finish one column each and publish beta_j v_j
        float *xS = reinterpret_cast<float *>(red + 2 * kFW * CB);
        if (threadIdx.x < CB) {
          const int q = threadIdx.x, jj = jb + q;
          float xq = 0.0f;
          if (jj < nt) {
            double z = 0.0;
            for (int w = 0; w < kFW; ++w) z += red[w * CB + q];
            const float vn = (float)(bt[jj] / ((double)sbS[jj] * z));
            xq = beS[jj] * vn;
            vS[jj] = vn;
          }
          xS[q] = xq;
        }
        __syncthreads();  // also orders the reuse of red by the next chunk
#pragma unroll
        for (int q = 0; q < CB; ++q) {
          const float xq = xS[q];
#pragma unroll
          for (int k = 0; k < R; ++k) y[k] += kv[q][k] * xq;
        }
      }
#pragma unroll
      for (int k = 0; k < R; ++k)
        u[k] = own[k] ? (float)(a[row_of(k)] / ((double)ra[k] * (double)y[k])) : 0.0f;
      ++iter;
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < nt; q += kFTB) Gout[c0i + q] = gS[q];
  if (threadIdx.x == 0) {
    iters[t] = iter;
    errs[t] = err;
    if (loss) {
      loss[2 * t] = cost;
      loss[2 * t + 1] = regloss;
    }
  }
}

// K = lowp(exp(-C/reg)), rows padded with zeros up to ldk.
template <typename KT> __device__ KT to_lowp(double x);
template <> __device__ float to_lowp<float>(double x) { return (float)x; }
template <> __device__ uint16_t to_lowp<uint16_t>(double x) {
  // round to nearest even from FP32
  const float fx = (float)x;
  uint32_t u = __float_as_uint(fx);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <typename KT>
__global__ void __launch_bounds__(kTB)
k_gibbs_lowp(const double *__restrict__ C, long ldc, int M, int N, double reg, KT *K,
             long ldk) {
  const long n = ldk * (long)N;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x) {
    const long i = k % ldk, j = k / ldk;
    K[k] = (i < M) ? to_lowp<KT>(exp(-C[i + j * ldc] / reg)) : to_lowp<KT>(0.0);
  }
}

}  // namespace

void gibbs_lowp(hipStream_t st, const double *C, long ldc, int M, int N, double reg,
                int precision, void *K, long ldk) {
  RWIG_REQUIRE(ldk >= M && ldk % 4 == 0, "ldk must be >= M and a multiple of 4");
  const long n = ldk * (long)N;
  if (precision == 1)
    hipLaunchKernelGGL(k_gibbs_lowp<float>, dim3(grid_for(n, kTB, 8192)), dim3(kTB), 0, st, C,
                       ldc, M, N, reg, (float *)K, ldk);
  else if (precision == 2)
    hipLaunchKernelGGL(k_gibbs_lowp<uint16_t>, dim3(grid_for(n, kTB, 8192)), dim3(kTB), 0, st,
                       C, ldc, M, N, reg, (uint16_t *)K, ldk);
  else
    throw Error(RWIG_ERR_INVALID, "precision must be 1 (fp32) or 2 (bf16)");
  check_launch();
}

void sinkhorn_batched_fast(hipStream_t st, const double *C, long ldc, const void *K, long ldk,
                           int precision, int M, const double *a, int B, const int *col_ptr,
                           const int *col_idx, const double *b, int max_nt, double reg,
                           int maxiter, double tol, int err_mode, int absorb_every, double *F,
                           double *G, int *iters, double *errs, double *loss) {
  RWIG_REQUIRE(M >= 1 && M <= 4 * 4 * kFTB, "batched solve supports 1..8192 source points");
  RWIG_REQUIRE(ldk >= M && ldk % 4 == 0, "ldk must be >= M and a multiple of 4");
  RWIG_REQUIRE(absorb_every >= 1, "absorb_every must be positive");
  constexpr int CBmax = 8;
  const size_t shmem = sizeof(double) * ((size_t)max_nt + 4 * kFW * CBmax) +
                       sizeof(float) * 3 * (size_t)max_nt;
  RWIG_REQUIRE(max_nt >= 1 && shmem <= 65536, "targets must have 1..2700 support points");
  const dim3 grid(B), block(kFTB);
#define RWIG_LAUNCH_FAST(RG_, CB_, KT_)                                                         \
  hipLaunchKernelGGL((k_sinkhorn_fast_resident<RG_, CB_, KT_>), grid, block, shmem, st, C, ldc, \
                     (const KT_ *)K, ldk, M, a, col_ptr, col_idx, b, reg, maxiter, tol,         \
                     err_mode, absorb_every, F, G, iters, errs, loss)
  if (precision == 1) {
    if (M <= 4 * kFTB) RWIG_LAUNCH_FAST(1, 8, float);
    else if (M <= 8 * kFTB) RWIG_LAUNCH_FAST(2, 8, float);
    else RWIG_LAUNCH_FAST(4, 4, float);
  } else if (precision == 2) {
    if (M <= 4 * kFTB) RWIG_LAUNCH_FAST(1, 8, uint16_t);
    else if (M <= 8 * kFTB) RWIG_LAUNCH_FAST(2, 8, uint16_t);
    else RWIG_LAUNCH_FAST(4, 4, uint16_t);
  } else {
    throw Error(RWIG_ERR_INVALID, "precision must be 1 (fp32) or 2 (bf16)");
  }
#undef RWIG_LAUNCH_FAST
  check_launch();
}

}  // namespace rwig
