// Batched Sinkhorn on MFMA: one source against B targets that share one cost
// matrix, with every Gibbs product of an iteration done as one dense GEMM
// over all targets (targets are the GEMM columns).
//
// Notation: K = exp(-C/reg) (M x Nc), target t supported on S_t subset of
// the Nc columns. Potentials f (M x B), g (on the supports) in FP64;
// alpha = exp(f/reg), beta = exp(g/reg); the absorbed plan of target t is
// P0_ij = alpha_i K_ij beta_j and the current plan P = diag(u) P0 diag(v).
//
// Per iteration (BF16 MFMA, FP32 accumulation; each product writes its FP32
// tile and an update kernel reads it: k_support_cycle at the support entries
// only for Z, k_rows4 four rows per thread for Y):
//   Z = K^T W  ->  v on the supports, X = beta .* (v or v - 1) (zero off S_t)
//   Y = K X    ->  u,                 W = alpha .* (u or u - 1)
// Two forms of the update, chosen per target at every absorption:
//   multiplicative (far from convergence):  v = b / (sb .* z), u = a / (ra .* y)
//     with W = alpha u, X = beta v and calibration ratios ra = r0 / (K beta),
//     sb = c0 / (K^T alpha) computed with the same BF16 arithmetic at u = v = 1,
//     so the rounding of K divides out; all sums are of positive terms.
//   additive (near convergence):  v = b / (c0 + beta .* z), u = a / (r0 + alpha .* y)
//     with W = alpha (u - 1), X = beta (v - 1): the exact FP64 marginals r0, c0
//     carry the O(1) part and the BF16 products only the corrections, whose
//     rounding vanishes with (u - 1, v - 1) -- the fixed point is the FP64 one.
// Each cycle folds an update's per-element inputs into one FP32 factor:
// alpha / ra or the defect a - r0 for Y, beta b / sb or the defect b - c0 for
// Z. Near convergence the defects are small, so their FP32 rounding moves the
// fixed point by a relative 6e-8 of the defect.
// Absorption (FP64, every absorb_every iterations): f += reg log u,
// g += reg log v (between absorptions u, v are kept in FP32, as u - 1 and
// v - 1 in the additive form); r0 = alpha .* (K beta), c0 = beta .* (K^T alpha) by FP64
// MFMA GEMMs; the L1 marginal error of the absorbed plan decides convergence
// (a target stops at the first absorption with error <= tol) and the form of
// the next cycle. Stopped targets leave the GEMM columns, and the products
// take smaller tiles as the columns thin out. At the end one FP64 GEMM with
// K o C gives each target's transport cost.
#include "dense.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace rwig {

namespace {

// per-target L1 (or reference) marginal error of the absorbed plan, stopping
// decision, form of the next cycle; one block per target. cnt[0] counts the
// targets stopped so far, cnt[1] the running ones in the multiplicative form.
__global__ void __launch_bounds__(256)
k_dense_err(int M, int B, const double *__restrict__ a, const double *__restrict__ r0,
            const int *__restrict__ col_ptr, const double *__restrict__ b,
            const double *__restrict__ c0, int err_mode, double tol, int iter, int maxiter,
            double switch_thr, int *done, int *mode, int *iters, double *errs,
            int *cnt) {
  __shared__ double sh[256 / kWave];
  const int t = blockIdx.x;
  if (done[t]) return;
  double e1 = 0.0, e2 = 0.0;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    const double r = r0[i + (long)t * M];
    if (err_mode == 0) {
      const double d = log(r) - log(a[i]);
      e1 += d * d;
    } else {
      e1 += fabs(r - a[i]);
    }
  }
  for (int p = col_ptr[t] + threadIdx.x; p < col_ptr[t + 1]; p += blockDim.x) {
    if (err_mode == 0) {
      const double d = log(c0[p]) - log(b[p]);
      e2 += d * d;
    } else {
      e2 += fabs(c0[p] - b[p]);
    }
  }
  e1 = block_sum(e1, sh);
  e2 = block_sum(e2, sh);
  if (threadIdx.x == 0) {
    const double err = (err_mode == 0) ? sqrt(e1) + sqrt(e2) : e1 + e2;
    errs[t] = err;
    iters[t] = iter;
    if (!(iter < maxiter && err >= tol)) {
      done[t] = 1;
      atomicAdd(&cnt[0], 1);
    }
    mode[t] = (err < switch_thr) ? 1 : 0;
    if (!done[t] && !mode[t]) atomicAdd(&cnt[1], 1);
  }
}

// transport cost and regularised loss of the absorbed plan P (marginals r0,
// c0), one block per target: <P, C> = sum_i pc_i with pc = alpha .* ((K o C) beta),
// and, since log P = (f + g - C) / reg,
//   <P, C> + reg sum P (log P - 1) = f . r0 + g . c0 - reg sum r0
__global__ void __launch_bounds__(256)
k_dense_loss(int M, const int *__restrict__ col_ptr, const double *__restrict__ F,
             const double *__restrict__ r0, const double *__restrict__ G,
             const double *__restrict__ c0, const double *__restrict__ pc, double reg,
             double *loss) {
  __shared__ double sh[256 / kWave];
  const int t = blockIdx.x;
  double cc = 0.0, fr = 0.0, rs = 0.0, gc = 0.0;
  for (int i = threadIdx.x; i < M; i += blockDim.x) {
    const long k = i + (long)t * M;
    cc += pc[k];
    fr += F[k] * r0[k];
    rs += r0[k];
  }
  for (int p = col_ptr[t] + threadIdx.x; p < col_ptr[t + 1]; p += blockDim.x) gc += G[p] * c0[p];
  cc = block_sum(cc, sh);
  fr = block_sum(fr, sh);
  rs = block_sum(rs, sh);
  gc = block_sum(gc, sh);
  if (threadIdx.x == 0) {
    loss[2 * t] = cc;
    loss[2 * t + 1] = (fr + gc) - reg * rs;
  }
}

// Update on the supports from the split partials of a product whose rows are
// the Nc columns of C and whose columns are the active targets (Z = K^T W):
// one thread per support entry p reads its own element, op(p, t, j, z). The
// product's own epilogue would visit all Nc x targets entries, 95% of them
// off the supports.
template <class CF, class W, class Op>
__global__ void __launch_bounds__(256)
k_support(const W *__restrict__ work, int gx, long ntile, int nnz,
          const int *__restrict__ tgt_of, const int *__restrict__ col_idx,
          const int *__restrict__ colpos, const int *__restrict__ done, Op op) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  const int t = tgt_of[p];
  const int n = colpos[t];
  if (n < 0 || done[t]) return;
  const int j = col_idx[p];
  op(p, t, j, work[tile_partial<CF>(j, n, 0, gx, ntile)]);
}

// The K^T W update of a cycle: as k_support, with the entry's GEMM column and
// form resolved once per cycle into wn[p] = (column << 1) | additive, or -1
// for a stopped target; op(p, additive, z).
template <class CF, class Op>
__global__ void __launch_bounds__(256)
k_support_cycle(const float *__restrict__ work, int gx, long ntile, int nnz,
                const int *__restrict__ col_idx, const int *__restrict__ wn, Op op) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= nnz) return;
  const int w = wn[p];
  if (w < 0) return;
  op(p, (w & 1) != 0, work[tile_partial<CF>(col_idx[p], w >> 1, 0, gx, ntile)]);
}

// Update over the rows from the partial tiles of a product whose rows are the
// M rows of C and whose columns are the active targets (Y = K X), for M a
// multiple of 4: one thread per 4 consecutive rows of one column, in the
// order the tiles are stored (16-byte loads of the tile), op(i, t, y) for
// rows i..i+3 of target act[n]. The product's own epilogue visits one
// element per thread.
template <class CF, class Op>
__global__ void __launch_bounds__(256)
k_rows4(const float *__restrict__ work, int M, int nact, int gx, long nq,
        const int *__restrict__ act, Op op) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= nq) return;
  constexpr int QT = CF::BM * CF::BN / 4;  // groups of 4 per tile
  const long tile = q / QT;
  const int pos = (int)(q % QT);
  const int nl = pos / (CF::BM / 4), ml = (pos % (CF::BM / 4)) * 4;
  const int m = (int)(tile % gx) * CF::BM + ml, n = (int)(tile / gx) * CF::BN + nl;
  if (m >= M || n >= nact) return;
  op(m, act[n], reinterpret_cast<const float4 *>(work)[q]);
}

// The Y = K X update of one element from the product y and the cycle's
// factor q of the element: W (the next product's operand) and, when kept, u
// (u - 1 in the additive form).
//   multiplicative: q = alpha / ra, W = alpha u = a q / y, u = a q / (alpha y)
//   additive:       q = a - r0 (FP32: the defect, small near convergence),
//                   u = a / ((a - q) + alpha y), W = alpha (u - 1)
// alpha is read by the additive form and, in the multiplicative one, only for
// the kept u.
__device__ inline void ky_update(bool add, bool keep, double ai, float qk, float alk, float y,
                                 uint16_t &w, float &luk) {
  if (add) {
    const double u = ai / ((ai - (double)qk) + (double)alk * (double)y);
    w = f2bf((float)((double)alk * (u - 1.0)));
    luk = (float)(u - 1.0);
  } else {
    const double wy = ai * (double)qk / (double)y;
    w = f2bf((float)wy);
    if (keep) luk = (float)(wy / (double)alk);
  }
}

template <typename T> T *alloc_into(std::vector<DBuf<char>> &pool, size_t n) {
  pool.emplace_back(n * sizeof(T));
  return reinterpret_cast<T *>(pool.back().get());
}

}  // namespace

void dense_kernel_build(hipStream_t st, const double *C, long ldc, int M, int Nc, double reg,
                        DenseKernel &dk) {
  dk.M = M;
  dk.Nc = Nc;
  dk.ldm = pad64(M);
  dk.ldn = pad64(Nc);
  dk.Kb.reset((size_t)dk.ldm * Nc);
  dk.KbT.reset((size_t)dk.ldn * M);
  dk.Kd.reset((size_t)dk.ldm * Nc);
  dk.KdT.reset((size_t)dk.ldn * M);
  uint16_t *Kb = dk.Kb.get(), *KbT = dk.KbT.get();
  double *Kd = dk.Kd.get(), *KdT = dk.KdT.get();
  const long ldm = dk.ldm, ldn = dk.ldn;
  launch_map(st, ldm * Nc, [=] __device__(long k) {
    const long i = k % ldm, j = k / ldm;
    const double v = (i < M) ? exp(-C[i + j * ldc] / reg) : 0.0;
    Kd[k] = v;
    Kb[k] = f2bf((float)v);
  });
  launch_map(st, ldn * M, [=] __device__(long k) {
    const long j = k % ldn, i = k / ldn;
    const double v = (j < Nc) ? exp(-C[i + j * ldc] / reg) : 0.0;
    KdT[k] = v;
    KbT[k] = f2bf((float)v);
  });
}

void dense_kernel_cost_build(hipStream_t st, const double *C, long ldc, double reg,
                             DenseKernel &dk) {
  const int M = dk.M, Nc = dk.Nc;
  const long ldn = dk.ldn;
  dk.KCT.reset((size_t)ldn * M);
  double *KCT = dk.KCT.get();
  // the same K(i,j) as KdT holds, times C(i,j)
  launch_map(st, ldn * M, [=] __device__(long k) {
    const long j = k % ldn, i = k / ldn;
    double v = 0.0;
    if (j < Nc) {
      const double c = C[i + j * ldc];
      v = exp(-c / reg) * c;
    }
    KCT[k] = v;
  });
}

void sinkhorn_batched_dense(hipStream_t st, const DenseKernel &dk, const double *a, int B,
                            const int *col_ptr, const int *col_idx, const int *tgt_of, int nnz, const double *b, double reg,
                            int maxiter, double tol, int err_mode, int absorb_every,
                            double switch_thr, double *F, double *G, int *iters, double *errs,
                            double *loss) {
  const int M = dk.M, Nc = dk.Nc;
  const long ldm = dk.ldm, ldn = dk.ldn;
  const long MB = (long)M * B;
  std::vector<DBuf<char>> pool;
  // per-target state
  float *lu = alloc_into<float>(pool, MB), *al = alloc_into<float>(pool, MB);
  float *qy = alloc_into<float>(pool, MB);  // the Y = K X update's factor (ky_update)
  double *r0 = alloc_into<double>(pool, MB);
  float *lv = alloc_into<float>(pool, nnz), *be = alloc_into<float>(pool, nnz);
  float *qz = alloc_into<float>(pool, nnz);  // the K^T W update's factor of a cycle
  int *wn = alloc_into<int>(pool, nnz);      // its map (k_support_cycle)
  float *sb = alloc_into<float>(pool, nnz);
  double *c0 = alloc_into<double>(pool, nnz);
  int *done = alloc_into<int>(pool, B), *mode = alloc_into<int>(pool, B);
  int *cnt = alloc_into<int>(pool, 2);  // targets stopped, multiplicative ones running
  int *act = alloc_into<int>(pool, B);  // targets still iterating (GEMM column map)
  int *colpos = alloc_into<int>(pool, B);  // target -> its column in act, or -1
  // GEMM operands
  uint16_t *W = alloc_into<uint16_t>(pool, (size_t)ldm * B);
  uint16_t *X = alloc_into<uint16_t>(pool, (size_t)ldn * B);
  double *Ad = alloc_into<double>(pool, (size_t)ldm * B);
  double *Bd = alloc_into<double>(pool, (size_t)ldn * B);
  float *Yc = alloc_into<float>(pool, MB);
  // deferred GEMM epilogues (the fused updates run in their own kernel, with
  // the occupancy their loads need); no k split: at 1024 targets the 128 x 128
  // tiles give one workgroup per CU already, and a split in two measured
  // slower (15.6k against 16.0k solves/s with the register-path tiles)
  SplitK<uint16_t> skb;
  skb.ks = 1;
  skb.work = alloc_into<float>(
      pool, std::max(tiles_work_elems<uint16_t>(M, B, 1), tiles_work_elems<uint16_t>(Nc, B, 1)));
  SplitK<double> skd;
  skd.work = alloc_into<double>(
      pool, std::max(tiles_work_elems<double>(M, B, 1), tiles_work_elems<double>(Nc, B, 1)));

  const uint16_t *Kb = dk.Kb.get(), *KbT = dk.KbT.get();
  const double *Kd = dk.Kd.get(), *KdT = dk.KdT.get();
  int nact = B;
  // products over the active targets, in the tiles their count takes (the
  // same arithmetic per target in every shape)
  auto bf16_product = [&](int rows, int kdim, const uint16_t *Am, long la, const uint16_t *Bm,
                          long lb, auto op) {
    with_tiles<uint16_t>(nact, [&](auto cfg) {
      gemm_tn<uint16_t, decltype(cfg)>(st, rows, nact, kdim, Am, la, Bm, lb, op, act, nullptr,
                                       skb);
    });
  };
  // Z = K^T W over the active targets, then op(p, t, j, z) on the support
  // entries only, from the product's partial tiles (no k split)
  auto ktw_support = [&](auto op) {
    auto run = [&](auto cfg) {
      using CF = decltype(cfg);
      gemm_tn<uint16_t, CF>(st, Nc, nact, M, Kb, ldm, W, ldm,
                            [] __device__(int, int, float) {}, act, nullptr, skb, false);
      const int gx = (Nc + CF::BM - 1) / CF::BM;
      const long ntile = (long)gx * ((nact + CF::BN - 1) / CF::BN);
      hipLaunchKernelGGL((k_support<CF, float, decltype(op)>), dim3((nnz + 255) / 256), dim3(256),
                         0, st, skb.work, gx, ntile, nnz, tgt_of, col_idx, colpos, done, op);
      check_launch();
    };
    with_tiles<uint16_t>(nact, run);
  };
  auto f64_product = [&](int rows, int kdim, const double *Am, long la, const double *Bm,
                         long lb, auto op) {
    with_tiles<double>(nact, [&](auto cfg) {
      gemm_tn<double, decltype(cfg)>(st, rows, nact, kdim, Am, la, Bm, lb, op, act, nullptr, skd);
    });
  };

  // ---- initial state
  launch_map(st, MB, [=] __device__(long k) {
    F[k] = 0.0;
    lu[k] = 1.0f;
  });
  launch_map(st, (long)nnz, [=] __device__(long p) {
    G[p] = 0.0;
    lv[p] = 1.0f;
  });
  launch_map(st, (long)ldm * B, [=] __device__(long k) {
    W[k] = 0;
    Ad[k] = 0.0;
  });
  launch_map(st, (long)ldn * B, [=] __device__(long k) {
    X[k] = 0;
    Bd[k] = 0.0;
  });
  launch_map(st, (long)B, [=] __device__(long t) {
    done[t] = 0;
    mode[t] = 0;
  });
  launch_map(st, 1L, [=] __device__(long) { cnt[0] = 0; });

  auto absorb = [&](int iter) {
    // fold the scalings into the potentials (targets still running)
    launch_map(st, MB, [=] __device__(long k) {
      const long t = k / M, i = k % M;
      if (!done[t]) {
        const double x = (double)lu[k];
        const double fk = F[k] + reg * (mode[t] ? log1p(x) : log(x));
        F[k] = fk;
        const double ak = exp(fk / reg);
        al[k] = (float)ak;
        Ad[i + t * ldm] = ak;
      }
    });
    launch_map(st, (long)nnz, [=] __device__(long p) {
      const long t = tgt_of[p];
      if (!done[t]) {
        const double x = (double)lv[p];
        const double gq = G[p] + reg * (mode[t] ? log1p(x) : log(x));
        G[p] = gq;
        const double bq = exp(gq / reg);
        be[p] = (float)bq;
        Bd[col_idx[p] + t * ldn] = bq;
      }
    });
    // exact marginals of the absorbed plan (FP64 MFMA)
    f64_product(M, Nc, KdT, ldn, Bd, ldn, [=] __device__(int i, int t, double v) {
      if (!done[t]) r0[i + (long)t * M] = v * Ad[i + (long)t * ldm];
    });
    // c0 on the supports only, from the product's partial tiles
    with_tiles<double>(nact, [&](auto cfg) {
      using CF = decltype(cfg);
      gemm_tn<double, CF>(st, Nc, nact, M, Kd, ldm, Ad, ldm, [] __device__(int, int, double) {},
                          act, nullptr, skd, false);
      const int gx = (Nc + CF::BM - 1) / CF::BM;
      const long ntile = (long)gx * ((nact + CF::BN - 1) / CF::BN);
      auto op = [=] __device__(int p, int t, int j, double v) { c0[p] = v * Bd[j + (long)t * ldn]; };
      hipLaunchKernelGGL((k_support<CF, double, decltype(op)>), dim3((nnz + 255) / 256), dim3(256),
                         0, st, skd.work, gx, ntile, nnz, tgt_of, col_idx, colpos, done, op);
      check_launch();
    });
    launch_map(st, 1L, [=] __device__(long) { cnt[1] = 0; });
    hipLaunchKernelGGL(k_dense_err, dim3(B), dim3(256), 0, st, M, B, a, r0, col_ptr, b, c0,
                       err_mode, tol, iter, maxiter, switch_thr, done, mode, iters, errs, cnt);
    check_launch();
  };
  // start of a cycle: u = a / r0 at v = 1, W for its first product. mult: some
  // running target iterates in the multiplicative form, whose calibration
  // ratios come from the same BF16 products at u = v = 1 (the additive form
  // needs none)
  auto start_cycle = [&](bool mult) {
    if (mult) {
      launch_map(st, MB, [=] __device__(long k) {
        const long t = k / M, i = k % M;
        W[i + t * ldm] = f2bf(al[k]);
      });
      launch_map(st, (long)nnz, [=] __device__(long p) {
        X[col_idx[p] + (long)tgt_of[p] * ldn] = f2bf(be[p]);
      });
      bf16_product(M, Nc, KbT, ldn, X, ldn,
                   [=] __device__(int i, int t, float y) { Yc[i + (long)t * M] = y; });
      ktw_support(
          [=] __device__(int p, int, int, float z) { sb[p] = (float)(c0[p] / (double)z); });
    }
    launch_map(st, MB, [=] __device__(long k) {
      const long t = k / M, i = k % M;
      if (done[t]) return;
      qy[k] = mode[t] ? (float)(a[i] - r0[k]) : (float)((double)al[k] * (double)Yc[k] / r0[k]);
      const double u = a[i] / r0[k];
      lu[k] = mode[t] ? (float)(u - 1.0) : (float)u;
      W[i + t * ldm] = f2bf(mode[t] ? (float)((double)al[k] * (u - 1.0)) : (float)((double)al[k] * u));
    });
    // the K^T W update's map and factor for the cycle: beta b / sb
    // (multiplicative, X = beta v = q / z) or the defect b - c0 (additive)
    launch_map(st, (long)nnz, [=] __device__(long p) {
      const int t = tgt_of[p];
      const int n = colpos[t];
      if (n < 0 || done[t]) {
        wn[p] = -1;
        return;
      }
      const bool add = mode[t] != 0;
      lv[p] = add ? 0.0f : 1.0f;
      wn[p] = (n << 1) | (add ? 1 : 0);
      qz[p] = add ? (float)(b[p] - c0[p]) : (float)((double)be[p] * b[p] / (double)sb[p]);
    });
  };

  // keep: the last iteration before an absorption, whose u the absorption
  // folds into f (the other iterations need only W, not u itself)
  auto iterate = [&](bool keep) {
    // Z = K^T W -> v, X (v itself only on the kept iteration)
    with_tiles<uint16_t>(nact, [&](auto cfg) {
      using CF = decltype(cfg);
      gemm_tn<uint16_t, CF>(st, Nc, nact, M, Kb, ldm, W, ldm,
                            [] __device__(int, int, float) {}, act, nullptr, skb, false);
      const int gx = (Nc + CF::BM - 1) / CF::BM;
      const long ntile = (long)gx * ((nact + CF::BN - 1) / CF::BN);
      auto op = [=] __device__(int p, bool add, float z) {
        const long x = col_idx[p] + (long)tgt_of[p] * ldn;
        const double q = (double)qz[p];
        if (add) {
          const double bp = b[p], bq = (double)be[p];
          const double v = bp / ((bp - q) + bq * (double)z);
          X[x] = f2bf((float)(bq * (v - 1.0)));
          if (keep) lv[p] = (float)(v - 1.0);
        } else {
          const double xv = q / (double)z;
          X[x] = f2bf((float)xv);
          if (keep) lv[p] = (float)(xv / (double)be[p]);
        }
      };
      hipLaunchKernelGGL((k_support_cycle<CF, decltype(op)>), dim3((nnz + 255) / 256),
                         dim3(256), 0, st, skb.work, gx, ntile, nnz, col_idx, wn, op);
      check_launch();
    });
    // Y = K X -> u, W
    if (M % 4 == 0) {
      // 4 rows per thread from the product's partial tiles (16-byte loads
      // of y, q and alpha, 8-byte stores of W)
      auto op = [=] __device__(int i, int t, float4 y4) {
        if (done[t]) return;
        const long k = i + (long)t * M;
        const bool add = mode[t] != 0;
        const float4 q4 = *reinterpret_cast<const float4 *>(qy + k);
        const float yv[4] = {y4.x, y4.y, y4.z, y4.w}, qv[4] = {q4.x, q4.y, q4.z, q4.w};
        float alv[4] = {1.0f, 1.0f, 1.0f, 1.0f};
        if (add || keep) {
          const float4 al4 = *reinterpret_cast<const float4 *>(al + k);
          alv[0] = al4.x, alv[1] = al4.y, alv[2] = al4.z, alv[3] = al4.w;
        }
        uint16_t w[4];
        float l[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) ky_update(add, keep, a[i + r], qv[r], alv[r], yv[r], w[r], l[r]);
        *reinterpret_cast<uint2 *>(W + i + (long)t * ldm) =
            make_uint2(w[0] | ((unsigned)w[1] << 16), w[2] | ((unsigned)w[3] << 16));
        if (keep) *reinterpret_cast<float4 *>(lu + k) = make_float4(l[0], l[1], l[2], l[3]);
      };
      with_tiles<uint16_t>(nact, [&](auto cfg) {
        using CF = decltype(cfg);
        gemm_tn<uint16_t, CF>(st, M, nact, Nc, KbT, ldn, X, ldn,
                              [] __device__(int, int, float) {}, act, nullptr, skb, false);
        const int gx = (M + CF::BM - 1) / CF::BM;
        const long nq = (long)gx * ((nact + CF::BN - 1) / CF::BN) * (CF::BM * CF::BN / 4);
        hipLaunchKernelGGL((k_rows4<CF, decltype(op)>), dim3((unsigned)((nq + 255) / 256)),
                           dim3(256), 0, st, skb.work, M, nact, gx, nq, act, op);
        check_launch();
      });
    } else {
      bf16_product(M, Nc, KbT, ldn, X, ldn, [=] __device__(int i, int t, float y) {
        if (done[t]) return;
        const long k = i + (long)t * M;
        const bool add = mode[t] != 0;
        uint16_t w;
        float l;
        ky_update(add, keep, a[i], qy[k], (add || keep) ? al[k] : 1.0f, y, w, l);
        W[i + (long)t * ldm] = w;
        if (keep) lu[k] = l;
      });
    }
  };

  // active list: every target at first; after each absorption the targets
  // that stopped leave the GEMMs
  std::vector<int> hact(B), hdone_v(B);
  std::vector<int> hcolpos(B);
  for (int t = 0; t < B; ++t) hact[t] = hcolpos[t] = t;
  h2d(act, hact.data(), B, st);
  h2d(colpos, hcolpos.data(), B, st);
  int iter = 0, hcnt[2] = {0, 0};
  absorb(iter);
  while (true) {
    d2h(hcnt, cnt, 2, st);
    RWIG_HIP_CHECK(hipStreamSynchronize(st));
    const int hdone = hcnt[0];
    if (hdone >= B || iter >= maxiter) break;
    if (hdone > B - nact) {
      d2h(hdone_v.data(), done, B, st);
      RWIG_HIP_CHECK(hipStreamSynchronize(st));
      nact = 0;
      for (int t = 0; t < B; ++t) {
        hcolpos[t] = -1;
        if (!hdone_v[t]) {
          hcolpos[t] = nact;
          hact[nact++] = t;
        }
      }
      h2d(act, hact.data(), nact, st);
      h2d(colpos, hcolpos.data(), B, st);
    }
    start_cycle(hcnt[1] > 0);
    const int n = std::min(absorb_every, maxiter - iter);
    for (int s = 0; s < n; ++s) iterate(s == n - 1);
    iter += n;
    absorb(iter);
  }
  if (loss) {
    // every target's final Ad, Bd, r0, c0 are those of its last absorption
    for (int t = 0; t < B; ++t) hact[t] = t;
    h2d(act, hact.data(), B, st);
    nact = B;
    double *pc = alloc_into<double>(pool, MB);
    f64_product(M, Nc, dk.KCT.get(), ldn, Bd, ldn, [=] __device__(int i, int t, double v) {
      pc[i + (long)t * M] = v * Ad[i + (long)t * ldm];
    });
    hipLaunchKernelGGL(k_dense_loss, dim3(B), dim3(256), 0, st, M, col_ptr, F, r0, G, c0, pc,
                       reg, loss);
    check_launch();
  }
  RWIG_HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace rwig

using namespace rwig;

extern "C" {

int rwig_dense_kernel_create(const double *C, long ldc, int M, int Nc, double reg, void *stream,
                             void **handle) {
  return guarded("rwig_dense_kernel_create", [&] {
    RWIG_REQUIRE(C && handle && M > 0 && Nc > 0 && reg > 0.0, "invalid argument");
    auto *dk = new DenseKernel();
    try {
      dense_kernel_build((hipStream_t)stream, C, ldc, M, Nc, reg, *dk);
      dense_kernel_cost_build((hipStream_t)stream, C, ldc, reg, *dk);
      RWIG_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    } catch (...) {
      delete dk;
      throw;
    }
    *handle = dk;
  });
}

int rwig_dense_kernel_destroy(void *handle) {
  return guarded("rwig_dense_kernel_destroy", [&] {
    RWIG_HIP_CHECK(hipDeviceSynchronize());
    delete static_cast<DenseKernel *>(handle);
  });
}

int rwig_sinkhorn_batched_mfma_dev(void *handle, const double *a, int B, const int *col_ptr,
                                   const int *col_idx, const int *tgt_of, int nnz,
                                   const double *b, double reg,
                                   int max_iter, double zero_tol, int err_mode,
                                   int absorb_every, double switch_thr, double *f, double *g,
                                   int *iters, double *errs, double *loss, void *stream) {
  return guarded("rwig_sinkhorn_batched_mfma_dev", [&] {
    RWIG_REQUIRE(handle != nullptr, "null kernel handle");
    RWIG_REQUIRE(B >= 1 && nnz >= B, "every target needs a support");
    RWIG_REQUIRE(reg > 0.0 && max_iter >= 0 && absorb_every >= 1, "invalid argument");
    RWIG_REQUIRE(err_mode == 0 || err_mode == 1, "err_mode must be 0 or 1");
    RWIG_REQUIRE(a && col_ptr && col_idx && tgt_of && b && f && g && iters && errs,
                 "null argument");
    const DenseKernel &dk = *static_cast<const DenseKernel *>(handle);
    sinkhorn_batched_dense((hipStream_t)stream, dk, a, B, col_ptr, col_idx, tgt_of, nnz,
                           b, reg, max_iter, zero_tol, err_mode, absorb_every, switch_thr, f, g,
                           iters, errs, loss);
  });
}

}  // extern "C"
