/* rwig_hip.h -- C ABI of librwig_hip.so, the MI355X (gfx950) entropic-OT
 * hot path behind rwig's sinkhorn() / barycenter() / wdl().
 *
 * Conventions (those of the reference's CUDA ABI,
 * inst/include/cuda_interface.cuh:13-39):
 *   - matrices are column-major (R layout), FP64, X(i,j) = X[i + j*rows];
 *   - plain entry points take HOST pointers; inputs are borrowed read-only
 *     for the call, outputs are caller-allocated and filled;
 *   - every entry point returns an int status (RWIG_OK on success) where the
 *     reference returned void; rwig_last_error() describes the last failure
 *     on the calling thread.
 * Entry points suffixed _dev take DEVICE pointers and a hipStream_t (passed
 * as void*); they do not synchronise the stream.
 *
 * Results follow the reference's CPU path (src/sinkhorn_impl_*.cpp and
 * src/barycenter_impl_*.cpp): solves run on the
 * support of a and b, the loop runs while (iter < max_iter && err >= zero_tol),
 * and return_status is 0 (err <= zero_tol), 1 (iter == max_iter) or 2.
 */
#ifndef RWIG_HIP_H
#define RWIG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define RWIG_OK 0
#define RWIG_ERR_INVALID 1  /* bad shape or argument */
#define RWIG_ERR_HIP 2      /* HIP runtime error */
#define RWIG_ERR_NOMEM 3    /* device allocation failed or over budget */
#define RWIG_ERR_NODEV 4    /* no usable gfx950 device */

/* Replaces cuda_available_cpp() (src/check_cuda.cpp:6-12): 1 when a HIP
 * device is usable by this library, else 0. */
int rwig_available(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char *rwig_last_error(void);

/* Vanilla Sinkhorn. Replaces cuda_sinkhorn_vanilla
 * (inst/include/cuda_interface.cuh:13-18, src/cuda_sinkhorn.cu:181-329),
 * same argument order; semantics of Sinkhorn::compute_vanilla
 * (src/sinkhorn_impl_vanilla.cpp:18-86). grad_a is read only when withgrad.
 * return_status may be null. Reduced problems up to 1024 x 1024 run their
 * whole loop in one launch over several workgroups that meet at grid-wide
 * barriers; those workgroups must be resident at once, so the library holds
 * up to 64 CUs for the loop's duration. A barrier wait is bounded (about a
 * second): past it the call returns RWIG_ERR_HIP. */
int rwig_sinkhorn_vanilla(double *P, double *grad_a, double *u, double *v,
                          double *loss, int *iter_out, double *err_out,
                          const double *a, const double *b, const double *C,
                          int m, int n, double reg, int withgrad, int max_iter,
                          double zero_tol, int *return_status);

/* Log-stabilised Sinkhorn: the GPU form of sinkhorn_log_cpp
 * (src/sinkhorn.cpp:142-170, Sinkhorn::compute_log,
 * src/sinkhorn_impl_log.cpp:225-294), which the reference runs on the CPU
 * only. f, g are -inf off the support, as in the reference. */
int rwig_sinkhorn_log(double *P, double *grad_a, double *f, double *g,
                      double *loss, int *iter_out, double *err_out,
                      const double *a, const double *b, const double *C,
                      int m, int n, double reg, int withgrad, int max_iter,
                      double zero_tol, int *return_status);

/* Parallel Wasserstein barycenter. Replaces cuda_barycenter_parallel
 * (inst/include/cuda_interface.cuh:20-26, src/cuda_barycenter.cu:268-404),
 * same argument order; semantics of Barycenter::compute_parallel
 * (src/barycenter_impl_parallel.cpp:19-53). b_ext is read only when
 * withgrad; grad_A, grad_w, loss are written only when withgrad. */
int rwig_barycenter_parallel(double *U, double *V, double *b, double *grad_A,
                             double *grad_w, double *loss, int *iter_out,
                             double *err_out, const double *A, const double *w,
                             const double *C, const double *b_ext, int M,
                             int N, int S, double reg, int withgrad,
                             int max_iter, double zero_tol,
                             int *return_status);

/* Log-stabilised barycenter: the GPU form of barycenter_log_cpp
 * (src/barycenter.cpp:168-207, Barycenter::compute_log,
 * src/barycenter_impl_log.cpp:23-48), CPU-only in the reference. F, G take
 * the place of U, V. The stopping error is the spectral norm of
 * -KV/reg - log A, as in src/barycenter_impl_log.cpp:252. */
int rwig_barycenter_log(double *F, double *G, double *b, double *grad_A,
                        double *grad_w, double *loss, int *iter_out,
                        double *err_out, const double *A, const double *w,
                        const double *C, const double *b_ext, int M, int N,
                        int S, double reg, int withgrad, int max_iter,
                        double zero_tol, int *return_status);

/* rwig_barycenter_log with a choice of arithmetic and a checkpointed adjoint.
 * Same iteration, outputs and stopping rule as rwig_barycenter_log
 * (src/barycenter_impl_log.cpp:189-382); the forward keeps (F, G, row
 * soft-min) every checkpoint_every iterations and the adjoint recomputes each
 * segment from its checkpoint, so the adjoint's memory is
 * max_iter/checkpoint_every + checkpoint_every + 2 slices of (M + N) x S
 * instead of max_iter + 1.
 *   precision 0: FP64 throughout (equal to rwig_barycenter_log up to
 *                rounding); when (max C - min C) / reg <= 600 the soft-min
 *                passes and the adjoint's weighted sums run as FP64 MFMA
 *                products with the scaled kernel exp(-(C - min C) / reg);
 *   precision 1: FP32 cost, iterates and soft-min passes, FP64 gradient
 *                accumulation, loss and adjoint seed;
 *   precision 2: FP64 cost, iterates and exponent arguments, the passes'
 *                exponentials in FP32 (the FP32 rounding of C alone moves
 *                every exponent by up to 2^-24 max|C| / reg).
 *   seed_mode 0: the reference's adjoint seed 2 (b - b_ext) .* b, b
 *                normalised (src/barycenter_impl_log.cpp:323-325);
 *   seed_mode 1: exact gradient of ||b_L - b_ext||^2, b_L = exp(logb_L)
 *                unnormalised (loss reported accordingly);
 *   seed_mode 2: exact gradient of ||b - b_ext||^2 through the normalisation.
 * Inputs and outputs are host FP64 arrays, column-major. */
int rwig_barycenter_log_ck(double *F, double *G, double *b, double *grad_A,
                           double *grad_w, double *loss, int *iter_out,
                           double *err_out, const double *A, const double *w,
                           const double *C, const double *b_ext, int M, int N,
                           int S, double reg, int withgrad, int max_iter,
                           double zero_tol, int checkpoint_every,
                           int seed_mode, int precision, int *return_status);

/* rwig_barycenter_log_ck on DEVICE buffers, enqueued on `stream`: C (M x N,
 * column-major, leading dimension M) in the working precision's type (FP32
 * for precision 1, FP64 for precisions 0 and 2), A (M x S), w (S) and b_ext
 * (N, with withgrad) in FP64; F (M x S), G (N x S), b (N), grad_A (M x S),
 * grad_w (S) and loss (1) are FP64 device outputs, each may be null (not
 * returned). iter_out, err_out and return_status are host values; the call
 * synchronises the stream before returning. For inputs already resident in
 * HBM (no counterpart in the reference, whose entry points take R vectors). */
int rwig_barycenter_log_ck_dev(double *F, double *G, double *b, double *grad_A,
                               double *grad_w, double *loss, int *iter_out,
                               double *err_out, const double *A, const double *w,
                               const void *C, const double *b_ext, int M, int N,
                               int S, double reg, int withgrad, int max_iter,
                               double zero_tol, int checkpoint_every,
                               int seed_mode, int precision, int *return_status,
                               void *stream);

/* D log barycenters without gradient that share A (M x S) and C (M x N),
 * document d weighted by W(:, d) (S x D): B (N x D) receives each document's
 * normalised barycenter as barycenter_log_cpp computes it
 * (src/barycenter_impl_log.cpp:23-48, 189-272; spectral-norm stop per
 * document), iters and errs (D each, may be null) their iteration counts and
 * errors. The WDL inference of sinkhorn_mode 2 (src/wdl_impl.cpp:354-377),
 * which the reference runs as one barycenter per document: here C stays on
 * the device and the soft-min passes of 32 / S documents share one sweep of
 * it. FP64, host buffers. */
int rwig_barycenter_log_batched(double *B, int *iters, double *errs,
                                const double *A, const double *W,
                                const double *C, int M, int N, int S, int D,
                                double reg, int max_iter, double zero_tol);

/* rwig_sinkhorn_log on DEVICE pointers of an already reduced problem (every
 * a_i and b_j > 0): the per-call plain FP64 form with inputs resident in HBM.
 * P (m x n), f (m), g (n), loss (1 scalar) and grad_a (m, withgrad) are
 * device outputs; the call synchronises the stream before returning. */
int rwig_sinkhorn_log_dev(const double *a, const double *b, const double *C,
                          int m, int n, double reg, int withgrad, int max_iter,
                          double zero_tol, double *P, double *grad_a, double *f,
                          double *g, double *loss, int *iter_out,
                          double *err_out, void *stream);

/* Wasserstein Dictionary Learning. Replaces cuda_wdl
 * (inst/include/cuda_interface.cuh:31-39, src/cuda_wdl.cu:539-818), same
 * argument order: Y (N x M documents), C (N x N cost); A (N x S topics),
 * W (S x M weights) and Yhat (N x M barycenters) are filled. Alpha, Lambda
 * start from std::mt19937(seed) standard normals (Alpha first), as in
 * src/cuda_wdl.cu:592-593, :677-689. The training step is
 * WassersteinDictionaryLearning::_train_batch_batched
 * (src/wdl_impl.cpp:87-247) with the reference's batch-size and g_Lambda
 * broadcast rules; Yhat comes from the parallel barycenter, batched over
 * documents, each stopping at its own iteration. FP64. */
int rwig_wdl(double *A, double *W, double *Yhat, const double *Y,
             const double *C, int N, int M, int S, double reg, int max_iter,
             double zero_tol, int B, int E, int opt, double eta, double gamma,
             double beta1, double beta2, double eps, int seed, int verbose);

/* rwig_wdl from explicit initial Alpha0 (N x S) and Lambda0 (S x M) -- the
 * values the reference's CPU path draws with R's RNG
 * (src/wdl_impl.cpp:277-278). precision 0: FP64, 1: FP32 Gibbs products and
 * recurrences (optimizer, softmax in FP64). checkpoint_every: 0 keeps the
 * iterates of all max_iter steps for the adjoint, as the reference does
 * (src/wdl_impl.cpp:111-121); k > 0 keeps VB every k steps and recomputes
 * each segment in the adjoint -- the same results bit for bit, with
 * k + max_iter/k + 1 slices of history instead of max_iter + 1 (each slice
 * 4 N x S x B values), for max_iter more forward steps per batch. */
int rwig_wdl_init(double *A, double *W, double *Yhat, const double *Y,
                  const double *C, int N, int M, int S, double reg,
                  int max_iter, double zero_tol, int B, int E, int opt,
                  double eta, double gamma, double beta1, double beta2,
                  double eps, const double *Alpha0, const double *Lambda0,
                  int precision, int checkpoint_every);

/* One batched WDL gradient on device pointers (src/wdl_impl.cpp:87-247):
 * K = exp(-C/reg) and KT its transpose, column-major, leading dimension ldk
 * (>= N, multiple of 8); A (N x S), Yb (N x D), Wb (S x D) in the precision
 * given (0 double, 1 float); outputs FP64: gAlpha (N x S), glambda (S) and
 * loss = sum ||b_L - y||^2 over the batch (may be null); checkpoint_every as
 * in rwig_wdl_init. */
int rwig_wdl_batch_grad_dev(const void *K, const void *KT, long ldk, int N,
                            int S, int D, int L, const void *A,
                            const void *Yb, const void *Wb, double *gAlpha,
                            double *glambda, double *loss, int precision,
                            int checkpoint_every, void *stream);

/* One WDL training step on device buffers, the body of the reference's batch
 * loop (src/wdl_impl.cpp:316-337, WDL::_train_batch_batched plus the
 * optimizer step and softmax): A = softmax(Alpha), W = softmax(Lambda) by
 * columns; the batched gradient of documents [first, first + D) of Y (as
 * rwig_wdl_batch_grad_dev); g_Lambda = g_lambda 1^T over all M documents;
 * optimizer step `step` (>= 1; optimizer 0 SGD, 1 Adam, 2 AdamW) on Alpha
 * (N x S) and Lambda (S x M) with their moments mA, vA (N x S) and mL, vL
 * (S x M), updated in place (zero before step 1). K, KT (leading dimension
 * ldk) and Y (N x M) in the precision given (0 double, 1 float); Alpha,
 * Lambda, the moments and loss (may be null) FP64. The sequence of steps over
 * the batches of E epochs gives rwig_wdl_init's A and W bit for bit. */
int rwig_wdl_step_dev(const void *K, const void *KT, long ldk, const void *Y,
                      int N, int M, int S, int L, int first, int D,
                      double *Alpha, double *Lambda, double *mA, double *vA,
                      double *mL, double *vL, int step, int optimizer,
                      double eta, double gamma, double beta1, double beta2,
                      double eps, double *loss, int precision,
                      int checkpoint_every, void *stream);

/* ---------------------------------------------------------------------------
 * Batched solves, device pointers (no counterpart in the reference, whose GPU
 * path solves one problem per call, inst/include/cuda_interface.cuh:13-39; the
 * reference solves a batch as repeated sinkhorn_log_cpp calls,
 * src/sinkhorn.cpp:142).
 *
 * One source measure a (M points, all a_i > 0) against B target measures.
 * C is the FP64 cost, column-major, leading dimension ldc; target t is
 * supported on the columns col_idx[col_ptr[t] .. col_ptr[t+1]) of C with
 * masses b[] at the same positions (all > 0); max_nt bounds the support size.
 * Outputs: f (B x M, target-major), g (at the positions of b), iters (B),
 * errs (B) and, if loss is non-null, loss[2t] = sum C.*P and
 * loss[2t+1] = sum C.*P + reg sum P.*(log P - 1).
 *
 * precision 0: FP64 log-domain iterations of Sinkhorn::_fwd_log
 *   (src/sinkhorn_impl_log.cpp:298-376), one kernel launch per batch;
 * precision 1 / 2: FP64 absorption of the scalings into (f, g) every
 *   absorb_every iterations plus defect-corrected iterations on a shared
 *   low-precision Gibbs kernel Klo (FP32 / BF16, see rwig_gibbs_lowp_dev);
 *   the stopping test runs at the absorptions on the exact FP64 plan.
 * err_mode 0: the reference's error ||log-marginal residuals||_2 (summed over
 *   rows and columns, src/sinkhorn_impl_log.cpp:360); 1: the L1 marginal error
 *   sum|P1 - a| + sum|P^T 1 - b|.
 */
int rwig_sinkhorn_batched_dev(const double *C, long ldc, int M, const double *a,
                              int B, const int *col_ptr, const int *col_idx,
                              const double *b, int max_nt, double reg,
                              int max_iter, double zero_tol, int err_mode,
                              int precision, const void *Klo, long ldk,
                              int absorb_every, double *f, double *g,
                              int *iters, double *errs, double *loss,
                              void *stream);

/* Klo = exp(-C/reg) rounded to FP32 (precision 1) or BF16 (precision 2),
 * column-major M x N with leading dimension ldk (>= M, multiple of 4; rows
 * M .. ldk-1 are zero). Built once per cost matrix and reg. */
int rwig_gibbs_lowp_dev(const double *C, long ldc, int M, int N, double reg,
                        int precision, void *Klo, long ldk, void *stream);

/* Batched Sinkhorn on MFMA (one source against B targets sharing one cost
 * matrix): every Gibbs product of an iteration is one BF16 MFMA GEMM over all
 * targets (FP32 accumulation), with the scaling update fused into its
 * epilogue; FP64 absorption every absorb_every iterations computes the exact
 * marginals by FP64 MFMA GEMMs and decides convergence (L1 or reference
 * error <= zero_tol). Targets whose absorbed error is above switch_thr
 * iterate in a multiplicative form (positive sums only), the others in an
 * additive defect-correction form whose fixed point is the FP64 one.
 * The kernel handle holds exp(-C/reg) in BF16 and FP64 in both orientations,
 * and exp(-C/reg) .* C in FP64 for the transport cost (rwig_dense_kernel_create,
 * one per cost matrix and reg); the solve reads the cost only through it.
 * tgt_of[p] is the target owning support entry p. Other arguments as
 * rwig_sinkhorn_batched_dev. A target whose error turns non-finite stops
 * there and reports that error (the reference's return_status 2). No
 * counterpart in the reference (it solves one problem per call). */
int rwig_dense_kernel_create(const double *C, long ldc, int M, int Nc,
                             double reg, void *stream, void **handle);
int rwig_dense_kernel_destroy(void *handle);
int rwig_sinkhorn_batched_mfma_dev(void *handle, const double *a, int B,
                                   const int *col_ptr,
                                   const int *col_idx, const int *tgt_of,
                                   int nnz, const double *b, double reg,
                                   int max_iter, double zero_tol, int err_mode,
                                   int absorb_every, double switch_thr,
                                   double *f, double *g, int *iters,
                                   double *errs, double *loss, void *stream);

/* MFMA GEMM D(m, n) = sum_k A(m, k) B(k, n) with A row-major (A[m*lda + k])
 * and B, D column-major (B[k + n*ldb], D[m + n*ldd]); lda and ldb multiples
 * of 16 bytes. The building block of the batched Gibbs products (no
 * counterpart in the reference, which calls cuBLAS dgemm,
 * src/cuda_kernels.cu:563). bf16: A, B are BF16 bit patterns, D is FP32
 * (v_mfma_f32_16x16x32_bf16); f64: all FP64 (v_mfma_f64_16x16x4_f64). */
/* ks >= 1 splits k over ks workgroups per tile; with a workspace (work,
 * rwig_gemm_workspace_bytes(dtype, M, N, ks) bytes, dtype 0 bf16, 1 f64)
 * the partial tiles are summed in split order by a second kernel, which also
 * writes D; without one (work null, ks 1) the product kernel writes D. With
 * a workspace and D null only the product runs: the partial tiles stay in
 * work, as the batched solvers run their products before their own updates.
 * The tile shape follows N (smaller tiles for fewer columns); for a given ks
 * every shape gives the same result bit for bit. */
int rwig_gemm_workspace_bytes(int dtype, int M, int N, int ks, long *bytes);
int rwig_gemm_tn_bf16_dev(int M, int N, int K, const void *A, long lda,
                          const void *B, long ldb, float *D, long ldd, int ks,
                          void *work, void *stream);
int rwig_gemm_tn_f64_dev(int M, int N, int K, const double *A, long lda,
                         const double *B, long ldb, double *D, long ldd,
                         int ks, void *work, void *stream);

/* Cost and histogram construction (src/utils.cpp:17-68).
 * rwig_euclidean: the V x V matrix of Euclidean distances between the rows
 * of E (V x d, column-major), as euclidean_cpp (src/utils.cpp:17-37);
 * squared = 1 gives the squared distances (the cost the WDL configuration
 * uses). Exactly symmetric with a zero diagonal. Host pointers;
 * rwig_euclidean_dev takes device pointers and enqueues on the stream.
 * rwig_doc2dist: documents as normalised bags of words over a dictionary, as
 * doc2dist_cpp (src/utils.cpp:39-68): out (V x D, column-major, host) has in
 * column c the counts of document c's tokens divided by their sum; a token
 * maps to the first equal dictionary entry, or to the last entry when it is
 * not in the dictionary (src/utils.cpp:53-57). Document c's tokens are
 * tokens[offsets[c] .. offsets[c+1]), NUL-terminated strings. */
int rwig_euclidean(const double *E, int V, int d, double *out, int squared);
int rwig_euclidean_dev(const double *E, int V, int d, long lde, double *out,
                       long ldo, int squared, void *stream);
int rwig_doc2dist(const char *const *tokens, const long *offsets, int D,
                  const char *const *dict, int V, double *out);

/* Truncated SVD, replacing tsvd_cpp(M, k, flip_sign) (src/tsvd.cpp:7-86):
 * out (m x kk, column-major, host) = U_k diag(s_k) for the kk = min(k, m, n)
 * leading singular triplets of M (m x n, column-major, host), with the sign
 * of each column set by flip_sign: 0 "auto" (the statistic of Bro, Acar &
 * Kolda (2008) as src/tsvd.cpp:39-71 computes it), 1 "sklearn" (the entry of
 * largest magnitude positive), 2 "none". k_out (optional) receives kk;
 * residual (optional) the relative Ritz residual at which the subspace
 * iteration stopped. k < 1 is an error (the reference's column range would
 * be out of bounds). */
int rwig_tsvd(double *out, int *k_out, double *residual, const double *M,
              int m, int n, int k, int flip_sign);
/* WIG document scores (R/wig.R:96-106) from the outputs of wdl: A (N x S
 * topics), W (S x M weights), Yhat (N x M reconstructions), host,
 * column-major. svd_method 0 ("topics"): scores = c(t(tsvd(t(A), 1)) %*% W);
 * 1 ("docs"): scores = c(tsvd(t(Yhat), 1)); both with the "auto" sign rule.
 * The unused matrix may be NULL. scores: M values. */
int rwig_wig_scores(double *scores, const double *A, const double *W,
                    const double *Yhat, int N, int S, int M, int svd_method);
/* WIG index (R/wig.R:2-6, :139-159): the M document scores summed per
 * calendar period of their dates (days since 1970-01-01, R's Date), periods
 * in ascending order; with standardize, (x - mean(x)) / sd(x) + 100 with R's
 * mean and sd (NaN for a single period, where R gives NA). group_unit (the
 * unit of lubridate::floor_date): 0 day, 1 week (from Sunday), 2 month,
 * 3 bimonth, 4 quarter, 5 halfyear, 6 year, 7 season (from December, March,
 * June, September). period_days and index must hold M values; n_periods
 * receives their count. Host only: no device work. */
int rwig_wig_index(int *period_days, double *index, int *n_periods,
                   const double *scores, const int *days, int M,
                   int group_unit, int standardize);

/* One fused soft-min pass on DEVICE pointers, the log-sum-exp reduction of
 * the log-domain solvers (src/sinkhorn_impl_log.cpp:48-50 and its use in
 * src/barycenter_impl_log.cpp:226-253), for S slices at once:
 *   rows: out(i, s) = softmin_j [ (C(i,j) - F(i,s)) - G(j,s) ]      (M x S)
 *   cols: out(j, s) = softmin_i [ (C(i,j) - G(j,s)) - F(i,s) ]      (N x S)
 * with softmin_x r = min r - reg log sum exp(-(r - min r) / reg). The column
 * pass reads CT = C^T (N x M, leading dimension ldct). precision 0: FP64,
 * 1: FP32 (hardware exp). work: rwig_softmin_workspace_bytes(which = 0 rows,
 * 1 cols, ...) bytes. Enqueued on the stream; no synchronisation. */
int rwig_softmin_workspace_bytes(int which, int M, int N, int S, int precision,
                                 long *bytes);
int rwig_softmin_rows_dev(const void *C, int M, int N, long ldc, const void *F,
                          long ldf, const void *G, long ldg, int S, double reg,
                          void *out, int precision, void *work, void *stream);
/* rwig_softmin_rows_dev without the workspace argument: it allocates
 * nothing. No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * The chunk partials live in a static 32 MiB device array, one per device,
 * held until the process ends; a pass whose partials
 * (rwig_softmin_workspace_bytes with which = 0) exceed 32 MiB is refused with
 * RWIG_ERR_INVALID. Every call on a device shares that array: calls that
 * overlap on one device (other streams, other threads) race. */
int rwig_softmin_rows_noalloc_dev(const void *C, int M, int N, long ldc,
                                  const void *F, long ldf, const void *G,
                                  long ldg, int S, double reg, void *out,
                                  int precision, void *stream);
int rwig_softmin_cols_dev(const void *CT, int M, int N, long ldct,
                          const void *F, long ldf, const void *G, long ldg,
                          int S, double reg, void *out, int precision,
                          void *work, void *stream);

/* ---------------------------------------------------------------- serving
 * A serving session: continuous batching of Sinkhorn and barycenter solves
 * on cost matrices registered once and kept in HBM. No counterpart in the
 * reference (one solve per call, every buffer allocated per call,
 * src/cuda_sinkhorn.cu:217-328); the algorithms are the log-domain Sinkhorn
 * (src/sinkhorn_impl_log.cpp:18-118) and the parallel barycenter
 * (src/barycenter_impl_parallel.cpp:60-151), run as BF16 MFMA Gibbs products
 * with FP64 absorption every absorb_every iterations (switch_thr as in
 * rwig_sinkhorn_batched_mfma_dev).
 *
 * One engine thread per session owns every GPU launch. Requests on the same
 * (cost, reg) share one GEMM per Gibbs product (one column per Sinkhorn
 * problem, S per barycenter); different (cost, reg) pairs run on their own
 * streams. Requests are admitted in arrival order while the session's HBM
 * budget (cost matrices, cached Gibbs kernels, group columns) holds them.
 * Stopping: L1 marginal error of the absorbed plans <= zero_tol, tested at
 * each absorption, or max_iter reached (iteration counts are multiples of
 * absorb_every, or max_iter). A request's result does not depend on the
 * other requests in flight. Entries of a, b and A must be finite and
 * non-negative, each measure with positive mass; zero entries are off the
 * support, as the reference's support reduction
 * (src/sinkhorn_impl_vanilla.cpp:29-35): their potentials are -inf and their
 * plan entries zero. Submission refuses (RWIG_ERR_INVALID) a reg so small
 * that a used row or column of the cost has min C / reg > 60, where the
 * shared low-precision Gibbs kernel would underflow; the log-domain entry
 * points solve those. Calls are thread-safe. */
#define RWIG_SERVE_EAGER 1 /* flags: launch every kernel of a cycle directly
                             instead of replaying recorded HIP graphs */
int rwig_serve_create(double hbm_budget_bytes, int max_group_columns,
                      int absorb_every, double switch_thr, int flags,
                      void **session);
int rwig_serve_destroy(void *session);
/* HBM bytes the session charges for a (cost, reg) pair of an M x N cost: its
 * cached Gibbs kernel, and its group's buffers at `columns` columns (groups
 * start at 32 columns and double as requests arrive). */
int rwig_serve_footprint(int M, int N, int columns, double *kernel_bytes,
                         double *group_bytes);
/* C: host, column-major M x N, copied to the device. */
int rwig_serve_register_cost(void *session, const double *C, int M, int N,
                             int *cost_id);
int rwig_serve_submit_sinkhorn(void *session, int cost_id, double reg,
                               const double *a, int M, const double *b, int N,
                               int max_iter, double zero_tol, long *request_id);
/* A: host, column-major M x S; w: S barycentric weights. */
int rwig_serve_submit_barycenter(void *session, int cost_id, double reg,
                                 const double *A, int M, int S, const double *w,
                                 int max_iter, double zero_tol,
                                 long *request_id);
/* Waits until the request has finished (timeout_s < 0: no limit); state:
 * 0 queued, 1 running, 2 done, 3 failed (engine failure, or a marginal
 * error that turned non-finite). */
int rwig_serve_wait(void *session, long request_id, double timeout_s,
                    int *state);
/* Results of a finished request, which is then forgotten. Sinkhorn: x = f
 * (M), y = g (N) potentials, loss = sum C.*P + reg sum P.*(log P - 1).
 * Barycenter: x = b (N, normalised); y (may be null, M*S + N*S values) = the
 * potentials F (M x S) then G (N x S) of the absorbed plans
 * P_s = exp((F_s 1^T + 1 G_s^T - C) / reg), as the reference's log barycenter
 * returns F and G. latency_ms: submit to completion. status (may be null): the reference's return_status
 * (src/sinkhorn.cpp:88-95): 0 err <= zero_tol, 1 max_iter reached, 2 the
 * solve broke down (non-finite error; the call still returns RWIG_OK and the
 * outputs as they stood). A request failed by the engine itself (HIP error,
 * session destroyed) returns RWIG_ERR_HIP. */
int rwig_serve_result(void *session, long request_id, double *x, double *y,
                      double *loss, int *iter, double *err, double *latency_ms,
                      int *status);
/* out[0..14]: completed, failed, queued, running, p50 latency (ms),
 * p99 latency (ms), mean latency (ms), solves/s (first submit to last
 * completion), HBM bytes in use, peak HBM bytes, budget, engine cycles,
 * groups, mean queueing delay (ms), max latency (ms). */
int rwig_serve_stats(void *session, double *out, int n_out);
int rwig_serve_reset_stats(void *session);

#ifdef __cplusplus
}
#endif

#endif /* RWIG_HIP_H */
