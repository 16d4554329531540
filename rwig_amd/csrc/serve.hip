// Serving engine: continuous batching of Sinkhorn and barycenter solves on
// cost matrices kept resident in HBM.
//
// Requests arrive from any thread (submit) and are solved by one engine
// thread that owns every GPU launch. Requests sharing a (cost, reg) pair run
// in one group: each request owns columns of the group's GEMM operands (one
// for a Sinkhorn problem, S for a barycenter of S measures) and every Gibbs
// product of an iteration is one BF16 MFMA GEMM over the group's running
// columns, with the same defect-corrected update and FP64 absorption as the
// batched solver (multiplicative form far from convergence, additive form
// near it; potentials, marginals and errors in FP64). Groups run on their own
// streams, so the work of different cost matrices overlaps on the device.
//
// The engine works in cycles. At the start of a cycle, queued requests are
// admitted while the HBM budget holds them: a new group costs its Gibbs
// kernel (cached per (cost, reg), kept while the budget allows) and its column
// buffers, a request its columns. Every group then absorbs (its requests'
// stopping test), finished requests are harvested, and the remaining columns
// run absorb_every iterations. A request's arithmetic depends only on its own
// columns (the GEMMs read and write columns independently and every request
// starts its absorption cycle at admission), so its result is the same
// alone and in any mix.
//
// Zero entries of a, b (Sinkhorn) or A (barycenter) are handled as the
// reference's support reduction (src/sinkhorn_impl_vanilla.cpp:29-35):
// those rows and columns keep a zero scaling, their potentials are -inf and
// their plan entries zero, so they drop out of every product. A request
// whose Gibbs kernel would underflow (a row or column of the cost with
// min C / reg above kMaxCostOverReg) is refused at submission, and one whose
// marginal error turns non-finite ends as failed with status 2.
//
// Algorithms: log-domain Sinkhorn (src/sinkhorn_impl_log.cpp:18-118) and the
// parallel barycenter (src/barycenter_impl_parallel.cpp:60-151), written in
// the absorbed form: potentials f, g in FP64 and scalings u, v relative to
// them. For a barycenter the update of the common target is
// b = prod_s (K^T U_s)^{w_s} with log(K^T U_s) = log q_s - g_s / reg, q_s the
// column marginal of diag(u_s) P0_s, and v_s = b / q_s.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "dense.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace rwig {

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kSinkhorn = 0, kBarycenter = 1;
enum ReqState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };
constexpr int kColPage = 32;  // column capacity granularity of a group
// Largest min_j C_ij / reg (over a row, and min_i over a column) a request may
// have: exp(-60) = 8.8e-27 is far inside the BF16/FP32 range of the shared
// Gibbs kernel (smallest normal 1.2e-38 = exp(-87.3)), with room for the
// scalings exp(f / reg) that absorb the shift.
constexpr double kMaxCostOverReg = 60.0;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// f(idx, row) for idx < cnt[which] (a count read on the device) and row <
// rows: the grid covers an upper bound of the count, so the launch can be
// recorded once in a graph and replayed over changing column lists.
template <class F>
__global__ void __launch_bounds__(kTB) k_cols(long rows, const int *cnt, int which, F f) {
  const long n = rows * (long)cnt[which];
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n;
       k += (long)gridDim.x * blockDim.x)
    f((int)(k / rows), k % rows);
}

template <class F>
void launch_cols(hipStream_t st, long rows, int bound, const int *cnt, int which, F f) {
  hipLaunchKernelGGL(k_cols<F>, dim3(grid_for(rows * bound, kTB, 8192)), dim3(kTB), 0, st, rows,
                     cnt, which, f);
  check_launch();
}

// indices into the device list header
constexpr int kNact = 0, kNreq = 1, kNbar = 2;
// per-list length after the 4-int header: act, areq, riter, abar, lim
constexpr int kLists = 5;

// Stopping test of each running request at an absorption, one block per
// request: L1 marginal error of the absorbed plans, sum over the request's
// columns; for a barycenter the common target is re-estimated from the exact
// column marginals, b = prod_s (c0_s e^{-g_s/reg})^{w_s}, and stored in the
// request's first target column.
//
// A request's rows and columns are split over gridDim.y blocks: each writes
// its partial sum to part[base * gridDim.y + y], and the block that arrives
// last (a counter per request, reset by that block) adds the partials in
// block order, so the error does not depend on which block finishes last.
__global__ void __launch_bounds__(256)
k_serve_err(int M, int N, double reg, const int *__restrict__ areq,
            const int *__restrict__ riter, const int *__restrict__ kind,
            const int *__restrict__ ns, const double *__restrict__ wcol,
            const int *__restrict__ maxit, const double *__restrict__ tol,
            const double *__restrict__ acol, const double *__restrict__ r0, double *bcol,
            const double *__restrict__ c0, const double *__restrict__ G, double switch_thr,
            int *done, int *mode, double *errs, const int *cnt, double *part, int *arrive) {
  __shared__ double sh[256 / kWave];
  __shared__ int last;
  if ((int)blockIdx.x >= cnt[kNreq]) return;
  const int base = areq[blockIdx.x];
  const int S = ns[base];
  const int nb = gridDim.y, y = blockIdx.y;
  const int cm = (M + nb - 1) / nb, cn = (N + nb - 1) / nb;
  const int i0 = y * cm, i1 = min(M, i0 + cm), j0 = y * cn, j1 = min(N, j0 + cn);
  double e = 0.0;
  for (int s = 0; s < S; ++s) {
    const long o = (long)(base + s) * M;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) e += fabs(r0[o + i] - acol[o + i]);
  }
  if (kind[base] == kSinkhorn) {
    const long o = (long)base * N;
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) e += fabs(c0[o + j] - bcol[o + j]);
  } else {
    for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
      double lb = 0.0;
      for (int s = 0; s < S; ++s) {
        const long q = j + (long)(base + s) * N;
        lb += wcol[base + s] * (log(c0[q]) - G[q] / reg);
      }
      const double bj = exp(lb);
      for (int s = 0; s < S; ++s) e += fabs(c0[j + (long)(base + s) * N] - bj);
      bcol[j + (long)base * N] = bj;
    }
  }
  e = block_sum(e, sh);
  if (threadIdx.x == 0) {
    part[(long)base * nb + y] = e;
    __threadfence();
    last = atomicAdd(&arrive[base], 1) == nb - 1;
  }
  __syncthreads();
  if (!last) return;
  if (threadIdx.x == 0) {
    __threadfence();
    const volatile double *pv = part + (long)base * nb;
    e = 0.0;
    for (int q = 0; q < nb; ++q) e += pv[q];
    arrive[base] = 0;
    const int it = riter[blockIdx.x];
    errs[base] = e;
    const int stop = (it < maxit[base] && e >= tol[base]) ? 0 : 1;
    const int md = (e < switch_thr) ? 1 : 0;
    for (int s = 0; s < S; ++s) {
      done[base + s] = stop;
      mode[base + s] = md;
    }
  }
}

struct Cost {
  int M = 0, N = 0;
  DBuf<double> C;
  std::vector<double> rowmin, colmin;  // min_j C_ij and min_i C_ij (host)
};

struct Gibbs {
  int cost = -1;
  double reg = 0.0;
  DenseKernel dk;
  size_t bytes = 0;
  int users = 0;      // groups using it
  uint64_t last = 0;  // engine cycle of the last use
};

struct Group;

struct Request {
  long id = 0;
  int kind = kSinkhorn, cost = 0, S = 1, max_iter = 0;
  double reg = 0.0, tol = 0.0;
  std::vector<double> in1, in2;  // Sinkhorn: a (M), b (N); barycenter: A (M x S), w (S)
  std::vector<double> out1, out2;  // Sinkhorn: f (M), g (N); barycenter: b (N)
  std::vector<double> r0, c0;      // Sinkhorn: marginals of the final plan (loss)
  double loss = 0.0, err = 0.0;
  int iter = 0;
  int state = kQueued;
  int status = 2;           // 0 err <= tol, 1 max_iter reached, 2 otherwise
  bool numeric = false;     // failed by a non-finite error (results kept)
  std::string error;
  Clock::time_point t_submit, t_admit, t_done;
  Group *group = nullptr;
  int col = -1;
};

// Column-major device array with a fixed number of bytes per column.
struct Slab {
  DBuf<char> buf;
  size_t per_col = 0;
  template <class T> T *p() const { return reinterpret_cast<T *>(buf.get()); }
};

// Requests of one (cost, reg) pair and their columns.
struct Group {
  int cost = -1;
  double reg = 0.0;
  Gibbs *gk = nullptr;
  int M = 0, N = 0;
  long ldm = 0, ldn = 0;
  int cap = 0;
  std::unique_ptr<Stream> stream;  // created with the group
  hipEvent_t ev = nullptr;
  bool graphs = true;  // replay each phase of a cycle as a HIP graph
  int absorb_every = 20;
  double switch_thr = 1e-2;
  hipGraphExec_t gx[2] = {nullptr, nullptr};  // absorption, iterations
  uint64_t last = 0;
  std::vector<Request *> owner;  // column -> request occupying it
  std::vector<Request *> running;
  // per-column device state
  Slab acol, F, r0, lu, al, ra, Yc, W, Ad;   // source side (M rows)
  Slab bcol, G, c0, lv, be, sb, Zq, X, Bd;   // target side (N rows)
  Slab kind, base, ns, done, mode, maxit;    // ints per column
  Slab wcol, tol, errs;                      // doubles per column
  // device lists: header (nact, nreq, nbar, -) | act | areq | riter | abar |
  // lim (per column: iterations to run in this cycle's iteration phase)
  DBuf<int> lists;
  // stopping test spread over err_blocks() blocks per request: partial sums
  // and per-request arrival counters (zero between launches)
  DBuf<double> errpart;
  DBuf<int> arrive;
  int *hl[2] = {nullptr, nullptr};  // pinned host copies, one per cycle phase
  int *hdone_pin = nullptr;
  double *herr_pin = nullptr;
  // split-k workspaces of the skinny products (sized for cap columns)
  DBuf<float> wk_b;
  DBuf<double> wk_d;
  int nact = 0, nreq = 0, nbar = 0;

  std::vector<Slab *> slabs() {
    return {&acol, &F, &r0, &lu, &al, &ra, &Yc, &W, &Ad, &bcol, &G, &c0, &lv, &be, &sb,
            &Zq, &X, &Bd, &kind, &base, &ns, &done, &mode, &maxit, &wcol, &tol, &errs};
  }

  void layout() {
    acol.per_col = F.per_col = r0.per_col = 8 * (size_t)M;
    lu.per_col = al.per_col = ra.per_col = Yc.per_col = 4 * (size_t)M;
    W.per_col = 2 * (size_t)ldm;
    Ad.per_col = 8 * (size_t)ldm;
    bcol.per_col = G.per_col = c0.per_col = 8 * (size_t)N;
    lv.per_col = be.per_col = sb.per_col = Zq.per_col = 4 * (size_t)N;
    X.per_col = 2 * (size_t)ldn;
    Bd.per_col = 8 * (size_t)ldn;
    kind.per_col = base.per_col = ns.per_col = done.per_col = mode.per_col = maxit.per_col = 4;
    wcol.per_col = tol.per_col = errs.per_col = 8;
  }

  size_t per_col_bytes() {
    size_t b = 0;
    for (Slab *s : slabs()) b += s->per_col;
    return b + 16;  // device lists
  }

  // k splits of the four products: Z = K^T W (rows N, k M), Y = K X (rows M,
  // k N) in BF16; r0 (rows M, k N), c0 (rows N, k M) in FP64
  int ks_z() const { return splitk_for<uint16_t>(N, M); }
  int ks_y() const { return splitk_for<uint16_t>(M, N); }
  int ks_r() const { return splitk_for<double>(M, N); }
  int ks_c() const { return splitk_for<double>(N, M); }
  // (the BF16 products take their tiles from the column capacity,
  // with_tiles; the k split above does not depend on the tiles, so neither
  // does a column's arithmetic)
  size_t wk_b_elems(int c) const {
    return std::max(tiles_work_elems<uint16_t>(N, c, ks_z()),
                    tiles_work_elems<uint16_t>(M, c, ks_y()));
  }
  size_t wk_d_elems(int c) const {
    return std::max(tiles_work_elems<double>(M, c, ks_r()),
                    tiles_work_elems<double>(N, c, ks_c()));
  }

  // Without a k split (small n) the product kernel runs the update itself:
  // one launch less per product, and the same value reaches the update.
  SplitK<uint16_t> skb(int ks) const {
    SplitK<uint16_t> k;
    k.ks = ks;
    k.work = ks > 1 ? wk_b.get() : nullptr;
    return k;
  }
  SplitK<double> skd(int ks) const {
    SplitK<double> k;
    k.ks = ks;
    k.work = ks > 1 ? wk_d.get() : nullptr;
    return k;
  }

  int err_blocks() const { return std::min(16, std::max(1, (std::max(M, N) + 1023) / 1024)); }

  // Device bytes of the group at c columns (layout() set).
  size_t bytes(int c) {
    return per_col_bytes() * (size_t)c + 4 * wk_b_elems(c) + 8 * wk_d_elems(c) +
           (8 * (size_t)err_blocks() + 4) * c;
  }

  void drop_graphs() {
    for (hipGraphExec_t &g : gx)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }

  ~Group() {
    drop_graphs();
    if (ev) (void)hipEventDestroy(ev);
    for (int *&h : hl)
      if (h) (void)hipHostFree(h);
    if (hdone_pin) (void)hipHostFree(hdone_pin);
    if (herr_pin) (void)hipHostFree(herr_pin);
  }

  // Grows every per-column array to ncap columns, keeping the first cap
  // columns (column-major arrays: a prefix copy). The stream is drained
  // first: the old buffers go back to the allocator cache.
  void grow(int ncap) {
    stream->sync();
    drop_graphs();  // they hold the old buffers' addresses
    for (Slab *s : slabs()) {
      DBuf<char> nb(s->per_col * (size_t)ncap);
      // zero first: the rows of W and X past M and N must read as zeros (the
      // BF16 products contract over the padded lengths)
      RWIG_HIP_CHECK(hipMemsetAsync(nb.get(), 0, s->per_col * (size_t)ncap, stream->get()));
      if (cap) d2d(nb.get(), s->buf.get(), s->per_col * (size_t)cap, stream->get());
      stream->sync();
      s->buf = std::move(nb);
    }
    lists.reset(4 + kLists * (size_t)ncap);
    errpart.reset((size_t)ncap * err_blocks());
    arrive.reset(ncap);
    int *ar = arrive.get();
    launch_map(stream->get(), (long)ncap, [=] __device__(long k) { ar[k] = 0; });
    wk_b.reset(wk_b_elems(ncap));
    wk_d.reset(wk_d_elems(ncap));
    for (int *&h : hl) {
      if (h) RWIG_HIP_CHECK(hipHostFree(h));
      RWIG_HIP_CHECK(hipHostMalloc((void **)&h, sizeof(int) * (4 + kLists * (size_t)ncap)));
    }
    if (hdone_pin) RWIG_HIP_CHECK(hipHostFree(hdone_pin));
    if (herr_pin) RWIG_HIP_CHECK(hipHostFree(herr_pin));
    RWIG_HIP_CHECK(hipHostMalloc((void **)&hdone_pin, sizeof(int) * (size_t)ncap));
    RWIG_HIP_CHECK(hipHostMalloc((void **)&herr_pin, sizeof(double) * (size_t)ncap));
    owner.resize(ncap, nullptr);
    cap = ncap;
  }

  // First run of S free columns, or -1.
  int find_cols(int S) const {
    for (int c = 0, run = 0; c < cap; ++c) {
      run = owner[c] ? 0 : run + 1;
      if (run == S) return c - S + 1;
    }
    return -1;
  }

  // Fills the pinned list buffer of a cycle phase (0: absorption, every
  // running request; 1: iterations, the requests still running after the
  // harvest). Each phase has its own buffer, rewritten only after a stream
  // synchronisation that follows its previous upload.
  void fill_lists(int phase) {
    int *h = hl[phase];
    int *hact = h + 4, *hreq = h + 4 + cap, *hit = h + 4 + 2 * cap, *hbar = h + 4 + 3 * cap,
        *hlim = h + 4 + 4 * cap;
    nact = nreq = nbar = 0;
    for (Request *r : running) {
      hreq[nreq] = r->col;
      hit[nreq] = r->iter;
      ++nreq;
      if (r->kind == kBarycenter) hbar[nbar++] = r->col;
      // the last cycle of a request stops at max_iter exactly
      const int lim = std::min(absorb_every, r->max_iter - r->iter);
      for (int s = 0; s < r->S; ++s) {
        hact[nact++] = r->col + s;
        hlim[r->col + s] = lim;
      }
    }
    h[kNact] = nact;
    h[kNreq] = nreq;
    h[kNbar] = nbar;
    h[3] = 0;
  }
  void upload_lists(int phase) {
    h2d(lists.get(), hl[phase], 4 + kLists * (size_t)cap, stream->get());
  }
  const int *d_cnt() const { return lists.get(); }
  const int *d_act() const { return lists.get() + 4; }
  const int *d_areq() const { return lists.get() + 4 + cap; }
  const int *d_riter() const { return lists.get() + 4 + 2 * cap; }
  const int *d_abar() const { return lists.get() + 4 + 3 * cap; }
  const int *d_lim() const { return lists.get() + 4 + 4 * cap; }

  // Runs one phase of a cycle on the stream: the list upload, then the
  // absorption (phase 0) or absorb_every iterations (phase 1), then for the
  // absorption the download of its stop flags and errors. With graphs, the
  // kernels of the phase are recorded once per column capacity and replayed:
  // every launch reads its column counts and lists from device memory. The
  // copies between page-locked host buffers and the device stay outside the
  // recorded graphs, as plain copies on the same stream before and after the
  // replay, so the graphs hold kernel nodes only.
  void run_phase(int phase) {
    hipStream_t st = stream->get();
    upload_lists(phase);
    auto body = [&] {
      if (phase == 0)
        absorb();
      else
        for (int s = 0; s < absorb_every; ++s) iterate(s);
    };
    if (!graphs) {
      body();
    } else {
      if (!gx[phase]) {
        hipGraph_t g = nullptr;
        RWIG_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        try {
          body();
        } catch (...) {
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        RWIG_HIP_CHECK(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&gx[phase], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        RWIG_HIP_CHECK(e);
      }
      RWIG_HIP_CHECK(hipGraphLaunch(gx[phase], st));
    }
    if (phase == 0) {
      download_flags();
      RWIG_HIP_CHECK(hipEventRecord(ev, st));
    }
  }

  // Column count for the tile policy: at most 32 columns take the 32 x 32
  // tiles here, not the 16 x 32 ones (on the register path they measured 7%
  // slower; with the padded contraction, on the direct-to-LDS path, no
  // faster).
  int tile_cols() const { return std::max(cap, 33); }

  // A product over the running columns (act, count from the device lists), in
  // the tiles the column capacity takes.
  template <typename T, class Op>
  void product(hipStream_t st, int rows, int kdim, const T *A, long lda, const T *Bm, long ldb,
               Op op, const SplitK<T> &sk) {
    with_tiles<T>(tile_cols(), [&](auto cfg) {
      gemm_tn<T, decltype(cfg)>(st, rows, cap, kdim, A, lda, Bm, ldb, op, d_act(), d_cnt() + kNact,
                                sk);
    });
  }

  // Fresh state of the columns [c0, c0 + S) of a newly admitted request.
  void init_request(const Request &r) {
    hipStream_t st = stream->get();
    const int c0i = r.col, S = r.S, Mi = M, Ni = N;
    const long ldmi = ldm, ldni = ldn;
    if (r.kind == kSinkhorn) {
      h2d(acol.p<double>() + (long)c0i * M, r.in1.data(), (size_t)M, st);
      h2d(bcol.p<double>() + (long)c0i * N, r.in2.data(), (size_t)N, st);
      const double one = 1.0;
      h2d(wcol.p<double>() + c0i, &one, 1, st);
    } else {
      h2d(acol.p<double>() + (long)c0i * M, r.in1.data(), (size_t)M * S, st);
      h2d(wcol.p<double>() + c0i, r.in2.data(), (size_t)S, st);
    }
    double *Fp = F.p<double>(), *Adp = Ad.p<double>(), *Gp = G.p<double>(), *Bdp = Bd.p<double>();
    float *lup = lu.p<float>(), *lvp = lv.p<float>();
    uint16_t *Wp = W.p<uint16_t>(), *Xp = X.p<uint16_t>();
    launch_map(st, ldmi * S, [=] __device__(long k) {
      const long i = k % ldmi, t = c0i + k / ldmi;
      Wp[i + t * ldmi] = 0;
      Adp[i + t * ldmi] = 0.0;
      if (i < Mi) {
        Fp[i + t * Mi] = 0.0;
        lup[i + t * Mi] = 1.0f;
      }
    });
    launch_map(st, ldni * S, [=] __device__(long k) {
      const long j = k % ldni, t = c0i + k / ldni;
      Xp[j + t * ldni] = 0;
      Bdp[j + t * ldni] = 0.0;
      if (j < Ni) {
        Gp[j + t * Ni] = 0.0;
        lvp[j + t * Ni] = 1.0f;
      }
    });
    int *kp = kind.p<int>(), *bp = base.p<int>(), *np = ns.p<int>(), *dp = done.p<int>(),
        *mp = mode.p<int>(), *xp = maxit.p<int>();
    double *tp = tol.p<double>();
    const int kd = r.kind, mx = r.max_iter;
    const double tl = r.tol;
    launch_map(st, (long)S, [=] __device__(long s) {
      const long c = c0i + s;
      kp[c] = kd;
      bp[c] = c0i;
      np[c] = S;
      dp[c] = 0;
      mp[c] = 0;
      xp[c] = mx;
      tp[c] = tl;
    });
  }

  // Fold the scalings into the potentials, exact FP64 marginals, stopping
  // test, calibration of the BF16 products and reset of the scalings; the
  // stop flags and errors go to pinned host memory.
  void absorb() {
    hipStream_t st = stream->get();
    const int Mi = M, Ni = N, cp = cap;
    const long ldmi = ldm, ldni = ldn;
    const double rg = reg;
    const int *act = d_act(), *cnt = d_cnt();
    double *Fp = F.p<double>(), *Gp = G.p<double>(), *Adp = Ad.p<double>(),
           *Bdp = Bd.p<double>(), *r0p = r0.p<double>(), *c0p = c0.p<double>(),
           *ap = acol.p<double>(), *bp = bcol.p<double>();
    float *lup = lu.p<float>(), *lvp = lv.p<float>(), *alp = al.p<float>(), *bep = be.p<float>(),
          *rap = ra.p<float>(), *sbp = sb.p<float>(), *Ycp = Yc.p<float>();
    uint16_t *Wp = W.p<uint16_t>(), *Xp = X.p<uint16_t>();
    int *dn = done.p<int>(), *md = mode.p<int>();
    const int *kd = kind.p<int>();
    const DenseKernel &dk = gk->dk;
    // rows with a_i = 0 (columns with b_j = 0) are off the support: potential
    // -inf, scaling 0, so they drop out of every product
    launch_cols(st, Mi, cp, cnt, kNact, [=] __device__(int c, long i) {
      const long t = act[c], q = i + t * Mi;
      if (ap[q] == 0.0) {
        Fp[q] = -INFINITY;
        alp[q] = 0.0f;
        Adp[i + t * ldmi] = 0.0;
        return;
      }
      const double x = (double)lup[q];
      const double fk = Fp[q] + rg * (md[t] ? log1p(x) : log(x));
      Fp[q] = fk;
      const double ak = exp(fk / rg);
      alp[q] = (float)ak;
      Adp[i + t * ldmi] = ak;
    });
    launch_cols(st, Ni, cp, cnt, kNact, [=] __device__(int c, long j) {
      const long t = act[c], q = j + t * Ni;
      if (kd[t] == kSinkhorn && bp[q] == 0.0) {
        Gp[q] = -INFINITY;
        bep[q] = 0.0f;
        Bdp[j + t * ldni] = 0.0;
        return;
      }
      const double x = (double)lvp[q];
      const double gq = Gp[q] + rg * (md[t] ? log1p(x) : log(x));
      Gp[q] = gq;
      const double bq = exp(gq / rg);
      bep[q] = (float)bq;
      Bdp[j + t * ldni] = bq;
    });
    // the FP64 products contract over the zero-padded lengths too (whole
    // k-steps: the direct-to-LDS tiles) and take their tiles from the column
    // capacity, up to the 64 x 64 tiles (the 128 x 128 ones, with the k
    // split fixed by the rows, measured slower on wide groups: 1,024 columns
    // of n = 2048 in 8 splits); the k split does not depend on the tiles
    with_tiles<double>(std::min(tile_cols(), 768), [&](auto cfg) {
      using CF = decltype(cfg);
      gemm_tn<double, CF>(
          st, Mi, cp, (int)ldni, dk.KdT.get(), ldni, Bdp, ldni,
          [=] __device__(int i, int t, double v) { r0p[i + (long)t * Mi] = v * Adp[i + (long)t * ldmi]; },
          act, cnt + kNact, skd(ks_r()));
      gemm_tn<double, CF>(
          st, Ni, cp, (int)ldmi, dk.Kd.get(), ldmi, Adp, ldmi,
          [=] __device__(int j, int t, double v) { c0p[j + (long)t * Ni] = v * Bdp[j + (long)t * ldni]; },
          act, cnt + kNact, skd(ks_c()));
    });
    hipLaunchKernelGGL(k_serve_err, dim3(cp, err_blocks()), dim3(256), 0, st, Mi, Ni, rg, d_areq(),
                       d_riter(), kind.p<int>(), ns.p<int>(), wcol.p<double>(), maxit.p<int>(),
                       tol.p<double>(), ap, r0p, bp, c0p, Gp, switch_thr, dn, md,
                       errs.p<double>(), cnt, errpart.get(), arrive.get());
    check_launch();
    // calibration: the same BF16 products at u = v = 1
    launch_cols(st, Mi, cp, cnt, kNact, [=] __device__(int c, long i) {
      const long t = act[c];
      Wp[i + t * ldmi] = f2bf(alp[i + t * Mi]);
    });
    launch_cols(st, Ni, cp, cnt, kNact, [=] __device__(int c, long j) {
      const long t = act[c];
      Xp[j + t * ldni] = f2bf(bep[j + t * Ni]);
    });
    with_tiles<uint16_t>(tile_cols(), [&](auto cfg) {
      using CF = decltype(cfg);
      gemm_tn<uint16_t, CF>(
          st, Mi, cp, (int)ldni, dk.KbT.get(), ldni, Xp, ldni,
          [=] __device__(int i, int t, float y) {
            if (!dn[t]) Ycp[i + (long)t * Mi] = y;
          },
          act, cnt + kNact, skb(ks_y()));
      gemm_tn<uint16_t, CF>(
          st, Ni, cp, (int)ldmi, dk.Kb.get(), ldmi, Wp, ldmi,
          [=] __device__(int j, int t, float z) {
            const long q = j + (long)t * Ni;
            if (!dn[t]) sbp[q] = (float)(c0p[q] / (double)z);
          },
          act, cnt + kNact, skb(ks_z()));
    });
    // u = a / r0 at v = 1; W for the first product of the cycle
    launch_cols(st, Mi, cp, cnt, kNact, [=] __device__(int c, long i) {
      const long t = act[c], q = i + t * Mi;
      if (dn[t]) return;
      rap[q] = (float)(r0p[q] / (double)Ycp[q]);
      const double u = ap[q] > 0.0 ? ap[q] / r0p[q] : 0.0;
      const double a_ = (double)alp[q];
      lup[q] = md[t] ? (float)(u - 1.0) : (float)u;
      Wp[i + t * ldmi] = f2bf(md[t] ? (float)(a_ * (u - 1.0)) : (float)(a_ * u));
    });
    launch_cols(st, Ni, cp, cnt, kNact, [=] __device__(int c, long j) {
      const long t = act[c];
      if (!dn[t]) lvp[j + t * Ni] = md[t] ? 0.0f : 1.0f;
    });
  }
  // the absorption's stop flags and errors into page-locked host buffers
  void download_flags() {
    hipStream_t st = stream->get();
    d2h(hdone_pin, done.p<int>(), (size_t)cap, st);
    d2h(herr_pin, errs.p<double>(), (size_t)cap, st);
  }

  // One iteration over the running columns: Z = K^T W -> v, X; for the
  // barycenters the common target and v; Y = K X -> u, W.
  // Columns whose request has run its max_iter iterations (lim <= s) are
  // left as they are. Rows and columns off the support keep zero scalings.
  void iterate(int s_it) {
    hipStream_t st = stream->get();
    const int Mi = M, Ni = N, cp = cap;
    const long ldmi = ldm, ldni = ldn;
    const double rg = reg;
    const int *abar = d_abar(), *cnt = d_cnt(), *lim = d_lim();
    const int *kd = kind.p<int>(), *md = mode.p<int>(), *np = ns.p<int>();
    const double *c0p = c0.p<double>(), *r0p = r0.p<double>(), *ap = acol.p<double>(),
                 *bp = bcol.p<double>(), *Gp = G.p<double>(), *wp = wcol.p<double>();
    const float *bep = be.p<float>(), *sbp = sb.p<float>(), *alp = al.p<float>(),
                *rap = ra.p<float>();
    float *lvp = lv.p<float>(), *lup = lu.p<float>(), *Zqp = Zq.p<float>();
    uint16_t *Wp = W.p<uint16_t>(), *Xp = X.p<uint16_t>();
    const DenseKernel &dk = gk->dk;
    product<uint16_t>(
        st, Ni, (int)ldmi, dk.Kb.get(), ldmi, Wp, ldmi,
        [=] __device__(int j, int t, float z) {
          if (s_it >= lim[t]) return;
          const long q = j + (long)t * Ni;
          if (kd[t] == kBarycenter) {
            Zqp[q] = z;
            return;
          }
          const double b_ = (double)bep[q];
          const bool on = bp[q] > 0.0;
          double v;
          if (md[t]) {
            v = on ? bp[q] / (c0p[q] + b_ * (double)z) : 0.0;
            Xp[j + (long)t * ldni] = f2bf((float)(b_ * (v - 1.0)));
            lvp[q] = (float)(v - 1.0);
          } else {
            v = on ? bp[q] / ((double)sbp[q] * (double)z) : 0.0;
            Xp[j + (long)t * ldni] = f2bf((float)(b_ * v));
            lvp[q] = (float)v;
          }
        },
        skb(ks_z()));
    launch_cols(st, Ni, cp, cnt, kNbar, [=] __device__(int c, long j) {
      const long r = abar[c];
      if (s_it >= lim[r]) return;
      const int S = np[r], m1 = md[r];
      double lb = 0.0;
      for (int s = 0; s < S; ++s) {
        const long q = j + (r + s) * Ni;
        const double z = (double)Zqp[q];
        const double qs = m1 ? c0p[q] + (double)bep[q] * z : (double)sbp[q] * z;
        lb += wp[r + s] * (log(qs) - Gp[q] / rg);
      }
      const double bj = exp(lb);
      for (int s = 0; s < S; ++s) {
        const long q = j + (r + s) * Ni;
        const double z = (double)Zqp[q];
        const double qs = m1 ? c0p[q] + (double)bep[q] * z : (double)sbp[q] * z;
        const double v = bj / qs;
        const double b_ = (double)bep[q];
        Xp[j + (r + s) * ldni] = f2bf(m1 ? (float)(b_ * (v - 1.0)) : (float)(b_ * v));
        lvp[q] = m1 ? (float)(v - 1.0) : (float)v;
      }
    });
    product<uint16_t>(
        st, Mi, (int)ldni, dk.KbT.get(), ldni, Xp, ldni,
        [=] __device__(int i, int t, float y) {
          if (s_it >= lim[t]) return;
          const long q = i + (long)t * Mi;
          const double a_ = (double)alp[q];
          const bool on = ap[q] > 0.0;
          double u;
          if (md[t]) {
            u = on ? ap[q] / (r0p[q] + a_ * (double)y) : 0.0;
            Wp[i + (long)t * ldmi] = f2bf((float)(a_ * (u - 1.0)));
            lup[q] = (float)(u - 1.0);
          } else {
            u = on ? ap[q] / ((double)rap[q] * (double)y) : 0.0;
            Wp[i + (long)t * ldmi] = f2bf((float)(a_ * u));
            lup[q] = (float)u;
          }
        },
        skb(ks_y()));
  }

  // Copies of a finished request's results to the host (enqueued; complete
  // after the next stream synchronisation).
  void harvest_enqueue(Request &r) {
    hipStream_t st = stream->get();
    const long c = r.col;
    r.err = herr_pin[c];
    // return_status as src/sinkhorn.cpp:88-95 computes it
    if (!std::isfinite(r.err)) {
      r.status = 2;
      r.numeric = true;
    } else {
      r.status = r.err <= r.tol ? 0 : (r.iter >= r.max_iter ? 1 : 2);
    }
    if (r.kind == kSinkhorn) {
      r.out1.resize(M);
      r.out2.resize(N);
      r.r0.resize(M);
      r.c0.resize(N);
      d2h(r.out1.data(), F.p<double>() + c * M, (size_t)M, st);
      d2h(r.out2.data(), G.p<double>() + c * N, (size_t)N, st);
      d2h(r.r0.data(), r0.p<double>() + c * M, (size_t)M, st);
      d2h(r.c0.data(), c0.p<double>() + c * N, (size_t)N, st);
    } else {
      // b, and the potentials F (M x S) then G (N x S) of the absorbed plans
      r.out1.resize(N);
      d2h(r.out1.data(), bcol.p<double>() + c * N, (size_t)N, st);
      const size_t ms = (size_t)M * r.S, ns = (size_t)N * r.S;
      r.out2.resize(ms + ns);
      d2h(r.out2.data(), F.p<double>() + c * M, ms, st);
      d2h(r.out2.data() + ms, G.p<double>() + c * N, ns, st);
    }
  }

  // Host-side completion of a harvested request (after the synchronisation).
  void harvest_finish(Request &r) {
    if (r.kind == kSinkhorn) {
      // regularised loss sum C.*P + reg sum P.*(log P - 1)
      // (src/sinkhorn_impl_log.cpp:95-101) of the absorbed plan
      // P = exp((f 1^T + 1 g^T - C) / reg): there C + reg (log P - 1) =
      // f_i + g_j - reg, so the loss is f.r + g.c - reg sum(r) with r, c the
      // plan's exact marginals from the last absorption
      // (entries off the support have f or g = -inf and a zero marginal:
      // they add nothing)
      double fr = 0.0, gc = 0.0, tr = 0.0;
      for (int i = 0; i < M; ++i) {
        if (r.r0[i] == 0.0) continue;
        fr += r.out1[i] * r.r0[i];
        tr += r.r0[i];
      }
      for (int j = 0; j < N; ++j)
        if (r.c0[j] != 0.0) gc += r.out2[j] * r.c0[j];
      r.loss = fr + gc - reg * tr;
      r.r0.clear();
      r.c0.clear();
    } else {
      double t = 0.0;
      for (double x : r.out1) t += x;
      for (double &x : r.out1) x /= t;  // src/barycenter_impl_parallel.cpp:39-43
    }
  }
};

}  // namespace

class ServeEngine {
 public:
  ServeEngine(double budget, int max_cols, int absorb_every, double switch_thr, bool graphs)
      : budget_(budget), max_cols_(max_cols), absorb_every_(absorb_every),
        switch_thr_(switch_thr), graphs_(graphs) {
    RWIG_HIP_CHECK(hipGetDevice(&device_));
    th_ = std::thread([this] { loop(); });
  }

  ~ServeEngine() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    th_.join();
  }

  int register_cost(const double *C, int M, int N) {
    const size_t bytes = (size_t)M * N * sizeof(double);
    auto c = std::make_unique<Cost>();
    c->M = M;
    c->N = N;
    c->rowmin.assign(M, INFINITY);
    c->colmin.assign(N, INFINITY);
    for (long j = 0; j < N; ++j)
      for (long i = 0; i < M; ++i) {
        const double x = C[i + j * M];
        RWIG_REQUIRE(std::isfinite(x), "the cost matrix must be finite");
        c->rowmin[i] = std::min(c->rowmin[i], x);
        c->colmin[j] = std::min(c->colmin[j], x);
      }
    {
      std::lock_guard<std::mutex> lk(mu_);
      RWIG_REQUIRE(fatal_.empty(), "serving session failed: " + fatal_);
      RWIG_REQUIRE(used_ + bytes <= budget_, "cost matrix exceeds the session's HBM budget");
      used_ += bytes;
      cost_bytes_ += bytes;
      peak_ = std::max(peak_, used_);
    }
    try {
      Stream s;
      c->C.reset((size_t)M * N);
      h2d(c->C.get(), C, (size_t)M * N, s.get());
      s.sync();
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      used_ -= bytes;
      cost_bytes_ -= bytes;
      throw;
    }
    std::lock_guard<std::mutex> lk(mu_);
    costs_.push_back(std::move(c));
    return (int)costs_.size() - 1;
  }

  long submit(std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> lk(mu_);
    RWIG_REQUIRE(fatal_.empty(), "serving session failed: " + fatal_);
    RWIG_REQUIRE(r->cost >= 0 && r->cost < (int)costs_.size(), "unknown cost id");
    const Cost &c = *costs_[r->cost];
    if (r->kind == kSinkhorn)
      RWIG_REQUIRE((int)r->in1.size() == c.M && (int)r->in2.size() == c.N,
                   "a and b must match the cost matrix");
    else
      RWIG_REQUIRE((int)r->in1.size() == c.M * r->S, "A must have M rows");
    RWIG_REQUIRE(r->S <= max_cols_, "request has more measures than a group holds");
    // the shared Gibbs kernel exp(-C / reg) must not underflow on a row or
    // column the request uses
    auto too_small = [&](const char *what, long k, double cmin) {
      throw Error(RWIG_ERR_INVALID,
                  "reg = " + std::to_string(r->reg) + " is too small for this cost: " + what +
                      " " + std::to_string(k) + " has min C / reg = " +
                      std::to_string(cmin / r->reg) + " > " + std::to_string(kMaxCostOverReg) +
                      ", where exp(-C / reg) underflows the low-precision Gibbs kernel; solve "
                      "it with the log-domain form (rwig_sinkhorn_log, rwig_barycenter_log)");
    };
    for (long i = 0; i < c.M; ++i) {
      bool used = false;
      for (int s = 0; s < r->S && !used; ++s) used = r->in1[i + (long)s * c.M] > 0.0;
      if (used && c.rowmin[i] / r->reg > kMaxCostOverReg) too_small("row", i, c.rowmin[i]);
    }
    for (long j = 0; j < c.N; ++j) {
      const bool used = r->kind == kBarycenter || r->in2[j] > 0.0;
      if (used && c.colmin[j] / r->reg > kMaxCostOverReg) too_small("column", j, c.colmin[j]);
    }
    const size_t need = alone_bytes(c.M, c.N, r->S);
    RWIG_REQUIRE(cost_bytes_ + need <= budget_,
                 "request does not fit the session's HBM budget even alone");
    r->id = next_id_++;
    r->t_submit = Clock::now();
    if (!started_) {
      t_first_ = r->t_submit;
      started_ = true;
    }
    const long id = r->id;
    Request *rp = r.get();
    reqs_[id] = std::move(r);
    incoming_.push_back(rp);
    ++queued_;
    cv_work_.notify_all();
    return id;
  }

  int wait(long id, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    RWIG_REQUIRE(it != reqs_.end(), "unknown request id");
    Request *r = it->second.get();
    auto pred = [&] { return r->state == kDone || r->state == kFailed; };
    if (timeout_s < 0)
      cv_done_.wait(lk, pred);
    else
      cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    return r->state;
  }

  std::unique_ptr<Request> take(long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(id);
    RWIG_REQUIRE(it != reqs_.end(), "unknown request id");
    Request *r = it->second.get();
    RWIG_REQUIRE(r->state == kDone || r->state == kFailed, "request has not finished");
    std::unique_ptr<Request> out = std::move(it->second);
    reqs_.erase(it);
    return out;
  }

  void stats(double *out, int n) {
    std::lock_guard<std::mutex> lk(mu_);
    std::vector<double> lat = lat_;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      const double x = p * (lat.size() - 1);
      const size_t lo = (size_t)x, hi = std::min(lo + 1, lat.size() - 1);
      return lat[lo] + (x - lo) * (lat[hi] - lat[lo]);
    };
    double mean = 0.0, qmean = 0.0;
    for (double x : lat) mean += x;
    for (double x : qwait_) qmean += x;
    if (!lat.empty()) mean /= lat.size();
    if (!qwait_.empty()) qmean /= qwait_.size();
    const double span = (completed_ > 0) ? ms_between(t_first_, t_last_) / 1e3 : 0.0;
    const double v[] = {(double)completed_,
                        (double)failed_,
                        (double)queued_,
                        (double)running_,
                        pct(0.50),
                        pct(0.99),
                        mean,
                        span > 0 ? completed_ / span : 0.0,
                        (double)used_,
                        (double)peak_,
                        budget_,
                        (double)cycles_,
                        (double)ngroups_,
                        qmean,
                        lat.empty() ? 0.0 : lat.back()};
    const int nv = (int)(sizeof(v) / sizeof(v[0]));
    for (int k = 0; k < n; ++k) out[k] = k < nv ? v[k] : 0.0;
  }

  void reset_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    lat_.clear();
    qwait_.clear();
    completed_ = failed_ = 0;
    started_ = false;
    peak_ = used_;
    cycles_ = 0;
  }

 private:
  // HBM bytes of a request running alone: its Gibbs kernel and one page of
  // group columns (at least S).
  size_t alone_bytes(int M, int N, int S) {
    Group probe;
    probe.M = M;
    probe.N = N;
    probe.ldm = pad64(M);
    probe.ldn = pad64(N);
    probe.layout();
    const int cols = std::max(kColPage, S);
    return dense_kernel_bytes(M, N) + probe.bytes(cols);
  }

  // Makes room for `need` bytes by releasing idle groups, then idle Gibbs
  // kernels, least recently used first (engine thread).
  bool make_room(size_t need, const Gibbs *keep) {
    auto used = [&] {
      std::lock_guard<std::mutex> lk(mu_);
      return used_;
    };
    while (used() + need > budget_) {
      Group *gv = nullptr;
      for (auto &g : groups_)
        if (g->running.empty() && (!gv || g->last < gv->last)) gv = g.get();
      if (gv) {
        release_group(gv);
        continue;
      }
      Gibbs *kv = nullptr;
      for (auto &k : gibbs_)
        if (k->users == 0 && k.get() != keep && (!kv || k->last < kv->last)) kv = k.get();
      if (!kv) return false;
      {
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= kv->bytes;
      }
      gibbs_.erase(std::find_if(gibbs_.begin(), gibbs_.end(),
                                [&](const std::unique_ptr<Gibbs> &p) { return p.get() == kv; }));
    }
    return true;
  }

  void charge(long delta) {
    std::lock_guard<std::mutex> lk(mu_);
    used_ = (size_t)((long)used_ + delta);
    peak_ = std::max(peak_, used_);
  }

  void release_group(Group *g) {
    g->stream->sync();
    charge(-(long)g->bytes(g->cap));
    g->gk->users -= 1;
    groups_.erase(std::find_if(groups_.begin(), groups_.end(),
                               [&](const std::unique_ptr<Group> &p) { return p.get() == g; }));
    ngroups_ = (int)groups_.size();
  }

  // Admits r if the budget holds it (engine thread). Returns false to keep
  // it queued.
  bool admit(Request *r) {
    const Cost *cp;
    {
      std::lock_guard<std::mutex> lk(mu_);  // costs_ may grow concurrently
      cp = costs_[r->cost].get();
    }
    const Cost &c = *cp;
    Group *g = nullptr;
    for (auto &p : groups_)
      if (p->cost == r->cost && p->reg == r->reg) g = p.get();
    Gibbs *k = nullptr;
    for (auto &p : gibbs_)
      if (p->cost == r->cost && p->reg == r->reg) k = p.get();
    if (g && g->find_cols(r->S) >= 0) {
      place(g, r);
      return true;
    }
    Group probe;
    probe.M = c.M;
    probe.N = c.N;
    probe.ldm = pad64(c.M);
    probe.ldn = pad64(c.N);
    probe.layout();
    if (g) {
      int ncap = g->cap;
      while (ncap < max_cols_ && g->find_cols(r->S) < 0) {
        ncap = std::min(max_cols_, std::max(ncap * 2, ncap + r->S));
        const size_t need = g->bytes(ncap) - g->bytes(g->cap);
        if (!make_room(need, g->gk)) return false;
        charge((long)need);
        g->grow(ncap);
      }
      if (g->find_cols(r->S) < 0) return false;
      place(g, r);
      return true;
    }
    const int cap0 = std::min(max_cols_, std::max(kColPage, (r->S + kColPage - 1) / kColPage * kColPage));
    const size_t kbytes = k ? 0 : dense_kernel_bytes(c.M, c.N);
    if (!make_room(kbytes + probe.bytes(cap0), k)) return false;
    auto ng = std::make_unique<Group>();
    ng->cost = r->cost;
    ng->reg = r->reg;
    ng->M = c.M;
    ng->N = c.N;
    ng->ldm = pad64(c.M);
    ng->ldn = pad64(c.N);
    ng->layout();
    ng->graphs = graphs_;
    ng->absorb_every = absorb_every_;
    ng->switch_thr = switch_thr_;
    ng->stream = std::make_unique<Stream>();
    RWIG_HIP_CHECK(hipEventCreateWithFlags(&ng->ev, hipEventDisableTiming));
    if (!k) {
      auto nk = std::make_unique<Gibbs>();
      nk->cost = r->cost;
      nk->reg = r->reg;
      nk->bytes = kbytes;
      dense_kernel_build(ng->stream->get(), c.C.get(), c.M, c.M, c.N, r->reg, nk->dk);
      charge((long)kbytes);
      k = nk.get();
      gibbs_.push_back(std::move(nk));
    }
    k->users += 1;
    ng->gk = k;
    charge((long)ng->bytes(cap0));
    ng->grow(cap0);
    g = ng.get();
    groups_.push_back(std::move(ng));
    ngroups_ = (int)groups_.size();
    place(g, r);
    return true;
  }

  void place(Group *g, Request *r) {
    const int col = g->find_cols(r->S);
    for (int s = 0; s < r->S; ++s) g->owner[col + s] = r;
    r->group = g;
    r->col = col;
    r->iter = 0;
    g->init_request(*r);
    g->running.push_back(r);
    std::lock_guard<std::mutex> lk(mu_);
    r->state = kRunning;
    r->t_admit = Clock::now();
    qwait_.push_back(ms_between(r->t_submit, r->t_admit));
    --queued_;
    ++running_;
  }

  void finish(Request *r, bool ok, const std::string &msg) {
    std::lock_guard<std::mutex> lk(mu_);
    r->state = ok ? kDone : kFailed;
    r->error = msg;
    r->t_done = Clock::now();
    if (ok) {
      lat_.push_back(ms_between(r->t_submit, r->t_done));
      ++completed_;
      t_last_ = r->t_done;
    } else {
      ++failed_;
    }
  }

  void cycle() {
    ++cycles_;
    std::vector<Group *> active;
    for (auto &g : groups_)
      if (!g->running.empty()) active.push_back(g.get());
    for (Group *g : active) {
      g->last = cycles_;
      g->gk->last = cycles_;
      g->fill_lists(0);
      g->run_phase(0);
    }
    // harvest the groups in the order their absorptions finish (polled), so
    // that a group of large problems does not hold back the iterations of the
    // groups behind it in the list; the next cycle starts once every group
    // has been harvested
    std::vector<Group *> waiting = active;
    while (!waiting.empty()) {
      Group *g = nullptr;
      for (auto it = waiting.begin(); it != waiting.end(); ++it) {
        const hipError_t q = hipEventQuery((*it)->ev);
        if (q == hipErrorNotReady) continue;
        RWIG_HIP_CHECK(q);
        g = *it;
        waiting.erase(it);
        break;
      }
      if (!g) {
        std::this_thread::sleep_for(std::chrono::microseconds(10));
        continue;
      }
      std::vector<Request *> keep, fin;
      for (Request *r : g->running) {
        if (g->hdone_pin[r->col]) {
          g->harvest_enqueue(*r);
          fin.push_back(r);
        } else {
          keep.push_back(r);
        }
      }
      if (!fin.empty()) {
        g->stream->sync();
        for (Request *r : fin) {
          g->harvest_finish(*r);
          for (int s = 0; s < r->S; ++s) g->owner[r->col + s] = nullptr;
          {
            std::lock_guard<std::mutex> lk(mu_);
            --running_;
          }
          if (r->numeric)
            finish(r, false,
                   "the marginal error turned non-finite (a scaling or the Gibbs kernel left the "
                   "floating-point range)");
          else
            finish(r, true, "");
        }
        cv_done_.notify_all();
      }
      g->running.swap(keep);
      if (g->running.empty()) continue;
      g->fill_lists(1);
      g->run_phase(1);
      for (Request *r : g->running) r->iter += std::min(absorb_every_, r->max_iter - r->iter);
    }
  }

  void fail_all(const std::string &msg) {
    std::vector<Request *> all;
    {
      std::lock_guard<std::mutex> lk(mu_);
      fatal_ = msg;
      for (auto &kv : reqs_)
        if (kv.second->state == kQueued || kv.second->state == kRunning) all.push_back(kv.second.get());
      incoming_.clear();
    }
    for (Request *r : all) finish(r, false, msg);
    {
      std::lock_guard<std::mutex> lk(mu_);
      queued_ = running_ = 0;
    }
    pending_.clear();
    cv_done_.notify_all();
  }

  void loop() {
    (void)hipSetDevice(device_);
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] {
          return stop_ || !incoming_.empty() || !pending_.empty() || running_ > 0;
        });
        if (stop_) break;
        for (Request *r : incoming_) pending_.push_back(r);
        incoming_.clear();
      }
      try {
        // admission in arrival order; the first request that does not fit
        // waits for memory, and so do the ones behind it
        while (!pending_.empty() && admit(pending_.front())) pending_.pop_front();
        cycle();
      } catch (const std::exception &e) {
        fail_all(e.what());
        break;
      }
    }
    try {
      for (auto &g : groups_) g->stream->sync();
      groups_.clear();
      gibbs_.clear();
    } catch (...) {
    }
    std::vector<Request *> left;
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (auto &kv : reqs_)
        if (kv.second->state == kQueued || kv.second->state == kRunning) left.push_back(kv.second.get());
    }
    for (Request *r : left) finish(r, false, "serving session destroyed");
    cv_done_.notify_all();
  }

  // configuration
  double budget_;
  int max_cols_, absorb_every_;
  double switch_thr_;
  bool graphs_;
  int device_ = 0;
  // shared state (mu_)
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::unique_ptr<Cost>> costs_;
  std::unordered_map<long, std::unique_ptr<Request>> reqs_;
  std::deque<Request *> incoming_;
  size_t used_ = 0, peak_ = 0, cost_bytes_ = 0;
  long next_id_ = 1;
  bool stop_ = false, started_ = false;
  std::string fatal_;
  int queued_ = 0, running_ = 0, ngroups_ = 0;
  long completed_ = 0, failed_ = 0;
  uint64_t cycles_ = 0;
  std::vector<double> lat_, qwait_;
  Clock::time_point t_first_, t_last_;
  // engine-thread state
  std::deque<Request *> pending_;
  std::vector<std::unique_ptr<Group>> groups_;
  std::vector<std::unique_ptr<Gibbs>> gibbs_;
  std::thread th_;
};

}  // namespace rwig

using namespace rwig;

namespace {

ServeEngine *engine_of(void *h) {
  RWIG_REQUIRE(h != nullptr, "null serving session");
  return static_cast<ServeEngine *>(h);
}

// A measure: finite non-negative entries with a positive total; zero entries
// are off its support.
void check_mass(const double *x, int n, const char *what) {
  double t = 0.0;
  for (int k = 0; k < n; ++k) {
    RWIG_REQUIRE(std::isfinite(x[k]) && x[k] >= 0.0,
                 std::string("entries of ") + what + " must be finite and non-negative");
    t += x[k];
  }
  RWIG_REQUIRE(t > 0.0, std::string(what) + " must have positive mass");
}

}  // namespace

extern "C" {

int rwig_serve_create(double hbm_budget_bytes, int max_group_columns, int absorb_every,
                      double switch_thr, int flags, void **session) {
  return guarded("rwig_serve_create", [&] {
    RWIG_REQUIRE(session != nullptr, "null session pointer");
    RWIG_REQUIRE(hbm_budget_bytes > 0, "HBM budget must be positive");
    RWIG_REQUIRE(max_group_columns >= 1 && max_group_columns <= 4096,
                 "max_group_columns must be in [1, 4096]");
    RWIG_REQUIRE(absorb_every >= 1, "absorb_every must be positive");
    ensure_device();
    RWIG_REQUIRE((flags & ~RWIG_SERVE_EAGER) == 0, "unknown flags");
    *session = new ServeEngine(hbm_budget_bytes, max_group_columns, absorb_every, switch_thr,
                               (flags & RWIG_SERVE_EAGER) == 0);
  });
}

int rwig_serve_footprint(int M, int N, int columns, double *kernel_bytes, double *group_bytes) {
  return guarded("rwig_serve_footprint", [&] {
    RWIG_REQUIRE(M > 0 && N > 0 && columns > 0, "invalid argument");
    Group probe;
    probe.M = M;
    probe.N = N;
    probe.ldm = pad64(M);
    probe.ldn = pad64(N);
    probe.layout();
    if (kernel_bytes) *kernel_bytes = (double)dense_kernel_bytes(M, N);
    if (group_bytes) *group_bytes = (double)probe.bytes(columns);
  });
}

int rwig_serve_destroy(void *session) {
  return guarded("rwig_serve_destroy", [&] { delete engine_of(session); });
}

int rwig_serve_register_cost(void *session, const double *C, int M, int N, int *cost_id) {
  return guarded("rwig_serve_register_cost", [&] {
    RWIG_REQUIRE(C && cost_id && M > 0 && N > 0, "invalid argument");
    *cost_id = engine_of(session)->register_cost(C, M, N);
  });
}

int rwig_serve_submit_sinkhorn(void *session, int cost_id, double reg, const double *a, int M,
                               const double *b, int N, int max_iter, double zero_tol,
                               long *request_id) {
  return guarded("rwig_serve_submit_sinkhorn", [&] {
    RWIG_REQUIRE(a && b && request_id && M > 0 && N > 0, "invalid argument");
    RWIG_REQUIRE(reg > 0.0 && max_iter >= 0, "reg must be positive, max_iter non-negative");
    check_mass(a, M, "a");
    check_mass(b, N, "b");
    auto r = std::make_unique<Request>();
    r->kind = kSinkhorn;
    r->cost = cost_id;
    r->reg = reg;
    r->S = 1;
    r->max_iter = max_iter;
    r->tol = zero_tol;
    r->in1.assign(a, a + M);
    r->in2.assign(b, b + N);
    *request_id = engine_of(session)->submit(std::move(r));
  });
}

int rwig_serve_submit_barycenter(void *session, int cost_id, double reg, const double *A, int M,
                                 int S, const double *w, int max_iter, double zero_tol,
                                 long *request_id) {
  return guarded("rwig_serve_submit_barycenter", [&] {
    RWIG_REQUIRE(A && w && request_id && M > 0 && S >= 1, "invalid argument");
    RWIG_REQUIRE(reg > 0.0 && max_iter >= 0, "reg must be positive, max_iter non-negative");
    for (int s = 0; s < S; ++s) check_mass(A + (long)s * M, M, "each column of A");
    auto r = std::make_unique<Request>();
    r->kind = kBarycenter;
    r->cost = cost_id;
    r->reg = reg;
    r->S = S;
    r->max_iter = max_iter;
    r->tol = zero_tol;
    r->in1.assign(A, A + (long)M * S);
    r->in2.assign(w, w + S);
    *request_id = engine_of(session)->submit(std::move(r));
  });
}

int rwig_serve_wait(void *session, long request_id, double timeout_s, int *state) {
  return guarded("rwig_serve_wait", [&] {
    const int s = engine_of(session)->wait(request_id, timeout_s);
    if (state) *state = s;
  });
}

int rwig_serve_result(void *session, long request_id, double *x, double *y, double *loss,
                      int *iter, double *err, double *latency_ms, int *status) {
  int code = RWIG_OK;
  std::string msg;
  const int rc = guarded("rwig_serve_result", [&] {
    std::unique_ptr<Request> r = engine_of(session)->take(request_id);
    if (status) *status = r->status;
    if (r->state == kFailed && !r->numeric) {
      code = RWIG_ERR_HIP;
      msg = "request failed: " + r->error;
      return;
    }
    if (x) std::copy(r->out1.begin(), r->out1.end(), x);
    if (y) std::copy(r->out2.begin(), r->out2.end(), y);
    if (loss) *loss = r->loss;
    if (iter) *iter = r->iter;
    if (err) *err = r->err;
    if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
  });
  if (rc != RWIG_OK) return rc;
  if (code != RWIG_OK) {
    set_last_error("rwig_serve_result: " + msg);
    return code;
  }
  return RWIG_OK;
}

int rwig_serve_stats(void *session, double *out, int n_out) {
  return guarded("rwig_serve_stats", [&] {
    RWIG_REQUIRE(out && n_out >= 0, "invalid argument");
    engine_of(session)->stats(out, n_out);
  });
}

int rwig_serve_reset_stats(void *session) {
  return guarded("rwig_serve_reset_stats", [&] { engine_of(session)->reset_stats(); });
}

}  // extern "C"
