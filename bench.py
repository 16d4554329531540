"""Benchmark of the entropic-OT hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): batched
log-stabilised Sinkhorn, one source against many targets -- n = 4096 points
uniform in [0,1]^2, squared Euclidean cost, uniform source, B = 1024 target
histograms with Dirichlet(0.5) mass on a random 5% support (205 points),
eps = 0.01, every solve run until its L1 marginal error
sum|P1 - a| + sum|P^T 1 - b| (FP64, on the GPU) is <= 1e-6.

A step is one batched solve of all 1024 targets with inputs resident in HBM.
The headline value is BF16 solves/s of the MFMA path (every Gibbs product of
an iteration one BF16 MFMA GEMM over all targets, FP32 accumulation, FP64
absorption every 20 iterations); the resident per-target kernels (BF16, FP32),
the batched FP64 log-domain form and the plain FP64 per-call form (one
rwig_sinkhorn_log_dev call per target) are measured beside it, and the CPU oracle (the reference algorithm restated in
C++) is timed on a bounded sample of targets on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Sinkhorn/barycenter solves/s at 1e-6 marginal err vs plain FP64; p50/p99 ms"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TF = 2500.0
ABSORB_MFMA = 20  # iterations between FP64 absorptions on the MFMA path (the workload's)
ABSORB_ALT = 30   # a second interval, reported beside the headline
GEMM_KSPLIT = 1   # k split of the MFMA path's BF16 products (deferred epilogue)


def pmc_traffic():
    """HBM bytes per launch of the dominant product from the committed counter
    runs (profiles/c2_gemm_hbm_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "c2_gemm_hbm_traffic.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)["traffic_bytes"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


C5_ARRIVAL_RATE = 1200.0  # requests/s of the paced serving run, fixed across runs


def host_info():
    """Threads the CPU baselines use (the process's CPU affinity, at most 16:
    the host share of one GPU), with nproc and the CPU model."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"threads": max(1, min(16, aff)), "affinity_cpus": aff,
            "nproc": os.cpu_count(), "cpu_model": model}


def c1_leg(log, cpu=True, reps=50):
    """Plain vanilla Sinkhorn (C1): n = m = 256 on the 1-D grid, eps 0.05, up
    to 1000 iterations, stopping on the reference's error at 1e-9. The GPU
    form is the drop-in rwig_sinkhorn_vanilla (host buffers in and out; the
    whole loop is one launch of 8 workgroups holding K and K^T in LDS); the
    CPU form the oracle's
    restatement of Sinkhorn::compute_vanilla on one thread (the reference
    runs BLAS single-threaded)."""
    import rwig_amd
    from rwig_amd.workloads import c1
    a, b, C, eps = c1()
    ctrl = {"reg": eps, "method": "vanilla", "max_iter": 1000, "zero_tol": 1e-9}
    out = {"workload": "vanilla Sinkhorn, n=m=256 on the 1-D grid, eps=0.05, <= 1000 "
                       "iterations, reference error (L2 residuals) < 1e-9"}
    for wg in (False, True):
        ctrl["with_grad"] = wg
        r = rwig_amd.sinkhorn(a, b, C, ctrl)
        t = time.perf_counter()
        for _ in range(reps):
            r = rwig_amd.sinkhorn(a, b, C, ctrl)
        dt = (time.perf_counter() - t) / reps
        key = "with_grad" if wg else "forward"
        out[key] = {"ms_per_solve": dt * 1e3, "iterations": int(r["iter"]),
                    "us_per_iteration": dt / max(1, r["iter"]) * 1e6,
                    "solves_per_s": 1.0 / dt, "err": r["err"],
                    "return_status": int(r["return_status"])}
        P = r["P"]
        out[key]["l1_marginal_err"] = float(np.abs(P.sum(1) - a).sum()
                                            + np.abs(P.sum(0) - b).sum())
        log(f"C1 GPU {key}: {dt * 1e3:.3f} ms per solve, {r['iter']} iterations, "
            f"{dt / max(1, r['iter']) * 1e6:.2f} us per iteration")
    if cpu:
        from oracle import oracle as O
        n = 20
        t = time.perf_counter()
        for _ in range(n):
            o = O.sinkhorn_vanilla(a, b, C, eps, maxiter=1000, zerotol=1e-9)
        dt = (time.perf_counter() - t) / n
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "solves/s", "cores": 1,
                               "kind": "port", "ms_per_solve": dt * 1e3,
                               "us_per_iteration": dt / max(1, o["iter"]) * 1e6,
                               "sample": f"{n} solves of the C1 problem, C++ oracle "
                                         "(Sinkhorn::compute_vanilla restated), one thread"}
        log(f"C1 CPU oracle: {dt * 1e3:.3f} ms per solve, {o['iter']} iterations")
    return out


def c2_gemm(torch, dev, stream, product_only=True):
    """The dominant kernel of the MFMA path at its full shape: the BF16 product
    with M = K = 4096 (the cost's points) and N = 1024 (the targets), FP32
    accumulation. product_only: the product kernel alone writes its FP32 tile
    into the workspace, as the path launches it before its own update kernels;
    otherwise the generic split-k epilogue follows and writes D. Returns
    (launch, flops per launch)."""
    from rwig_amd._lib import check, lib
    gM, gN, gK = 4096, 1024, 4096
    Wt = torch.randn((gN, gK), device=dev).to(torch.bfloat16)
    Kt = torch.randn((gM, gK), device=dev).to(torch.bfloat16)
    Dt = torch.empty((gN, gM), dtype=torch.float32, device=dev)
    nb = ctypes.c_long()
    check(lib().rwig_gemm_workspace_bytes(0, gM, gN, GEMM_KSPLIT, ctypes.byref(nb)))
    wk = torch.empty(nb.value, dtype=torch.uint8, device=dev)

    def launch():
        check(lib().rwig_gemm_tn_bf16_dev(gM, gN, gK, Kt.data_ptr(), gK, Wt.data_ptr(), gK,
                                          None if product_only else Dt.data_ptr(), gM,
                                          GEMM_KSPLIT, wk.data_ptr(), stream.cuda_stream))
    return launch, 2.0 * gM * gN * gK


def c2_gemm_workload(torch, dev, stream, Cd, reg, sup, wts, reps=50):
    """The same product kernel on the workload's own operands: A = the BF16
    Gibbs kernel of the C2 cost (symmetric, so its column-major storage is the
    row-major A the kernel reads), B = X, the targets' masses on their
    supports (95% zeros, like the X of every iteration's K X product).
    Returns the average launch time in ms, from HIP events over `reps`
    back-to-back launches on the stream."""
    from rwig_amd import batched as RB
    from rwig_amd._lib import check, lib
    Kb, ldk = RB.gibbs_lowp(Cd, reg, "bf16")
    M, B = Cd.shape[0], len(sup)
    X = torch.zeros((B, M), dtype=torch.bfloat16, device=dev)
    rows = torch.as_tensor(np.concatenate([np.full(len(s), t) for t, s in enumerate(sup)]),
                           device=dev, dtype=torch.long)
    cols = torch.as_tensor(np.concatenate(sup), device=dev, dtype=torch.long)
    X[rows, cols] = torch.as_tensor(np.concatenate(wts), device=dev).to(torch.bfloat16)
    nb = ctypes.c_long()
    check(lib().rwig_gemm_workspace_bytes(0, M, B, GEMM_KSPLIT, ctypes.byref(nb)))
    wk = torch.empty(nb.value, dtype=torch.uint8, device=dev)

    def launch():
        check(lib().rwig_gemm_tn_bf16_dev(M, B, M, Kb.data_ptr(), ldk, X.data_ptr(), M, None, M,
                                          GEMM_KSPLIT, wk.data_ptr(), stream.cuda_stream))
    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def gemm_only():
    """The counter runs' workload: 5 BF16 products at the C2 shape, 5 copies of
    256 MiB (the calibration of the byte counters), then 5 FP64 products of the
    same shape (the absorption's exact marginals at 1024 targets)."""
    import torch
    from rwig_amd._lib import check, lib
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    launch, _ = c2_gemm(torch, dev, stream, product_only=False)
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(5):
        launch()
    for _ in range(5):
        dst.copy_(src)
    gM, gN, gK = 4096, 1024, 4096
    Ad = torch.rand((gM, gK), dtype=torch.float64, device=dev)
    Bd = torch.rand((gN, gK), dtype=torch.float64, device=dev)
    Dd = torch.empty((gN, gM), dtype=torch.float64, device=dev)
    nb = ctypes.c_long()
    check(lib().rwig_gemm_workspace_bytes(1, gM, gN, 1, ctypes.byref(nb)))
    wk = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    for _ in range(5):
        check(lib().rwig_gemm_tn_f64_dev(gM, gN, gK, Ad.data_ptr(), gK, Bd.data_ptr(), gK,
                                         Dd.data_ptr(), gM, 1, wk.data_ptr(), stream.cuda_stream))
    torch.cuda.synchronize()
    print("gemm-only: 5 BF16 products 4096 x 1024 x 4096, 5 copies of 256 MiB, "
          "5 FP64 products 4096 x 1024 x 4096", flush=True)


def c3_resident(A, C, w, bext, eps, L, log, reps=3):
    """C3 with its inputs resident in HBM (rwig_barycenter_log_ck_dev): C in the
    working precision's type, outputs left on the device; ms per forward +
    adjoint, timed on the host around calls that end with a stream
    synchronisation."""
    import torch

    from rwig_amd._lib import check, lib
    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    M, S = A.shape
    N = C.shape[1]
    Cf = np.asfortranarray(C).ravel(order="F")
    C64 = torch.from_numpy(Cf).to(dev)
    C32 = C64.to(torch.float32)
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).to(dev)
    wd = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    bd = torch.from_numpy(np.ascontiguousarray(bext)).to(dev)
    o = {k: torch.empty(n, dtype=torch.float64, device=dev)
         for k, n in [("F", M * S), ("G", N * S), ("b", N), ("gA", M * S), ("gw", S),
                      ("loss", 1)]}
    it, er, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    res = {}
    for name, prec in (("fp32", 1), ("fp32exp", 2), ("fp64", 0)):
        Cd = C32 if prec == 1 else C64

        def call():
            check(lib().rwig_barycenter_log_ck_dev(
                o["F"].data_ptr(), o["G"].data_ptr(), o["b"].data_ptr(), o["gA"].data_ptr(),
                o["gw"].data_ptr(), o["loss"].data_ptr(), ctypes.byref(it), ctypes.byref(er),
                Ad.data_ptr(), wd.data_ptr(), Cd.data_ptr(), bd.data_ptr(), M, N, S, eps, 1, L,
                0.0, 20, 1, prec, ctypes.byref(rs), st))
        call()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            call()
        res[f"{name}_ms_per_gradient_resident"] = (time.perf_counter() - t) / reps * 1e3
        log(f"C3 {name}, inputs resident: {res[f'{name}_ms_per_gradient_resident']:.1f} ms "
            "per forward + adjoint")
    return res


def c3_leg(log, cpu=False):
    """Barycenter with unrolled gradient (64 x 64 grid, S = 8, eps 0.004,
    L = 200, checkpoints every 20): one forward + adjoint in FP32, in the
    mixed form (FP32 exponentials, FP64 arguments) and in FP64."""
    import rwig_amd
    from rwig_amd.workloads import c3
    A, C, w, bext, eps, L = c3()
    out = {"workload": "log barycenter + checkpointed adjoint, 64x64 grid (n=4096), S=8, "
                       "eps=0.004, L=200, checkpoint every 20, seed unnormalised",
           "precisions": {"fp32": "FP32 cost, iterates and passes",
                          "fp32exp": "FP64 cost, iterates and exponent arguments, FP32 "
                                     "exponentials (within 1e-5 of FP64 on b, grad_w and grad_A "
                                     "where A >= 1e-12)",
                          "fp64": "FP64 throughout"}}
    for prec in ("fp32", "fp32exp", "fp64"):
        kw = dict(reg=eps, max_iter=L, zero_tol=0.0, with_grad=True, seed="unnormalised",
                  checkpoint_every=20, precision=prec)
        rwig_amd.barycenter_log_ck(A, C, w, bext, **kw)  # warm-up
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            rwig_amd.barycenter_log_ck(A, C, w, bext, **kw)
        out[f"{prec}_ms_per_gradient"] = (time.perf_counter() - t) / reps * 1e3
        log(f"C3 {prec}: {out[f'{prec}_ms_per_gradient']:.1f} ms per forward + adjoint "
            "(host FP64 inputs and outputs included)")
    out.update(c3_resident(A, C, w, bext, eps, L, log))
    if cpu:
        from oracle import oracle as O
        th = host_info()["threads"]

        def run(nt, LL):
            t1 = time.perf_counter()
            O.barycenter_log(A, C, w, bext, eps, withgrad=True, n_threads=nt, maxiter=LL,
                             zerotol=0.0)
            return time.perf_counter() - t1
        t1, t2 = run(th, 1), run(th, 2)
        per_it = t2 - t1
        est = (t1 - per_it) + L * per_it
        t1s = run(0, 1)
        est_s = t1s * est / t1
        out["cpu_baseline"] = {
            "value": 1.0 / est, "unit": "gradients/s", "cores": th, "kind": "port",
            "s_per_gradient": est, "s_per_gradient_n_threads_0": est_s,
            "sample": f"C++ oracle log barycenter with gradient at the full size, L = 1 and "
                      f"L = 2 with n_threads={th} ({t1:.1f} s, {t2:.1f} s): the difference is "
                      f"one iteration, extrapolated to L = {L}; n_threads=0 (serial) timed "
                      f"at L = 1 ({t1s:.1f} s) and scaled by the same ratio"}
        log(f"C3 CPU oracle: {est:.0f} s per gradient at {th} threads, {est_s:.0f} s serial "
            "(extrapolated)")
    return out


def lse_leg(torch, dev, stream, log):
    """The fused soft-min pass alone at the C3 shape (M = N = 4096, S = 8,
    eps 0.004, FP32), row and column forms, timed with HIP events on their
    stream: achieved HBM bandwidth of the algorithmic bytes (C once, F, G,
    the output) against the 8 TB/s peak, and the exp rate (M N S exps)."""
    from rwig_amd._lib import check, lib
    M = N = 4096
    S, reg = 8, 0.004
    g = torch.Generator(device="cpu").manual_seed(1003)
    C = torch.rand((N, M), generator=g).to(dev)            # column-major M x N
    CT = C.t().contiguous()                                # column-major N x M
    F = (0.01 * torch.randn((S, M), generator=g)).to(dev)  # column-major M x S
    G = (0.01 * torch.randn((S, N), generator=g)).to(dev)
    out_r = torch.empty((S, M), device=dev)
    out_c = torch.empty((S, N), device=dev)
    res = {"workload": f"soft-min pass M=N={M}, S={S}, eps={reg}, FP32 (C3 shape)",
           "bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    for which, name, ptr, ld, out in ((0, "rows", C, M, out_r), (1, "cols", CT, N, out_c)):
        nb = ctypes.c_long()
        check(lib().rwig_softmin_workspace_bytes(which, M, N, S, 1, ctypes.byref(nb)))
        wk = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        fn = lib().rwig_softmin_rows_dev if which == 0 else lib().rwig_softmin_cols_dev

        def launch():
            check(fn(ptr.data_ptr(), M, N, ld, F.data_ptr(), M, G.data_ptr(), N, S, reg,
                     out.data_ptr(), 1, wk.data_ptr(), stream.cuda_stream))
        launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 4 * (M * N + (M + N) * S + (M if which == 0 else N) * S)
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "achieved": gbs, "frac": gbs / HBM_PEAK_GBS,
                     "alg_bytes_per_launch": nbytes, "exps_per_launch": M * N * S,
                     "gexp_per_s": M * N * S / (ms * 1e-3) / 1e9}
        log(f"soft-min {name}: {ms * 1e3:.1f} us, {gbs:.0f} GB/s of C ({gbs / HBM_PEAK_GBS:.3f} "
            f"of HBM peak), {M * N * S / (ms * 1e-3) / 1e9:.0f} G exp/s")
    return res


def c4_leg(torch, dev, stream, log, cpu=False):
    """WDL training step (V = 2000, K = 16 atoms, 512 documents, eps 0.1, L = 20):
    the batched gradient over all documents alone, and the whole step
    (rwig_wdl_step_dev: softmax, gradient, Adam lr 1e-3 on Alpha and Lambda),
    FP32 and FP64, inputs resident in HBM; each also with the adjoint's history
    checkpointed every 5 iterations."""
    from rwig_amd._lib import check, lib
    from rwig_amd.workloads import c4
    C, Y, Alpha0, Lambda0, eps, L = c4()
    V, D = Y.shape
    S = Alpha0.shape[1]

    def softmax_cols(X):
        e = np.exp(X - X.max(0, keepdims=True))
        return e / e.sum(0, keepdims=True)

    K = np.exp(-C / eps)
    A = softmax_cols(Alpha0)
    Wb = softmax_cols(Lambda0)
    ldk = (V + 7) // 8 * 8
    out = {"workload": f"WDL training step (batched gradient; whole step with Adam lr 1e-3), "
                       f"V={V}, K={S} atoms, {D} documents, eps={eps}, "
                       f"L={L}"}
    def history_bytes(ck, size):
        # adjoint history of the step: UB, VB, KTUB (SL + 1 slices), KVB (SL),
        # bB, and the VB checkpoints when ck > 0
        ldn = (V + 7) // 8 * 8
        sl = ck if 0 < ck < L else L
        slices = 3 * (sl + 1) + sl + (((L - 1) // ck + 1) if 0 < ck < L else 0)
        return (slices * ldn * S * D + (sl + 1) * V * D) * size

    for prec, dt, ck in (("fp32", torch.float32, 0), ("fp32", torch.float32, 5),
                         ("fp64", torch.float64, 0), ("fp64", torch.float64, 5)):
        Kt = torch.zeros((V, ldk), dtype=dt)
        Kt[:, :V] = torch.from_numpy(K.T.copy())
        KTt = torch.zeros((V, ldk), dtype=dt)
        KTt[:, :V] = torch.from_numpy(K.copy())
        Kd, KTd = Kt.to(dev), KTt.to(dev)
        At = torch.from_numpy(A.T.copy()).to(dt).to(dev)
        Yt = torch.from_numpy(Y.T.copy()).to(dt).to(dev)
        Wt = torch.from_numpy(Wb.T.copy()).to(dt).to(dev)
        gA = torch.zeros((S, V), dtype=torch.float64, device=dev)
        gl = torch.zeros(S, dtype=torch.float64, device=dev)
        loss = torch.zeros(1, dtype=torch.float64, device=dev)
        p = 0 if prec == "fp64" else 1

        def step():
            check(lib().rwig_wdl_batch_grad_dev(Kd.data_ptr(), KTd.data_ptr(), ldk, V, S, D, L,
                                                At.data_ptr(), Yt.data_ptr(), Wt.data_ptr(),
                                                gA.data_ptr(), gl.data_ptr(), loss.data_ptr(), p,
                                                ck, stream.cuda_stream))
        step()
        torch.cuda.synchronize()
        reps = 5
        t = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) / reps * 1e3
        key = prec if ck == 0 else f"{prec}_checkpoint{ck}"
        hb = history_bytes(ck, 8 if prec == "fp64" else 4)
        out[f"{key}_ms_per_step"] = ms
        out[f"{key}_docs_per_s"] = D / ms * 1e3
        out[f"{key}_history_bytes"] = hb
        log(f"C4 {key}: {ms:.2f} ms per batched gradient ({D / ms * 1e3:.0f} documents/s), "
            f"adjoint history {hb / 2**30:.2f} GiB")
        # the whole training step: softmax of Alpha and Lambda, the batched
        # gradient, g_Lambda, Adam (lr 1e-3) on Alpha and Lambda
        Al = torch.from_numpy(Alpha0.T.copy()).to(dev)
        La = torch.from_numpy(Lambda0.T.copy()).to(dev)
        mom = [torch.zeros_like(Al), torch.zeros_like(Al), torch.zeros_like(La),
               torch.zeros_like(La)]
        nstep = [0]

        def train_step():
            nstep[0] += 1
            check(lib().rwig_wdl_step_dev(Kd.data_ptr(), KTd.data_ptr(), ldk, Yt.data_ptr(), V, D,
                                          S, L, 0, D, Al.data_ptr(), La.data_ptr(),
                                          *[m.data_ptr() for m in mom], nstep[0], 1, 1e-3, 0.01,
                                          0.9, 0.999, 1e-8, loss.data_ptr(), p, ck,
                                          stream.cuda_stream))
        train_step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            train_step()
        torch.cuda.synchronize()
        ms_t = (time.perf_counter() - t) / reps * 1e3
        out[f"{key}_train_step_ms"] = ms_t
        out[f"{key}_train_docs_per_s"] = D / ms_t * 1e3
        log(f"C4 {key}: {ms_t:.2f} ms per training step with Adam ({D / ms_t * 1e3:.0f} "
            "documents/s)")
    if cpu:
        from oracle import oracle as O
        d = 2
        t1 = time.perf_counter()
        O.wdl_batch_grad(K, A, Y[:, :d], Wb[:, :d], L)
        dt = time.perf_counter() - t1
        out["cpu_baseline"] = {
            "value": d / dt, "unit": "documents/s", "cores": 1, "kind": "port",
            "s_per_step_extrapolated": dt * D / d,
            "sample": f"C++ oracle batched gradient (WDL::_train_batch_batched restated, one "
                      f"thread as the reference's BLAS) on {d} of the {D} documents: "
                      f"{dt:.1f} s, linear in the documents"}
        log(f"C4 CPU oracle: {d / dt:.2f} documents/s ({dt * D / d:.0f} s per 512-document "
            "step, extrapolated)")
    return out


def bary_leg(log, eps=0.01, sample=4):
    """Batched barycenters at L1 marginal error 1e-6 against the plain FP64
    form: one cost (2048 points uniform in [0,1]^2, squared Euclidean), 256
    problems of S = 4 histograms, eps 0.01. The batched form is the serving
    engine with all 256 submitted at once, after one untimed batch (BF16 MFMA,
    FP64 absorption every 20, stop at L1 marginal error 1e-6, checked per
    result). The plain form is the
    reference's log barycenter in FP64 (rwig_barycenter_log, host buffers) on
    a sample of the same problems, run for the iterations the batched form
    needed on each, so that both do the same iterations per solve."""
    import rwig_amd
    from rwig_amd import workloads as W
    from rwig_amd.serve import ServeSession
    C, probs = W.bary_batch()
    n, S = C.shape[0], probs[0][0].shape[1]
    out = {"workload": f"batched barycenters: n={n} points (squared Euclidean), {len(probs)} "
                       f"problems of S={S} Dirichlet(1) histograms, eps={eps}, L1 marginal "
                       "err <= 1e-6"}
    with ServeSession(hbm_budget_bytes=64 << 30, max_group_columns=1024, absorb_every=20) as s:
        cid = s.register_cost(C)
        # warm-up: one full batch (the group grows to its capacity and records
        # its cycle graphs), then the timed batch
        for r in [s.submit_barycenter(cid, A, w, eps) for A, w in probs]:
            s.result(r)
        t = time.perf_counter()
        rids = [s.submit_barycenter(cid, A, w, eps) for A, w in probs]
        res = [s.result(r) for r in rids]
        sps = len(probs) / (time.perf_counter() - t)
    its = np.array([r["iter"] for r in res])
    ok = all(r["err"] <= 1e-6 for r in res)
    log(f"batched barycenters: {sps:.1f} solves/s, iterations min/mean/max {its.min()}/"
        f"{its.mean():.1f}/{its.max()}, all converged {ok}")
    t_plain = 0.0
    for q in range(sample):
        A, w = probs[q]
        ctrl = {"method": "log", "reg": eps, "max_iter": int(its[q]), "zero_tol": 0.0}
        t1 = time.perf_counter()
        rwig_amd.barycenter(A, C, w, barycenter_control=ctrl)
        t_plain += time.perf_counter() - t1
    plain_sps = sample / t_plain
    log(f"plain FP64 log barycenter per call: {plain_sps:.2f} solves/s over {sample} problems")
    out.update(bf16_mfma_solves_per_s=sps, all_converged=ok,
               iterations_mean=float(its.mean()), fp64_plain_per_call_solves_per_s=plain_sps,
               plain_sample=sample, speedup_vs_plain_fp64=sps / plain_sps,
               plain_note="rwig_barycenter_log per call (host buffers, C uploaded each call), "
                          "max_iter = the batched form's iterations for the same problem")
    return out


def c5_leg(log, paced_s=8.0, graphs=True):
    """Serving mix: 32 cost matrices (n log-uniform in [64, 8192]) resident,
    70% Sinkhorn / 30% barycenter (S in {2, 4, 8}), eps in {0.1, 0.01}, stop at
    L1 marginal error 1e-6. Saturated throughput (1024 requests at once), then
    Poisson arrivals at a fixed rate (C5_ARRIVAL_RATE) for paced_s seconds:
    p50/p99 latency comparable across runs."""
    from rwig_amd import workloads as W
    from rwig_amd.serve import ServeSession
    clouds = W.c5_clouds()
    sizes = [C.shape[0] for C in clouds]

    def submit(s, ids, q):
        if q["kind"] == "sinkhorn":
            return s.submit_sinkhorn(ids[q["cloud"]], q["a"], q["b"], q["eps"])
        return s.submit_barycenter(ids[q["cloud"]], q["A"], q["w"], q["eps"])

    out = {"workload": "serving mix: 32 cost matrices n in [64, 8192] (log-uniform), 70% "
                       "Sinkhorn / 30% barycenter S in {2,4,8}, eps in {0.1, 0.01}, "
                       "L1 marginal err <= 1e-6, BF16 MFMA + FP64 absorption every 20"}
    with ServeSession(hbm_budget_bytes=64 << 30, max_group_columns=512, absorb_every=20,
                      graphs=graphs) as s:
        ids = [s.register_cost(C) for C in clouds]
        warm = [s.submit_sinkhorn(ids[k], np.full(n, 1.0 / n), np.full(n, 1.0 / n), eps)
                for k, n in enumerate(sizes) for eps in (0.1, 0.01)]
        for r in warm:
            s.result(r)
        reqs = list(W.c5_requests(sizes, 1024, seed=2005))
        s.reset_stats()
        t = time.perf_counter()
        rids = [submit(s, ids, q) for q in reqs]
        res = [s.result(r) for r in rids]
        sat = len(reqs) / (time.perf_counter() - t)
        ok = all(r["err"] <= 1e-6 and r["return_status"] == 0 for r in res)
        # the stopping error recomputed on the host in FP64 for a sample of
        # Sinkhorn results (the returned potentials' plan)
        rng_chk = np.random.default_rng(5)
        small = [k for k, q in enumerate(reqs) if q["kind"] == "sinkhorn"
                 and sizes[q["cloud"]] <= 2048]
        worst = 0.0
        for k in rng_chk.choice(small, min(16, len(small)), replace=False):
            q, r = reqs[k], res[k]
            Ck = clouds[q["cloud"]]
            P = np.exp((r["f"][:, None] + r["g"][None, :] - Ck) / q["eps"])
            e = np.abs(P.sum(1) - q["a"]).sum() + np.abs(P.sum(0) - q["b"]).sum()
            worst = max(worst, e)
        ok = bool(ok and worst <= 1e-6 * 1.0001)
        st = s.stats()
        out.update(host_checked_sample=16, host_checked_max_err=float(worst))
        out.update(saturated_solves_per_s=sat, saturated_p50_ms=st["p50_ms"],
                   saturated_p99_ms=st["p99_ms"], all_converged_saturated=ok)
        log(f"C5 saturated: {sat:.0f} solves/s, p50 {st['p50_ms']:.0f} ms, "
            f"p99 {st['p99_ms']:.0f} ms, all converged {ok}")
        rate = C5_ARRIVAL_RATE
        rng = np.random.default_rng(2006)
        reqs = list(W.c5_requests(sizes, int(rate * paced_s), seed=2007))
        s.reset_stats()
        t0 = time.perf_counter()
        tnext = t0
        rids = []
        for q in reqs:
            tnext += rng.exponential(1.0 / rate)
            d = tnext - time.perf_counter()
            if d > 0:
                time.sleep(d)
            rids.append(submit(s, ids, q))
        res = [s.result(r) for r in rids]
        st = s.stats()
        ok = all(r["err"] <= 1e-6 and r["return_status"] == 0 for r in res)
        out.update(arrival_rate_per_s=rate, requests=len(reqs),
                   paced_solves_per_s=st["solves_per_s"], p50_ms=st["p50_ms"],
                   p99_ms=st["p99_ms"], mean_queue_ms=st["mean_queue_ms"],
                   peak_hbm_bytes=st["peak_bytes"], all_converged=ok)
        log(f"C5 paced at {rate:.0f}/s: {st['solves_per_s']:.0f} solves/s, "
            f"p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms, all converged {ok}")
        # unloaded latency: requests of the same mix, one at a time (each
        # submitted after the previous result), submit to completion
        reqs = list(W.c5_requests(sizes, 48, seed=2008))
        lat, by_size = [], {"n<=512": [], "512<n<=2048": [], "n>2048": []}
        for q in reqs:
            r = s.result(submit(s, ids, q))
            n = sizes[q["cloud"]]
            key = "n<=512" if n <= 512 else ("512<n<=2048" if n <= 2048 else "n>2048")
            lat.append(r["latency_ms"])
            by_size[key].append(r["latency_ms"])
        lat = np.array(lat)
        out["unloaded"] = {
            "requests": len(reqs), "p50_ms": float(np.percentile(lat, 50)),
            "p90_ms": float(np.percentile(lat, 90)), "max_ms": float(lat.max()),
            "p50_ms_by_size": {k: (float(np.median(v)) if v else None)
                               for k, v in by_size.items()}}
        log(f"C5 unloaded (one request at a time, {len(reqs)} requests): p50 "
            f"{out['unloaded']['p50_ms']:.1f} ms, p90 {out['unloaded']['p90_ms']:.1f} ms, max "
            f"{out['unloaded']['max_ms']:.1f} ms; p50 by size {out['unloaded']['p50_ms_by_size']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--targets", type=int, default=1024)
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the barycenter-gradient, WDL and serving measurements")
    ap.add_argument("--gemm-only", action="store_true",
                    help="only the dominant product (5 launches) and a 256 MiB copy (5 "
                         "launches): the workload of the HBM-traffic counter runs")
    ap.add_argument("--only", choices=("c1", "c3", "lse", "c4", "c5", "bary"),
                    help="run one secondary measurement alone and print its JSON object "
                         "(the workload of a profiler run)")
    ap.add_argument("--serve-eager", action="store_true",
                    help="serving launches its kernels one by one instead of replaying "
                         "recorded HIP graphs (same kernels and results; for kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the batched "
              "solves shard by target with no exchange between GPUs", file=sys.stderr)
        sys.exit(2)
    if args.gemm_only:
        gemm_only()
        return
    if args.only:
        import torch
        dev = torch.device("cuda:0")
        stream = torch.cuda.current_stream()
        legs = {"c1": lambda: c1_leg(log, cpu=False), "c3": lambda: c3_leg(log),
                "lse": lambda: lse_leg(torch, dev, stream, log),
                "c4": lambda: c4_leg(torch, dev, stream, log),
                "c5": lambda: c5_leg(log, graphs=not args.serve_eager),
                "bary": lambda: bary_leg(log)}
        print(json.dumps(legs[args.only]()), flush=True)
        return

    import torch

    from rwig_amd import batched as RB
    from rwig_amd import workloads
    from rwig_amd._lib import check, lib

    tol = 1e-6
    log("generating the workload")
    C, a, sup, wts, reg = workloads.c2(B=args.targets)
    M = C.shape[0]
    B = len(sup)
    nts = np.array([len(s) for s in sup])
    dev = torch.device("cuda:0")
    Cd = torch.from_numpy(C).to(dev)
    assert Cd.stride() == (1, M)
    ad = torch.from_numpy(a).to(dev)
    tb = RB.TargetBatch(sup, wts, device=dev)
    stream = torch.cuda.current_stream()

    def run(precision, Klo=None, out=None):
        return RB.sinkhorn_batched(Cd, ad, tb, reg, precision=precision, max_iter=1000,
                                   zero_tol=tol, err="l1", absorb_every=20, Klo=Klo, out=out)

    def check_out(out, name):
        torch.cuda.synchronize()
        err = out["err"].cpu().numpy()
        it = out["iter"].cpu().numpy()
        ok = bool(np.all(err <= tol))
        log(f"{name}: iterations min/mean/max {it.min()}/{it.mean():.1f}/{it.max()}, "
            f"max err {err.max():.3e}, all converged {ok}")
        return ok, it, err

    # ---------------------------------------------------------------- BF16 MFMA headline
    log("building the Gibbs kernels (BF16 and FP64, both orientations)")
    kern = RB.DenseKernel(Cd, reg)

    def run_mfma(out=None, absorb=ABSORB_MFMA):
        return RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, kernel=kern, zero_tol=tol,
                                        max_iter=1000, absorb_every=absorb, out=out)

    out = run_mfma()
    for _ in range(args.warmup):
        run_mfma(out)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record(stream)
        run_mfma(out)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([e0.elapsed_time(e1) for e0, e1 in ev])
    ok_mfma, it_mfma, _ = check_out(out, "bf16-mfma")
    value = B * args.steps / wall

    # the same path with absorptions every ABSORB_ALT iterations (fewer FP64
    # GEMMs, more iterations past the stopping point)
    out_alt = run_mfma(absorb=ABSORB_ALT)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        run_mfma(out_alt, absorb=ABSORB_ALT)
    torch.cuda.synchronize()
    alt_sps = 2 * B / (time.perf_counter() - t0)
    ok_alt, it_alt, _ = check_out(out_alt, f"bf16-mfma absorb every {ABSORB_ALT}")

    # dominant kernel: the BF16 MFMA GEMM K^T W (M x B x M); its average launch
    # time measured here with HIP events on the same stream, same shape
    gemm, gemm_flops = c2_gemm(torch, dev, stream)

    gemm()
    torch.cuda.synchronize()
    g0, g1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    g0.record(stream)
    for _ in range(reps):
        gemm()
    g1.record(stream)
    torch.cuda.synchronize()
    gemm_ms = g0.elapsed_time(g1) / reps
    achieved_tf = gemm_flops / (gemm_ms * 1e-3) / 1e12
    kx_ms = c2_gemm_workload(torch, dev, stream, Cd, reg, sup, wts)
    log(f"BF16 product 4096 x 1024 x 4096: {gemm_ms * 1e3:.1f} us on dense random operands, "
        f"{kx_ms * 1e3:.1f} us as K X on the workload's operands (X on the supports)")

    # ---------------------------------------------------------------- other forms
    def timed(precision, Klo=None, reps=1):
        o = RB.sinkhorn_batched(Cd, ad, tb, reg, precision=precision, max_iter=1000,
                                zero_tol=tol, err="l1", absorb_every=20, Klo=Klo)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            RB.sinkhorn_batched(Cd, ad, tb, reg, precision=precision, max_iter=1000,
                                zero_tol=tol, err="l1", absorb_every=20, Klo=Klo, out=o)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        ok, it, _ = check_out(o, precision + "-resident")
        return B / dt, ok, it

    bf16r_sps, okr, itr = timed("bf16", RB.gibbs_lowp(Cd, reg, "bf16"), reps=2)
    fp32_sps, ok32, it32 = timed("fp32", RB.gibbs_lowp(Cd, reg, "fp32"), reps=2)
    fp64_sps, ok64, it64 = timed("fp64", reps=1)

    # ---------------------------------------------------------------- plain FP64 per call
    log("plain FP64 per-call form on a sample of targets")
    L = lib()
    ns = min(args.plain_sample, B)
    t_plain = 0.0
    for t in range(ns):
        idx = torch.as_tensor(sup[t], device=dev, dtype=torch.long)
        Ct = Cd.index_select(1, idx).t().contiguous().t()
        bt = torch.as_tensor(wts[t], device=dev)
        P = torch.empty((M, len(sup[t])), dtype=torch.float64, device=dev)
        f = torch.empty(M, dtype=torch.float64, device=dev)
        g = torch.empty(len(sup[t]), dtype=torch.float64, device=dev)
        lo = torch.empty(1, dtype=torch.float64, device=dev)
        it = ctypes.c_int()
        er = ctypes.c_double()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        check(L.rwig_sinkhorn_log_dev(ad.data_ptr(), bt.data_ptr(), Ct.data_ptr(), M,
                                      len(sup[t]), reg, 0, 1000, tol, P.data_ptr(), None,
                                      f.data_ptr(), g.data_ptr(), lo.data_ptr(),
                                      ctypes.byref(it), ctypes.byref(er),
                                      stream.cuda_stream))
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t1
    plain_sps = ns / t_plain
    log(f"plain FP64 per call: {plain_sps:.2f} solves/s over {ns} targets")

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu and args.cpu_sample > 0:
        from oracle import oracle as O
        hinfo = host_info()
        cores = hinfo["threads"]
        nc = min(args.cpu_sample, B)
        log(f"CPU oracle (log Sinkhorn, {cores} threads) on {nc} targets")
        t1 = time.perf_counter()
        for t in range(nc):
            Ct = np.asfortranarray(C[:, sup[t]])
            O.sinkhorn_log(a, wts[t], Ct, reg, n_threads=cores, maxiter=1000, zerotol=tol)
        dt = time.perf_counter() - t1
        cpu = {"value": nc / dt, "unit": "solves/s", "cores": cores, "kind": "port",
               "host": hinfo,
               "sample": f"{nc} of the {B} targets, reference log-Sinkhorn algorithm "
                         f"(C++ oracle, n_threads={cores}, zero_tol=1e-6 on the "
                         "reference error), n=4096 x 205 reduced problems"}

    extra = {}
    if not args.no_extra:
        extra["c1_vanilla"] = c1_leg(log, cpu=not args.no_cpu)
        extra["c3_barycenter_gradient"] = c3_leg(log, cpu=not args.no_cpu)
        extra["lse_roofline"] = lse_leg(torch, dev, stream, log)
        extra["c4_wdl_step"] = c4_leg(torch, dev, stream, log, cpu=not args.no_cpu)
        extra["c5_serving"] = c5_leg(log)
        extra["barycenter_batched"] = bary_leg(log)

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(step_ms, 50)),
        "p99_ms": float(np.percentile(step_ms, 99)),
        "latency_note": "p50/p99 of the batch completion time (every target of a step "
                        "completes within it)",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {
            "workload": "batched log-stabilised Sinkhorn, one source vs many targets",
            "n": M, "targets": B, "support": int(nts.max()), "eps": reg,
            "stop": "L1 marginal error <= 1e-6 (FP64 on GPU)",
            "precision": "BF16 MFMA products (FP32 accumulation), FP64 absorption every "
                         f"{ABSORB_MFMA} iterations (FP64 MFMA marginals)",
        },
        "converged": {"bf16_mfma": ok_mfma, f"bf16_mfma_absorb{ABSORB_ALT}": ok_alt,
                      "bf16_resident": okr, "fp32_resident": ok32, "fp64_batched": ok64},
        "iterations_mean": {"bf16_mfma": float(it_mfma.mean()),
                            f"bf16_mfma_absorb{ABSORB_ALT}": float(it_alt.mean()),
                            "bf16_resident": float(itr.mean()),
                            "fp32_resident": float(it32.mean()),
                            "fp64_batched": float(it64.mean())},
        "modes_solves_per_s": {"bf16_mfma": value, f"bf16_mfma_absorb{ABSORB_ALT}": alt_sps,
                               "bf16_resident": bf16r_sps,
                               "fp32_resident": fp32_sps, "fp64_batched": fp64_sps,
                               "fp64_plain_per_call": plain_sps},
        "speedup_vs_plain_fp64": value / plain_sps,
        "speedup_vs_fp64_batched": value / fp64_sps,
        "roofline": {"bound": "mfma", "achieved": achieved_tf, "peak": MFMA_BF16_PEAK_TF,
                     "unit": "TFLOP/s", "frac": achieved_tf / MFMA_BF16_PEAK_TF,
                     "traffic": pmc_traffic(),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the same "
                                       "product, copy-calibrated "
                                       "(profiles/c2_gemm_hbm_traffic.md)",
                     "alg_bytes_per_launch": (4096 * 4096 + 1024 * 4096) * 2 + 4096 * 1024 * 4,
                     "kernel": "k_gemm_tn<bf16> (4096 x 1024 x 4096, 128 x 128 tiles, "
                               "FP32 tile into the workspace, as the MFMA path launches its "
                               "products at 1024 targets before its update kernels)",
                     "kernel_ms": gemm_ms, "alg_flops_per_launch": gemm_flops,
                     "operands": "dense Gaussian BF16 A and B",
                     "workload_kx": {
                         "kernel_ms": kx_ms, "achieved": gemm_flops / (kx_ms * 1e-3) / 1e12,
                         "frac": gemm_flops / (kx_ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TF,
                         "note": "the same kernel and shape as the solve's K X: A = the BF16 "
                                 "Gibbs kernel of this cost, B = the targets' masses on their "
                                 "supports (95% zeros)"}},
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
