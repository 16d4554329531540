"""Host-side mirror of rwig's R entry points sinkhorn() and barycenter().

Same control lists, defaults, "auto" method rule, argument checks and output
element names as the reference (R/sinkhorn.R:85-177, R/barycenter.R:90-218,
R/check_args.R:108-271, src/sinkhorn.cpp:28-43, :154-169,
src/barycenter.cpp:43-59, :190-206). Every solve runs in librwig_hip.so on the
GPU; ``use_cuda=False`` (the reference's CPU path) is not offered.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._lib import check, lib

_SINKHORN_ARGS = ("reg", "with_grad", "use_cuda", "n_threads", "method", "threshold",
                  "max_iter", "zero_tol", "verbose")


def _fill_defaults(ctrl, kind):
    """check_sinkhorn_args / check_barycenter_args (R/check_args.R:108-271)."""
    ctrl = dict(ctrl or {})
    bad = [k for k in ctrl if k not in _SINKHORN_ARGS]
    if bad:
        raise ValueError(f"{', '.join(bad)} not matching one of the {kind} arguments: "
                         + ", ".join(f'"{a}"' for a in _SINKHORN_ARGS))
    if kind == "barycenter":
        ctrl.setdefault("reg", 0.1)
    ctrl.setdefault("with_grad", False)
    ctrl.setdefault("use_cuda", True)
    ctrl.setdefault("n_threads", 0)
    ctrl.setdefault("method", "auto")
    ctrl.setdefault("threshold", 0.1)
    ctrl.setdefault("max_iter", 1000)
    ctrl.setdefault("zero_tol", 1e-6)
    ctrl.setdefault("verbose", 10)
    allowed = ("auto", "vanilla", "log") if kind == "sinkhorn" else ("auto", "parallel", "log")
    if ctrl["method"] not in allowed:
        raise ValueError("method must be from: " + ", ".join(f'"{m}"' for m in allowed))
    if not ctrl["use_cuda"]:
        raise ValueError("this build runs every solve on the GPU; use_cuda=False is not offered")
    reg = ctrl.get("reg")
    if reg is None or not np.isscalar(reg):
        raise ValueError("reg must be a single value")
    return ctrl


def _auto_method(C, reg, threshold, small):
    """R/sinkhorn.R:122-142: log iff min(exp(-min C/reg), exp(-max C/reg)) < threshold."""
    k1 = math.exp(-float(np.min(C)) / reg)
    k2 = math.exp(-float(np.max(C)) / reg)
    return "log" if min(k1, k2) < threshold else small


def _f64(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _p(x):
    return None if x is None else x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sinkhorn(a, b, C, sinkhorn_control=None):
    """Entropic OT between histograms a (M) and b (N) under cost C (M x N).

    Mirrors R/sinkhorn.R:85-177. Returns a dict with P, [grad_a], u|f, v|g,
    loss, iter, err, return_status, method.
    """
    a = _f64(a).ravel()
    b = _f64(b).ravel()
    C = _f64(C)
    if C.ndim != 2:
        raise ValueError("C must be a numeric matrix!")
    if a.size != C.shape[0] or b.size != C.shape[1]:
        raise ValueError("a, b, C don't match in dimension!")
    ctrl = _fill_defaults(sinkhorn_control, "sinkhorn")
    reg = float(ctrl["reg"])
    method = ctrl["method"]
    if method == "auto":
        method = _auto_method(C, reg, ctrl["threshold"], "vanilla")
    M, N = C.shape
    wg = bool(ctrl["with_grad"])
    P = np.zeros((M, N), order="F")
    grad = np.zeros(M)
    u = np.zeros(M)
    v = np.zeros(N)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    fn = lib().rwig_sinkhorn_vanilla if method == "vanilla" else lib().rwig_sinkhorn_log
    check(fn(_p(P), _p(grad), _p(u), _p(v), ctypes.byref(loss), ctypes.byref(it),
             ctypes.byref(err), _p(a), _p(b), _p(C), M, N, reg, int(wg),
             int(ctrl["max_iter"]), float(ctrl["zero_tol"]), ctypes.byref(rc)))
    out = {"P": P}
    if wg:
        out["grad_a"] = grad
    names = ("u", "v") if method == "vanilla" else ("f", "g")
    out[names[0]] = u
    out[names[1]] = v
    out.update(loss=loss.value, iter=it.value, err=err.value, return_status=rc.value,
               method=method)
    return out


def barycenter(A, C, w, b_ext=None, barycenter_control=None):
    """Wasserstein barycenter of the columns of A (M x S) with weights w.

    Mirrors R/barycenter.R:90-218. Returns a dict with b, [grad_A, grad_w,
    loss], U|F, V|G, iter, err, return_status, method.
    """
    A = _f64(A)
    C = _f64(C)
    w = _f64(w).ravel()
    if A.ndim != 2:
        raise ValueError("A must be a numeric matrix!")
    if C.ndim != 2:
        raise ValueError("C must be a numeric matrix!")
    if A.shape[0] != C.shape[0]:
        raise ValueError("number of rows of A and C must equal!")
    ctrl = _fill_defaults(barycenter_control, "barycenter")
    reg = float(ctrl["reg"])
    wg = bool(ctrl["with_grad"])
    if wg and b_ext is None:
        raise ValueError("you must supply the `b_ext` with `with_grad = TRUE`!")
    method = ctrl["method"]
    if method == "auto":
        method = _auto_method(C, reg, ctrl["threshold"], "parallel")
    M, S = A.shape
    N = C.shape[1]
    if w.size != S:
        raise ValueError("length of w must equal the number of columns of A")
    bext = _f64(b_ext).ravel() if wg else None
    if wg and bext.size != N:
        raise ValueError("b_ext must have one entry per column of C")
    b = np.zeros(N)
    U = np.zeros((M, S), order="F")
    V = np.zeros((N, S), order="F")
    gA = np.zeros((M, S), order="F")
    gw = np.zeros(S)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    fn = lib().rwig_barycenter_parallel if method == "parallel" else lib().rwig_barycenter_log
    check(fn(_p(U), _p(V), _p(b), _p(gA), _p(gw), ctypes.byref(loss), ctypes.byref(it),
             ctypes.byref(err), _p(A), _p(w), _p(C), _p(bext), M, N, S, reg, int(wg),
             int(ctrl["max_iter"]), float(ctrl["zero_tol"]), ctypes.byref(rc)))
    out = {"b": b}
    if wg:
        out.update(grad_A=gA, grad_w=gw, loss=loss.value)
    names = ("U", "V") if method == "parallel" else ("F", "G")
    out[names[0]] = U
    out[names[1]] = V
    out.update(iter=it.value, err=err.value, return_status=rc.value, method=method)
    return out


_SEEDS = {"reference": 0, "unnormalised": 1, "normalised": 2}


def barycenter_log_ck(A, C, w, b_ext=None, reg=0.1, with_grad=False, max_iter=1000,
                      zero_tol=1e-6, checkpoint_every=20, seed="reference", precision="fp64"):
    """Log barycenter with a checkpointed adjoint.

    Same iteration and outputs as barycenter(method="log")
    (src/barycenter_impl_log.cpp:189-382). seed picks the adjoint's starting
    gradient: "reference" (2 (b - b_ext) .* b, b normalised, as the
    reference), "unnormalised" (exact gradient of ||exp(logb_L) - b_ext||^2)
    or "normalised" (exact gradient of ||b - b_ext||^2 through the
    normalisation). precision: "fp64"; "fp32" (C, iterates and soft-min
    passes in FP32); "fp32exp" (FP64 C, iterates and exponent arguments, the
    passes' exponentials in FP32).
    """
    A = _f64(A)
    C = _f64(C)
    w = _f64(w).ravel()
    if A.ndim != 2 or C.ndim != 2 or A.shape[0] != C.shape[0]:
        raise ValueError("A (M x S) and C (M x N) must be matrices with equal row counts")
    if with_grad and b_ext is None:
        raise ValueError("you must supply the `b_ext` with `with_grad = TRUE`!")
    if seed not in _SEEDS:
        raise ValueError("seed must be one of " + ", ".join(_SEEDS))
    prec = {"fp64": 0, "fp32": 1, "fp32exp": 2}[precision]
    M, S = A.shape
    N = C.shape[1]
    if w.size != S:
        raise ValueError("length of w must equal the number of columns of A")
    bext = _f64(b_ext).ravel() if with_grad else None
    b = np.zeros(N)
    F = np.zeros((M, S), order="F")
    G = np.zeros((N, S), order="F")
    gA = np.zeros((M, S), order="F")
    gw = np.zeros(S)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    check(lib().rwig_barycenter_log_ck(
        _p(F), _p(G), _p(b), _p(gA), _p(gw), ctypes.byref(loss), ctypes.byref(it),
        ctypes.byref(err), _p(A), _p(w), _p(C), _p(bext), M, N, S, float(reg), int(with_grad),
        int(max_iter), float(zero_tol), int(checkpoint_every), _SEEDS[seed], prec,
        ctypes.byref(rc)))
    out = {"b": b, "F": F, "G": G, "iter": it.value, "err": err.value,
           "return_status": rc.value}
    if with_grad:
        out.update(grad_A=gA, grad_w=gw, loss=loss.value)
    return out


def euclidean(embedding, squared=False):
    """Distance matrix between the rows of an embedding (V x d): mirrors
    euclidean() (R/utils.R:4-7, euclidean_cpp src/utils.cpp:17-37);
    squared=True gives squared distances."""
    E = _f64(embedding)
    if E.ndim != 2:
        raise ValueError("embedding must be a numeric matrix!")
    V, d = E.shape
    out = np.zeros((V, V), order="F")
    check(lib().rwig_euclidean(_p(E), V, d, _p(out), int(bool(squared))))
    return out


def doc2dist(doc_tokens, dictionary):
    """Normalised bag-of-words matrix (V x D) of tokenised documents over a
    dictionary: mirrors doc2dist() (R/utils.R:11-14, doc2dist_cpp
    src/utils.cpp:39-68), including its rule for tokens not in the
    dictionary (they count for the last entry)."""
    docs = [list(d) for d in doc_tokens]
    words = [str(w).encode() for w in dictionary]
    if not words:
        raise ValueError("the dictionary is empty")
    toks = [str(t).encode() for d in docs for t in d]
    offsets = np.zeros(len(docs) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(d) for d in docs])
    tok_arr = (ctypes.c_char_p * max(1, len(toks)))(*toks) if toks else None
    dict_arr = (ctypes.c_char_p * len(words))(*words)
    out = np.zeros((len(words), len(docs)), order="F")
    check(lib().rwig_doc2dist(tok_arr, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                              len(docs), dict_arr, len(words), _p(out)))
    return out


def cuda_available():
    """cuda_available_cpp() (src/check_cuda.cpp:6-12): is a HIP device usable."""
    return bool(lib().rwig_available())


_OPTIMIZERS = {"sgd": 0, "adam": 1, "adamw": 2}


def wdl(Y, C, reg, S, n_threads=0, batch_size=64, epochs=2, sinkhorn_mode=1, usecuda=True,
        max_iter=20, zero_tol=1e-6, optimizer=2, eta=0.001, gamma=0.01, beta1=0.9,
        beta2=0.999, eps=1e-8, verbose=False, seed=42, init=None, precision="fp64",
        checkpoint_every=0):
    """Wasserstein Dictionary Learning with the argument contract of wdl_cpp
    (src/wdl.cpp:108-170): Y (N x M) normalised documents, C (N x N) cost.

    Returns dict(A=N x S topics, W=S x M weights, Yhat=N x M barycenters).
    With init=(Alpha0, Lambda0) the explicit starting parameters replace the
    seeded draw (the reference's CPU path draws them with R's RNG);
    precision "fp32" runs the Gibbs products and recurrences in FP32;
    checkpoint_every k > 0 keeps every k-th iterate for the adjoint and
    recomputes the others (same results, less memory; with init only).
    sinkhorn_mode selects the inference of Yhat only, as in the reference
    (src/wdl_impl.cpp:354-377): 1 batches the parallel barycenters of all
    documents on the GPU; 2 batches their log-domain barycenters (FP64, GPU,
    rwig_barycenter_log_batched) with the reference's spectral-norm stopping
    rule per document."""
    if sinkhorn_mode not in (1, 2):
        raise ValueError("Sinkhorn mode not supported")
    if isinstance(optimizer, str):
        optimizer = _OPTIMIZERS[optimizer]
    if optimizer not in (0, 1, 2):
        raise ValueError("optimizer must be: 0, 1, 2!")
    Y = _f64(Y)
    C = _f64(C)
    N, M = Y.shape
    if N <= S:
        raise ValueError("Number of topics S must be smaller than the vocab size N!")
    if not usecuda:
        raise ValueError("this build runs every solve on the GPU; usecuda=False is not offered")
    A = np.zeros((N, S), order="F")
    W = np.zeros((S, M), order="F")
    Yhat = np.zeros((N, M), order="F")
    Yp = _p(Yhat) if sinkhorn_mode == 1 else None  # mode 2: inferred below
    args = (float(reg), int(max_iter), float(zero_tol), int(batch_size), int(epochs),
            int(optimizer), float(eta), float(gamma), float(beta1), float(beta2), float(eps))
    if init is None:
        if precision != "fp64":
            raise ValueError("the seeded entry point runs in FP64")
        if checkpoint_every:
            raise ValueError("checkpoint_every needs init: the seeded entry point keeps the "
                             "argument list of cuda_wdl")
        check(lib().rwig_wdl(_p(A), _p(W), Yp, _p(Y), _p(C), N, M, int(S), *args,
                             int(seed), int(bool(verbose))))
    else:
        Alpha0 = _f64(init[0])
        Lambda0 = _f64(init[1])
        if Alpha0.shape != (N, S) or Lambda0.shape != (S, M):
            raise ValueError("init must be (Alpha0 N x S, Lambda0 S x M)")
        prec = {"fp64": 0, "fp32": 1}[precision]
        check(lib().rwig_wdl_init(_p(A), _p(W), Yp, _p(Y), _p(C), N, M, int(S), *args,
                                  _p(Alpha0), _p(Lambda0), prec, int(checkpoint_every)))
    if sinkhorn_mode == 2:
        # every document's log barycenter in one call, C resident on the device
        check(lib().rwig_barycenter_log_batched(_p(Yhat), None, None, _p(A), _p(W), _p(C), N, N,
                                                int(S), M, float(reg), int(max_iter),
                                                float(zero_tol)))
    return {"A": A, "W": W, "Yhat": Yhat}
