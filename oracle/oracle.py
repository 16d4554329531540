"""ctypes front end of the FP64 CPU oracle (oracle/librwig_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU
baseline timed beside it. Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module. The product package
(``rwig_amd``) never imports it.

Every function returns a dict with the element names of the reference's Rcpp
output lists (src/sinkhorn.cpp:28-43, :154-169; src/barycenter.cpp:43-59,
:190-206; src/wdl.cpp:48-49). Matrices are numpy arrays in column-major
(Fortran) order, as R hands them to Rcpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librwig_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_spectral_norm.restype = ctypes.c_double
    return _lib


def _f64(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _ptr(x):
    if x is None:
        return None
    return x.ctypes.data_as(_dp)


def sinkhorn_vanilla(a, b, C, reg, withgrad=False, maxiter=1000, zerotol=1e-6):
    """Sinkhorn::compute_vanilla (src/sinkhorn_impl_vanilla.cpp:18-86)."""
    a, b, C = _f64(a).ravel(), _f64(b).ravel(), _f64(C)
    M, N = C.shape
    P = np.zeros((M, N), order="F")
    grad = np.zeros(M)
    u = np.zeros(M)
    v = np.zeros(N)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    lib().oracle_sinkhorn_vanilla(
        _ptr(a), _ptr(b), _ptr(C), M, N, ctypes.c_double(reg), int(withgrad),
        int(maxiter), ctypes.c_double(zerotol), _ptr(P), _ptr(grad), _ptr(u),
        _ptr(v), ctypes.byref(loss), ctypes.byref(it), ctypes.byref(err),
        ctypes.byref(rc))
    out = {"P": P}
    if withgrad:
        out["grad_a"] = grad
    out.update(u=u, v=v, loss=loss.value, iter=it.value, err=err.value,
               return_status=rc.value)
    return out


def sinkhorn_log(a, b, C, reg, withgrad=False, n_threads=0, maxiter=1000,
                 zerotol=1e-6):
    """Sinkhorn::compute_log (src/sinkhorn_impl_log.cpp:225-294)."""
    a, b, C = _f64(a).ravel(), _f64(b).ravel(), _f64(C)
    M, N = C.shape
    P = np.zeros((M, N), order="F")
    grad = np.zeros(M)
    f = np.zeros(M)
    g = np.zeros(N)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    st = lib().oracle_sinkhorn_log(
        _ptr(a), _ptr(b), _ptr(C), M, N, ctypes.c_double(reg), int(withgrad),
        int(n_threads), int(maxiter), ctypes.c_double(zerotol), _ptr(P),
        _ptr(grad), _ptr(f), _ptr(g), ctypes.byref(loss), ctypes.byref(it),
        ctypes.byref(err), ctypes.byref(rc))
    if st != 0:
        raise ValueError(f"oracle_sinkhorn_log failed with status {st}")
    out = {"P": P}
    if withgrad:
        out["grad_a"] = grad
    out.update(f=f, g=g, loss=loss.value, iter=it.value, err=err.value,
               return_status=rc.value)
    return out


def barycenter_parallel(A, C, w, b_ext=None, reg=0.1, withgrad=False,
                        maxiter=1000, zerotol=1e-6):
    """Barycenter::compute_parallel (src/barycenter_impl_parallel.cpp:19-53)."""
    A, C, w = _f64(A), _f64(C), _f64(w).ravel()
    M, S = A.shape
    N = C.shape[1]
    bext = _f64(b_ext).ravel() if withgrad else np.zeros(N)
    b = np.zeros(N)
    U = np.zeros((M, S), order="F")
    V = np.zeros((N, S), order="F")
    gA = np.zeros((M, S), order="F")
    gw = np.zeros(S)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    lib().oracle_barycenter_parallel(
        _ptr(A), _ptr(C), _ptr(w), _ptr(bext), M, N, S, ctypes.c_double(reg),
        int(withgrad), int(maxiter), ctypes.c_double(zerotol), _ptr(b), _ptr(U),
        _ptr(V), _ptr(gA), _ptr(gw), ctypes.byref(loss), ctypes.byref(it),
        ctypes.byref(err), ctypes.byref(rc))
    out = {"b": b}
    if withgrad:
        out.update(grad_A=gA, grad_w=gw, loss=loss.value)
    out.update(U=U, V=V, iter=it.value, err=err.value, return_status=rc.value)
    return out


def barycenter_log(A, C, w, b_ext=None, reg=0.1, withgrad=False, n_threads=0,
                   maxiter=1000, zerotol=1e-6):
    """Barycenter::compute_log (src/barycenter_impl_log.cpp:23-48)."""
    A, C, w = _f64(A), _f64(C), _f64(w).ravel()
    M, S = A.shape
    N = C.shape[1]
    bext = _f64(b_ext).ravel() if withgrad else np.zeros(N)
    b = np.zeros(N)
    F = np.zeros((M, S), order="F")
    G = np.zeros((N, S), order="F")
    gA = np.zeros((M, S), order="F")
    gw = np.zeros(S)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    err = ctypes.c_double()
    rc = ctypes.c_int()
    st = lib().oracle_barycenter_log(
        _ptr(A), _ptr(C), _ptr(w), _ptr(bext), M, N, S, ctypes.c_double(reg),
        int(withgrad), int(n_threads), int(maxiter), ctypes.c_double(zerotol),
        _ptr(b), _ptr(F), _ptr(G), _ptr(gA), _ptr(gw), ctypes.byref(loss),
        ctypes.byref(it), ctypes.byref(err), ctypes.byref(rc))
    if st != 0:
        raise ValueError(f"oracle_barycenter_log failed with status {st}")
    out = {"b": b}
    if withgrad:
        out.update(grad_A=gA, grad_w=gw, loss=loss.value)
    out.update(F=F, G=G, iter=it.value, err=err.value, return_status=rc.value)
    return out


def softmax_cols(X):
    """Column softmax with max shift (inst/include/wdl_impl.hpp:80-96)."""
    X = _f64(X)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    Y = np.zeros_like(X, order="F")
    lib().oracle_softmax_cols(_ptr(X), _ptr(Y), X.shape[0], X.shape[1])
    return Y


def optimizer_step(opt, theta, g, m_state, v_state, step, eta=1e-3, gamma=0.01,
                   beta1=0.9, beta2=0.999, eps=1e-8):
    """One step of inst/include/optimizer.hpp:72-99, in place on the arrays.

    Returns the new step counter."""
    st = ctypes.c_int(step)
    lib().oracle_optimizer_step(
        int(opt), _ptr(theta), _ptr(_f64(g)), theta.size, _ptr(m_state),
        _ptr(v_state), ctypes.byref(st), ctypes.c_double(eta),
        ctypes.c_double(gamma), ctypes.c_double(beta1), ctypes.c_double(beta2),
        ctypes.c_double(eps))
    return st.value


def wdl_batch_grad(K, A, Yb, Wb, L):
    """WDL::_train_batch_batched gradient (src/wdl_impl.cpp:87-247).

    Returns (g_Alpha N x S, g_lambda S, bB N x D unnormalised barycenters)."""
    K, A, Yb, Wb = _f64(K), _f64(A), _f64(Yb), _f64(Wb)
    N, S = A.shape
    D = Yb.shape[1]
    gA = np.zeros((N, S), order="F")
    gl = np.zeros(S)
    bB = np.zeros((N, D), order="F")
    lib().oracle_wdl_batch_grad(_ptr(K), _ptr(A), _ptr(Yb), _ptr(Wb), N, S, D,
                                int(L), _ptr(gA), _ptr(gl), _ptr(bB))
    return gA, gl, bB


def wdl(Y, C, reg, S, Alpha0, Lambda0, batch_size, epochs, sinkhorn_mode=1,
        n_threads=0, max_iter=20, zero_tol=1e-6, optimizer=2, eta=1e-3,
        gamma=0.01, beta1=0.9, beta2=0.999, eps=1e-8):
    """WassersteinDictionaryLearning::_compute_serial (src/wdl_impl.cpp:254-378)
    with explicit initial Alpha0 (N x S) and Lambda0 (S x M)."""
    Y, C = _f64(Y), _f64(C)
    N, M = Y.shape
    Alpha0, Lambda0 = _f64(Alpha0), _f64(Lambda0)
    A = np.zeros((N, S), order="F")
    W = np.zeros((S, M), order="F")
    Yhat = np.zeros((N, M), order="F")
    st = lib().oracle_wdl(
        _ptr(Y), _ptr(C), N, M, S, ctypes.c_double(reg), int(batch_size),
        int(epochs), int(sinkhorn_mode), int(n_threads), int(max_iter),
        ctypes.c_double(zero_tol), int(optimizer), ctypes.c_double(eta),
        ctypes.c_double(gamma), ctypes.c_double(beta1), ctypes.c_double(beta2),
        ctypes.c_double(eps), _ptr(Alpha0), _ptr(Lambda0), _ptr(A), _ptr(W),
        _ptr(Yhat))
    if st != 0:
        raise ValueError(f"oracle_wdl failed with status {st}")
    return {"A": A, "W": W, "Yhat": Yhat}


def wdl_init_mt19937(seed, N, S, M):
    """Initial (Alpha0 N x S, Lambda0 S x M) of the reference's CUDA WDL path
    (src/cuda_wdl.cu:592-593, :677-689): std::mt19937(seed) standard normals,
    Alpha first, column-major."""
    Alpha0 = np.zeros((N, S), order="F")
    Lambda0 = np.zeros((S, M), order="F")
    lib().oracle_wdl_init_mt19937(int(seed), int(N), int(S), int(M), _ptr(Alpha0),
                                  _ptr(Lambda0))
    return Alpha0, Lambda0


def spectral_norm(X):
    X = _f64(X)
    if X.ndim == 1:
        X = X.reshape(-1, 1, order="F")
    return lib().oracle_spectral_norm(_ptr(X), X.shape[0], X.shape[1])


def euclidean(E):
    """euclidean_cpp (src/utils.cpp:17-37): D(i, j) = sqrt(sum_k (E(i,k) -
    E(j,k))^2) for i < j, mirrored, zero diagonal. Pure Python loops over the
    pairs (small inputs only)."""
    E = np.asarray(E, dtype=np.float64)
    V = E.shape[0]
    D = np.zeros((V, V))
    for i in range(V):
        for j in range(i + 1, V):
            c = float(np.sqrt(np.sum((E[i] - E[j]) ** 2)))
            D[i, j] = c
            D[j, i] = c
    return D


def doc2dist(docs, dictionary):
    """doc2dist_cpp (src/utils.cpp:39-68): token -> index of the first equal
    dictionary entry (std::find), the last index when absent (:53-57); counts
    per document, each column divided by its sum (:61-63)."""
    V = len(dictionary)
    out = np.zeros((V, len(docs)))
    for j, doc in enumerate(docs):
        for tok in doc:
            idx = dictionary.index(tok) if tok in dictionary else V
            idx = V - 1 if idx == V else idx
            out[idx, j] += 1.0
        with np.errstate(invalid="ignore", divide="ignore"):
            out[:, j] /= out[:, j].sum()
    return out


def tsvd(M, k=1, flip_sign=0):
    """tsvd_cpp (src/tsvd.cpp:7-86) on LAPACK's full SVD (numpy.linalg.svd,
    gesdd: the divide-and-conquer routine Armadillo's svd() calls).

    Returns dict(Msvd=U_k diag(s_k) (m x kk), sk_left, sk_right) where the
    two lists hold the auto-mode statistics of :46-58 (zeros otherwise).
    The auto statistic is taken on Y = M - U_k S_k V_k^T with all kk
    components removed (:41), so u_K^T Y and Y v_K vanish up to rounding and
    its sign is a property of the arithmetic, not of M."""
    X = np.asarray(M, dtype=np.float64)
    U, s, Vt = np.linalg.svd(X, full_matrices=True)
    V = Vt.T.copy()
    kk = min(int(k), s.size)                                   # :37
    if kk < 1:
        raise ValueError("k must be at least 1")
    sk_l, sk_r = [0.0] * kk, [0.0] * kk
    if flip_sign == 0:
        Y = X - U[:, :kk] @ np.diag(s[:kk]) @ V[:, :kk].T      # :41
        for K in range(kk):
            sk_left = 0.0
            for j in range(Y.shape[1]):                        # :48-52
                uTy = float(np.sum(U[:, K] * Y[:, j]))
                sk_left += np.sign(uTy) * uTy ** 2
            sk_right = 0.0
            for i in range(Y.shape[0]):                        # :54-58
                vTy = float(np.sum(V[:, K] * Y[i, :]))
                sk_right += np.sign(vTy) * vTy ** 2
            if sk_left * sk_right < 0:                         # :60-66
                if abs(sk_left) < abs(sk_right):
                    sk_left = -sk_left
                else:
                    sk_right = -sk_right
            U[:, K] *= np.sign(sk_left)                        # :68-69
            V[:, K] *= np.sign(sk_right)
            sk_l[K], sk_r[K] = sk_left, sk_right
        Msvd = U[:, :kk] * s[:kk]                              # :71
    elif flip_sign == 1:
        Msvd = U[:, :kk] * s[:kk]                              # :74-80
        for c in range(kk):
            i = int(np.argmax(np.abs(Msvd[:, c])))
            if Msvd[i, c] < 0:
                Msvd[:, c] *= -1
    else:
        Msvd = U[:, :kk] * s[:kk]                              # :82
    return {"Msvd": Msvd, "sk_left": sk_l, "sk_right": sk_r}


def wig_scores(A, W, Yhat, svd_method="topics"):
    """Document scores of wig.data.frame (R/wig.R:96-106): "topics" gives
    c(t(tsvd(t(A), k = 1)) %*% W), "docs" gives c(tsvd(t(Yhat), k = 1)),
    both with tsvd's default "auto" sign rule."""
    if svd_method == "topics":
        t = tsvd(np.asarray(A, dtype=np.float64).T, 1, 0)["Msvd"].T   # 1 x S
        return (t @ np.asarray(W, dtype=np.float64)).ravel()
    if svd_method == "docs":
        return tsvd(np.asarray(Yhat, dtype=np.float64).T, 1, 0)["Msvd"].ravel()
    raise ValueError("`svd_method` not implemented!")


_UNITS = ("day", "week", "month", "bimonth", "quarter", "halfyear", "year", "season")


def floor_period(day, unit):
    """lubridate::floor_date of a Date (days since 1970-01-01) to the first
    day of its period (R/wig.R:3), with Python's calendar (datetime.date):
    weeks start on Sunday (lubridate's week_start = 7), seasons in December,
    March, June and September."""
    import datetime
    epoch = datetime.date(1970, 1, 1)
    dt = epoch + datetime.timedelta(days=int(day))
    if unit == "day":
        return int(day)
    if unit == "week":
        return int(day) - dt.isoweekday() % 7
    if unit == "season":
        y, m = dt.year, dt.month
        start = datetime.date(y - (m <= 2), 12, 1) if m in (12, 1, 2) else \
            datetime.date(y, 3 * (m // 3), 1)
        return (start - epoch).days
    span = {"month": 1, "bimonth": 2, "quarter": 3, "halfyear": 6, "year": 12}[unit]
    start = datetime.date(dt.year, (dt.month - 1) // span * span + 1, 1)
    return (start - epoch).days


def wig_index(days, scores, group_unit="month", standardize=True):
    """aggregate_by_period + standardisation of wig.data.frame (R/wig.R:2-6,
    :139-159): scores summed per period in document order (R's sum() adds in
    extended precision: numpy.longdouble here), periods ascending, then
    (x - mean) / sd + 100 with R's two-pass mean and n - 1 sd (NaN for one
    period). Returns (period start days, index)."""
    sums = {}
    for d, v in zip(days, scores):
        key = floor_period(d, group_unit)
        sums[key] = sums.get(key, np.longdouble(0.0)) + np.longdouble(float(v))
    keys = sorted(sums)
    x = np.array([float(sums[k]) for k in keys])
    if standardize:
        n = len(x)
        s = np.longdouble(0.0)
        for v in x:
            s += np.longdouble(v)
        s /= n
        if np.isfinite(float(s)):
            t = np.longdouble(0.0)
            for v in x:
                t += np.longdouble(v) - s
            s += t / n
        mean = float(s)
        if n > 1:
            q = np.longdouble(0.0)
            for v in x:
                dv = float(v) - mean
                q += np.longdouble(dv * dv)
            sd = float(np.sqrt(float(q / (n - 1))))
        else:
            sd = float("nan")
        x = (x - mean) / sd + 100.0
    return np.array(keys, dtype=np.int64), x
