"""Host-side mirror of rwig's tsvd() and wig() (R/tsvd.R:38-61,
R/wig.R:53-170): truncated SVD, WIG document scores and the WIG index.

The SVD products run in librwig_hip.so on the GPU (FP64 MFMA GEMMs inside a
subspace iteration); the period aggregation is host code in the same
library. Tokenisation and word2vec training (R/wdl.R:76-85, third-party R
packages) are not part of this path: wig() takes tokenised documents, a
dictionary and its embedding, the inputs wdl.character builds from them.
"""
from __future__ import annotations

import ctypes
import datetime

import numpy as np

from ._lib import check, lib
from .api import doc2dist, euclidean, wdl

_FLIP = {"auto": 0, "sklearn": 1, "none": 2}
_UNITS = {"day": 0, "week": 1, "month": 2, "bimonth": 3, "quarter": 4, "halfyear": 5,
          "year": 6, "season": 7}
_EPOCH = np.datetime64("1970-01-01", "D")


def _f64(x):
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _p(x):
    return None if x is None else x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tsvd(M, k=1, flip_sign="auto"):
    """Truncated SVD (R/tsvd.R:38-61 -> tsvd_cpp, src/tsvd.cpp:7-86):
    U_k diag(s_k), m x min(k, m, n), with the sign rule flip_sign in
    "auto" (Bro et al. 2008 as the reference computes it), "sklearn" (the
    entry of largest magnitude of each column positive) or "none"."""
    if flip_sign not in _FLIP:
        raise ValueError("'arg' should be one of \"auto\", \"sklearn\", \"none\"")
    X = _f64(M)
    if X.ndim != 2:
        raise ValueError("M must be a numeric matrix")
    m, n = X.shape
    if n < 2:
        raise ValueError("Features less than 2! Check number of columns of 'M'")
    if n <= k:
        raise ValueError("Dimension 'k' should be less than number of columns of 'M'")
    kk = min(int(k), m, n)
    out = np.zeros((m, max(kk, 1)), order="F")
    k_out = ctypes.c_int()
    check(lib().rwig_tsvd(_p(out), ctypes.byref(k_out), None, _p(X), m, n, int(k),
                          _FLIP[flip_sign]))
    return out[:, :k_out.value]


def wig_scores(A, W, Yhat=None, svd_method="topics"):
    """One score per document (R/wig.R:96-106): "topics" projects each
    document's topic weights W (S x M) on the leading singular vector of the
    topics A^T (A: N x S); "docs" takes the leading left singular vector of
    the reconstructed documents Yhat^T (Yhat: N x M), scaled by its singular
    value."""
    if svd_method not in ("topics", "docs"):
        raise ValueError("`svd_method` not implemented!")
    W = _f64(W)
    S, M = W.shape
    if svd_method == "topics":
        A = _f64(A)
        N = A.shape[0]
        if A.shape[1] != S:
            raise ValueError("A must have one column per row of W")
        Y = None
    else:
        if Yhat is None:
            raise ValueError("svd_method \"docs\" needs Yhat")
        Y = _f64(Yhat)
        N = Y.shape[0]
        if Y.shape[1] != M:
            raise ValueError("Yhat must have one column per document")
        A = None
    scores = np.zeros(M)
    check(lib().rwig_wig_scores(_p(scores), _p(A), _p(W), _p(Y), N, S, M,
                                0 if svd_method == "topics" else 1))
    return scores


def _parse_unit(unit):
    u = str(unit).strip().lower()
    if " " in u:
        num, u = u.split(None, 1)
        if num not in ("1", "1.0"):
            raise ValueError("only single-period units (e.g. \"month\") are supported")
    u = u[:-1] if u.endswith("s") and u[:-1] in _UNITS else u
    if u not in _UNITS:
        raise ValueError(f"unknown group_unit {unit!r}; one of {', '.join(_UNITS)}")
    return _UNITS[u]


def _to_days(dates):
    """Dates (datetime64, datetime.date/datetime or ISO strings) as days since
    1970-01-01, R's Date representation."""
    out = []
    for d in dates:
        if isinstance(d, (int, np.integer)):
            raise ValueError("`date_col` is not of date/datetime!")
        if isinstance(d, datetime.datetime):
            d = d.date()
        out.append(int((np.datetime64(d, "D") - _EPOCH).astype(np.int64)))
    return np.asarray(out, dtype=np.int32)


def wig_index(dates, scores, group_unit="month", standardize=True):
    """The WIG index (R/wig.R:2-6, :139-159): document scores summed per
    calendar period (lubridate::floor_date of each date), periods ascending,
    then (x - mean) / sd + 100 when standardize. Returns dict(ref_date
    (datetime64[D] period starts), WIG)."""
    return _index(_to_days(dates), scores, group_unit, standardize)


def _index(days, scores, group_unit, standardize):
    scores = np.ascontiguousarray(np.asarray(scores, dtype=np.float64).ravel())
    days = np.ascontiguousarray(days, dtype=np.int32)
    if days.size != scores.size:
        raise ValueError("one date per document score is required")
    if scores.size == 0:
        raise ValueError("no documents")
    M = scores.size
    period = np.zeros(M, dtype=np.int32)
    index = np.zeros(M)
    n = ctypes.c_int()
    check(lib().rwig_wig_index(period.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(index),
                               ctypes.byref(n), _p(scores),
                               days.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), M,
                               _parse_unit(group_unit), int(bool(standardize))))
    P = n.value
    return {"ref_date": _EPOCH + period[:P].astype("timedelta64[D]"), "WIG": index[:P].copy()}


def wig(dates, docs, dictionary, embedding, num_topics=4, batch_size=64, epochs=2,
        shuffle=True, permutation=None, rng_seed=42, reg=0.1, max_iter=20, zero_tol=1e-6,
        method="auto", threshold=0.1, optimizer="adamw", lr=0.005, decay=0.01, beta1=0.9,
        beta2=0.999, eps=1e-8, svd_method="topics", group_unit="month", standardize=True,
        init=None, precision="fp64"):
    """wig.data.frame (R/wig.R:53-170) after tokenisation and word2vec:
    docs are token lists, dictionary the embedding's row names, embedding its
    vectors (V x dim). Steps as wdl.character (R/wdl.R:91-146): cost =
    euclidean(embedding), documents = doc2dist(docs, dictionary), optional
    column shuffle, "auto" barycenter method by the threshold rule, WDL; then
    scores and index as wig.

    As in the reference, a shuffled fit returns its weights in shuffled
    document order and the scores are paired with the dates in the input
    order (R/wig.R:91-143). permutation (0-based) fixes the shuffle; without
    it the shuffle is drawn from numpy's default_rng(rng_seed) (the
    reference draws it with R's sample.int)."""
    if len(dates) != len(docs):
        raise ValueError("one date per document is required")
    days = _to_days(dates)
    C = euclidean(embedding)
    Y = doc2dist(docs, dictionary)
    if shuffle:
        perm = (np.asarray(permutation) if permutation is not None
                else np.random.default_rng(rng_seed).permutation(Y.shape[1]))
        Y = np.asfortranarray(Y[:, perm])
    if method == "auto":
        k1 = np.exp(-float(C.min()) / reg)
        k2 = np.exp(-float(C.max()) / reg)
        mode = 2 if min(k1, k2) < threshold else 1
    else:
        mode = {"parallel": 1, "log": 2}[method]
    fit = wdl(Y, C, reg, num_topics, batch_size=batch_size, epochs=epochs, sinkhorn_mode=mode,
              max_iter=max_iter, zero_tol=zero_tol, optimizer=optimizer, eta=lr, gamma=decay,
              beta1=beta1, beta2=beta2, eps=eps, seed=rng_seed, init=init, precision=precision)
    scores = wig_scores(fit["A"], fit["W"], fit["Yhat"], svd_method)
    index = _index(days, scores, group_unit, standardize)
    return {"index": index, "wdl_model": {"topics": fit["A"], "weights": fit["W"],
                                          "docs_pred": fit["Yhat"], "docs_dist": Y}}
