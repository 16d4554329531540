"""GPU tests of the truncated SVD (rwig_tsvd), the WIG document scores
(rwig_wig_scores), the WIG composition and WDL's log-domain inference,
against the oracle (src/tsvd.cpp:7-86, R/wig.R:53-170,
src/wdl_impl.cpp:354-377).

Bars: columns of U_k diag(s_k) to 1e-11 of s_1 (the reference's full LAPACK
SVD and the device subspace iteration differ in rounding only, for leading
singular values separated from the rest). Signs: "sklearn" must agree; the
"none" sign is LAPACK's convention and the "auto" sign is decided by a
statistic that is rounding noise in the reference (src/tsvd.cpp:41 removes
every kept component), so for those two modes each column is compared up to
its sign (parity unpinned for the sign)."""
import ctypes

import numpy as np
import pytest

import rwig_amd
from oracle import oracle as O
from rwig_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _tsvd(X, k, flip):
    X = np.asfortranarray(X, dtype=np.float64)
    m, n = X.shape
    kk = min(k, m, n)
    out = np.zeros((m, kk), order="F")
    k_out, res = ctypes.c_int(), ctypes.c_double()
    dp = ctypes.POINTER(ctypes.c_double)
    print(f"rwig_tsvd m={m} n={n} k={k} flip={flip}", flush=True)
    check(lib().rwig_tsvd(out.ctypes.data_as(dp), ctypes.byref(k_out), ctypes.byref(res),
                          X.ctypes.data_as(dp), m, n, k, flip))
    assert k_out.value == kk
    return out, res.value


def _matrices():
    rng = np.random.default_rng(11)
    yield "vignette", np.array([[1.0, 3.0], [2.0, -4.0]]), 1
    yield "tall", rng.random((300, 40)), 3
    yield "wide", rng.standard_normal((12, 500)), 2
    yield "square", rng.standard_normal((64, 64)) * np.linspace(2, 1, 64), 4
    yield "k clamped", rng.random((5, 9)), 8
    yield "rank 2", rng.standard_normal((90, 2)) @ rng.standard_normal((2, 30)), 3
    yield "histograms", rng.dirichlet(np.ones(400), 150).T, 1
    yield "one row", rng.random((1, 7)), 1
    yield "zero", np.zeros((6, 4)), 2


@pytest.mark.parametrize("flip", [0, 1, 2])
def test_tsvd_matches_oracle(flip):
    for name, X, k in _matrices():
        out, res = _tsvd(X, k, flip)
        ref = O.tsvd(X, k, flip)["Msvd"]
        s1 = max(np.linalg.norm(X, 2), 1e-300)
        tol = 1e-11 * s1 + 1e-300
        assert res <= 1e-12 or s1 == 1e-300, (name, res)
        for c in range(ref.shape[1]):
            d_same = np.abs(out[:, c] - ref[:, c]).max()
            d_flip = np.abs(out[:, c] + ref[:, c]).max()
            if flip == 1:
                assert d_same <= tol, (name, c, d_same)
            else:
                assert min(d_same, d_flip) <= tol, (name, c, d_same, d_flip)


def test_tsvd_sklearn_sign_rule_holds():
    X = np.random.default_rng(5).standard_normal((200, 30))
    out, _ = _tsvd(X, 6, 1)
    for c in range(6):
        assert out[np.argmax(np.abs(out[:, c])), c] > 0


@pytest.mark.parametrize("m,n,k", [(6000, 1500, 1), (1200, 5000, 3)])
def test_tsvd_large_known_spectrum(m, n, k):
    # sizes past the oracle: X = P diag(s) Q^T with orthonormal P, Q and
    # s_i = 100 * 0.8^i, so U_k diag(s_k) = P_k diag(s_k) up to column signs
    rng = np.random.default_rng(m + n)
    r = 40
    P, _ = np.linalg.qr(rng.standard_normal((m, r)))
    Q, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = 100.0 * 0.8 ** np.arange(r)
    X = (P * s) @ Q.T
    out, res = _tsvd(X, k, 1)
    assert res <= 1e-12
    for c in range(k):
        ref = P[:, c] * s[c]
        ref = ref if ref[np.argmax(np.abs(ref))] > 0 else -ref
        assert np.abs(out[:, c] - ref).max() <= 1e-11 * s[0], c


@pytest.mark.parametrize("svd_method", ["topics", "docs"])
def test_wig_scores_match_oracle(svd_method):
    rng = np.random.default_rng(3)
    N, S, M = 300, 4, 120
    A = rng.dirichlet(np.ones(N), S).T
    W = rng.dirichlet(np.ones(S), M).T
    Yhat = rng.dirichlet(np.ones(N), M).T
    got = rwig_amd.wig_scores(A, W, Yhat, svd_method)
    ref = O.wig_scores(A, W, Yhat, svd_method)
    scale = np.abs(ref).max()
    # one rank-1 SVD: the whole score vector shares one sign ("auto" rule)
    assert min(np.abs(got - ref).max(), np.abs(got + ref).max()) <= 1e-11 * scale


def test_wdl_log_mode_inference_matches_oracle():
    # sinkhorn_mode 2 infers Yhat with every document's log-domain barycenter
    # (rwig_barycenter_log_batched)
    rng = np.random.default_rng(8)
    N, S, M = 40, 3, 10
    x = np.sort(rng.random(N))
    C = (x[:, None] - x[None, :]) ** 2
    Y = rng.dirichlet(np.ones(N), M).T
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    fit = rwig_amd.wdl(Y, C, 0.05, S, batch_size=5, epochs=1, sinkhorn_mode=2, max_iter=20,
                       init=(Alpha0, Lambda0))
    for m in range(M):
        ref = O.barycenter_log(fit["A"], C, fit["W"][:, m], reg=0.05, maxiter=20)["b"]
        assert np.abs(fit["Yhat"][:, m] - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("S,D", [(4, 9), (5, 7), (32, 2)])
def test_batched_log_barycenters_match_oracle_per_document(S, D):
    # rwig_barycenter_log_batched packs 32 // S documents per soft-min sweep;
    # each must stop at its own iteration (spectral-norm error, tol 1e-6) with
    # the oracle's barycenter_log result
    import ctypes
    from rwig_amd._lib import check, lib
    rng = np.random.default_rng(100 + S)
    M, N = 50, 45
    x, y = rng.random((M, 2)), rng.random((N, 2))
    C = np.asfortranarray(((x[:, None] - y[None]) ** 2).sum(-1))
    A = np.asfortranarray(rng.dirichlet(np.ones(M), size=S).T)
    W = np.asfortranarray(rng.dirichlet(np.ones(S), size=D).T)
    B = np.zeros((N, D), order="F")
    iters = np.zeros(D, dtype=np.int32)
    errs = np.zeros(D)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    check(lib().rwig_barycenter_log_batched(
        dp(B), iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dp(errs), dp(A), dp(W), dp(C),
        M, N, S, D, 0.05, 1000, 1e-6))
    for d in range(D):
        o = O.barycenter_log(A, C, W[:, d], reg=0.05, maxiter=1000, zerotol=1e-6)
        assert iters[d] == o["iter"], (d, iters[d], o["iter"])
        assert abs(errs[d] - o["err"]) <= 1e-6 * o["err"] + 1e-14  # a near-cancelling norm
        assert np.abs(B[:, d] - o["b"]).max() <= 1e-10 * np.abs(o["b"]).max()
    print(f"S={S} D={D}: iterations {iters.tolist()}")


def test_wig_composition():
    rng = np.random.default_rng(21)
    V, dim, D = 60, 5, 40
    words = [f"w{i}" for i in range(V)]
    E = rng.standard_normal((V, dim))
    docs = [list(rng.choice(words, size=rng.integers(3, 12))) for _ in range(D)]
    dates = np.datetime64("2015-01-01") + rng.integers(0, 400, D).astype("timedelta64[D]")
    perm = rng.permutation(D)
    Alpha0 = rng.standard_normal((V, 3))
    Lambda0 = rng.standard_normal((3, D))
    out = rwig_amd.wig(dates, docs, words, E, num_topics=3, batch_size=8, epochs=1,
                       permutation=perm, reg=0.5, init=(Alpha0, Lambda0))
    idx = out["index"]
    assert np.all(np.diff(idx["ref_date"].astype(np.int64)) > 0)
    assert np.all(np.isfinite(idx["WIG"]))
    assert abs(np.mean(idx["WIG"]) - 100) < 1e-9 and abs(np.std(idx["WIG"], ddof=1) - 1) < 1e-9
    # the same pieces composed by hand (the shuffled scores against the dates
    # in input order, as R/wig.R:91-143 pairs them)
    model = out["wdl_model"]
    scores = O.wig_scores(model["topics"], model["weights"], model["docs_pred"], "topics")
    keys, ref = O.wig_index((dates - np.datetime64("1970-01-01")).astype(np.int64), scores)
    assert np.array_equal(idx["ref_date"].astype(np.int64), keys)
    assert min(np.abs(idx["WIG"] - ref).max(), np.abs(idx["WIG"] - (200 - ref)).max()) < 1e-9
