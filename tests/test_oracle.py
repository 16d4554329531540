"""Pins the FP64 CPU oracle (oracle/rwig_oracle.cpp) before it is trusted.

Checks, per oracle function:
  * the reference's golden vectors (inst/tinytest/test_sinkhorn.R,
    inst/tinytest/test_barycenter.R) at the reference's own tolerance;
  * closed forms (C = 0, M = 1, S = 1);
  * invariants (P = diag(u) K diag(v), vanilla/log iterates coincide,
    grad_a = f + c at convergence, threaded == serial);
  * central finite differences of the fixed-L unrolled maps for every
    gradient (vanilla, log, parallel barycenter, WDL batch step);
  * optimizer and softmax closed forms, spectral norm against LAPACK.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import mean_rel

TOL = 10 * np.sqrt(np.finfo(float).eps)


def _sk(golden):
    g = golden["sinkhorn"]
    return np.array(g["a"]), np.array(g["b"]), np.array(g["C"]), g["reg"], g


def _bc(golden):
    g = golden["barycenter"]
    return np.array(g["A"]), np.array(g["C"]), np.array(g["w"]), np.array(g["b_ext"]), g["reg"], g


# ---------------------------------------------------------------- golden


def test_golden_sinkhorn_vanilla(golden):
    a, b, C, reg, g = _sk(golden)
    r = O.sinkhorn_vanilla(a, b, C, reg, withgrad=True)
    for k in ["loss", "P", "u", "v", "grad_a"]:
        assert mean_rel(g["vanilla"][k], r[k]) < TOL, k
    assert r["return_status"] == 0


@pytest.mark.parametrize("n_threads", [0, 4])
def test_golden_sinkhorn_log(golden, n_threads):
    a, b, C, reg, g = _sk(golden)
    r = O.sinkhorn_log(a, b, C, reg, withgrad=True, n_threads=n_threads)
    for k in ["loss", "P", "f", "g", "grad_a"]:
        assert mean_rel(g["log"][k], r[k]) < TOL, k


def test_golden_barycenter_parallel(golden):
    A, C, w, bext, reg, g = _bc(golden)
    r = O.barycenter_parallel(A, C, w, bext, reg, withgrad=True)
    for k in ["loss", "U", "V", "grad_A", "grad_w", "b"]:
        assert mean_rel(g["parallel"][k], r[k]) < TOL, k


@pytest.mark.parametrize("n_threads", [0, 4])
def test_golden_barycenter_log(golden, n_threads):
    A, C, w, bext, reg, g = _bc(golden)
    r = O.barycenter_log(A, C, w, bext, reg, withgrad=True, n_threads=n_threads)
    for k in ["loss", "F", "G", "grad_A", "grad_w", "b"]:
        assert mean_rel(g["log"][k], r[k]) < TOL, k


# ---------------------------------------------------------------- closed forms


def _rand_simplex(rng, n):
    x = rng.random(n) + 0.1
    return x / x.sum()


@pytest.mark.parametrize("mode", ["vanilla", "log"])
def test_zero_cost_gives_independent_coupling(mode):
    # C = 0: K = 1 1^T, one iteration gives u = a / N, v = b / sum(u) ... and
    # P = a b^T exactly with err = 0.
    rng = np.random.default_rng(1)
    a, b = _rand_simplex(rng, 7), _rand_simplex(rng, 5)
    C = np.zeros((7, 5))
    fn = O.sinkhorn_vanilla if mode == "vanilla" else O.sinkhorn_log
    r = fn(a, b, C, 0.3)
    assert r["iter"] == 1
    np.testing.assert_allclose(r["P"], np.outer(a, b), rtol=1e-13, atol=0)


@pytest.mark.parametrize("mode", ["vanilla", "log"])
def test_single_source_point(mode):
    # M = 1: the only coupling with marginals (1, b) is P = b^T.
    rng = np.random.default_rng(2)
    b = _rand_simplex(rng, 6)
    C = rng.random((1, 6))
    fn = O.sinkhorn_vanilla if mode == "vanilla" else O.sinkhorn_log
    r = fn(np.array([1.0]), b, C, 0.2)
    np.testing.assert_allclose(r["P"][0], b, rtol=1e-12)


def test_barycenter_single_measure_one_step():
    # S = 1, w = (1): b = K^T (A / K 1) after the first iteration, and that
    # is already the fixed point (U * KV = A exactly on the next pass).
    rng = np.random.default_rng(3)
    A = _rand_simplex(rng, 6).reshape(-1, 1)
    C = rng.random((6, 6))
    reg = 0.5
    K = np.exp(-C / reg)
    r = O.barycenter_parallel(A, C, [1.0], reg=reg, maxiter=1)
    b = K.T @ (A[:, 0] / K.sum(axis=1))
    np.testing.assert_allclose(r["b"], b / b.sum(), rtol=1e-13)


# ---------------------------------------------------------------- invariants


def _problem(rng, M, N, zeros=0):
    a = _rand_simplex(rng, M)
    b = _rand_simplex(rng, N)
    if zeros:
        a[rng.choice(M, zeros, replace=False)] = 0
        b[rng.choice(N, zeros, replace=False)] = 0
        a /= a.sum()
        b /= b.sum()
    x = rng.random((M, 2))
    y = rng.random((N, 2))
    C = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return a, b, C


def test_plan_is_diag_u_K_diag_v_and_support_scatter():
    rng = np.random.default_rng(4)
    a, b, C = _problem(rng, 12, 9, zeros=2)
    reg = 0.2
    r = O.sinkhorn_vanilla(a, b, C, reg)
    K = np.exp(-C / reg)
    np.testing.assert_allclose(r["P"], r["u"][:, None] * K * r["v"][None, :], rtol=1e-14)
    assert np.all(r["P"][a == 0] == 0) and np.all(r["P"][:, b == 0] == 0)
    assert np.all(r["u"][a == 0] == 0) and np.all(r["v"][b == 0] == 0)
    rl = O.sinkhorn_log(a, b, C, reg)
    assert np.all(np.isneginf(rl["f"][a == 0])) and np.all(np.isneginf(rl["g"][b == 0]))


def test_vanilla_and_log_iterates_coincide():
    rng = np.random.default_rng(5)
    a, b, C = _problem(rng, 10, 8)
    reg = 0.3
    for L in [1, 3, 17]:
        rv = O.sinkhorn_vanilla(a, b, C, reg, maxiter=L, zerotol=0)
        rl = O.sinkhorn_log(a, b, C, reg, maxiter=L, zerotol=0)
        np.testing.assert_allclose(reg * np.log(rv["u"]), rl["f"], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(reg * np.log(rv["v"]), rl["g"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("mode", ["vanilla", "log"])
def test_grad_is_dual_potential_at_convergence(mode):
    # at convergence the gradient of the regularised cost wrt a is f + c 1
    rng = np.random.default_rng(6)
    a, b, C = _problem(rng, 9, 7)
    reg = 0.25
    fn = O.sinkhorn_vanilla if mode == "vanilla" else O.sinkhorn_log
    r = fn(a, b, C, reg, withgrad=True, zerotol=1e-13)
    f = reg * np.log(r["u"]) if mode == "vanilla" else r["f"]
    d = r["grad_a"] - f
    assert np.ptp(d) < 1e-9


def test_threaded_log_equals_serial():
    rng = np.random.default_rng(7)
    a, b, C = _problem(rng, 33, 21)
    r0 = O.sinkhorn_log(a, b, C, 0.05, withgrad=True, n_threads=0)
    r3 = O.sinkhorn_log(a, b, C, 0.05, withgrad=True, n_threads=3)
    for k in ["P", "f", "g", "grad_a"]:
        np.testing.assert_array_equal(r0[k], r3[k])
    assert r0["iter"] == r3["iter"]


# ---------------------------------------------------------------- finite differences


def _fd(fun, x, h=1e-6):
    # C-ordered work arrays: reshape(-1) of a Fortran-ordered array is a copy,
    # and writes through it would be lost
    g = np.zeros(x.shape)
    flat = np.ascontiguousarray(x).reshape(-1)
    for k in range(flat.size):
        xp = flat.copy(); xp[k] += h
        xm = flat.copy(); xm[k] -= h
        g.reshape(-1)[k] = (fun(xp.reshape(x.shape)) - fun(xm.reshape(x.shape))) / (2 * h)
    return g


@pytest.mark.parametrize("mode", ["vanilla", "log"])
def test_sinkhorn_grad_matches_fd_of_fixed_L_map(mode):
    rng = np.random.default_rng(8)
    a, b, C = _problem(rng, 6, 5)
    reg, L = 0.2, 7
    fn = O.sinkhorn_vanilla if mode == "vanilla" else O.sinkhorn_log
    g = fn(a, b, C, reg, withgrad=True, maxiter=L, zerotol=0)["grad_a"]
    fd = _fd(lambda x: fn(x, b, C, reg, maxiter=L, zerotol=0)["loss"], a)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)


def test_parallel_barycenter_grad_matches_fd_of_fixed_L_map():
    rng = np.random.default_rng(9)
    M, N, S = 6, 5, 3
    A = np.stack([_rand_simplex(rng, M) for _ in range(S)], 1)
    x = rng.random((M, 2)); y = rng.random((N, 2))
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    w = _rand_simplex(rng, S)
    bext = _rand_simplex(rng, N)
    reg, L = 0.3, 6
    K = np.exp(-C / reg)

    def loss(A_, w_):
        r = O.barycenter_parallel(A_, C, w_, bext, reg, maxiter=L, zerotol=0)
        b_unnorm = r["V"][:, 0] * (K.T @ r["U"])[:, 0]
        return np.sum((b_unnorm - bext) ** 2)

    r = O.barycenter_parallel(A, C, w, bext, reg, withgrad=True, maxiter=L, zerotol=0)
    np.testing.assert_allclose(r["grad_A"], _fd(lambda X: loss(X, w), A), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(r["grad_w"], _fd(lambda v: loss(A, v), w), rtol=1e-6, atol=1e-9)


def test_log_barycenter_grad_matches_fd_at_convergence():
    # the log adjoint seeds logb_bar with the normalised b, i.e. it is the
    # gradient of ||exp(logb_L) - b_ext||^2 only where sum(exp(logb_L)) = 1;
    # freeze L at the converged count and compare there.
    rng = np.random.default_rng(10)
    M, N, S = 6, 6, 2
    A = np.stack([_rand_simplex(rng, M) for _ in range(S)], 1)
    x = rng.random((M, 2))
    C = ((x[:, None] - x[None]) ** 2).sum(-1)
    w = _rand_simplex(rng, S)
    bext = _rand_simplex(rng, N)
    reg = 0.2
    L = O.barycenter_log(A, C, w, bext, reg, maxiter=5000, zerotol=1e-13)["iter"]

    def loss(A_, w_):
        # unnormalised b_L: the column soft-min of R(F_L, G_L) is -reg*logb_L
        r = O.barycenter_log(A_, C, w_, bext, reg, maxiter=L, zerotol=0)
        Z = -(C - r["F"][:, :1] - r["G"][:, 0][None, :]) / reg
        mx = Z.max(0)
        logb = mx + np.log(np.exp(Z - mx).sum(0))
        return np.sum((np.exp(logb) - bext) ** 2)

    r = O.barycenter_log(A, C, w, bext, reg, withgrad=True, maxiter=L, zerotol=0)
    np.testing.assert_allclose(r["grad_A"], _fd(lambda X: loss(X, w), A), rtol=1e-5, atol=1e-8)
    # grad_w is the derivative along the unnormalised weights
    np.testing.assert_allclose(r["grad_w"], _fd(lambda v: loss(A, v), w), rtol=1e-5, atol=1e-8)


def test_wdl_batch_grad_matches_fd():
    rng = np.random.default_rng(11)
    N, S, D, L, reg = 7, 3, 4, 5, 0.4
    x = rng.random((N, 2))
    C = ((x[:, None] - x[None]) ** 2).sum(-1)
    K = np.exp(-C / reg)
    Alpha = rng.standard_normal((N, S))
    Lam = rng.standard_normal((S, D))
    Y = np.stack([_rand_simplex(rng, N) for _ in range(D)], 1)

    def batch_loss(Alpha_, Lam_):
        A = O.softmax_cols(Alpha_)
        W = O.softmax_cols(Lam_)
        _, _, bB = O.wdl_batch_grad(K, A, Y, W, L)
        return np.sum((bB - Y) ** 2) / D

    gA, gl, _ = O.wdl_batch_grad(K, O.softmax_cols(Alpha), Y, O.softmax_cols(Lam), L)
    np.testing.assert_allclose(gA, _fd(lambda X: batch_loss(X, Lam), Alpha), rtol=1e-6, atol=1e-9)
    # g_lambda = mean over documents of d loss_d / d lambda_d: the derivative
    # along a shift of row s of every column of Lambda at once
    fd = np.zeros(S)
    for s in range(S):
        e = np.zeros((S, D)); e[s] = 1e-6
        fd[s] = (batch_loss(Alpha, Lam + e) - batch_loss(Alpha, Lam - e)) / 2e-6
    np.testing.assert_allclose(gl, fd, rtol=1e-6, atol=1e-9)


# ---------------------------------------------------------------- WDL pieces


def test_softmax_cols():
    rng = np.random.default_rng(12)
    X = rng.standard_normal((9, 4)) * 20
    Y = O.softmax_cols(X)
    E = np.exp(X - X.max(0))
    np.testing.assert_allclose(Y, E / E.sum(0), rtol=1e-14)
    np.testing.assert_allclose(O.softmax_cols(X + 3.0), Y, rtol=1e-13)


@pytest.mark.parametrize("opt", [0, 1, 2])
def test_optimizer_constant_gradient_closed_form(opt):
    # constant g: bias-corrected moments are exactly g and g^2, so every
    # Adam step moves theta by eta * g / (|g| + eps)
    rng = np.random.default_rng(13)
    theta0 = rng.standard_normal(5)
    g = rng.standard_normal(5)
    eta, gamma, eps, T = 1e-2, 0.1, 1e-8, 6
    th = theta0.copy(); m = np.zeros(5); v = np.zeros(5); step = 0
    for _ in range(T):
        step = O.optimizer_step(opt, th, g, m, v, step, eta=eta, gamma=gamma, eps=eps)
    assert step == T
    d = g if opt == 0 else g / (np.abs(g) + eps)
    exp = theta0.copy()
    for _ in range(T):
        exp = ((1 - eta * gamma) * exp if opt == 2 else exp) - eta * d
    np.testing.assert_allclose(th, exp, rtol=1e-10, atol=1e-12)


def test_spectral_norm_matches_lapack():
    rng = np.random.default_rng(14)
    for s in [1, 2, 5, 8]:
        X = rng.standard_normal((40, s))
        assert abs(O.spectral_norm(X) - np.linalg.norm(X, 2)) < 1e-12 * np.linalg.norm(X, 2)


def test_wdl_end_to_end_runs_and_is_normalised():
    rng = np.random.default_rng(15)
    N, M, S = 8, 6, 2
    x = rng.random((N, 2))
    C = ((x[:, None] - x[None]) ** 2).sum(-1)
    Y = np.stack([_rand_simplex(rng, N) for _ in range(M)], 1)
    out = O.wdl(Y, C, 0.3, S, rng.standard_normal((N, S)), rng.standard_normal((S, M)),
                batch_size=4, epochs=2, max_iter=10)
    np.testing.assert_allclose(out["A"].sum(0), 1, rtol=1e-13)
    np.testing.assert_allclose(out["W"].sum(0), 1, rtol=1e-13)
    np.testing.assert_allclose(out["Yhat"].sum(0), 1, rtol=1e-12)


# ------------------------------------------------- the WDL loop, pinned piece by piece
#
# The loop of _compute_serial (src/wdl_impl.cpp:254-347) is checked against
# pieces pinned elsewhere: the batch loss comes from the golden-pinned
# parallel barycenter (unnormalised b_L = V_L .* K^T U_L after exactly L
# iterations at zero_tol 0), its gradient from central differences, and each
# step from the closed-form-pinned optimizer. A wrong batch slice, a wrong
# batch size for the last batch, a g_Lambda broadcast over the batch's columns
# only, or a mean over B instead of D fails these.


def _wdl_problem(seed, N, M):
    rng = np.random.default_rng(seed)
    x = rng.random((N, 2))
    C = ((x[:, None] - x[None]) ** 2).sum(-1)
    Y = np.stack([_rand_simplex(rng, N) for _ in range(M)], 1)
    return C, Y, rng


def _wdl_batch_loss(C, reg, L, Yb, Alpha, Lam_b):
    """mean_d ||b_L,d - y_d||^2 with b_L,d the unnormalised barycenter after
    L parallel-barycenter iterations (the loss whose gradient
    src/wdl_impl.cpp:87-247 returns)."""
    A = O.softmax_cols(Alpha)
    W = O.softmax_cols(Lam_b)
    K = np.exp(-C / reg)
    tot = 0.0
    for d in range(Yb.shape[1]):
        r = O.barycenter_parallel(A, C, W[:, d], reg=reg, maxiter=L, zerotol=0.0)
        assert r["iter"] == L
        bL = r["V"][:, 0] * (K.T @ r["U"][:, 0])
        tot += np.sum((bL - Yb[:, d]) ** 2)
    return tot / Yb.shape[1]


def _wdl_fd_grads(C, reg, L, Yb, Alpha, Lam_b, h=1e-6):
    """Central differences: d/dAlpha, and d/dlambda along a shift of row s of
    every batch column of Lambda at once (the reference's g_lambda)."""
    f = lambda X: _wdl_batch_loss(C, reg, L, Yb, X, Lam_b)  # noqa: E731
    gA = _fd(f, Alpha, h)
    S = Lam_b.shape[0]
    gl = np.zeros(S)
    for s in range(S):
        e = np.zeros_like(Lam_b)
        e[s] = h
        gl[s] = (_wdl_batch_loss(C, reg, L, Yb, Alpha, Lam_b + e)
                 - _wdl_batch_loss(C, reg, L, Yb, Alpha, Lam_b - e)) / (2 * h)
    return gA, gl


def _wdl_by_pieces(C, Y, reg, S, Alpha0, Lambda0, B, E, L, opt, eta):
    """The epoch x batch loop composed from the pinned pieces."""
    N, M = Y.shape
    Alpha, Lam = Alpha0.copy(), Lambda0.copy()
    mA, vA, mL, vL = (np.zeros(N * S), np.zeros(N * S), np.zeros(S * M), np.zeros(S * M))
    sA = sL = 0
    batches = -(-M // B)
    sizes = []
    for _ in range(E):
        for bid in range(batches):
            lo, hi = bid * B, min(M, bid * B + B)
            sizes.append(hi - lo)
            gA, gl = _wdl_fd_grads(C, reg, L, Y[:, lo:hi], Alpha, Lam[:, lo:hi])
            a = Alpha.ravel(order="F").copy()
            sA = O.optimizer_step(opt, a, gA.ravel(order="F"), mA, vA, sA, eta=eta)
            Alpha = a.reshape((N, S), order="F")
            lam = Lam.ravel(order="F").copy()
            gL = np.repeat(gl[:, None], M, axis=1)  # g_Lambda = g_lambda 1_M^T
            sL = O.optimizer_step(opt, lam, gL.ravel(order="F"), mL, vL, sL, eta=eta)
            Lam = lam.reshape((S, M), order="F")
    return Alpha, Lam, sizes


@pytest.mark.parametrize("opt,eta", [(0, 0.5), (1, 0.05)])
def test_wdl_one_batch_one_epoch_is_one_step_on_the_fd_gradient(opt, eta):
    N, M, S, L, reg = 7, 4, 3, 5, 0.4
    C, Y, rng = _wdl_problem(21, N, M)
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    gA, gl = _wdl_fd_grads(C, reg, L, Y, Alpha0, Lambda0)
    assert np.min(np.abs(gA)) > 1e-6 and np.min(np.abs(gl)) > 1e-6  # Adam's step is resolved
    step = (lambda g: g) if opt == 0 else (lambda g: g / (np.abs(g) + 1e-8))
    A1 = O.softmax_cols(Alpha0 - eta * step(gA))
    W1 = O.softmax_cols(Lambda0 - eta * step(gl)[:, None])
    out = O.wdl(Y, C, reg, S, Alpha0, Lambda0, batch_size=M, epochs=1, max_iter=L,
                zero_tol=1e-9, optimizer=opt, eta=eta)
    np.testing.assert_allclose(out["A"], A1, rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(out["W"], W1, rtol=1e-7, atol=1e-12)
    # inference: each document's normalised barycenter of the trained atoms
    for m in range(M):
        r = O.barycenter_parallel(A1, C, W1[:, m], reg=reg, maxiter=L, zerotol=1e-9)
        np.testing.assert_allclose(out["Yhat"][:, m], r["b"], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("opt", [0, 2])
def test_wdl_ragged_last_batch_and_lambda_broadcast_over_two_epochs(opt):
    # M % B != 0: batches of 3 and 2 documents; every column of Lambda moves
    # by the same g_lambda after every batch, whether or not it is in it
    N, M, S, L, reg, B, E = 6, 5, 2, 4, 0.3, 3, 2
    C, Y, rng = _wdl_problem(22, N, M)
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    eta = 0.5 if opt == 0 else 0.05
    Alpha, Lam, sizes = _wdl_by_pieces(C, Y, reg, S, Alpha0, Lambda0, B, E, L, opt, eta)
    assert sizes == [3, 2, 3, 2]
    out = O.wdl(Y, C, reg, S, Alpha0, Lambda0, batch_size=B, epochs=E, max_iter=L,
                optimizer=opt, eta=eta)
    np.testing.assert_allclose(out["A"], O.softmax_cols(Alpha), rtol=1e-7, atol=1e-12)
    np.testing.assert_allclose(out["W"], O.softmax_cols(Lam), rtol=1e-7, atol=1e-12)
    # the broadcast: Lambda moved by one vector for all columns (SGD only:
    # AdamW's decay scales each column by the same factor, which keeps the
    # per-column differences, so check the differences instead)
    d = Lam - Lambda0 if opt == 0 else Lam - (1 - eta * 0.01) ** len(sizes) * Lambda0
    np.testing.assert_allclose(d, np.repeat(d[:, :1], M, axis=1), rtol=1e-12, atol=1e-14)


def test_wdl_init_mt19937_is_the_polar_method_on_mt19937():
    # numpy's legacy RandomState(seed) is MT19937 seeded by init_genrand, as
    # std::mt19937(seed); the C++ standard fixes its 10000th output at seed 5489
    raw = np.random.RandomState(5489)._bit_generator.random_raw(10000)
    assert int(raw[-1]) == 4123659995
    seed, N, S, M = 42, 9, 3, 5
    bits = iter(np.random.RandomState(seed)._bit_generator.random_raw(4 * (N * S + S * M) + 64))

    def canonical():  # 53-bit uniform from two 32-bit draws, low word first
        lo, hi = next(bits), next(bits)
        return (float(lo) + float(hi) * 2.0 ** 32) / 2.0 ** 64

    saved = []

    def normal():  # Marsaglia's polar method, second value of a pair kept
        if saved:
            return saved.pop()
        while True:
            x = 2.0 * canonical() - 1.0
            y = 2.0 * canonical() - 1.0
            r2 = x * x + y * y
            if 0.0 < r2 <= 1.0:
                break
        mult = np.sqrt(-2.0 * np.log(r2) / r2)
        saved.append(x * mult)
        return y * mult

    draws = np.array([normal() for _ in range(N * S + S * M)])
    Alpha0, Lambda0 = O.wdl_init_mt19937(seed, N, S, M)
    np.testing.assert_allclose(Alpha0.ravel(order="F"), draws[:N * S], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(Lambda0.ravel(order="F"), draws[N * S:], rtol=1e-14, atol=1e-15)


# ------------------------------------------------------------- cost / histograms


def test_oracle_euclidean_against_scipy():
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(11)
    E = rng.standard_normal((23, 5))
    D = O.euclidean(E)
    assert np.allclose(D, cdist(E, E), rtol=1e-14, atol=1e-14)
    assert np.array_equal(D, D.T) and np.all(np.diag(D) == 0.0)


def test_oracle_doc2dist_known_answer():
    # "zzz" is not in the dictionary: it counts for the last entry; "a"
    # appears twice in the dictionary: the first one counts
    docs = [["a", "b", "a"], ["c", "zzz"], ["b"]]
    dictionary = ["a", "b", "c", "a"]
    out = O.doc2dist(docs, dictionary)
    expect = np.zeros((4, 3))
    expect[0, 0], expect[1, 0] = 2 / 3, 1 / 3
    # "zzz" goes to the last index, 3 (the duplicate "a"), "c" to index 2
    expect[2, 1], expect[3, 1] = 0.5, 0.5
    expect[1, 2] = 1.0
    assert np.array_equal(out, expect)
