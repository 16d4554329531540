"""GPU tests of Wasserstein Dictionary Learning (rwig_wdl*, librwig_hip.so)
against the FP64 oracle (src/wdl_impl.cpp restated).

FP64: batched gradient and the whole fit (A, W, Yhat) to 1e-9 relative, from
explicit initial parameters and from the seeded mt19937 draw.
FP32: batched gradient to 1e-4 relative (FP32 Gibbs products, stated
tolerance)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import rwig_amd  # noqa: E402
from oracle import oracle as O  # noqa: E402
from rwig_amd._lib import check, lib  # noqa: E402


# FP32 Gibbs products and recurrences at the WDL configuration (eps 0.1, C in
# [0, 4]): the rounding of K = exp(-C/eps) in FP32 is ~6e-8 relative, and the
# 2L - 1 products of the adjoint carry it
TOL_C4_FP32 = 1e-4


def wdl_problem(seed, N, M, S):
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((N, 5))
    E /= np.linalg.norm(E, axis=1, keepdims=True)
    C = ((E[:, None] - E[None]) ** 2).sum(-1)
    Y = rng.dirichlet(np.ones(N), size=M).T.copy()
    return np.asfortranarray(C), np.asfortranarray(Y), rng


def rel(x, y):
    return float(np.max(np.abs(np.asarray(x) - np.asarray(y))) / np.max(np.abs(y)))


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-4)])
def test_batch_grad_matches_oracle(precision, tol):
    N, S, D, L, reg = 70, 4, 6, 12, 0.3
    C, Y, rng = wdl_problem(1, N, D, S)
    K = np.exp(-C / reg)
    A = O.softmax_cols(rng.standard_normal((N, S)))
    Wb = O.softmax_cols(rng.standard_normal((S, D)))
    gA_o, gl_o, bB_o = O.wdl_batch_grad(K, A, Y, Wb, L)
    dt = torch.float64 if precision == 0 else torch.float32
    ldk = (N + 7) // 8 * 8
    Kt = torch.zeros((N, ldk), dtype=dt)
    Kt[:, :N] = torch.from_numpy(K.T.copy())  # column j of K as row j: column-major K
    KTt = torch.zeros((N, ldk), dtype=dt)
    KTt[:, :N] = torch.from_numpy(K.copy())  # column-major K^T
    dev = [x.cuda() for x in (Kt, KTt)]
    At = torch.from_numpy(A.T.copy()).to(dt).cuda()   # column-major N x S
    Yt = torch.from_numpy(Y.T.copy()).to(dt).cuda()
    Wt = torch.from_numpy(Wb.T.copy()).to(dt).cuda()
    gA = torch.zeros((S, N), dtype=torch.float64, device="cuda")
    gl = torch.zeros(S, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    check(lib().rwig_wdl_batch_grad_dev(dev[0].data_ptr(), dev[1].data_ptr(), ldk, N, S, D, L,
                                        At.data_ptr(), Yt.data_ptr(), Wt.data_ptr(),
                                        gA.data_ptr(), gl.data_ptr(), loss.data_ptr(),
                                        precision, 0, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rel(gA.cpu().numpy().T, gA_o) < tol
    assert rel(gl.cpu().numpy(), gl_o) < tol
    assert abs(loss.item() - np.sum((bB_o - Y) ** 2)) < tol * np.sum((bB_o - Y) ** 2)


def test_wdl_fit_matches_oracle_fp64():
    # M % batch_size != 0 exercises the reference's last-batch rule
    N, M, S = 60, 12, 3
    C, Y, rng = wdl_problem(2, N, M, S)
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    kw = dict(batch_size=5, epochs=2, max_iter=10, optimizer=2, eta=0.01)
    r = rwig_amd.wdl(Y, C, 0.3, S, init=(Alpha0, Lambda0), **kw)
    o = O.wdl(Y, C, 0.3, S, Alpha0, Lambda0, **kw)
    for k in ["A", "W", "Yhat"]:
        assert rel(r[k], o[k]) < 1e-9, k


def _step_sequence(Y, C, reg, S, Alpha0, Lambda0, batch_size, epochs, max_iter, optimizer,
                   eta, precision):
    """The fit as a sequence of rwig_wdl_step_dev calls on device buffers;
    returns the final Alpha, Lambda and the per-step losses."""
    N, M = Y.shape
    dt = torch.float64 if precision == 0 else torch.float32
    K = np.exp(-C / reg)
    ldk = (N + 7) // 8 * 8
    Kt = torch.zeros((N, ldk), dtype=dt)
    Kt[:, :N] = torch.from_numpy(K.T.copy())
    KTt = torch.zeros((N, ldk), dtype=dt)
    KTt[:, :N] = torch.from_numpy(K.copy())
    Kd, KTd = Kt.cuda(), KTt.cuda()
    Yd = torch.from_numpy(Y.T.copy()).to(dt).cuda()
    Al = torch.from_numpy(Alpha0.T.copy()).cuda()      # column-major N x S
    La = torch.from_numpy(Lambda0.T.copy()).cuda()     # column-major S x M
    mom = [torch.zeros_like(Al), torch.zeros_like(Al), torch.zeros_like(La),
           torch.zeros_like(La)]
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    batches = -(-M // batch_size)
    losses, step = [], 0
    for _ in range(epochs):
        for bid in range(batches):
            D = M % batch_size if bid == M // batch_size else batch_size
            step += 1
            check(lib().rwig_wdl_step_dev(Kd.data_ptr(), KTd.data_ptr(), ldk, Yd.data_ptr(), N, M,
                                          S, max_iter, bid * batch_size, D, Al.data_ptr(),
                                          La.data_ptr(), *[m.data_ptr() for m in mom], step,
                                          optimizer, eta, 0.01, 0.9, 0.999, 1e-8,
                                          loss.data_ptr(), precision, 0, st))
            torch.cuda.synchronize()
            losses.append(loss.item())
    return Al.cpu().numpy().T, La.cpu().numpy().T, losses


@pytest.mark.parametrize("precision,opt", [(0, 2), (0, 1), (1, 2)])
def test_step_entry_point_sequence_is_the_fit(precision, opt):
    # the fit's batch loop (src/wdl_impl.cpp:316-337) rebuilt from the device
    # step: same A and W as rwig_wdl_init (same kernels; only the final
    # softmax is taken on the host here), and in FP64 the oracle's fit
    N, M, S, reg = 50, 11, 3, 0.3
    C, Y, rng = wdl_problem(6, N, M, S)
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    kw = dict(batch_size=4, epochs=2, max_iter=9, optimizer=opt, eta=0.02)
    Al, La, losses = _step_sequence(Y, C, reg, S, Alpha0, Lambda0, precision=precision, **kw)
    A, W = O.softmax_cols(Al), O.softmax_cols(La)
    r = rwig_amd.wdl(Y, C, reg, S, init=(Alpha0, Lambda0),
                     precision="fp64" if precision == 0 else "fp32", **kw)
    assert rel(A, r["A"]) < 1e-13 and rel(W, r["W"]) < 1e-13
    assert len(losses) == 2 * 3 and all(np.isfinite(losses))
    if precision == 0:
        o = O.wdl(Y, C, reg, S, Alpha0, Lambda0, **kw)
        assert rel(A, o["A"]) < 1e-9 and rel(W, o["W"]) < 1e-9


def test_step_entry_point_refuses_bad_batches():
    N, M, S = 20, 6, 2
    z = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = z.data_ptr()
    st = torch.cuda.current_stream().cuda_stream
    for first, D, step in ((4, 3, 1), (0, 0, 1), (-1, 2, 1), (0, 2, 0)):
        rc = lib().rwig_wdl_step_dev(p, p, 24, p, N, M, S, 5, first, D, p, p, p, p, p, p, step,
                                     2, 0.01, 0.01, 0.9, 0.999, 1e-8, None, 0, 0, st)
        assert rc != 0, (first, D, step)


@pytest.mark.parametrize("opt", [0, 1])
def test_wdl_fit_other_optimizers(opt):
    N, M, S = 40, 8, 2
    C, Y, rng = wdl_problem(3, N, M, S)
    Alpha0 = rng.standard_normal((N, S))
    Lambda0 = rng.standard_normal((S, M))
    kw = dict(batch_size=4, epochs=1, max_iter=8, optimizer=opt, eta=0.05)
    r = rwig_amd.wdl(Y, C, 0.5, S, init=(Alpha0, Lambda0), **kw)
    o = O.wdl(Y, C, 0.5, S, Alpha0, Lambda0, **kw)
    for k in ["A", "W", "Yhat"]:
        assert rel(r[k], o[k]) < 1e-9, k


@pytest.mark.parametrize("seed,opt", [(7, 2), (123, 1)])
def test_wdl_seeded_entry_point_matches_oracle(seed, opt):
    # rwig_wdl draws Alpha0, Lambda0 as the reference's CUDA path does
    # (src/cuda_wdl.cu:592-593, :677-689); the oracle's mt19937 helper is
    # pinned to the polar method on MT19937 in the CPU tests
    N, M, S = 30, 7, 3
    C, Y, _ = wdl_problem(5 + seed, N, M, S)
    kw = dict(batch_size=3, epochs=2, max_iter=6, optimizer=opt, eta=0.02)
    r = rwig_amd.wdl(Y, C, 0.4, S, seed=seed, **kw)
    Alpha0, Lambda0 = O.wdl_init_mt19937(seed, N, S, M)
    o = O.wdl(Y, C, 0.4, S, Alpha0, Lambda0, **kw)
    for k in ["A", "W", "Yhat"]:
        assert rel(r[k], o[k]) < 1e-9, k


def test_wdl_seeded_entry_point():
    N, M, S = 30, 6, 2
    C, Y, _ = wdl_problem(4, N, M, S)
    r1 = rwig_amd.wdl(Y, C, 0.4, S, batch_size=3, epochs=1, max_iter=5, seed=7)
    r2 = rwig_amd.wdl(Y, C, 0.4, S, batch_size=3, epochs=1, max_iter=5, seed=7)
    for k in ["A", "W", "Yhat"]:
        assert np.array_equal(r1[k], r2[k])
    np.testing.assert_allclose(r1["A"].sum(0), 1, rtol=1e-12)
    np.testing.assert_allclose(r1["W"].sum(0), 1, rtol=1e-12)
    np.testing.assert_allclose(r1["Yhat"].sum(0), 1, rtol=1e-12)


# ------------------------------------------------------ C4 at its workload size


class C4Batch:
    """The WDL configuration's batch on the device: V = 2000 words, K = 16
    atoms, D = 512 documents, eps 0.1, L = 20."""

    def __init__(self):
        from rwig_amd import workloads
        C, Y, Alpha0, Lambda0, eps, L = workloads.c4()
        self.N, self.D = Y.shape
        self.S = Alpha0.shape[1]
        self.L = L
        self.Y, self.Alpha0, self.Lambda0 = Y, Alpha0, Lambda0
        K = np.exp(-C / eps)
        self.ldk = (self.N + 7) // 8 * 8
        self.dev = {}
        for prec, dt in ((0, torch.float64), (1, torch.float32)):
            Kt = torch.zeros((self.N, self.ldk), dtype=dt)
            Kt[:, :self.N] = torch.from_numpy(K.T.copy())
            KTt = torch.zeros((self.N, self.ldk), dtype=dt)
            KTt[:, :self.N] = torch.from_numpy(K.copy())
            self.dev[prec] = (Kt.cuda(), KTt.cuda(),
                              torch.from_numpy(Y.T.copy()).to(dt).cuda())

    def grad(self, Alpha, Lambda, precision=0, ck=0):
        Kd, KTd, Yd = self.dev[precision]
        dt = Yd.dtype
        A = O.softmax_cols(Alpha)
        W = O.softmax_cols(Lambda)
        At = torch.from_numpy(A.T.copy()).to(dt).cuda()
        Wt = torch.from_numpy(W.T.copy()).to(dt).cuda()
        gA = torch.zeros((self.S, self.N), dtype=torch.float64, device="cuda")
        gl = torch.zeros(self.S, dtype=torch.float64, device="cuda")
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        check(lib().rwig_wdl_batch_grad_dev(Kd.data_ptr(), KTd.data_ptr(), self.ldk, self.N,
                                            self.S, self.D, self.L, At.data_ptr(),
                                            Yd.data_ptr(), Wt.data_ptr(), gA.data_ptr(),
                                            gl.data_ptr(), loss.data_ptr(), precision, ck,
                                            torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        return gA.cpu().numpy().T, gl.cpu().numpy(), loss.item()


@pytest.fixture(scope="module")
def c4batch():
    return C4Batch()


def test_c4_fp64_gradient_matches_finite_differences(c4batch):
    # the gradient is of the batch's mean loss (1/D) sum_d ||b_L,d - y_d||^2
    # (src/wdl_impl.cpp:236-246); g_lambda along a shift of row s of every
    # column of Lambda at once
    b, h = c4batch, 1e-6
    gA, gl, _ = b.grad(b.Alpha0, b.Lambda0)
    mean = lambda Al, La: b.grad(Al, La)[2] / b.D  # noqa: E731
    rng = np.random.default_rng(40)
    coords = [(int(rng.integers(b.N)), s) for s in rng.choice(b.S, 8, replace=False)]
    fd, an = [], []
    for i, s in coords:
        Ap, Am = b.Alpha0.copy(), b.Alpha0.copy()
        Ap[i, s] += h
        Am[i, s] -= h
        fd.append((mean(Ap, b.Lambda0) - mean(Am, b.Lambda0)) / (2 * h))
        an.append(gA[i, s])
    for s in range(b.S):
        Lp, Lm = b.Lambda0.copy(), b.Lambda0.copy()
        Lp[s] += h
        Lm[s] -= h
        fd.append((mean(b.Alpha0, Lp) - mean(b.Alpha0, Lm)) / (2 * h))
        an.append(gl[s])
    fd, an = np.array(fd), np.array(an)
    err = np.linalg.norm(fd - an) / np.linalg.norm(fd)
    print(f"C4 FD check: |fd - adjoint| / |fd| = {err:.2e} over {len(fd)} coordinates "
          f"(|g| {np.abs(fd).min():.1e} .. {np.abs(fd).max():.1e})", flush=True)
    assert err < 1e-5


def test_c4_fp32_against_fp64(c4batch):
    b = c4batch
    gA64, gl64, l64 = b.grad(b.Alpha0, b.Lambda0, 0)
    gA32, gl32, l32 = b.grad(b.Alpha0, b.Lambda0, 1)
    eA = np.linalg.norm(gA32 - gA64) / np.linalg.norm(gA64)
    el = np.linalg.norm(gl32 - gl64) / np.linalg.norm(gl64)
    eloss = abs(l32 - l64) / abs(l64)
    print(f"C4 FP32 vs FP64: g_Alpha {eA:.2e}, g_lambda {el:.2e}, loss {eloss:.2e}", flush=True)
    assert max(eA, el, eloss) < TOL_C4_FP32


# ------------------------------------------------------ checkpointed adjoint


@pytest.mark.parametrize("precision", [0, 1])
def test_checkpointed_batch_grad_is_bit_identical(precision):
    # the adjoint's segments are recomputed with the forward's own kernels
    N, S, D, L, reg = 70, 4, 6, 12, 0.3
    C, Y, rng = wdl_problem(6, N, D, S)
    K = np.exp(-C / reg)
    A = O.softmax_cols(rng.standard_normal((N, S)))
    Wb = O.softmax_cols(rng.standard_normal((S, D)))
    dt = torch.float64 if precision == 0 else torch.float32
    ldk = (N + 7) // 8 * 8
    Kt = torch.zeros((N, ldk), dtype=dt)
    Kt[:, :N] = torch.from_numpy(K.T.copy())
    KTt = torch.zeros((N, ldk), dtype=dt)
    KTt[:, :N] = torch.from_numpy(K.copy())
    Kd, KTd = Kt.cuda(), KTt.cuda()
    At = torch.from_numpy(A.T.copy()).to(dt).cuda()
    Yt = torch.from_numpy(Y.T.copy()).to(dt).cuda()
    Wt = torch.from_numpy(Wb.T.copy()).to(dt).cuda()
    outs = []
    for ck in [0, 1, 3, 5, 7, 11, 12, 40]:
        gA = torch.zeros((S, N), dtype=torch.float64, device="cuda")
        gl = torch.zeros(S, dtype=torch.float64, device="cuda")
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        check(lib().rwig_wdl_batch_grad_dev(Kd.data_ptr(), KTd.data_ptr(), ldk, N, S, D, L,
                                            At.data_ptr(), Yt.data_ptr(), Wt.data_ptr(),
                                            gA.data_ptr(), gl.data_ptr(), loss.data_ptr(),
                                            precision, ck, torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        outs.append((ck, gA.cpu(), gl.cpu(), loss.cpu()))
    for ck, gA, gl, loss in outs[1:]:
        assert torch.equal(gA, outs[0][1]), ck
        assert torch.equal(gl, outs[0][2]), ck
        assert torch.equal(loss, outs[0][3]), ck


def test_checkpointed_fit_is_bit_identical():
    N, M, S = 60, 12, 3
    C, Y, rng = wdl_problem(7, N, M, S)
    init = (rng.standard_normal((N, S)), rng.standard_normal((S, M)))
    kw = dict(batch_size=5, epochs=2, max_iter=10, optimizer=1, eta=0.01, init=init)
    r0 = rwig_amd.wdl(Y, C, 0.3, S, **kw)
    r4 = rwig_amd.wdl(Y, C, 0.3, S, checkpoint_every=4, **kw)
    for k in ["A", "W", "Yhat"]:
        assert np.array_equal(r0[k], r4[k]), k


def test_c4_checkpointed_gradient_is_bit_identical(c4batch):
    b = c4batch
    for prec in (0, 1):
        full = b.grad(b.Alpha0, b.Lambda0, prec, ck=0)
        ck5 = b.grad(b.Alpha0, b.Lambda0, prec, ck=5)
        assert np.array_equal(full[0], ck5[0]) and np.array_equal(full[1], ck5[1])
        assert full[2] == ck5[2]
