"""GPU tests of the batched one-source-many-targets Sinkhorn
(rwig_sinkhorn_batched_dev).

FP64 mode: every target equals the oracle's log Sinkhorn on the target's
reduced problem (src/sinkhorn_impl_log.cpp), same iteration count, f and g to
1e-10 relative. FP32 / BF16 modes: the returned potentials define a plan whose
L1 marginal error, recomputed here in FP64 on the host, is below the stated
tolerance, and whose transport cost agrees with the FP64 mode's to 1e-5
relative; the returned transport cost and regularised loss are those of that
plan, evaluated here on the host. Every mode gives bit-identical results for a
target solved alone and inside a batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rwig_amd import batched as RB  # noqa: E402
from rwig_amd import workloads  # noqa: E402


def make(seed, M, B, nt_lo, nt_hi, reg):
    rng = np.random.default_rng(seed)
    X = rng.random((M, 2))
    C = np.asfortranarray(((X[:, None] - X[None]) ** 2).sum(-1))
    a = rng.dirichlet(np.ones(M))
    sup, wts = [], []
    for _ in range(B):
        k = int(rng.integers(nt_lo, nt_hi + 1))
        sup.append(np.sort(rng.choice(M, k, replace=False)))
        wts.append(rng.dirichlet(np.full(k, 0.5)))
    return C, a, sup, wts, reg


def to_dev(C, a):
    Ct = torch.from_numpy(np.asfortranarray(C)).cuda()
    Ct = Ct.t().contiguous().t()  # keep column-major strides (1, M)
    assert Ct.stride() == (1, C.shape[0])
    return Ct, torch.from_numpy(a).cuda()


def host(out, tb):
    torch.cuda.synchronize()
    f = out["f"].cpu().numpy()
    g = out["g"].cpu().numpy()
    gs = [g[tb.offsets[t]:tb.offsets[t + 1]] for t in range(tb.B)]
    return f, gs, out["iter"].cpu().numpy(), out["err"].cpu().numpy(), out["loss"].cpu().numpy()


def plan_l1(C, a, sup, w, f, g, reg):
    Ct = C[:, sup]
    P = np.exp(-((Ct - f[:, None]) - g[None, :]) / reg)
    return np.abs(P.sum(1) - a).sum() + np.abs(P.sum(0) - w).sum(), (Ct * P).sum()


def plan_loss(C, sup, f, g, reg):
    """Transport cost and regularised loss <P, C> + reg sum P (log P - 1) of the
    plan the potentials define, with its terms' magnitude (for the tolerance)."""
    Ct = C[:, sup]
    lp = -((Ct - f[:, None]) - g[None, :]) / reg
    P = np.exp(lp)
    ent = reg * (P * (lp - 1.0))
    cost = (Ct * P).sum()
    return cost, cost + ent.sum(), (Ct * P).sum() + np.abs(ent).sum()


@pytest.mark.parametrize("seed,M,B,lo,hi,reg", [
    (1, 300, 5, 8, 60, 0.05),
    (2, 1000, 3, 40, 90, 0.02),
    (3, 64, 4, 1, 64, 0.1),
])
def test_fp64_batched_matches_oracle_per_target(seed, M, B, lo, hi, reg):
    C, a, sup, wts, reg = make(seed, M, B, lo, hi, reg)
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    out = RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64", err="reference")
    f, gs, it, err, loss = host(out, tb)
    for t in range(B):
        print("target", t, "nt", len(sup[t]), flush=True)
        bfull = np.zeros(M)
        bfull[sup[t]] = wts[t]
        o = O.sinkhorn_log(a, bfull, C, reg)
        assert it[t] == o["iter"]
        scale = max(np.abs(o["f"]).max(), 1e-3)
        assert np.abs(f[t] - o["f"]).max() / scale < 1e-10
        og = o["g"][sup[t]]
        assert np.abs(gs[t] - og).max() / scale < 1e-10
        assert abs(loss[t, 1] - o["loss"]) <= 1e-10 * abs(o["loss"])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_fast_modes_reach_tolerance_and_match_fp64(precision):
    C, a, sup, wts, _ = make(7, 2048, 6, 80, 120, 0.01)
    reg, tol = 0.01, 1e-6
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    ref = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64", err="l1",
                                   zero_tol=tol, max_iter=5000), tb)
    out = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision=precision, err="l1",
                                   zero_tol=tol, max_iter=5000, absorb_every=20), tb)
    for t in range(tb.B):
        e_ref, c_ref = plan_l1(C, a, sup[t], wts[t], ref[0][t], ref[1][t], reg)
        e, c = plan_l1(C, a, sup[t], wts[t], out[0][t], out[1][t], reg)
        print(precision, t, "iters", out[2][t], "vs", ref[2][t], "err", e, flush=True)
        assert e_ref <= tol * 1.0001
        assert e <= tol * 1.0001, (t, e)
        assert abs(out[3][t] - e) <= 1e-9 + 1e-6 * e  # reported error is the plan's
        assert abs(c - c_ref) <= 1e-5 * abs(c_ref)
        assert abs(out[4][t, 0] - c) <= 1e-9 * abs(c)
        _, rl, mag = plan_loss(C, sup[t], out[0][t], out[1][t], reg)
        assert abs(out[4][t, 1] - rl) <= 1e-12 * mag, (t, out[4][t, 1], rl)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_result_alone_equals_result_in_mix(precision):
    C, a, sup, wts, reg = make(11, 512, 7, 5, 50, 0.02)
    Cd, ad = to_dev(C, a)
    mix = host(RB.sinkhorn_batched(Cd, ad, RB.TargetBatch(sup, wts), reg, precision=precision,
                                   err="l1"), RB.TargetBatch(sup, wts))
    for t in [0, 3, 6]:
        tb1 = RB.TargetBatch([sup[t]], [wts[t]])
        one = host(RB.sinkhorn_batched(Cd, ad, tb1, reg, precision=precision, err="l1"), tb1)
        assert np.array_equal(one[0][0], mix[0][t])
        assert np.array_equal(one[1][0], mix[1][t])
        assert one[2][0] == mix[2][t]
        assert one[3][0] == mix[3][t]


def test_c2_scaled_down_fp64_against_oracle():
    # the C2 generator at a quarter of its size, two targets through the oracle
    C, a, sup, wts, reg = workloads.c2(n=1024, B=4)
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    f, gs, it, err, loss = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64",
                                                    err="reference"), tb)
    for t in range(2):
        bfull = np.zeros(C.shape[0])
        bfull[sup[t]] = wts[t]
        o = O.sinkhorn_log(a, bfull, C, reg)
        assert it[t] == o["iter"]
        assert np.abs(f[t] - o["f"]).max() / np.abs(o["f"]).max() < 1e-10


@pytest.mark.parametrize("absorb_every", [10, 20])
def test_mfma_path_reaches_tolerance_and_matches_fp64(absorb_every):
    C, a, sup, wts, _ = make(21, 1024, 9, 30, 70, 0.01)
    reg, tol = 0.01, 1e-6
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    ref = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64", err="l1",
                                   zero_tol=tol, max_iter=5000), tb)
    out = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, zero_tol=tol, max_iter=5000,
                                        absorb_every=absorb_every), tb)
    for t in range(tb.B):
        e_ref, c_ref = plan_l1(C, a, sup[t], wts[t], ref[0][t], ref[1][t], reg)
        e, c = plan_l1(C, a, sup[t], wts[t], out[0][t], out[1][t], reg)
        print("mfma", t, "iters", out[2][t], "vs", ref[2][t], "err", e, flush=True)
        assert e <= tol * 1.0001, (t, e)
        assert abs(out[3][t] - e) <= 1e-9 + 1e-6 * e
        assert abs(c - c_ref) <= 1e-5 * abs(c_ref)
        assert abs(out[4][t, 0] - c) <= 1e-9 * abs(c)
        _, rl, mag = plan_loss(C, sup[t], out[0][t], out[1][t], reg)
        assert abs(out[4][t, 1] - rl) <= 1e-12 * mag, (t, out[4][t, 1], rl)


def test_mfma_path_rows_not_a_multiple_of_4():
    """M = 301 takes the one-row-per-thread form of the Y = K X update (the
    4-row form needs M % 4 == 0). Every target must reach the tolerance, with
    the error, transport cost and regularised loss of the plan recomputed on
    the host, and the cost within 1e-5 of the FP64 batched form's."""
    C, a, sup, wts, _ = make(43, 301, 7, 10, 60, 0.02)
    reg, tol = 0.02, 1e-6
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    ref = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64", err="l1",
                                   zero_tol=tol, max_iter=5000), tb)
    out = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, zero_tol=tol, max_iter=5000,
                                        absorb_every=10), tb)
    for t in range(tb.B):
        _, c_ref = plan_l1(C, a, sup[t], wts[t], ref[0][t], ref[1][t], reg)
        e, c = plan_l1(C, a, sup[t], wts[t], out[0][t], out[1][t], reg)
        print("mfma M=301", t, "iters", out[2][t], "err", e, flush=True)
        assert e <= tol * 1.0001, (t, e)
        assert abs(out[3][t] - e) <= 1e-9 + 1e-6 * e
        assert abs(c - c_ref) <= 1e-5 * abs(c_ref)
        _, rl, mag = plan_loss(C, sup[t], out[0][t], out[1][t], reg)
        assert abs(out[4][t, 1] - rl) <= 1e-12 * mag, (t, out[4][t, 1], rl)


def test_mfma_path_alone_equals_in_mix():
    C, a, sup, wts, reg = make(23, 600, 6, 10, 60, 0.02)
    Cd, ad = to_dev(C, a)
    kern = RB.DenseKernel(Cd, reg)
    tb = RB.TargetBatch(sup, wts)
    mix = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, kernel=kern), tb)
    for t in [0, 5]:
        tb1 = RB.TargetBatch([sup[t]], [wts[t]])
        one = host(RB.sinkhorn_batched_mfma(Cd, ad, tb1, reg, kernel=kern), tb1)
        assert np.array_equal(one[0][0], mix[0][t])
        assert np.array_equal(one[1][0], mix[1][t])
        assert one[2][0] == mix[2][t]


@pytest.mark.parametrize("M", [256, 300])
def test_mfma_path_same_result_in_every_tile_shape(M):
    """900 targets pass through every tile shape of the products as they stop
    (128 x 128 above 768 or 512 targets, then the smaller shapes); a target
    solved alone runs in the smallest ones from its first iteration. Its
    result must be the same bit for bit. M = 256 takes the direct-to-LDS
    tiles, M = 300 (k not a multiple of the k-step) the register-path ones."""
    C, a, sup, wts, reg = make(31, M, 900, 8, 40, 0.05)
    Cd, ad = to_dev(C, a)
    kern = RB.DenseKernel(Cd, reg)
    tb = RB.TargetBatch(sup, wts)
    mix = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, kernel=kern, absorb_every=10), tb)
    order = np.argsort(mix[2])
    print("iterations min/median/max", mix[2].min(), np.median(mix[2]), mix[2].max(), flush=True)
    assert mix[2].max() > mix[2].min()  # targets stop at different absorptions
    for t in [int(order[0]), int(order[len(order) // 2]), int(order[-1])]:
        tb1 = RB.TargetBatch([sup[t]], [wts[t]])
        one = host(RB.sinkhorn_batched_mfma(Cd, ad, tb1, reg, kernel=kern, absorb_every=10), tb1)
        assert one[2][0] == mix[2][t]
        assert np.array_equal(one[0][0], mix[0][t]), t
        assert np.array_equal(one[1][0], mix[1][t]), t
        assert np.array_equal(one[4][0], mix[4][t]), t


def test_mfma_path_at_the_full_c2_size():
    """The headline workload at its full size: n = 4096, 1024 targets on 5%
    supports, eps 0.01, absorption every 20, stop at L1 marginal error 1e-6.
    Every target must report an error <= 1e-6; for 16 of them (spread over
    the batch) the error, transport cost and regularised loss are recomputed
    on the host from the returned potentials in FP64 and must agree."""
    C, a, sup, wts, reg = workloads.c2()
    tol = 1e-6
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    out = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, zero_tol=tol, max_iter=1000,
                                        absorb_every=20), tb)
    f, gs, it, err, loss = out
    print("iterations min/mean/max", it.min(), it.mean(), it.max(), "max err", err.max(),
          flush=True)
    assert np.all(err <= tol)
    assert np.all(it % 20 == 0) and it.max() < 1000
    picks = np.linspace(0, len(sup) - 1, 16).astype(int)
    for t in picks:
        e, c = plan_l1(C, a, sup[t], wts[t], f[t], gs[t], reg)
        assert e <= tol * 1.0001, (t, e)
        assert abs(err[t] - e) <= 1e-9 + 1e-6 * e, (t, err[t], e)
        assert abs(loss[t, 0] - c) <= 1e-9 * abs(c), (t, loss[t, 0], c)
        _, rl, mag = plan_loss(C, sup[t], f[t], gs[t], reg)
        assert abs(loss[t, 1] - rl) <= 1e-12 * mag, (t, loss[t, 1], rl)
    # the same 16 targets in the FP64 batched form (the reference's FP64
    # log-domain iteration), run to the same L1 tolerance: transport cost and
    # regularised loss within 1e-5
    tb16 = RB.TargetBatch([sup[t] for t in picks], [wts[t] for t in picks])
    ref = host(RB.sinkhorn_batched(Cd, ad, tb16, reg, precision="fp64", err="l1",
                                   zero_tol=tol, max_iter=1000), tb16)
    for q, t in enumerate(picks):
        assert ref[3][q] <= tol
        for k in (0, 1):
            d = abs(loss[t, k] - ref[4][q, k]) / abs(ref[4][q, k])
            assert d <= 1e-5, (t, k, loss[t, k], ref[4][q, k], d)


def test_fp64_batched_at_the_full_c2_size_against_oracle():
    """Two targets of the full C2 workload (4096 x 205 reduced problems) in
    the FP64 batched form against the oracle's log Sinkhorn
    (src/sinkhorn_impl_log.cpp:225-376): same iteration count, f and g to
    1e-10 relative."""
    C, a, sup, wts, reg = workloads.c2()
    picks = [0, 777]
    tb = RB.TargetBatch([sup[t] for t in picks], [wts[t] for t in picks])
    Cd, ad = to_dev(C, a)
    f, gs, it, err, loss = host(RB.sinkhorn_batched(Cd, ad, tb, reg, precision="fp64",
                                                    err="reference"), tb)
    for q, t in enumerate(picks):
        bfull = np.zeros(C.shape[0])
        bfull[sup[t]] = wts[t]
        o = O.sinkhorn_log(a, bfull, C, reg, n_threads=8)
        print(f"target {t}: iterations {it[q]} (oracle {o['iter']})", flush=True)
        assert it[q] == o["iter"]
        assert np.abs(f[q] - o["f"]).max() / np.abs(o["f"]).max() < 1e-10
        og = o["g"][sup[t]]
        assert np.abs(gs[q] - og).max() / np.abs(og).max() < 1e-10


def ref_err(C, a, sup, w, f, g, reg):
    """The reference's error formula (src/sinkhorn_impl_log.cpp): L2 norms of
    the log-marginal mismatches of the plan the potentials define."""
    P = np.exp(-((C[:, sup] - f[:, None]) - g[None, :]) / reg)
    return (np.linalg.norm(np.log(P.sum(1)) - np.log(a))
            + np.linalg.norm(np.log(P.sum(0)) - np.log(w)))


def test_mfma_path_edge_cases():
    """Supports from 1 point to all M points; a cap on the iterations; the
    reference error formula as the stopping test."""
    C, a, sup, wts, reg = make(41, 64, 12, 1, 64, 0.1)
    sup[0] = np.array([17])  # a one-point target: its plan is the source itself
    wts[0] = np.array([1.0])
    tb = RB.TargetBatch(sup, wts)
    Cd, ad = to_dev(C, a)
    tol = 1e-6
    out = host(RB.sinkhorn_batched_mfma(Cd, ad, tb, reg, zero_tol=tol, max_iter=2000), tb)
    for t in range(tb.B):
        e, _ = plan_l1(C, a, sup[t], wts[t], out[0][t], out[1][t], reg)
        assert out[3][t] <= tol and e <= tol * 1.0001, (t, out[3][t], e)
    P0 = np.exp(-((C[:, sup[0]] - out[0][0][:, None]) - out[1][0][None, :]) / reg)
    assert np.abs(P0[:, 0] - a).max() <= 1e-6 * a.max()
    # a cap: every target stops at max_iter (a multiple of the absorption
    # interval) with the error of its absorbed plan
    C2, a2, sup2, wts2, _ = make(42, 300, 5, 20, 60, 0.01)
    tb2 = RB.TargetBatch(sup2, wts2)
    Cd2, ad2 = to_dev(C2, a2)
    capped = host(RB.sinkhorn_batched_mfma(Cd2, ad2, tb2, 0.01, zero_tol=1e-12, max_iter=40,
                                           absorb_every=20), tb2)
    assert np.all(capped[2] == 40)
    for t in range(tb2.B):
        e, _ = plan_l1(C2, a2, sup2[t], wts2[t], capped[0][t], capped[1][t], 0.01)
        assert e > 1e-12 and abs(capped[3][t] - e) <= 1e-9 + 1e-6 * e, (t, capped[3][t], e)
    # the reference formula decides the stop and is what is reported
    refm = host(RB.sinkhorn_batched_mfma(Cd2, ad2, tb2, 0.01, zero_tol=tol, max_iter=3000,
                                         err="reference"), tb2)
    for t in range(tb2.B):
        e = ref_err(C2, a2, sup2[t], wts2[t], refm[0][t], refm[1][t], 0.01)
        assert refm[3][t] <= tol and abs(refm[3][t] - e) <= 1e-9 + 1e-6 * e, (t, refm[3][t], e)
