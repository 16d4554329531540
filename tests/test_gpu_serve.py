"""GPU tests of the serving engine (rwig_serve_*).

Bars:
- every result meets its stopping rule when checked on the host in FP64
  (L1 marginal error of the returned potentials' plan <= zero_tol);
- Sinkhorn losses and barycenters agree with the FP64 oracle run to a tight
  tolerance within TOL_LOSS / TOL_BARY (the engine stops at L1 error 1e-6);
- a request's result is bit-identical alone and in a mix of other requests
  (same and other cost matrices, submitted from two threads while it runs);
- the HBM footprint never exceeds the session's budget, and requests that do
  not fit wait in the queue and still complete;
- each result carries the reference's return_status (0 converged, 1 max_iter,
  2 breakdown); max_iter is honoured exactly; zero-mass entries are solved on
  the support; a reg too small for the cost's range is refused.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from rwig_amd import RwigError
from rwig_amd.serve import DONE, ServeSession, footprint

pytestmark = pytest.mark.gpu

TOL_LOSS = 1e-5
TOL_BARY = 1e-5


def cloud(rng, n):
    X = rng.random((n, 2))
    sq = (X * X).sum(1)
    C = np.maximum(sq[:, None] + sq[None] - 2.0 * X @ X.T, 0.0)
    np.fill_diagonal(C, 0.0)
    return C


def marginal_err(C, eps, f, g, a, b):
    P = np.exp((f[:, None] + g[None] - C) / eps)
    return np.abs(P.sum(1) - a).sum() + np.abs(P.sum(0) - b).sum()


@pytest.fixture(scope="module")
def clouds():
    rng = np.random.default_rng(2005)
    return [cloud(rng, 200), cloud(rng, 700), cloud(rng, 64)]


def test_sinkhorn_accuracy_against_oracle(clouds):
    rng = np.random.default_rng(1)
    with ServeSession(hbm_budget_bytes=2 << 30) as s:
        ids = [s.register_cost(C) for C in clouds]
        jobs = []
        for k, eps in [(0, 0.1), (0, 0.01), (1, 0.1), (1, 0.01), (2, 0.01)]:
            n = clouds[k].shape[0]
            a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
            jobs.append((k, eps, a, b, s.submit_sinkhorn(ids[k], a, b, eps, zero_tol=1e-6)))
        for k, eps, a, b, rid in jobs:
            r = s.result(rid)
            C = clouds[k]
            assert r["err"] <= 1e-6
            e_host = marginal_err(C, eps, r["f"], r["g"], a, b)
            assert abs(e_host - r["err"]) <= 1e-9, (e_host, r["err"])
            o = O.sinkhorn_log(a, b, C, eps, maxiter=5000, zerotol=1e-12)
            dl = abs(r["loss"] - o["loss"]) / abs(o["loss"])
            print(f"n={C.shape[0]} eps={eps}: iter {r['iter']} err {r['err']:.2e} "
                  f"loss rel diff {dl:.2e}")
            assert dl < TOL_LOSS


def test_barycenter_accuracy_against_oracle(clouds):
    rng = np.random.default_rng(2)
    with ServeSession(hbm_budget_bytes=2 << 30) as s:
        ids = [s.register_cost(C) for C in clouds]
        for k, S, eps in [(0, 2, 0.1), (1, 4, 0.1), (0, 8, 0.05)]:
            n = clouds[k].shape[0]
            A = rng.dirichlet(np.ones(n), size=S).T.copy()
            w = rng.dirichlet(np.ones(S))
            r = s.result(s.submit_barycenter(ids[k], A, w, eps, zero_tol=1e-6))
            o = O.barycenter_parallel(A, clouds[k], w, None, eps, maxiter=20000, zerotol=1e-13)
            d = np.abs(r["b"] - o["b"]).sum()
            print(f"barycenter n={n} S={S} eps={eps}: iter {r['iter']} err {r['err']:.2e} "
                  f"|b - b_oracle|_1 {d:.2e}")
            assert r["err"] <= 1e-6
            assert abs(r["b"].sum() - 1.0) < 1e-12
            assert d < TOL_BARY


def test_alone_equals_mix(clouds):
    rng = np.random.default_rng(3)
    n0, n1 = clouds[0].shape[0], clouds[1].shape[0]
    a, b = rng.dirichlet(np.ones(n0)), rng.dirichlet(np.ones(n0))
    A = rng.dirichlet(np.ones(n1), size=4).T.copy()
    w = rng.dirichlet(np.ones(4))
    with ServeSession(hbm_budget_bytes=2 << 30) as s:
        ids = [s.register_cost(C) for C in clouds]
        alone_s = s.result(s.submit_sinkhorn(ids[0], a, b, 0.01))
        alone_b = s.result(s.submit_barycenter(ids[1], A, w, 0.1))

    def flood(sess, seed, out):
        r = np.random.default_rng(seed)
        for k in range(12):
            c = int(r.integers(3))
            n = clouds[c].shape[0]
            eps = [0.1, 0.01][k % 2]
            if k % 4 == 3:
                S = int(r.choice([2, 4, 8]))
                out.append(sess.submit_barycenter(ids[c], r.dirichlet(np.ones(n), size=S).T.copy(),
                                                  r.dirichlet(np.ones(S)), eps))
            else:
                out.append(sess.submit_sinkhorn(ids[c], r.dirichlet(np.ones(n)),
                                                r.dirichlet(np.ones(n)), eps))

    with ServeSession(hbm_budget_bytes=2 << 30) as s:
        ids = [s.register_cost(C) for C in clouds]
        others = []
        flood(s, 10, others)  # already running when the probes arrive
        rs = s.submit_sinkhorn(ids[0], a, b, 0.01)
        t = threading.Thread(target=flood, args=(s, 11, others))
        t.start()
        rb = s.submit_barycenter(ids[1], A, w, 0.1)
        t.join()
        mix_s, mix_b = s.result(rs), s.result(rb)
        for rid in others:
            assert s.wait(rid) == DONE
            s.result(rid)
        st = s.stats()
    assert mix_s["iter"] == alone_s["iter"] and mix_b["iter"] == alone_b["iter"]
    assert np.array_equal(mix_s["f"], alone_s["f"]) and np.array_equal(mix_s["g"], alone_s["g"])
    assert mix_s["loss"] == alone_s["loss"] and mix_s["err"] == alone_s["err"]
    assert np.array_equal(mix_b["b"], alone_b["b"]) and mix_b["err"] == alone_b["err"]
    assert st["completed"] == 26 and st["failed"] == 0


def test_admission_by_hbm_budget(clouds):
    # budget for the three cost matrices and not all three groups at once:
    # requests queue until an idle group is released, within the budget
    rng = np.random.default_rng(4)
    cost_bytes = sum(C.nbytes for C in clouds)
    k1, g1 = footprint(700, 700, 32)
    k0, g0 = footprint(200, 200, 32)
    # the 700-point group fits, not together with the 200-point one
    budget = int(cost_bytes + k1 + g1 + 0.5 * (k0 + g0))
    with ServeSession(hbm_budget_bytes=budget, max_group_columns=64) as s:
        ids = [s.register_cost(C) for C in clouds]
        jobs = []
        for k in range(9):
            c = k % 3
            n = clouds[c].shape[0]
            a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
            jobs.append((c, a, b, s.submit_sinkhorn(ids[c], a, b, 0.1)))
        for c, a, b, rid in jobs:
            r = s.result(rid)
            assert r["err"] <= 1e-6
            assert marginal_err(clouds[c], 0.1, r["f"], r["g"], a, b) <= 1e-6
        st = s.stats()
    print("admission stats:", {k: st[k] for k in ("peak_bytes", "budget_bytes", "mean_queue_ms",
                                                   "p50_ms", "p99_ms")})
    assert st["completed"] == 9
    assert st["peak_bytes"] <= budget
    assert st["mean_queue_ms"] > 0


def test_request_larger_than_budget_is_rejected(clouds):
    with ServeSession(hbm_budget_bytes=clouds[1].nbytes + (1 << 20)) as s:
        cid = s.register_cost(clouds[1])
        n = clouds[1].shape[0]
        with pytest.raises(RwigError):
            s.submit_sinkhorn(cid, np.full(n, 1.0 / n), np.full(n, 1.0 / n), 0.1)


def test_max_iter_and_invalid_input(clouds):
    rng = np.random.default_rng(5)
    n = clouds[0].shape[0]
    with ServeSession(hbm_budget_bytes=1 << 30, absorb_every=10) as s:
        cid = s.register_cost(clouds[0])
        a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
        r = s.result(s.submit_sinkhorn(cid, a, b, 0.01, max_iter=30, zero_tol=1e-12))
        assert r["iter"] == 30 and r["err"] > 1e-12
        assert r["return_status"] == 1
        r0 = s.result(s.submit_sinkhorn(cid, a, b, 0.01, max_iter=0))
        assert r0["iter"] == 0 and r0["return_status"] == 1
        for v in (-1e-3, np.nan, np.inf):
            bad = a.copy()
            bad[3] = v
            with pytest.raises(RwigError):
                s.submit_sinkhorn(cid, bad, b, 0.1)
        with pytest.raises(RwigError, match="positive mass"):
            s.submit_sinkhorn(cid, np.zeros(n), b, 0.1)
        with pytest.raises(RwigError):
            s.submit_sinkhorn(cid + 5, a, b, 0.1)


@pytest.mark.parametrize("kind", ["sinkhorn", "barycenter"])
def test_max_iter_not_a_multiple_of_absorb_every(clouds, kind):
    # the last cycle runs only the iterations left: 50 = 20 + 20 + 10
    rng = np.random.default_rng(7)
    n = clouds[0].shape[0]
    with ServeSession(hbm_budget_bytes=1 << 30, absorb_every=20) as s:
        cid = s.register_cost(clouds[0])
        if kind == "sinkhorn":
            a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
            r = s.result(s.submit_sinkhorn(cid, a, b, 0.01, max_iter=50, zero_tol=1e-300))
            # the reported error is that of the plan after exactly 50 iterations
            assert abs(marginal_err(clouds[0], 0.01, r["f"], r["g"], a, b) - r["err"]) <= 1e-9
        else:
            A = rng.dirichlet(np.ones(n), size=3).T.copy()
            r = s.result(s.submit_barycenter(cid, A, rng.dirichlet(np.ones(3)), 0.05,
                                             max_iter=50, zero_tol=1e-300))
        assert r["iter"] == 50 and r["return_status"] == 1
        ok = s.result(s.submit_sinkhorn(cid, rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n)),
                                        0.1))
        assert ok["return_status"] == 0 and ok["err"] <= 1e-6


def test_reg_too_small_for_the_cost_is_refused(clouds):
    # every entry of a shifted cost is >= 0.9: at reg 0.01 exp(-C/reg) <= e^-90
    # underflows the BF16 Gibbs kernel, so the request is refused up front
    n = clouds[0].shape[0]
    u = np.full(n, 1.0 / n)
    with ServeSession(hbm_budget_bytes=1 << 30) as s:
        cid = s.register_cost(clouds[0] + 0.9)
        with pytest.raises(RwigError, match="too small for this cost"):
            s.submit_sinkhorn(cid, u, u, 0.01)
        with pytest.raises(RwigError, match="too small for this cost"):
            s.submit_barycenter(cid, np.stack([u, u], 1), np.array([0.5, 0.5]), 0.01)
        r = s.result(s.submit_sinkhorn(cid, u, u, 0.1))  # min C / reg = 9: served
        assert r["return_status"] == 0 and r["err"] <= 1e-6


def test_zero_mass_sinkhorn_is_solved_on_the_support(clouds):
    # the reference reduces the problem to the supports of a and b
    # (src/sinkhorn_impl_log.cpp:225-294): potentials -inf off them
    rng = np.random.default_rng(8)
    C = clouds[1]
    n = C.shape[0]
    a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))
    a[rng.choice(n, n // 3, replace=False)] = 0.0
    b[rng.choice(n, n // 2, replace=False)] = 0.0
    a /= a.sum()
    b /= b.sum()
    with ServeSession(hbm_budget_bytes=1 << 30) as s:
        cid = s.register_cost(C)
        for eps in (0.1, 0.01):
            r = s.result(s.submit_sinkhorn(cid, a, b, eps))
            assert r["return_status"] == 0 and r["err"] <= 1e-6
            assert np.all(np.isneginf(r["f"][a == 0])) and np.all(np.isfinite(r["f"][a > 0]))
            assert np.all(np.isneginf(r["g"][b == 0])) and np.all(np.isfinite(r["g"][b > 0]))
            ia, ib = a > 0, b > 0
            e_host = marginal_err(C[np.ix_(ia, ib)], eps, r["f"][ia], r["g"][ib], a[ia], b[ib])
            assert abs(e_host - r["err"]) <= 1e-9
            o = O.sinkhorn_log(a, b, C, eps, maxiter=5000, zerotol=1e-12)
            assert np.all(np.isneginf(o["f"][a == 0]))
            assert abs(r["loss"] - o["loss"]) / abs(o["loss"]) < TOL_LOSS


def test_zero_mass_barycenter_inputs(clouds):
    # rows where A_s is zero keep a zero scaling, as U = A / KV gives in the
    # reference's parallel barycenter (src/barycenter_impl_parallel.cpp:102-150)
    rng = np.random.default_rng(9)
    C = clouds[0]
    n, S = C.shape[0], 3
    A = rng.dirichlet(np.ones(n), size=S).T.copy()
    for s_ in range(S):
        A[rng.choice(n, n // 4, replace=False), s_] = 0.0
    A /= A.sum(0)
    w = rng.dirichlet(np.ones(S))
    with ServeSession(hbm_budget_bytes=1 << 30) as s:
        cid = s.register_cost(C)
        r = s.result(s.submit_barycenter(cid, A, w, 0.1))
    o = O.barycenter_parallel(A, C, w, None, 0.1, maxiter=20000, zerotol=1e-13)
    assert r["return_status"] == 0 and r["err"] <= 1e-6
    assert np.abs(r["b"] - o["b"]).sum() < TOL_BARY


def test_breakdown_ends_as_failed_with_status_2():
    # supports far apart on a line: every cost entry between them is >= 0.27,
    # so at reg 0.002 the BF16 Gibbs kernel is zero on all of them although
    # each row and column has a zero-cost entry (the guard passes). The
    # scalings overflow, the error turns non-finite, and the request ends
    # failed with the reference's status 2 instead of "converged"
    from rwig_amd.serve import FAILED
    x = np.linspace(0.0, 1.0, 64)
    C = (x[:, None] - x[None]) ** 2
    a = np.where(x <= 0.24, 1.0, 0.0)
    b = np.where(x >= 0.76, 1.0, 0.0)
    a /= a.sum()
    b /= b.sum()
    with ServeSession(hbm_budget_bytes=1 << 30) as s:
        cid = s.register_cost(C)
        rid = s.submit_sinkhorn(cid, a, b, 0.002, max_iter=200)
        assert s.wait(rid) == FAILED
        r = s.result(rid)
        assert r["return_status"] == 2 and not np.isfinite(r["err"])
        ok = s.result(s.submit_sinkhorn(cid, np.full(64, 1 / 64), np.full(64, 1 / 64), 0.05))
        assert ok["return_status"] == 0
        st = s.stats()
    assert st["failed"] == 1 and st["completed"] == 1


def test_graph_replay_equals_eager(clouds):
    # the cycle phases replayed as HIP graphs (default) give the bits of the
    # same launches made directly, over many replays with changing column sets
    rng = np.random.default_rng(6)
    reqs = []
    for k in range(10):
        c = k % 2
        n = clouds[c].shape[0]
        eps = 0.01 if k % 3 == 0 else 0.1
        if k % 4 == 1:
            S = 2 + 2 * (k % 2)
            reqs.append(("b", c, eps, rng.dirichlet(np.ones(n), size=S).T.copy(),
                         rng.dirichlet(np.ones(S))))
        else:
            reqs.append(("s", c, eps, rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(n))))
    outs = []
    for graphs in (False, True):
        with ServeSession(hbm_budget_bytes=2 << 30, graphs=graphs) as s:
            ids = [s.register_cost(C) for C in clouds]
            rids = []
            for kind, c, eps, x, y in reqs:
                if kind == "s":
                    rids.append(s.submit_sinkhorn(ids[c], x, y, eps))
                else:
                    rids.append(s.submit_barycenter(ids[c], x, y, eps))
            outs.append([s.result(r) for r in rids])
            assert s.stats()["cycles"] > 10
    for e, g in zip(*outs):
        assert e["iter"] == g["iter"] and e["err"] == g["err"]
        for key in ("f", "g", "b"):
            if key in e:
                assert np.array_equal(e[key], g[key]), key


# ------------------------------------------------------ C5 at its workload sizes


def bary_host_err(C, eps, F, G, A, w):
    """The engine's barycenter stopping error, recomputed on the host in FP64
    from the returned potentials: sum_s |P_s 1 - A_s|_1 + |P_s^T 1 - b|_1 with
    b = prod_s (P_s^T 1 .* exp(-G_s / eps))^{w_s} (the re-estimated common
    target); also b normalised."""
    S = A.shape[1]
    err, cols = 0.0, []
    for s in range(S):
        P = np.exp((F[:, s, None] + G[None, :, s] - C) / eps)
        err += np.abs(P.sum(1) - A[:, s]).sum()
        cols.append(P.sum(0))
    lb = sum(w[s] * (np.log(cols[s]) - G[:, s] / eps) for s in range(S))
    b = np.exp(lb)
    err += sum(np.abs(c - b).sum() for c in cols)
    return err, b / b.sum()


@pytest.fixture(scope="module")
def c5_pick():
    from rwig_amd import workloads
    sizes = workloads.c5_sizes()
    big = int(np.argmax(sizes))  # n = 7317, the largest of the 32 clouds
    mid = next(k for k, n in enumerate(sizes) if 1000 <= n <= 2000)  # n = 1418
    clouds = workloads.c5_clouds(only={big, mid})
    return clouds, big, mid


def test_c5_largest_clouds_host_recomputed(c5_pick):
    clouds, big, mid = c5_pick
    rng = np.random.default_rng(2008)
    jobs = []
    with ServeSession(hbm_budget_bytes=32 << 30) as s:
        ids = {k: s.register_cost(clouds[k]) for k in (big, mid)}
        nb, nm = clouds[big].shape[0], clouds[mid].shape[0]
        for eps in (0.1, 0.01):
            a, b = rng.dirichlet(np.ones(nb)), rng.dirichlet(np.ones(nb))
            jobs.append(("s", big, eps, a, b, s.submit_sinkhorn(ids[big], a, b, eps)))
            A = np.asfortranarray(rng.dirichlet(np.ones(nb), size=8).T)
            w = rng.dirichlet(np.ones(8))
            jobs.append(("b", big, eps, A, w, s.submit_barycenter(ids[big], A, w, eps)))
        A = np.asfortranarray(rng.dirichlet(np.ones(nm), size=8).T)
        w = rng.dirichlet(np.ones(8))
        jobs.append(("b", mid, 0.01, A, w, s.submit_barycenter(ids[mid], A, w, 0.01)))
        a, b = rng.dirichlet(np.ones(nm)), rng.dirichlet(np.ones(nm))
        jobs.append(("s", mid, 0.01, a, b, s.submit_sinkhorn(ids[mid], a, b, 0.01)))
        res = [s.result(j[-1]) for j in jobs]
    for (kind, k, eps, x, y, _), r in zip(jobs, res):
        C = clouds[k]
        assert r["return_status"] == 0 and r["err"] <= 1e-6
        if kind == "s":
            e = marginal_err(C, eps, r["f"], r["g"], x, y)
        else:
            e, bn = bary_host_err(C, eps, r["F"], r["G"], x, y)
            assert np.abs(bn - r["b"]).sum() <= 1e-12
        print(f"C5 {kind} n={C.shape[0]} eps={eps}: iter {r['iter']} err {r['err']:.3e} "
              f"host {e:.3e}", flush=True)
        assert e <= 1e-6 * 1.0001 and abs(e - r["err"]) <= 1e-9
    # the mid-size Sinkhorn against the oracle run to 1e-12
    (_, k, eps, a, b, _), r = jobs[-1], res[-1]
    o = O.sinkhorn_log(a, b, clouds[k], eps, maxiter=5000, zerotol=1e-12, n_threads=8)
    assert abs(r["loss"] - o["loss"]) / abs(o["loss"]) < TOL_LOSS


def test_c5_alone_equals_mix_at_the_largest_n(c5_pick):
    clouds, big, mid = c5_pick
    rng = np.random.default_rng(2009)
    nb, nm = clouds[big].shape[0], clouds[mid].shape[0]
    a, b = rng.dirichlet(np.ones(nb)), rng.dirichlet(np.ones(nb))
    with ServeSession(hbm_budget_bytes=32 << 30) as s:
        ids = {k: s.register_cost(clouds[k]) for k in (big, mid)}
        alone = s.result(s.submit_sinkhorn(ids[big], a, b, 0.01))
    with ServeSession(hbm_budget_bytes=32 << 30) as s:
        ids = {k: s.register_cost(clouds[k]) for k in (big, mid)}
        others = [s.submit_sinkhorn(ids[big], rng.dirichlet(np.ones(nb)),
                                    rng.dirichlet(np.ones(nb)), 0.01),
                  s.submit_barycenter(ids[big], np.asfortranarray(
                      rng.dirichlet(np.ones(nb), size=4).T), rng.dirichlet(np.ones(4)), 0.01)]
        probe = s.submit_sinkhorn(ids[big], a, b, 0.01)
        others += [s.submit_sinkhorn(ids[mid], rng.dirichlet(np.ones(nm)),
                                     rng.dirichlet(np.ones(nm)), eps) for eps in (0.1, 0.01)]
        mix = s.result(probe)
        for rid in others:
            assert s.result(rid)["return_status"] == 0
    assert mix["iter"] == alone["iter"] and mix["err"] == alone["err"]
    assert np.array_equal(mix["f"], alone["f"]) and np.array_equal(mix["g"], alone["g"])
    assert mix["loss"] == alone["loss"]
