"""GPU parity of the plain FP64 form (librwig_hip.so) against the reference's
golden vectors and the FP64 oracle.

Bar: the reference's tolerance (1.49e-7 mean-relative) on its golden
vectors, and 1e-10 relative (max-abs over the array, scaled by its max) against
the oracle on the same seeded inputs, with identical iteration counts.
"""
import numpy as np
import pytest

import rwig_amd
from oracle import oracle as O
from tests.conftest import mean_rel

pytestmark = pytest.mark.gpu

TOL_REF = 10 * np.sqrt(np.finfo(float).eps)
TOL_FP64 = 1e-10


def rel(x, y, floor=1e-6):
    """max |x - y| over the finite entries, relative to max |y|; arrays whose
    entries are all below `floor` (e.g. potentials that are zero up to
    rounding) are compared on the scale `floor`."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    fin = np.isfinite(y)
    assert np.array_equal(np.isfinite(x), fin)
    if not fin.any():
        return 0.0
    scale = max(np.max(np.abs(y[fin])), floor)
    return float(np.max(np.abs(x[fin] - y[fin])) / scale)


def problem(seed, M, N, zeros=0, reg_scale=1.0):
    rng = np.random.default_rng(seed)
    a = rng.dirichlet(np.ones(M))
    b = rng.dirichlet(np.ones(N))
    if zeros:
        a[rng.choice(M, zeros, replace=False)] = 0
        b[rng.choice(N, zeros, replace=False)] = 0
        a /= a.sum()
        b /= b.sum()
    x = rng.random((M, 2))
    y = rng.random((N, 2))
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    return a, b, C


def test_device_available():
    assert rwig_amd.cuda_available()


# ------------------------------------------------------------- golden vectors


def test_golden_sinkhorn(golden):
    g = golden["sinkhorn"]
    a, b, C = np.array(g["a"]), np.array(g["b"]), np.array(g["C"])
    sv = rwig_amd.sinkhorn(a, b, C, {"reg": g["reg"], "method": "vanilla", "with_grad": True,
                                     "verbose": 0})
    for k in ["loss", "P", "u", "v", "grad_a"]:
        assert mean_rel(g["vanilla"][k], sv[k]) < TOL_REF, k
    sl = rwig_amd.sinkhorn(a, b, C, {"reg": g["reg"], "method": "log", "with_grad": True,
                                     "verbose": 0})
    for k in ["loss", "P", "f", "g", "grad_a"]:
        assert mean_rel(g["log"][k], sl[k]) < TOL_REF, k
    assert sv["method"] == "vanilla" and sl["method"] == "log"


def test_golden_barycenter(golden):
    g = golden["barycenter"]
    A, C, w, bext = (np.array(g[k]) for k in ["A", "C", "w", "b_ext"])
    bp = rwig_amd.barycenter(A, C, w, bext, {"reg": g["reg"], "method": "parallel",
                                             "with_grad": True, "verbose": 0})
    for k in ["loss", "U", "V", "grad_A", "grad_w"]:
        assert mean_rel(g["parallel"][k], bp[k]) < TOL_REF, k
    bl = rwig_amd.barycenter(A, C, w, bext, {"reg": g["reg"], "method": "log",
                                             "with_grad": True, "verbose": 0})
    for k in ["loss", "grad_A", "grad_w"]:
        assert mean_rel(g["log"][k], bl[k]) < TOL_REF, k


# ------------------------------------------------------------- oracle parity


@pytest.mark.parametrize("seed,M,N,zeros,reg,wg", [
    (0, 5, 3, 0, 0.1, True),
    (1, 40, 33, 3, 0.05, True),
    (2, 300, 257, 10, 0.05, True),
    (3, 1, 9, 0, 0.2, False),
    (4, 700, 90, 0, 0.02, False),
    # M > 1024 rows in the one-workgroup form (rows looped, no column groups)
    (5, 1500, 80, 0, 0.05, True),
    # past the one-workgroup limit (M N > 2^17): the multi-kernel form
    (6, 600, 450, 5, 0.05, True),
])
def test_sinkhorn_vanilla_matches_oracle(seed, M, N, zeros, reg, wg):
    a, b, C = problem(seed, M, N, zeros)
    r = rwig_amd.sinkhorn(a, b, C, {"reg": reg, "method": "vanilla", "with_grad": wg})
    o = O.sinkhorn_vanilla(a, b, C, reg, withgrad=wg)
    assert r["iter"] == o["iter"]
    assert r["return_status"] == o["return_status"]
    keys = ["P", "u", "v", "loss"] + (["grad_a"] if wg else [])
    for k in keys:
        assert rel(r[k], o[k]) < TOL_FP64, k
    assert abs(r["err"] - o["err"]) <= 1e-9 * max(o["err"], 1e-300) + 1e-16


@pytest.mark.parametrize("seed,M,N,zeros,reg,wg", [
    (10, 5, 3, 0, 0.1, True),
    (11, 37, 41, 4, 0.05, True),
    (12, 200, 300, 7, 0.01, True),
    (13, 1000, 64, 0, 0.01, False),
])
def test_sinkhorn_log_matches_oracle(seed, M, N, zeros, reg, wg):
    a, b, C = problem(seed, M, N, zeros)
    r = rwig_amd.sinkhorn(a, b, C, {"reg": reg, "method": "log", "with_grad": wg})
    o = O.sinkhorn_log(a, b, C, reg, withgrad=wg)
    assert r["iter"] == o["iter"]
    assert r["return_status"] == o["return_status"]
    keys = ["P", "f", "g", "loss"] + (["grad_a"] if wg else [])
    for k in keys:
        assert rel(r[k], o[k]) < TOL_FP64, k


def test_sinkhorn_maxiter_exhaustion_status():
    a, b, C = problem(20, 30, 20)
    r = rwig_amd.sinkhorn(a, b, C, {"reg": 0.01, "method": "vanilla", "max_iter": 3})
    assert r["iter"] == 3 and r["return_status"] == 1
    for M, N in [(30, 20), (600, 450)]:  # one-workgroup and multi-kernel forms
        a, b, C = problem(21, M, N)
        r = rwig_amd.sinkhorn(a, b, C, {"reg": 0.05, "method": "vanilla", "max_iter": 0})
        o = O.sinkhorn_vanilla(a, b, C, 0.05, maxiter=0)
        assert r["iter"] == o["iter"] == 0
        for k in ["P", "u", "v", "loss"]:
            assert rel(r[k], o[k]) < TOL_FP64, k


def bary_problem(seed, M, N, S):
    rng = np.random.default_rng(seed)
    A = rng.dirichlet(np.ones(M), size=S).T.copy()
    w = rng.dirichlet(np.ones(S))
    bext = rng.dirichlet(np.ones(N))
    x = rng.random((M, 2))
    y = rng.random((N, 2))
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    return A, C, w, bext


@pytest.mark.parametrize("seed,M,N,S,reg,wg", [
    (30, 5, 5, 2, 0.1, True),
    (31, 50, 40, 3, 0.1, True),
    (32, 256, 256, 8, 0.05, True),
    (33, 100, 120, 5, 0.1, False),
    (34, 60, 64, 32, 0.1, True),   # 32 measures: the widest slice kernels
])
def test_barycenter_parallel_matches_oracle(seed, M, N, S, reg, wg):
    A, C, w, bext = bary_problem(seed, M, N, S)
    r = rwig_amd.barycenter(A, C, w, bext if wg else None,
                            {"reg": reg, "method": "parallel", "with_grad": wg})
    o = O.barycenter_parallel(A, C, w, bext, reg, withgrad=wg)
    assert r["iter"] == o["iter"]
    keys = ["b", "U", "V"] + (["grad_A", "grad_w", "loss"] if wg else [])
    for k in keys:
        assert rel(r[k], o[k]) < TOL_FP64, k


@pytest.mark.parametrize("seed,M,N,S,reg,wg", [
    (40, 5, 5, 2, 0.1, True),
    (41, 40, 50, 3, 0.05, True),
    (42, 200, 150, 4, 0.02, True),
    (43, 64, 64, 1, 0.1, False),
    (44, 60, 64, 32, 0.1, True),   # 32 measures: one-column blocks at S = 32
    (45, 70, 50, 16, 0.05, True),
])
def test_barycenter_log_matches_oracle(seed, M, N, S, reg, wg):
    A, C, w, bext = bary_problem(seed, M, N, S)
    r = rwig_amd.barycenter(A, C, w, bext if wg else None,
                            {"reg": reg, "method": "log", "with_grad": wg})
    o = O.barycenter_log(A, C, w, bext, reg, withgrad=wg)
    assert r["iter"] == o["iter"]
    keys = ["b", "F", "G"] + (["grad_A", "grad_w", "loss"] if wg else [])
    for k in keys:
        # the potentials are differences of terms of size reg * log(.): an
        # array of them that is zero up to rounding (G at S = 1) is compared
        # on the scale reg, not 1e-6
        floor = reg if k in ("F", "G") else 1e-6
        assert rel(r[k], o[k], floor=floor) < TOL_FP64, k


def test_c1_vanilla_256_config():
    # 1-D grid x_i = i/255, squared distance (max 1), Dirichlet(1) floored at
    # 1e-12, eps 0.05, tol 1e-9, up to 1000 iterations
    rng = np.random.default_rng(1001)
    x = np.arange(256) / 255.0
    C = (x[:, None] - x[None]) ** 2
    a = np.maximum(rng.dirichlet(np.ones(256)), 1e-12); a /= a.sum()
    b = np.maximum(rng.dirichlet(np.ones(256)), 1e-12); b /= b.sum()
    r = rwig_amd.sinkhorn(a, b, C, {"reg": 0.05, "method": "vanilla", "zero_tol": 1e-9,
                                    "with_grad": True})
    o = O.sinkhorn_vanilla(a, b, C, 0.05, withgrad=True, zerotol=1e-9)
    assert r["iter"] == o["iter"]
    for k in ["P", "u", "v", "loss", "grad_a"]:
        assert rel(r[k], o[k]) < TOL_FP64, k


@pytest.mark.parametrize("seed,M,N,zeros,reg,wg", [
    (20, 256, 256, 0, 0.05, True),   # C1's size: 8 workgroups
    (21, 300, 257, 10, 0.05, True),  # ragged rows per workgroup, zero-mass entries
    (22, 40, 33, 3, 0.05, False),
    (23, 1000, 64, 0, 0.02, False),  # the most workgroups
])
def test_vanilla_cooperative_loop_equals_the_one_workgroup_loop(seed, M, N, zeros, reg, wg,
                                                                monkeypatch):
    """Small vanilla problems run over several workgroups (one launch, K and
    K^T split over their LDS, grid-wide barriers through global memory) with
    the one-workgroup loop's arithmetic: the same bits. The adjoint of the
    multi-workgroup form sums its products in another order than the
    multi-kernel adjoint, so grad_a agrees to 1e-12 relative."""
    a, b, C = problem(seed, M, N, zeros)
    ctrl = {"reg": reg, "method": "vanilla", "with_grad": wg}
    r = rwig_amd.sinkhorn(a, b, C, ctrl)
    monkeypatch.setenv("RWIG_VANILLA_FORM", "tiny")
    t = rwig_amd.sinkhorn(a, b, C, ctrl)
    assert r["iter"] == t["iter"] and r["err"] == t["err"]
    for k in ["P", "u", "v", "loss"]:
        assert np.array_equal(r[k], t[k]), k
    if wg:
        g, h = r["grad_a"], t["grad_a"]
        assert np.abs(g - h).max() <= 1e-12 * np.abs(h).max(), np.abs(g - h).max()
