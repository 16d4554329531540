"""GPU tests of the checkpointed log barycenter (rwig_barycenter_log_ck).

Bars:
- FP64, reference seed: equal to the plain FP64 log barycenter and to the
  oracle to 1e-10 relative, for every checkpoint interval (the recomputation
  replays the forward's own kernels, so the history it rebuilds is the
  forward's).
- FP64, exact seeds: gradients within 1e-5 relative of central finite
  differences (h = 1e-6) of the loss each seed differentiates.
- FP32 (iterates and soft-min passes; FP64 seed, loss and accumulation):
  within TOL_FP32_* of the FP64 result on the same inputs.
- FP32 exponentials with FP64 arguments: within 1e-5 of FP64 on b, the loss,
  grad_w and grad_A where A >= 1e-12; below that the gradient is
  ill-conditioned in FP64 itself (a test shows one ulp of C moving it by
  ~1e-8).
- the log form's "unnormalised" gradient equals the golden-pinned parallel
  barycenter gradient of the oracle (1e-9).
- FP64 soft-min passes and the adjoint's weighted sums: as products with the
  scaled kernel exp(-(C - min C) / reg) when (max C - min C) / reg <= 600, as
  log-sum-exp passes otherwise; the two forms agree at the C3 size (1e-12 on
  b, F, G and the loss, 1e-10 on grad_w and on grad_A where A >= 1e-9), and
  the wide-range fallback matches the oracle to 1e-10.
"""
import numpy as np
import pytest

import rwig_amd
from oracle import oracle as O
from rwig_amd.workloads import c3
from tests.test_gpu_plain import bary_problem, rel

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_FD = 1e-5
H_FD = 1e-6
TOL_FP32_B = 1e-5
TOL_FP32_GRAD = 1e-4  # FP32 cost: bounded by its rounding at eps 0.004 (see the test)
TOL_FP32_EXP = 1e-5


def ck(A, C, w, bext, reg, L, **kw):
    kw.setdefault("zero_tol", 0.0)
    return rwig_amd.barycenter_log_ck(A, C, w, bext, reg=reg, max_iter=L, **kw)


@pytest.mark.parametrize("seed,M,N,S,reg", [(50, 40, 50, 3, 0.05), (51, 120, 100, 5, 0.02)])
@pytest.mark.parametrize("every", [1, 7, 20, 1000])
def test_fp64_reference_seed_matches_plain_and_oracle(seed, M, N, S, reg, every):
    _reference_seed_case(seed, M, N, S, reg, every)


@pytest.mark.parametrize("seed,M,N,S,reg", [(58, 60, 70, 2, 0.05), (59, 64, 64, 16, 0.05),
                                            (60, 50, 40, 32, 0.03)])
def test_fp64_slice_counts_match_plain_and_oracle(seed, M, N, S, reg):
    # the 4-, 16- and 32-slice forms of the scaled passes' column update;
    # M = N = 64 with a cost that is not symmetric (random points)
    _reference_seed_case(seed, M, N, S, reg, 9)


def _reference_seed_case(seed, M, N, S, reg, every):
    A, C, w, bext = bary_problem(seed, M, N, S)
    r = rwig_amd.barycenter_log_ck(A, C, w, bext, reg=reg, with_grad=True, max_iter=1000,
                                   zero_tol=1e-6, checkpoint_every=every)
    p = rwig_amd.barycenter(A, C, w, bext, {"reg": reg, "method": "log", "with_grad": True})
    o = O.barycenter_log(A, C, w, bext, reg, withgrad=True)
    assert r["iter"] == p["iter"] == o["iter"]
    assert r["return_status"] == p["return_status"]
    for k in ["b", "F", "G", "grad_A", "grad_w", "loss"]:
        assert rel(r[k], p[k]) < 1e-12, k
        assert rel(r[k], o[k]) < TOL_FP64, k


def test_c3_fp64_scaled_products_equal_the_log_sum_exp_passes():
    # C3's cost spans 1 / 0.004 = 250 units of reg: the checkpointed form runs
    # its soft-min passes as scaled FP64 products, the plain log barycenter as
    # log-sum-exp passes; 200 fixed iterations
    A, C, w, bext, eps, L = c3_inputs()
    r = rwig_amd.barycenter_log_ck(A, C, w, bext, reg=eps, max_iter=L, zero_tol=0.0,
                                   with_grad=True)
    p = rwig_amd.barycenter(A, C, w, bext, {"reg": eps, "method": "log", "max_iter": L,
                                             "zero_tol": 0.0, "with_grad": True})
    assert r["iter"] == p["iter"] == L
    errs = {k: rel(r[k], p[k]) for k in ["b", "F", "G", "loss", "grad_w"]}
    # grad_A where it is resolvable in FP64 (A >= 1e-9: one ulp of C moves it
    # by ~1e-14 there, by ~1e-8 where A < 1e-20, as the test below shows)
    big = A >= 1e-9
    d = r["grad_A"] - p["grad_A"]
    errs["grad_A(A>=1e-9)"] = np.max(np.abs(d[big])) / np.max(np.abs(p["grad_A"][big]))
    bands = _band_errors(A, r["grad_A"], p["grad_A"])
    print("C3 FP64 scaled products against log-sum-exp passes: "
          + ", ".join(f"{k} {e:.2e}" for k, e in errs.items()) + "; grad_A by band of A: "
          + ", ".join(f"[{lo:g},{hi:g}) {e:.1e}" for (lo, hi), e in bands.items()))
    for k in ["b", "F", "G", "loss"]:
        assert errs[k] < 1e-12, k
    assert errs["grad_w"] < 1e-10 and errs["grad_A(A>=1e-9)"] < 1e-10


def test_fp64_wide_cost_range_keeps_the_log_sum_exp_passes():
    # (max C - min C) / reg = 1,000 or more: the scaled kernel would leave the
    # FP64 range, so the log-sum-exp passes run; results against the oracle
    A, C, w, bext = bary_problem(57, 40, 50, 3)
    C = C * (40.0 / C.max())
    reg = 0.04
    r = rwig_amd.barycenter_log_ck(A, C, w, bext, reg=reg, with_grad=True, max_iter=300,
                                   zero_tol=1e-8, checkpoint_every=9)
    o = O.barycenter_log(A, C, w, bext, reg, withgrad=True, maxiter=300, zerotol=1e-8)
    assert r["iter"] == o["iter"]
    for k in ["b", "F", "G", "grad_A", "grad_w", "loss"]:
        assert rel(r[k], o[k]) < TOL_FP64, k


def test_no_grad_and_zero_iterations():
    A, C, w, bext = bary_problem(52, 30, 30, 2)
    r0 = rwig_amd.barycenter_log_ck(A, C, w, bext, reg=0.1, with_grad=True, max_iter=0)
    p0 = rwig_amd.barycenter(A, C, w, bext, {"reg": 0.1, "method": "log", "with_grad": True,
                                             "max_iter": 0})
    assert r0["iter"] == p0["iter"] == 0
    assert np.all(r0["grad_A"] == 0) and np.all(r0["grad_w"] == 0)
    np.testing.assert_allclose(r0["b"], np.full(30, 1 / 30), rtol=1e-15)
    assert r0["loss"] == pytest.approx(((1 / 30 - bext) ** 2).sum(), rel=1e-14)
    assert rel(r0["loss"], p0["loss"]) < 1e-14
    r1 = rwig_amd.barycenter_log_ck(A, C, w, reg=0.1, max_iter=50)
    p = rwig_amd.barycenter(A, C, w, None, {"reg": 0.1, "method": "log", "max_iter": 50})
    assert r1["iter"] == p["iter"]
    assert rel(r1["b"], p["b"]) < 1e-12


def test_invalid_arguments_raise():
    A, C, w, bext = bary_problem(53, 10, 10, 2)
    with pytest.raises(rwig_amd.RwigError):
        rwig_amd.barycenter_log_ck(A, C, w, bext, reg=0.1, with_grad=True, checkpoint_every=0)
    with pytest.raises(ValueError):
        rwig_amd.barycenter_log_ck(A, C, w, bext, reg=0.1, with_grad=True, seed="other")


def fd_check(A, C, w, bext, reg, L, seed, coords, every=20):
    """Central differences of the seed's loss over A[coords] and all of w."""
    g = ck(A, C, w, bext, reg, L, with_grad=True, seed=seed, checkpoint_every=every)

    def loss(A_, w_):
        if seed == "normalised":
            b = ck(A_, C, w_, None, reg, L)["b"]
            return float(((b - bext) ** 2).sum())
        return ck(A_, C, w_, bext, reg, L, with_grad=True, seed=seed)["loss"]

    assert abs(loss(A, w) - g["loss"]) <= 1e-14 * max(1.0, abs(g["loss"]))
    fdA, an_A = [], []
    for (i, s) in coords:
        Ap, Am = A.copy(order="F"), A.copy(order="F")
        Ap[i, s] += H_FD
        Am[i, s] -= H_FD
        fdA.append((loss(Ap, w) - loss(Am, w)) / (2 * H_FD))
        an_A.append(g["grad_A"][i, s])
    fdw = []
    for s in range(len(w)):
        wp, wm = w.copy(), w.copy()
        wp[s] += H_FD
        wm[s] -= H_FD
        fdw.append((loss(A, wp) - loss(A, wm)) / (2 * H_FD))
    fd = np.array(fdA + fdw)
    an = np.array(an_A + list(g["grad_w"]))
    err = np.linalg.norm(fd - an) / np.linalg.norm(fd)
    print(f"FD check seed={seed} L={L}: |fd - adjoint| / |fd| = {err:.2e}")
    return err


@pytest.mark.parametrize("seed", ["unnormalised", "normalised"])
def test_fp64_gradient_matches_finite_differences(seed):
    A, C, w, bext = bary_problem(54, 64, 64, 4)
    rng = np.random.default_rng(5)
    coords = [(int(rng.integers(64)), int(rng.integers(4))) for _ in range(12)]
    assert fd_check(A, C, w, bext, 0.02, 60, seed, coords, every=7) < TOL_FD


def c3_inputs():
    A, C, w, bext, eps, L = c3()
    return A, C, w, bext, eps, L


def test_c3_fp64_gradient_matches_finite_differences():
    # Gaussian-mixture barycenter on the 64 x 64 grid, S = 8, eps 0.004,
    # L = 200, checkpoints every 20; A coordinates in the bulk of each
    # histogram (A > 1e-4) so that A +- h stays positive
    A, C, w, bext, eps, L = c3_inputs()
    rng = np.random.default_rng(6)
    coords = []
    for s in range(A.shape[1]):
        bulk = np.flatnonzero(A[:, s] > 1e-4)
        coords.append((int(rng.choice(bulk)), s))
    assert fd_check(A, C, w, bext, eps, L, "unnormalised", coords, every=20) < TOL_FD


def _band_errors(A, g, g64):
    out = {}
    for lo, hi in [(1e-9, 1.0), (1e-12, 1e-9), (1e-20, 1e-12), (0.0, 1e-20)]:
        m = (A >= lo) & (A < hi)
        if m.any():
            out[(lo, hi)] = np.linalg.norm((g - g64)[m]) / np.linalg.norm(g64[m])
    return out


def test_c3_fp32exp_against_fp64():
    # FP64 cost, iterates and exponent arguments, FP32 exponentials: within
    # 1e-5 of FP64 on b, the loss, grad_w and every entry of grad_A with
    # A >= 1e-12 (measured: 2e-7 for A >= 1e-9, 2.8e-6 in [1e-12, 1e-9)).
    # Entries below are ill-conditioned in FP64 itself (next test)
    A, C, w, bext, eps, L = c3_inputs()
    r64 = ck(A, C, w, bext, eps, L, with_grad=True, seed="unnormalised", checkpoint_every=20)
    rmx = ck(A, C, w, bext, eps, L, with_grad=True, seed="unnormalised", checkpoint_every=20,
             precision="fp32exp")
    assert rmx["iter"] == r64["iter"] == L
    eb = rel(rmx["b"], r64["b"])
    big = A >= 1e-12
    d = rmx["grad_A"] - r64["grad_A"]
    egA = np.linalg.norm(d[big]) / np.linalg.norm(r64["grad_A"][big])
    egA_max = np.max(np.abs(d[big])) / np.max(np.abs(r64["grad_A"][big]))
    egw = np.linalg.norm(rmx["grad_w"] - r64["grad_w"]) / np.linalg.norm(r64["grad_w"])
    el = abs(rmx["loss"] - r64["loss"]) / abs(r64["loss"])
    bands = _band_errors(A, rmx["grad_A"], r64["grad_A"])
    print(f"C3 FP32-exp vs FP64: b {eb:.2e} grad_A(A>=1e-12) {egA:.2e} (max {egA_max:.2e}) "
          f"grad_w {egw:.2e} loss {el:.2e}; grad_A by band of A: "
          + ", ".join(f"[{lo:g},{hi:g}) {e:.1e}" for (lo, hi), e in bands.items()))
    assert max(eb, el, egA, egA_max, egw) < TOL_FP32_EXP


def test_c3_tail_gradient_is_ill_conditioned_in_fp64():
    # One ulp of C (relative 2.2e-16, alternating sign) moves the FP64
    # gradient by ~1e-14 where A >= 1e-9 but by ~1e-8 where A < 1e-20: a
    # condition number near 1e8 that no arithmetic with FP32 exponentials
    # (argument rounding ~1e-7) can resolve. The FP64 mode is the one for
    # those entries
    A, C, w, bext, eps, L = c3_inputs()
    kw = dict(with_grad=True, seed="unnormalised", checkpoint_every=20)
    r = ck(A, C, w, bext, eps, L, **kw)
    sgn = np.sign(np.sin(np.arange(C.size))).reshape(C.shape)
    rp = ck(A, np.asfortranarray(C * (1 + 2.2e-16 * sgn)), w, bext, eps, L, **kw)
    bands = _band_errors(A, rp["grad_A"], r["grad_A"])
    print("C3 FP64 gradient moved by one ulp of C, by band of A: "
          + ", ".join(f"[{lo:g},{hi:g}) {e:.1e}" for (lo, hi), e in bands.items()))
    assert bands[(1e-9, 1.0)] < 1e-12
    assert bands[(0.0, 1e-20)] > 1e-10


def test_log_form_gradient_equals_the_parallel_barycenter_gradient():
    # In exact arithmetic the log iteration (F = reg log U, G = reg log V) is
    # the parallel barycenter's map, and the "unnormalised" seed is the
    # reference parallel adjoint's seed 2 (b_L - b_ext) with b_L unnormalised
    # (src/barycenter_impl_parallel.cpp:154-207). At a regularisation where
    # the scaling form is stable the two gradients agree; the oracle's
    # parallel form is pinned by the reference's golden vectors
    A, C, w, bext = bary_problem(55, 80, 80, 4)
    for reg, L in [(0.05, 40), (0.02, 120)]:
        r = ck(A, C, w, bext, reg, L, with_grad=True, seed="unnormalised", checkpoint_every=9)
        o = O.barycenter_parallel(A, C, w, bext, reg, withgrad=True, maxiter=L, zerotol=0.0)
        assert o["iter"] == r["iter"] == L
        assert rel(r["b"], o["b"]) < 1e-10
        assert rel(r["grad_A"], o["grad_A"]) < 1e-9
        assert rel(r["grad_w"], o["grad_w"]) < 1e-9


def test_c3_fp32_against_fp64():
    # FP32 storage of C rounds every exponent C / reg by up to
    # 2^-24 max|C| / reg = 1.5e-5 at eps 0.004, which bounds this form's
    # gradient near 1e-5: measured 2.5e-5 for A in [1e-12, 1e-6), 2e-6 above.
    # Below A = 1e-12 its potentials (reg log A) carry the same rounding. The
    # mixed form above meets 1e-5 on every entry.
    A, C, w, bext, eps, L = c3_inputs()
    r64 = ck(A, C, w, bext, eps, L, with_grad=True, seed="unnormalised", checkpoint_every=20)
    r32 = ck(A, C, w, bext, eps, L, with_grad=True, seed="unnormalised", checkpoint_every=20,
             precision="fp32")
    assert r32["iter"] == r64["iter"] == L
    eb = rel(r32["b"], r64["b"])
    big = A >= 1e-12
    d = r32["grad_A"] - r64["grad_A"]
    egA = np.linalg.norm(d[big]) / np.linalg.norm(r64["grad_A"][big])
    egw = np.linalg.norm(r32["grad_w"] - r64["grad_w"]) / np.linalg.norm(r64["grad_w"])
    el = abs(r32["loss"] - r64["loss"]) / abs(r64["loss"])
    print(f"C3 FP32 vs FP64: b {eb:.2e}  grad_A(A>=1e-12) {egA:.2e}  grad_w {egw:.2e}  "
          f"loss {el:.2e}")
    assert eb < TOL_FP32_B and el < TOL_FP32_B
    assert egA < TOL_FP32_GRAD and egw < TOL_FP32_GRAD


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_device_entry_point_equals_the_host_one(precision):
    # rwig_barycenter_log_ck_dev runs the same solve on device buffers (C in
    # the working precision's type): every output equals the host entry
    # point's bit for bit, gradient included, and null outputs are allowed
    import ctypes

    import torch

    from rwig_amd._lib import check, lib

    A, C, w, bext = bary_problem(77, 300, 260, 3)
    reg, L = 0.02, 50
    prec_name = {0: "fp64", 1: "fp32", 2: "fp32exp"}[precision]
    r = ck(A, C, w, bext, reg, L, with_grad=True, seed="unnormalised", checkpoint_every=7,
           precision=prec_name)
    M, S = A.shape
    N = C.shape[1]
    dev = torch.device("cuda:0")
    cdt = torch.float32 if precision == 1 else torch.float64
    Cd = torch.from_numpy(np.asfortranarray(C).ravel(order="F")).to(cdt).to(dev)
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).to(dev)
    wd = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    bd = torch.from_numpy(np.ascontiguousarray(bext)).to(dev)
    out = {k: torch.empty(n, dtype=torch.float64, device=dev)
           for k, n in [("F", M * S), ("G", N * S), ("b", N), ("gA", M * S), ("gw", S),
                        ("loss", 1)]}
    it, er, rs = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    st = torch.cuda.current_stream().cuda_stream
    print(f"rwig_barycenter_log_ck_dev precision {precision}", flush=True)
    check(lib().rwig_barycenter_log_ck_dev(
        out["F"].data_ptr(), out["G"].data_ptr(), out["b"].data_ptr(), out["gA"].data_ptr(),
        out["gw"].data_ptr(), out["loss"].data_ptr(), ctypes.byref(it), ctypes.byref(er),
        Ad.data_ptr(), wd.data_ptr(), Cd.data_ptr(), bd.data_ptr(), M, N, S, reg, 1, L, 0.0, 7,
        1, precision, ctypes.byref(rs), st))
    h = {k: v.cpu().numpy() for k, v in out.items()}
    assert it.value == r["iter"] == L
    assert er.value == r["err"]
    assert np.array_equal(h["b"], r["b"])
    assert np.array_equal(h["F"].reshape(S, M).T, r["F"])
    assert np.array_equal(h["G"].reshape(S, N).T, r["G"])
    assert np.array_equal(h["gA"].reshape(S, M).T, r["grad_A"])
    assert np.array_equal(h["gw"], r["grad_w"])
    assert h["loss"][0] == r["loss"]
    # outputs may be null
    check(lib().rwig_barycenter_log_ck_dev(
        None, None, out["b"].data_ptr(), None, None, None, ctypes.byref(it), ctypes.byref(er),
        Ad.data_ptr(), wd.data_ptr(), Cd.data_ptr(), bd.data_ptr(), M, N, S, reg, 1, L, 0.0, 7,
        1, precision, ctypes.byref(rs), st))
    assert np.array_equal(out["b"].cpu().numpy(), r["b"])
