"""CPU tests of the TSVD / WIG oracle and of the host-only WIG aggregation in
librwig_hip.so (rwig_wig_index needs no device).

Pins of the oracle (src/tsvd.cpp:7-86, R/wig.R:2-6, :139-159):
  * tsvd magnitudes: the closed form of the 2 x 2 example of the reference's
    TSVD vignette, and scipy's gesvd driver (another LAPACK routine);
  * "sklearn" signs: scikit-learn's own svd_flip (u-based decision);
  * "auto" statistic: it is rounding noise (Y has every kept component
    removed, src/tsvd.cpp:41), which a restatement that removed only the
    other components would fail;
  * calendar floors: numpy's month/year datetime units and a day-by-day walk
    back to Sunday;
  * the index: hand-computed sums and standardisations.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

SQ5 = np.sqrt(5.0)


def test_tsvd_vignette_closed_form():
    # A = [[1, 3], [2, -4]]: A^T A = [[5, -5], [-5, 25]], lambda_1 = 15 + 5 sqrt 5,
    # v_1 ~ (1, -2 - sqrt 5), Msvd = A v_1 / |v_1|
    A = np.array([[1.0, 3.0], [2.0, -4.0]])
    v = np.array([1.0, -2.0 - SQ5])
    expect = A @ v / np.linalg.norm(v)   # (-2.690, 4.353): largest entry positive
    for flip in (0, 1, 2):
        out = O.tsvd(A, 1, flip)["Msvd"][:, 0]
        assert min(np.abs(out - expect).max(), np.abs(out + expect).max()) < 1e-14
    assert np.abs(O.tsvd(A, 1, 1)["Msvd"][:, 0] - expect).max() < 1e-14
    assert np.isclose(np.linalg.norm(expect), np.sqrt(15 + 5 * SQ5), rtol=1e-15)


@pytest.mark.parametrize("m,n,k", [(30, 8, 1), (8, 30, 3), (25, 25, 5), (40, 6, 9)])
def test_tsvd_sklearn_mode_matches_svd_flip(m, n, k):
    scipy_linalg = pytest.importorskip("scipy.linalg")
    extmath = pytest.importorskip("sklearn.utils.extmath")
    X = np.random.default_rng(m * n + k).standard_normal((m, n))
    U, s, Vt = scipy_linalg.svd(X, lapack_driver="gesvd")
    kk = min(k, m, n)
    U, Vt = extmath.svd_flip(U[:, :kk].copy(), Vt[:kk].copy(), u_based_decision=True)
    ref = U * s[:kk]
    out = O.tsvd(X, k, 1)["Msvd"]
    assert out.shape == (m, kk)
    assert np.abs(out - ref).max() < 1e-12 * s[0]


@pytest.mark.parametrize("m,n,k", [(30, 8, 1), (8, 30, 2), (50, 12, 3)])
def test_tsvd_auto_statistic_is_rounding_noise(m, n, k):
    X = np.random.default_rng(7 + m).random((m, n))
    r = O.tsvd(X, k, 0)
    s1 = np.linalg.norm(X, 2)
    # removing only the other components would leave sk of order s_K^4 / ...,
    # far above this bound
    for sl, sr in zip(r["sk_left"], r["sk_right"]):
        assert abs(sl) < (1e-12 * s1) ** 2 and abs(sr) < (1e-12 * s1) ** 2


def test_tsvd_rank_deficient_columns_are_zero():
    x = np.arange(1.0, 7.0)
    X = np.outer(x, [1.0, 2.0, 0.5])   # rank 1
    out = O.tsvd(X, 3, 1)["Msvd"]
    assert np.abs(out[:, 1:]).max() < 1e-13
    expect = X @ np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5])
    assert np.abs(out[:, 0] - expect).max() < 1e-13


def _walk_to_sunday(day):
    import datetime
    d = datetime.date(1970, 1, 1) + datetime.timedelta(days=int(day))
    while d.weekday() != 6:
        d -= datetime.timedelta(days=1)
    return (d - datetime.date(1970, 1, 1)).days


def test_floor_period_against_numpy_units():
    rng = np.random.default_rng(3)
    for day in list(rng.integers(-40000, 40000, 300)) + [0, -1, 59, 60, 365, 10956, 11016]:
        dd = np.datetime64(int(day), "D")
        month = int(dd.astype("datetime64[M]").astype("datetime64[D]").astype(np.int64))
        year = int(dd.astype("datetime64[Y]").astype("datetime64[D]").astype(np.int64))
        assert O.floor_period(day, "day") == day
        assert O.floor_period(day, "month") == month
        assert O.floor_period(day, "year") == year
        assert O.floor_period(day, "week") == _walk_to_sunday(day)
        m = int(dd.astype("datetime64[M]").astype(np.int64)) % 12   # 0-based month
        for unit, span in (("bimonth", 2), ("quarter", 3), ("halfyear", 6)):
            start = np.datetime64(year, "D").astype("datetime64[M]") + (m // span) * span
            assert O.floor_period(day, unit) == int(start.astype("datetime64[D]").astype(np.int64))
        # seasons: Dec-Feb from 1 December, then March, June, September
        ym = dd.astype("datetime64[M]")
        back = (m + 1) % 3              # months since the season started
        start = (ym - back).astype("datetime64[D]")
        assert O.floor_period(day, "season") == int(start.astype(np.int64))


def test_wig_index_known_answers():
    jan5, jan20, feb1, mar15 = 15344, 15359, 15371, 15414   # 2012
    keys, x = O.wig_index([jan5, jan20, feb1, mar15], [1.0, 2.0, 3.0, 5.0], "month")
    assert list(keys) == [15340, 15371, 15400]
    # sums 3, 3, 5: mean 11/3, sd sqrt(4/3)
    expect = (np.array([3.0, 3.0, 5.0]) - 11 / 3) / np.sqrt(4 / 3) + 100
    assert np.allclose(x, expect, rtol=0, atol=1e-13)
    keys, x = O.wig_index([jan5, feb1], [4.0, -1.0], "month")
    assert np.allclose(x, [100 + np.sqrt(0.5), 100 - np.sqrt(0.5)], atol=1e-13)
    keys, x = O.wig_index([jan5, jan20], [4.0, -1.0], "month")
    assert list(keys) == [15340] and np.isnan(x[0])
    keys, x = O.wig_index([jan5, mar15, jan20], [1.0, 2.0, 3.0], "year", standardize=False)
    assert list(keys) == [15340] and x[0] == 6.0


# ---------------------------------------------------------------- library, host only

def _lib_index(days, scores, unit, standardize):
    from rwig_amd._lib import check, lib
    M = len(days)
    days = np.ascontiguousarray(days, dtype=np.int32)
    scores = np.ascontiguousarray(scores, dtype=np.float64)
    period = np.zeros(M, dtype=np.int32)
    index = np.zeros(M)
    n = ctypes.c_int()
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().rwig_wig_index(period.ctypes.data_as(ip), index.ctypes.data_as(dp),
                               ctypes.byref(n), scores.ctypes.data_as(dp),
                               days.ctypes.data_as(ip), M, unit, int(standardize)))
    return period[:n.value], index[:n.value]


@pytest.mark.parametrize("unit", range(8))
@pytest.mark.parametrize("standardize", [True, False])
def test_library_wig_index_matches_oracle(unit, standardize):
    names = ("day", "week", "month", "bimonth", "quarter", "halfyear", "year", "season")
    rng = np.random.default_rng(unit)
    days = rng.integers(-3000, 20000, 500)
    scores = rng.standard_normal(500) * 10 ** rng.uniform(-3, 3, 500)
    keys, x = _lib_index(days, scores, unit, standardize)
    rkeys, rx = O.wig_index(days, scores, names[unit], standardize)
    assert np.array_equal(keys, rkeys)
    assert np.allclose(x, rx, rtol=1e-15, atol=0)


def test_library_wig_index_errors():
    from rwig_amd import RwigError, wig_index
    with pytest.raises(Exception):
        _lib_index([1, 2], [1.0, 2.0], 8, True)
    with pytest.raises(ValueError):
        wig_index(np.array(["2012-01-01"], dtype="datetime64[D]"), [1.0], "fortnight")
    with pytest.raises(ValueError):
        wig_index([1, 2], [1.0, 2.0])   # not dates
    out = wig_index(np.array(["2012-01-01"], dtype="datetime64[D]"), [1.0])
    assert np.isnan(out["WIG"][0])
    assert issubclass(RwigError, RuntimeError)


def test_tsvd_wrapper_argument_checks():
    from rwig_amd import tsvd
    with pytest.raises(ValueError, match="Features less than 2"):
        tsvd(np.ones((4, 1)))
    with pytest.raises(ValueError, match="should be less than"):
        tsvd(np.ones((4, 3)), k=3)
    with pytest.raises(ValueError):
        tsvd(np.ones((4, 3)), flip_sign="max")
