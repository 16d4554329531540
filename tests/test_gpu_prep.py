"""GPU tests of cost and histogram construction (rwig_euclidean*,
rwig_doc2dist) against the oracle's restatements of src/utils.cpp:17-68.

Bars: distances to 1e-14 relative (summation order only), exact symmetry and
zero diagonal; bags of words bit-identical (integer counts over their sum)."""
import numpy as np
import pytest

import rwig_amd
from oracle import oracle as O
from rwig_amd.workloads import c4

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("V,d", [(1, 3), (17, 5), (100, 50), (33, 1)])
def test_euclidean_matches_oracle(V, d):
    rng = np.random.default_rng(V + d)
    E = rng.standard_normal((V, d))
    D = rwig_amd.euclidean(E)
    ref = O.euclidean(E)
    assert np.array_equal(D, D.T)
    assert np.all(np.diag(D) == 0.0)
    assert np.max(np.abs(D - ref)) <= 1e-14 * max(1.0, np.max(ref))
    D2 = rwig_amd.euclidean(E, squared=True)
    assert np.allclose(D2, ref ** 2, rtol=1e-13, atol=1e-14)


def test_euclidean_feeds_the_wdl_cost():
    # the WDL configuration's cost: squared distances of L2-normalised
    # embeddings (2000 words, 50 dimensions)
    C, *_ = c4()
    rng = np.random.default_rng(1004)
    E = rng.standard_normal((2000, 50))
    E /= np.linalg.norm(E, axis=1, keepdims=True)
    D2 = rwig_amd.euclidean(E, squared=True)
    assert np.max(np.abs(D2 - C)) < 1e-12


def test_doc2dist_matches_oracle():
    rng = np.random.default_rng(3)
    dictionary = [f"w{k}" for k in range(40)] + ["w3"]  # a duplicate entry
    docs = []
    for _ in range(25):
        n = int(rng.integers(1, 30))
        docs.append([f"w{int(k)}" if k < 45 else "unknown" for k in rng.integers(0, 50, n)])
    out = rwig_amd.doc2dist(docs, dictionary)
    ref = O.doc2dist(docs, dictionary)
    assert np.array_equal(out, ref)
    assert np.allclose(out.sum(0), 1.0)


def test_doc2dist_empty_document_is_nan_like_the_reference():
    out = rwig_amd.doc2dist([["a"], []], ["a", "b"])
    assert out[0, 0] == 1.0 and np.all(np.isnan(out[:, 1]))
