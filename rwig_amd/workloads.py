"""Seeded synthetic inputs of the benchmark configurations (seed = 1000 + id).

C1: vanilla Sinkhorn, n = m = 256 on the 1-D grid x_i = i/255, squared
    distance (max 1), a, b ~ Dirichlet(1) floored at 1e-12, eps 0.05.
C2: one source against many targets: n = 4096 points uniform in [0,1]^2,
    squared Euclidean cost, a uniform, B = 1024 targets each with
    Dirichlet(0.5) mass on a random 5% support (205 points), eps 0.01.
"""
from __future__ import annotations

import numpy as np


def c1(seed=1001):
    rng = np.random.default_rng(seed)
    x = np.arange(256) / 255.0
    C = (x[:, None] - x[None]) ** 2
    a = np.maximum(rng.dirichlet(np.ones(256)), 1e-12)
    b = np.maximum(rng.dirichlet(np.ones(256)), 1e-12)
    return a / a.sum(), b / b.sum(), np.asfortranarray(C), 0.05


def c2(n=4096, B=1024, frac=0.05, seed=1002):
    """Returns (C (n x n, Fortran order), a, supports, weights, eps)."""
    rng = np.random.default_rng(seed)
    X = rng.random((n, 2))
    sq = (X * X).sum(1)
    C = sq[:, None] + sq[None, :] - 2.0 * X @ X.T
    np.maximum(C, 0.0, out=C)
    np.fill_diagonal(C, 0.0)
    a = np.full(n, 1.0 / n)
    k = max(1, int(round(frac * n)))
    supports, weights = [], []
    for _ in range(B):
        s = np.sort(rng.choice(n, k, replace=False))
        w = rng.dirichlet(np.full(k, 0.5))
        w = np.maximum(w, 1e-300)
        supports.append(s)
        weights.append(w / w.sum())
    return np.asfortranarray(C), a, supports, weights, 0.01


def c3(grid=64, S=8, seed=1003):
    """Barycenter with unrolled gradient: S Gaussian-mixture histograms on a
    grid x grid lattice of [0,1]^2 (3 isotropic components each, means in
    [0.2, 0.8]^2, sigma in [0.03, 0.1]), weights ~ Dirichlet(1), squared
    Euclidean cost normalised to max 1, eps 0.004, L = 200 iterations.

    Returns (A (n x S, Fortran order), C (n x n, Fortran order), w, b_ext, eps,
    L); b_ext is a further mixture histogram, the target of the loss.
    """
    rng = np.random.default_rng(seed)
    t = (np.arange(grid) + 0.5) / grid
    X = np.stack(np.meshgrid(t, t, indexing="ij"), -1).reshape(-1, 2)
    n = X.shape[0]

    def mixture():
        h = np.zeros(n)
        for _ in range(3):
            mu = rng.uniform(0.2, 0.8, 2)
            sig = rng.uniform(0.03, 0.1)
            h += np.exp(-((X - mu) ** 2).sum(1) / (2.0 * sig * sig))
        h = np.maximum(h, 1e-300)
        return h / h.sum()

    A = np.stack([mixture() for _ in range(S)], 1)
    w = rng.dirichlet(np.ones(S))
    b_ext = mixture()
    sq = (X * X).sum(1)
    C = np.maximum(sq[:, None] + sq[None, :] - 2.0 * X @ X.T, 0.0)
    np.fill_diagonal(C, 0.0)
    C /= C.max()
    return np.asfortranarray(A), np.asfortranarray(C), w, b_ext, 0.004, 200


def c5_sizes(n_clouds=32, seed=1005):
    """The point counts of c5_clouds (same seed, no matrices built)."""
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(64), np.log(8192), n_clouds))).astype(int)


def c5_clouds(n_clouds=32, seed=1005, only=None):
    """Serving mix, cost side: n_clouds point clouds of n_k points uniform in
    [0,1]^2, n_k = round(exp(U[ln 64, ln 8192])), squared Euclidean cost.
    Returns the list of cost matrices (n_k x n_k, Fortran order); with only
    (a set of indices) the other entries are None, the random stream the same."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(64), np.log(8192), n_clouds))).astype(int)
    out = []
    for k, n in enumerate(sizes):
        X = rng.random((int(n), 2))
        if only is not None and k not in only:
            out.append(None)
            continue
        sq = (X * X).sum(1)
        C = np.maximum(sq[:, None] + sq[None, :] - 2.0 * X @ X.T, 0.0)
        np.fill_diagonal(C, 0.0)
        out.append(np.asfortranarray(C))
    return out


def bary_batch(n=2048, B=256, S=4, seed=1006):
    """Batched barycenters on one cost: n points uniform in [0,1]^2, squared
    Euclidean cost (as the serving clouds), B problems of S histograms
    ~ Dirichlet(1) with weights ~ Dirichlet(1).
    Returns (C (n x n, Fortran order), list of (A (n x S, Fortran order), w))."""
    rng = np.random.default_rng(seed)
    X = rng.random((n, 2))
    sq = (X * X).sum(1)
    C = np.maximum(sq[:, None] + sq[None, :] - 2.0 * X @ X.T, 0.0)
    np.fill_diagonal(C, 0.0)
    probs = [(np.asfortranarray(rng.dirichlet(np.ones(n), size=S).T), rng.dirichlet(np.ones(S)))
             for _ in range(B)]
    return np.asfortranarray(C), probs


def c5_requests(sizes, count, seed=2005):
    """Serving mix, request side: 70% Sinkhorn (a, b ~ Dirichlet(1) on a
    random cloud), 30% barycenter (S in {2, 4, 8}, columns of A and the
    weights ~ Dirichlet(1)), eps in {0.1, 0.01}. Yields dicts."""
    rng = np.random.default_rng(seed)
    for _ in range(count):
        k = int(rng.integers(len(sizes)))
        n = int(sizes[k])
        eps = float(rng.choice([0.1, 0.01]))
        if rng.random() < 0.7:
            yield {"kind": "sinkhorn", "cloud": k, "eps": eps,
                   "a": rng.dirichlet(np.ones(n)), "b": rng.dirichlet(np.ones(n))}
        else:
            S = int(rng.choice([2, 4, 8]))
            yield {"kind": "barycenter", "cloud": k, "eps": eps,
                   "A": np.asfortranarray(rng.dirichlet(np.ones(n), size=S).T),
                   "w": rng.dirichlet(np.ones(S))}


def c4(V=2000, dim=50, K=16, D=512, seed=1004):
    """WDL training step: V words with dim-d embeddings ~ N(0, I), rows L2
    normalised, squared Euclidean cost; K atoms; D documents with lengths
    max(1, Poisson(60)) and word ids from a Zipf(1.1) law truncated to
    1..V, as normalised bags of words; eps 0.1, L = 20.

    Returns (C (V x V), Y (V x D), Alpha0 (V x K), Lambda0 (K x D), eps, L),
    Fortran order."""
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((V, dim))
    E /= np.linalg.norm(E, axis=1, keepdims=True)
    sq = (E * E).sum(1)
    C = np.maximum(sq[:, None] + sq[None, :] - 2.0 * E @ E.T, 0.0)
    np.fill_diagonal(C, 0.0)
    p = np.arange(1, V + 1, dtype=float) ** -1.1
    p /= p.sum()
    Y = np.zeros((V, D))
    for d in range(D):
        n = max(1, int(rng.poisson(60)))
        ids = rng.choice(V, size=n, p=p)
        np.add.at(Y[:, d], ids, 1.0)
    Y /= Y.sum(0, keepdims=True)
    Alpha0 = rng.standard_normal((V, K))
    Lambda0 = rng.standard_normal((K, D))
    return (np.asfortranarray(C), np.asfortranarray(Y), np.asfortranarray(Alpha0),
            np.asfortranarray(Lambda0), 0.1, 20)
