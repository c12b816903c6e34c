"""Synthetic inputs shared by the tests (seeded)."""
import numpy as np


def ring_test_data(seed=0, n=50, n_features=20, k=10, n_components=2):
    """The reference's generate_test_data (tests.rs:14-38), seeded: ring
    neighbours (i+j+1) % n with no self slot, unsorted U[0,2) distances,
    init U[0,10)."""
    rng = np.random.default_rng(seed)
    data = rng.random((n, n_features), dtype=np.float32)
    idx = np.array([[(i + j + 1) % n for j in range(k)] for i in range(n)], np.uint32)
    dists = (rng.random((n, k), dtype=np.float32) * np.float32(2.0)).astype(np.float32)
    init = (rng.random((n, n_components), dtype=np.float32) * np.float32(10.0)).astype(np.float32)
    return data, idx, dists, init


def gmm(n, dim, n_components, seed=0, mean_scale=5.0, std=1.0):
    """Isotropic Gaussian mixture (means ~ N(0, mean_scale^2 I))."""
    rng = np.random.default_rng(seed)
    means = rng.normal(0.0, mean_scale, (n_components, dim)).astype(np.float32)
    lab = rng.integers(0, n_components, n)
    x = means[lab] + rng.normal(0.0, std, (n, dim)).astype(np.float32)
    return x.astype(np.float32), lab


def knn_exact(x, k):
    """Exact Euclidean kNN on the CPU: self at slot 0, then the k-1 nearest
    other points by (distance, index)."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    sq = (x * x).sum(1)
    idx = np.zeros((n, k), np.uint32)
    dst = np.zeros((n, k), np.float32)
    for s in range(0, n, 1024):
        e = min(n, s + 1024)
        d2 = sq[s:e, None] + sq[None, :] - 2.0 * x[s:e] @ x.T
        d2[np.arange(e - s), np.arange(s, e)] = -1.0
        part = np.argsort(d2, axis=1, kind="stable")[:, :k]
        idx[s:e] = part
        dd = np.sqrt(np.maximum(np.take_along_axis(d2, part, 1), 0.0))
        dd[:, 0] = 0.0
        dst[s:e] = dd
    return idx, dst


def uniform_init(n, dim=2, seed=1, lo=-10.0, hi=10.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (n, dim)).astype(np.float32)
