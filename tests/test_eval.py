"""Evaluators: fuzzy cross-entropy, kNN preservation, trustworthiness.

CPU tests pin the host forms (closed forms, sklearn); GPU tests check the
device forms against them on the same inputs.
"""
import numpy as np
import pytest

import umap_eval as E
from helpers import gmm, knn_exact


# ---------------------------------------------------------------------------
# CPU

def test_counter_pairs_match_oracle_rng(oracle):
    """The pair stream is the optimizer's index function applied to the key."""
    n, m, seed = 9973, 300, 11
    i, j = E.counter_pairs(n, m, seed)
    key = int(E._mix(np.uint64(seed) + E._G2))
    L = oracle.lib()
    want = np.array([L.orc_rng_index(key, c, n) for c in range(2 * m)], np.int64)
    assert np.array_equal(i, want[0::2]) and np.array_equal(j, want[1::2])
    assert i.min() >= 0 and max(i.max(), j.max()) < n


def test_fuzzy_ce_two_point_closed_form():
    """Two points joined both ways with weight w at distance d: every sampled
    pair with i != j is that pair, so CE = 2 [w log(w/v) + (1-w) log((1-w)/(1-v))]
    + (4 - 2) (-log(1 - v)), v = 1/(1 + a d^(2b))."""
    a, b, w, d = 1.577, 0.895, 0.6, 1.7
    y = np.array([[0.0, 0.0], [d, 0.0]], np.float32)
    v = 1.0 / (1.0 + a * (float(np.float32(d)) ** 2) ** b)
    want = 2 * (w * np.log(w / v) + (1 - w) * np.log((1 - w) / (1 - v))) + 2 * (-np.log(1 - v))
    got = E.fuzzy_cross_entropy([0, 1, 2], [1, 0], np.float32([w, w]), y, a, b, n_random=1000)
    assert got == pytest.approx(want, rel=1e-6)


def test_fuzzy_ce_attractive_term_grows_with_distance():
    """With no repulsive sample, moving a connected pair apart raises the CE."""
    a, b = 1.577, 0.895
    ip, ix, dv = [0, 1, 2, 2], [1, 0], np.float32([1.0, 1.0])
    ce = [E.fuzzy_cross_entropy(ip, ix, dv, np.float32([[0, 0], [d, 0], [50, 0]]), a, b, n_random=0)
          for d in (0.1, 1.0, 5.0)]
    assert ce[0] < ce[1] < ce[2]


def test_trustworthiness_matches_sklearn():
    from sklearn.manifold import trustworthiness as sk_trust
    rng = np.random.default_rng(5)
    x = rng.normal(size=(300, 10))
    y = x[:, :2] + 0.3 * rng.normal(size=(300, 2))
    ours = E.trustworthiness(x, y, k=7, m=10_000)
    assert ours == pytest.approx(sk_trust(x, y, n_neighbors=7), abs=1e-12)


# ---------------------------------------------------------------------------
# GPU

def _f32_sqdist(y, q):
    """Squared distances from row q to every row, accumulated over dimensions
    in order in f32 (the device's order)."""
    y = np.asarray(y, np.float32)
    acc = np.zeros(y.shape[0], np.float32)
    for d in range(y.shape[1]):
        df = (y[q, d] - y[:, d]).astype(np.float32)
        acc = (acc + df * df).astype(np.float32)
    return acc


@pytest.mark.gpu
def test_gpu_fuzzy_ce_matches_host(gpu):
    U = gpu
    x, _ = gmm(2000, 16, 6, seed=2)
    idx, dst = knn_exact(x, 10)
    cfg = U.UmapConfig()
    cfg.graph.n_neighbors = 10
    m = U.Umap(cfg).learn_manifold(x, idx, dst)
    g = m.graph()
    a, b = m.curve_params()
    y = np.random.default_rng(4).uniform(-10, 10, (2000, 2)).astype(np.float32)
    y[[3, 77, 1500]] = np.nan  # rows skipped by both forms
    dy = U.DeviceArray.from_host(y)
    got = m.fuzzy_cross_entropy(dy, n_random=50_000, seed=9)
    want = E.fuzzy_cross_entropy(g.indptr(), g.indices(), g.data(), y, a, b, n_random=50_000, seed=9)
    assert got["total"] == pytest.approx(want, rel=1e-9)
    assert got["edges"] < g.nnz() and got["pairs"] < 50_000
    dy.free()


@pytest.mark.gpu
@pytest.mark.parametrize("dim,k", [(2, 15), (3, 10), (8, 30)])
def test_gpu_knn_query_exact(gpu, dim, k):
    U = gpu
    n = 5000
    y = np.random.default_rng(dim).normal(size=(n, dim)).astype(np.float32)
    q = np.sort(np.random.default_rng(1).choice(n, 400, replace=False)).astype(np.uint32)
    dy, dq = U.DeviceArray.from_host(y), U.DeviceArray.from_host(q)
    got = U.knn_query_device(dy, dq, k).to_host()
    for r, qi in enumerate(q):
        d2 = _f32_sqdist(y, int(qi))
        d2[qi] = np.inf
        order = np.lexsort((np.arange(n), d2))[:k]
        assert np.array_equal(got[r], order), r


@pytest.mark.gpu
@pytest.mark.parametrize("dim", [16, 24])
def test_gpu_trustworthiness_exact_ranks(gpu, dim):
    """Penalty equals the rank sum from f32 keys (distance, index) on the host,
    and the trustworthiness matches the f64 host evaluator closely."""
    U = gpu
    n, k = 3000, 15
    x, _ = gmm(n, dim, 5, seed=3)
    y = (x[:, :2] + np.random.default_rng(0).normal(0, 2.0, (n, 2))).astype(np.float32)
    q = np.arange(0, n, 7, dtype=np.uint32)
    dx, dy, dq = U.DeviceArray.from_host(x), U.DeviceArray.from_host(y), U.DeviceArray.from_host(q)
    low = U.knn_query_device(dy, dq, k)
    pen = U.trustworthiness_penalty_device(dx, dq, low)
    nb = low.to_host()
    want = 0
    for r, qi in enumerate(q):
        d2 = _f32_sqdist(x, int(qi))
        for l in nb[r]:
            closer = (d2 < d2[l]) | ((d2 == d2[l]) & (np.arange(n) < l))
            closer[qi] = False
            rank = 1 + int(closer.sum())
            want += max(rank - k, 0)
    assert pen == want
    t_gpu = E.trustworthiness_gpu(dx, dy, k=k, m=500, seed=3)
    t_cpu = E.trustworthiness(x, y, k=k, m=500, seed=3)
    assert t_gpu == pytest.approx(t_cpu, abs=1e-4)


@pytest.mark.gpu
def test_gpu_knn_preservation_matches_host(gpu):
    U = gpu
    n = 4000
    x, _ = gmm(n, 8, 5, seed=6)
    idx, _ = knn_exact(x, 16)
    y = (x[:, :2] + np.random.default_rng(2).normal(0, 0.5, (n, 2))).astype(np.float32)
    dy = U.DeviceArray.from_host(y)
    got = E.knn_preservation_gpu(idx, dy, k=15, m=600, seed=3)
    want = E.knn_preservation(idx, y, k=15, m=600, seed=3)
    assert got == pytest.approx(want, abs=2e-3)
