"""The kNN producer against a host restatement of its own definition: squared
distances summed in float32 in dimension order, ranked by (distance, index),
slot 0 = the row itself.  The matrix-core path (dim >= 8) re-ranks its
candidates with exactly that sum, so indices and distances must match bit for
bit -- with duplicated points (exact ties), tiny inputs (fewer points than
the candidate list) and every supported dimension."""
import numpy as np
import pytest

from helpers import gmm

pytestmark = pytest.mark.gpu
F = np.float32


def _host_knn_f32(x, k):
    """Sequential-in-dimension float32 sums, then a stable (distance, index) sort."""
    x = np.asarray(x, F)
    n, dim = x.shape
    acc = np.zeros((n, n), F)
    for d in range(dim):
        df = (x[:, None, d] - x[None, :, d]).astype(F)
        acc = (acc + (df * df).astype(F)).astype(F)
    np.fill_diagonal(acc, -1.0)  # the row itself first
    order = np.lexsort((np.broadcast_to(np.arange(n), (n, n)), acc), axis=1)[:, :k]
    d2 = np.take_along_axis(acc, order, 1)
    d2[:, 0] = 0.0
    return order.astype(np.uint32), np.sqrt(d2).astype(F)


@pytest.mark.parametrize("dim", [8, 16, 20, 32, 64, 128])
@pytest.mark.parametrize("k", [1, 15, 30])
def test_knn_matches_f32_restatement(gpu, dim, k):
    n = 1500
    x, _ = gmm(n, dim, 6, seed=dim + k)
    x[100:130] = x[200:230]  # duplicated points: exact ties, broken by index
    x[7] = x[8] = x[9] = x[10]
    di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
    gi, gd = di.to_host(), dd.to_host()
    fb = gpu.knn_last_fallback_rows()
    hi, hd = _host_knn_f32(x, k)
    bad = int((gi != hi).sum())
    print(f"[knn] dim={dim} k={k} fallback rows={fb} differing={bad}", flush=True)
    assert bad == 0
    assert np.array_equal(gd.view(np.uint32), hd.view(np.uint32))


@pytest.mark.parametrize("dim", [1, 5, 12, 50, 100, 127])
def test_knn_any_width_up_to_128(gpu, dim):
    """Widths without their own kernel are zero-padded to the next one; the
    padding adds +0.0 to every sum, so the result equals the restatement at
    the original width bit for bit."""
    n, k = 900, 15
    x, _ = gmm(n, dim, 4, seed=dim)
    x[50:60] = x[70:80]  # exact ties
    di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
    hi, hd = _host_knn_f32(x, k)
    assert np.array_equal(di.to_host(), hi)
    assert np.array_equal(dd.to_host().view(np.uint32), hd.view(np.uint32))


@pytest.mark.parametrize("dim,k", [(129, 15), (300, 1), (784, 15), (784, 30), (1000, 33)])
def test_knn_wide_rows(gpu, dim, k):
    """Widths above 128 (e.g. 784, MNIST's): one lane per query summing each
    pair in dimension order, so equal to the restatement bit for bit."""
    n = 700
    x, _ = gmm(n, dim, 4, seed=dim + k)
    x[5:9] = x[20:24]  # exact ties
    di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
    hi, hd = _host_knn_f32(x, k)
    assert np.array_equal(di.to_host(), hi)
    assert np.array_equal(dd.to_host().view(np.uint32), hd.view(np.uint32))
    assert gpu.knn_last_fallback_rows() == 0


@pytest.mark.parametrize("n", [1, 5, 20, 40])
def test_knn_tiny_inputs(gpu, n):
    x = np.random.default_rng(n).normal(size=(n, 16)).astype(F)
    k = min(n, 15)
    di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
    hi, hd = _host_knn_f32(x, k)
    assert np.array_equal(di.to_host(), hi) and np.array_equal(dd.to_host().view(np.uint32), hd.view(np.uint32))


def _host_rows_f32(x, rows, k, self_slot):
    """Row-by-row form of the restatement for large n: float32 sums in
    dimension order, (distance, index) order; self_slot puts the row first."""
    x = np.asarray(x, F)
    n, dim = x.shape
    out_i, out_d = [], []
    for q in rows:
        acc = np.zeros(n, F)
        for d in range(dim):
            df = (x[:, d] - x[q, d]).astype(F)
            acc = (acc + (df * df).astype(F)).astype(F)
        acc[q] = np.inf
        order = np.lexsort((np.arange(n), acc))[: k - 1 if self_slot else k]
        d2 = acc[order]
        if self_slot:
            order = np.concatenate([[q], order])
            d2 = np.concatenate([[F(0)], d2]).astype(F)
        out_i.append(order.astype(np.uint32))
        out_d.append(np.sqrt(d2).astype(F))
    return np.array(out_i), np.array(out_d)


@pytest.mark.parametrize("k", [15, 30])
def test_knn_unproven_rows_split_over_spans(gpu, k):
    """Points far from the origin and close together: the matrix-core error
    bound cannot prove any row, so every row is rerun by the scalar kernel,
    whose points are then split into spans (9000 points: 3 spans) and the
    spans' lists merged.  Rows still equal the restatement bit for bit,
    exact ties included."""
    n, dim = 9000, 8
    x, _ = gmm(n, dim, 5, seed=k)
    x = (1000.0 + 0.01 * x).astype(F)
    x[4490:4510] = x[10:30]  # duplicates on both sides of a span boundary
    x[8990:9000] = x[3000:3010]
    di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
    fb = gpu.knn_last_fallback_rows()
    print(f"[knn] unproven rows: fallback rows={fb}", flush=True)
    assert fb > 1000
    rows = np.unique(np.concatenate([np.arange(0, n, 37), np.arange(10, 30), np.arange(4490, 4510),
                                     np.arange(8990, 9000), np.arange(3000, 3010)]))
    hi, hd = _host_rows_f32(x, rows, k, True)
    gi, gd = di.to_host()[rows], dd.to_host()[rows]
    assert np.array_equal(gi, hi)
    assert np.array_equal(gd.view(np.uint32), hd.view(np.uint32))


@pytest.mark.parametrize("dim,k", [(2, 15), (8, 30)])
def test_knn_query_split_over_spans(gpu, dim, k):
    """The query form with few queries against 40,000 points runs 10 spans
    per query block; the merged rows equal the restatement, with exact ties
    placed on both sides of span boundaries."""
    n = 40_000
    y = np.random.default_rng(dim + k).normal(size=(n, dim)).astype(F)
    y[4090:4100] = y[0:10]
    y[39_990:40_000] = y[4090:4100]
    q = np.unique(np.concatenate([np.random.default_rng(2).choice(n, 250, replace=False), np.arange(0, 10),
                                  np.arange(4090, 4100)])).astype(np.uint32)
    got = gpu.knn_query_device(gpu.DeviceArray.from_host(y), gpu.DeviceArray.from_host(q), k).to_host()
    hi, _ = _host_rows_f32(y, q, k, False)
    assert np.array_equal(got, hi)
