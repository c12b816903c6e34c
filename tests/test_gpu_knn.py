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
