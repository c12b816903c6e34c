"""Parity at the 1M-point benchmark size (n = 1,000,000, D = 64, 50-component
mixture, exact kNN k = 15, 2-D, min_dist 0.1, 500 epochs), where the CPU
restatement of the SGD itself would take minutes:

* the graph (CSR row pointers, column indices and weights, sigma and rho) is
  bit-identical to the oracle's, from the same GPU kNN;
* after all 500 epochs of the default parallel epoch, every edge's schedule
  state (epoch_of_next_sample, epoch_of_next_negative_sample) and the
  positive / negative counts equal a CPU replay of the reference's f64
  schedule recurrence exactly (it draws no random numbers and does not depend
  on the embedding), and the embedding is finite.
"""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _log(msg):
    print(f"[fullsize] {msg}", flush=True)


def test_graph_and_schedule_at_one_million_points(gpu, oracle):
    sys.path.insert(0, REPO)
    import bench
    t = time.time()
    n, dim, k, epochs = 1_000_000, 64, 15, 500
    x, _ = bench.make_data(n, dim, 50, seed=0)
    dx = gpu.DeviceArray.from_host(x)
    di, dd = gpu.knn_bruteforce_device(dx, k)
    idx, d = di.to_host(), dd.to_host()
    _log(f"data + GPU kNN {time.time() - t:.1f}s")
    cfg = gpu.UmapConfig()
    cfg.graph.n_neighbors = k
    cfg.manifold.min_dist = 0.1
    m = gpu.learn_manifold_device(cfg, di, dd)
    g = m.graph()
    t = time.time()
    om = oracle.learn_manifold(idx, d)
    _log(f"oracle graph {time.time() - t:.1f}s, nnz {g.nnz()} vs {om.indices.size}")
    assert np.array_equal(g.indptr(), om.indptr)
    assert np.array_equal(g.indices(), om.indices)
    assert np.array_equal(g.data().view(np.uint32), om.data.view(np.uint32))
    s, r = m.sigmas_rhos()
    assert np.array_equal(s.view(np.uint32), om.sigmas.view(np.uint32))
    assert np.array_equal(r.view(np.uint32), om.rhos.view(np.uint32))

    init = np.random.default_rng(1).uniform(-10, 10, (n, 2)).astype(np.float32)
    opt = gpu.Optimizer.new(m, init, epochs, cfg, gpu.MetricType.Euclidean, seed=2)
    st0 = opt.export_state()
    t = time.time()
    opt.step_epochs(epochs)
    _log(f"500 epochs {time.time() - t:.2f}s")
    st = opt.export_state()
    t = time.time()
    a, c, p, q = oracle.schedule_replay(st0["epochs_per_sample"], st0["epochs_per_negative_sample"],
                                        st0["epoch_of_next_sample"], st0["epoch_of_next_negative_sample"], 0, epochs)
    _log(f"schedule replay {time.time() - t:.1f}s: {p} positive, {q} negative")
    assert np.array_equal(st["epoch_of_next_sample"].view(np.uint64), a.view(np.uint64))
    assert np.array_equal(st["epoch_of_next_negative_sample"].view(np.uint64), c.view(np.uint64))
    assert opt.stats()[:2] == (p, q)
    assert np.isfinite(st["embedding"]).all()
