"""The reference's own test suite (umap-rs/src/tests.rs:40-215) ported to the
GPU build's API, plus the API contract around it (ownership, fit's NaN rows,
serde-style round trips).  The fixture is tests.rs:14-38 with a fixed seed."""
import numpy as np
import pytest

from helpers import ring_test_data

pytestmark = pytest.mark.gpu


def _cfg(U, **opt):
    cfg = U.UmapConfig()
    cfg.graph.n_neighbors = 10
    for k, v in opt.items():
        setattr(cfg.optimization, k, v)
    return cfg


def test_high_level_fit(gpu):  # tests.rs:40-63
    data, idx, d, init = ring_test_data(seed=1)
    fitted = gpu.Umap.new(_cfg(gpu, n_epochs=10)).fit(data, idx, d, init)
    e = fitted.embedding()
    assert e.shape == (50, 2)
    assert np.all(np.isfinite(e) | np.isnan(e))


def test_learn_manifold(gpu):  # tests.rs:65-79
    data, idx, d, _ = ring_test_data(seed=2)
    m = gpu.Umap.new(_cfg(gpu)).learn_manifold(data, idx, d)
    assert m.n_vertices() == 50
    assert m.graph().nnz() > 0
    a, b = m.curve_params()
    assert a > 0 and b > 0


def test_optimizer_step_epochs(gpu):  # tests.rs:81-110
    data, idx, d, init = ring_test_data(seed=3)
    cfg = _cfg(gpu)
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    metric = gpu.EuclideanMetric()
    opt = gpu.Optimizer.new(m, init, 20, cfg, metric.metric_type())
    assert opt.current_epoch() == 0 and opt.remaining_epochs() == 20
    opt.step_epochs(5, metric)
    assert opt.current_epoch() == 5 and opt.remaining_epochs() == 15
    opt.step_epochs(15, metric)
    assert opt.current_epoch() == 20 and opt.remaining_epochs() == 0
    assert opt.into_fitted(cfg).embedding().shape == (50, 2)


def test_optimizer_overstep_panics(gpu):  # tests.rs:112-130
    data, idx, d, init = ring_test_data(seed=4)
    cfg = _cfg(gpu)
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    metric = gpu.EuclideanMetric()
    opt = gpu.Optimizer.new(m, init, 20, cfg, metric.metric_type())
    opt.step_epochs(18, metric)
    with pytest.raises(gpu.UmapPanic, match=r"Cannot step 5 epochs: would exceed total_epochs 20 \(current: 18\)"):
        opt.step_epochs(5, metric)


def test_checkpoint_serialization(gpu):  # tests.rs:132-162
    data, idx, d, init = ring_test_data(seed=5)
    cfg = _cfg(gpu)
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    metric = gpu.EuclideanMetric()
    opt = gpu.Optimizer.new(m, init, 50, cfg, metric.metric_type())
    opt.step_epochs(20, metric)
    serialized = opt.to_bincode()  # bincode::serialize(&opt)
    opt2 = gpu.Optimizer.from_bincode(serialized)
    assert opt2.current_epoch() == 20 and opt2.remaining_epochs() == 30
    opt2.step_epochs(30, metric)
    assert opt2.into_fitted(cfg).embedding().shape == (50, 2)
    opt3 = gpu.Optimizer.from_state(opt.to_state())  # the host-array form
    assert opt3.current_epoch() == 20 and opt3.remaining_epochs() == 30


def test_manifold_serialization(gpu):  # tests.rs:164-188
    data, idx, d, _ = ring_test_data(seed=6)
    m = gpu.Umap.new(_cfg(gpu)).learn_manifold(data, idx, d)
    m2 = gpu.LearnedManifold.from_bincode(m.to_bincode())
    assert gpu.LearnedManifold.from_state(m.to_state()).graph().nnz() == m.graph().nnz()
    assert m2.curve_params() == m.curve_params()
    assert m2.n_vertices() == m.n_vertices()
    assert m2.graph().nnz() == m.graph().nnz()
    assert np.array_equal(m2.graph().data(), m.graph().data())


def test_fitted_umap_serialization(gpu):  # tests.rs:190-215
    data, idx, d, init = ring_test_data(seed=7)
    fitted = gpu.Umap.new(_cfg(gpu, n_epochs=10)).fit(data, idx, d, init)
    f2 = gpu.FittedUmap.from_bincode(fitted.to_bincode())
    assert gpu.FittedUmap.from_state(fitted.to_state()).config() == fitted.config()
    assert f2.embedding().shape == fitted.embedding().shape
    assert f2.manifold().n_vertices() == fitted.manifold().n_vertices()
    assert f2.config() == fitted.config()


# --------------------------------------------------------------- contract
def test_fit_equals_low_level_path_and_nans_disconnected_rows(gpu):
    data, idx, d, init = ring_test_data(seed=8, n=200, k=10)
    d[17, :] = np.inf  # vertex 17 loses its own edges...
    idx[:, :] = np.where(idx == 17, 18, idx)  # ...and nobody points at it
    cfg = _cfg(gpu, n_epochs=30)
    fitted = gpu.Umap.new(cfg).fit(data, idx, d, init, seed=4)
    e = fitted.embedding()
    assert np.all(np.isnan(e[17])) and np.all(np.isfinite(np.delete(e, 17, 0)))
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    opt = gpu.Optimizer.new(m, init, 30, cfg, seed=4)
    opt.set_mode(gpu.SGD_SERIAL)
    opt.step_epochs(30)
    low = opt.into_fitted(cfg).embedding()
    assert np.all(np.isfinite(low[17]))  # into_fitted alone does not mask (embedding.rs:327-334 is in fit)
    assert fitted.manifold().disconnections()[1] == 1


def test_consumed_handles(gpu):
    data, idx, d, init = ring_test_data(seed=9)
    cfg = _cfg(gpu)
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    keep = m.clone()
    opt = gpu.Optimizer.new(m, init, 10, cfg)
    with pytest.raises(gpu.UmapPanic, match="moved"):
        m.n_vertices()
    assert keep.n_vertices() == 50
    opt.step_epochs(10)
    f = opt.into_fitted(cfg)
    with pytest.raises(gpu.UmapPanic, match="moved"):
        opt.current_epoch()
    e = f.into_embedding()
    assert e.shape == (50, 2)


def test_optimizer_new_rejects_wrong_init(gpu):
    data, idx, d, init = ring_test_data(seed=10)
    cfg = _cfg(gpu)
    m = gpu.Umap.new(cfg).learn_manifold(data, idx, d)
    with pytest.raises(gpu.UmapPanic, match="init has 49 samples"):
        gpu.Optimizer.new(m, init[:49], 10, cfg)
    assert m.n_vertices() == 50  # not consumed on failure


def test_knn_bruteforce_matches_cpu(gpu):
    from helpers import gmm, knn_exact
    for dim, k in ((16, 15), (32, 30), (64, 15)):
        x, _ = gmm(3000, dim, 8, seed=dim)
        di, dd = gpu.knn_bruteforce_device(gpu.DeviceArray.from_host(x), k)
        gi, gd = di.to_host(), dd.to_host()
        ci, cd = knn_exact(x, k)
        assert np.all(gi[:, 0] == np.arange(3000)) and np.all(gd[:, 0] == 0)
        assert np.allclose(gd, cd, rtol=1e-4, atol=1e-4)
        assert (gi == ci).mean() > 0.999  # differences only at near-ties
