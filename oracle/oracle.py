"""ctypes wrapper of the CPU restatement in ``umap_oracle.c``.

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
library never imports it.  Each function below states which reference stage it
restates (paths are relative to the reference tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64, u32, i32, f32, f64 = C.c_uint64, C.c_uint32, C.c_int32, C.c_float, C.c_double
P = C.c_void_p


def build() -> None:
    """Compile the restatement with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_free.argtypes = [P]
        L.orc_unrolled_sum_f32.argtypes = [P, u64]
        L.orc_unrolled_sum_f32.restype = f32
        L.orc_ndarray_mean_2d.argtypes = [P, u64, u64, u64]
        L.orc_ndarray_mean_2d.restype = f32
        L.orc_find_ab_params.argtypes = [f32, f32, i32, C.POINTER(f32), C.POINTER(f32), C.POINTER(i32)]
        L.orc_smooth_knn_dist.argtypes = [P, u64, u64, u64, u64, i32, f32, f32, P, P]
        L.orc_membership_csr.argtypes = [u64, u64, P, P, u64, f32, P, P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]
        L.orc_membership_csr.restype = u64
        L.orc_set_operations.argtypes = [u64, P, P, P, f32, C.POINTER(P), C.POINTER(P), C.POINTER(P)]
        L.orc_set_operations.restype = u64
        L.orc_zero_rows.argtypes = [u64, P, P, P]
        L.orc_zero_rows.restype = u64
        L.orc_optimizer_prepare.argtypes = [u64, P, P, P, u64, u64, P, u64] + [C.POINTER(P)] * 6 + [C.POINTER(f32)]
        L.orc_optimizer_prepare.restype = u64
        L.orc_rng_epoch_key.argtypes = [u64, u64]
        L.orc_rng_epoch_key.restype = u64
        L.orc_rng_edge_key.argtypes = [u64, u64]
        L.orc_rng_edge_key.restype = u64
        L.orc_rng_index.argtypes = [u64, u64, u32]
        L.orc_rng_index.restype = u32
        epoch_args = [P, P, u64, P, P, u64, u32, P, P, P, P, f32, f32, f32, f32, u64, u64]
        L.orc_sgd_epoch_serial.argtypes = epoch_args + [P]
        L.orc_sgd_epoch_hogwild.argtypes = epoch_args + [P]
        L.orc_sgd_epoch_generic_euclid.argtypes = epoch_args
        L.orc_schedule_replay.argtypes = [P, P, P, P, u64, u64, u64, P]
        L.orc_set_force_denominator.argtypes = [i32]
        L.orc_step_epochs.argtypes = [P, u64, u64, P, P, u64, P, P, P, P, f32, f32, f32, f32, u64, u64, u64, u64, i32, P,
                                      i32]
        L.orc_num_threads.restype = i32
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(P)


def _take(p, n, dtype):
    """Copy a malloc'ed C array into numpy and free it."""
    L = lib()
    if n == 0:
        L.orc_free(p)
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(p.value)
    out = np.frombuffer(bytes(buf), dtype=dtype).copy()
    L.orc_free(p)
    return out


def num_threads() -> int:
    return int(lib().orc_num_threads())


def unrolled_sum(x) -> np.float32:
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.float32(lib().orc_unrolled_sum_f32(_ptr(x), x.size))


def ndarray_mean(x) -> np.float32:
    """ndarray 0.17 ``mean()`` (smooth_knn_dist.rs:74)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, k = x.shape
    return np.float32(lib().orc_ndarray_mean_2d(_ptr(x), n, k, k))


def find_ab_params(spread: float, min_dist: float, n_iterations: int = 1000):
    """umap/find_ab_params.rs:9-82.  Returns (a, b, iterations_run)."""
    a, b, it = f32(), f32(), i32()
    lib().orc_find_ab_params(spread, min_dist, n_iterations, C.byref(a), C.byref(b), C.byref(it))
    return np.float32(a.value), np.float32(b.value), it.value


def smooth_knn_dist(dists, k=None, n_iter=64, local_connectivity=1.0, bandwidth=1.0):
    """umap/smooth_knn_dist.rs:60-198.  Returns (sigmas, rhos) as float32."""
    d = np.ascontiguousarray(dists, dtype=np.float32)
    n, kk = d.shape
    k = kk if k is None else k
    sig = np.zeros(n, np.float32)
    rho = np.zeros(n, np.float32)
    lib().orc_smooth_knn_dist(_ptr(d), n, kk, kk, k, n_iter, local_connectivity, bandwidth, _ptr(sig), _ptr(rho))
    return sig, rho


def membership_csr(knn_indices, knn_dists, sigmas, rhos, disconnection_distance=np.inf):
    """umap/fuzzy_simplicial_set.rs:171-306.  Returns (indptr u64, indices u32, data f32)."""
    idx = np.ascontiguousarray(knn_indices, dtype=np.uint32)
    d = np.ascontiguousarray(knn_dists, dtype=np.float32)
    sig = np.ascontiguousarray(sigmas, dtype=np.float32)
    rho = np.ascontiguousarray(rhos, dtype=np.float32)
    n, k = d.shape
    pi, pj, pd = P(), P(), P()
    nnz = lib().orc_membership_csr(n, k, _ptr(idx), _ptr(d), k, disconnection_distance, _ptr(sig), _ptr(rho),
                                   C.byref(pi), C.byref(pj), C.byref(pd))
    return _take(pi, n + 1, np.uint64), _take(pj, nnz, np.uint32), _take(pd, nnz, np.float32)


def set_operations(indptr, indices, data, set_op_mix_ratio=1.0):
    """umap/fuzzy_simplicial_set.rs:392-552 (with the CSC structure :310-370)."""
    ip = np.ascontiguousarray(indptr, dtype=np.uint64)
    ix = np.ascontiguousarray(indices, dtype=np.uint32)
    dv = np.ascontiguousarray(data, dtype=np.float32)
    n = ip.size - 1
    pi, pj, pd = P(), P(), P()
    nnz = lib().orc_set_operations(n, _ptr(ip), _ptr(ix), _ptr(dv), set_op_mix_ratio, C.byref(pi), C.byref(pj),
                                   C.byref(pd))
    return _take(pi, n + 1, np.uint64), _take(pj, nnz, np.uint32), _take(pd, nnz, np.float32)


def zero_rows(indptr, data):
    """Rows with sum(w) == 0 (embedding.rs:216-222).  Returns (count, flags)."""
    ip = np.ascontiguousarray(indptr, dtype=np.uint64)
    dv = np.ascontiguousarray(data, dtype=np.float32)
    n = ip.size - 1
    flags = np.zeros(max(n, 1), np.uint8)
    c = lib().orc_zero_rows(n, _ptr(ip), _ptr(dv), _ptr(flags))
    return int(c), flags[:n].astype(bool)


class Manifold:
    """LearnedManifold (manifold.rs:15-38) as numpy arrays."""

    def __init__(self, indptr, indices, data, sigmas, rhos, n_vertices, a, b, edges_removed=0,
                 vertices_disconnected=0):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.sigmas, self.rhos = sigmas, rhos
        self.n_vertices, self.a, self.b = n_vertices, np.float32(a), np.float32(b)
        self.edges_removed, self.vertices_disconnected = edges_removed, vertices_disconnected

    @property
    def nnz(self):
        return int(self.indptr[-1])


def learn_manifold(knn_indices, knn_dists, n_neighbors=None, local_connectivity=1.0, set_op_mix_ratio=1.0,
                   symmetrize=True, disconnection_distance=np.inf, min_dist=0.1, spread=1.0, a=None, b=None,
                   timings=None):
    """Umap::learn_manifold (embedding.rs:149-240) without the validation panics.
    timings (a dict), if given, receives each stage's wall seconds under the
    reference's tracing span names."""
    import time
    clock = [time.perf_counter()]

    def stage(name):
        now = time.perf_counter()
        if timings is not None:
            timings[name] = now - clock[0]
        clock[0] = now

    d = np.ascontiguousarray(knn_dists, dtype=np.float32)
    n, k = d.shape
    n_neighbors = k if n_neighbors is None else n_neighbors
    if a is not None and b is not None:
        aa, bb = np.float32(a), np.float32(b)
    else:
        aa, bb, _ = find_ab_params(spread, min_dist)
    stage("find_ab_params")
    edges_removed = int(np.count_nonzero(d >= np.float32(disconnection_distance)))
    stage("disconnection detection")
    sig, rho = smooth_knn_dist(d, n_neighbors, 64, local_connectivity, 1.0)
    stage("smooth_knn_dist")
    ip, ix, dv = membership_csr(knn_indices, d, sig, rho, disconnection_distance)
    stage("build_membership_csr")
    if symmetrize:
        ip, ix, dv = set_operations(ip, ix, dv, set_op_mix_ratio)
    stage("set_operations")
    vd, _ = zero_rows(ip, dv)
    stage("disconnected vertices")
    return Manifold(ip, ix, dv, sig, rho, n, aa, bb, edges_removed, vd)


class OptimizerState:
    """Optimizer (optimizer.rs:28-56) state after Optimizer::new (:74-290)."""

    def __init__(self, manifold: Manifold, init, total_epochs, negative_sample_rate=5, learning_rate=1.0,
                 repulsion_strength=1.0, seed=0, semantics=0):
        emb = np.array(init, dtype=np.float32, order="C", copy=True)
        n, dim = emb.shape
        ph, pt, pe, po, pn, pon = (P() for _ in range(6))
        thr = f32()
        m = manifold
        ne = lib().orc_optimizer_prepare(n, _ptr(m.indptr), _ptr(m.indices), _ptr(m.data), total_epochs,
                                         negative_sample_rate, _ptr(emb), dim, C.byref(ph), C.byref(pt),
                                         C.byref(pe), C.byref(po), C.byref(pn), C.byref(pon), C.byref(thr))
        self.head = _take(ph, ne, np.uint32)
        self.tail = _take(pt, ne, np.uint32)
        self.eps = _take(pe, ne, np.float64)
        self.eons = _take(po, ne, np.float64)
        self.epns = _take(pn, ne, np.float64)
        self.eonns = _take(pon, ne, np.float64)
        self.threshold = np.float32(thr.value)
        self.embedding = emb
        self.n, self.dim = n, dim
        self.a, self.b = m.a, m.b
        self.gamma = np.float32(repulsion_strength)
        self.initial_alpha = np.float32(learning_rate)
        self.total_epochs = total_epochs
        self.current_epoch = 0
        self.seed = seed
        self.semantics = semantics  # bit 1: shared array; bit 2: umap-learn force denominators
        self.counts = np.zeros(2, np.uint64)

    @property
    def n_edges(self):
        return self.head.size

    def step_epochs(self, n_epochs: int, mode: str = "serial"):
        """Optimizer::step_epochs (optimizer.rs:297-366); mode 'serial', 'hogwild', or 'sync' (the
        build's row-owner synchronous epoch, restated for bit-level checks of the device default);
        'generic_serial' and 'generic_sync' are the same two epochs with MetricType::Generic and the
        Euclidean metric (optimize_layout_generic.rs:35-124, distances.rs:22-38)."""
        if self.current_epoch + n_epochs > self.total_epochs:
            raise ValueError(f"Cannot step {n_epochs} epochs: would exceed total_epochs {self.total_epochs} "
                             f"(current: {self.current_epoch})")
        lib().orc_step_epochs(_ptr(self.embedding), self.n, self.dim, _ptr(self.head), _ptr(self.tail),
                              self.n_edges, _ptr(self.eps), _ptr(self.eons), _ptr(self.epns), _ptr(self.eonns),
                              self.a, self.b, self.gamma, self.initial_alpha, self.current_epoch, n_epochs,
                              self.total_epochs, self.seed, _STEP_MODES[mode], _ptr(self.counts),
                              self.semantics)
        self.current_epoch += n_epochs

    def step_epochs_generic_euclid(self, n_epochs: int):
        """Optimizer::step_epochs with MetricType::Generic and the Euclidean metric
        (optimize_layout_generic.rs:35-124)."""
        copy = self.embedding.copy()
        for e in range(self.current_epoch, self.current_epoch + n_epochs):
            alpha = np.float32(self.initial_alpha * (np.float32(1.0) - np.float32(e) / np.float32(self.total_epochs)))
            lib().orc_sgd_epoch_generic_euclid(_ptr(self.embedding), _ptr(copy), self.dim, _ptr(self.head),
                                               _ptr(self.tail), self.n_edges, self.n, _ptr(self.eps),
                                               _ptr(self.eons), _ptr(self.epns), _ptr(self.eonns), self.a, self.b,
                                               self.gamma, alpha, e, self.seed)
        self.current_epoch += n_epochs


_STEP_MODES = {"serial": 0, "hogwild": 1, "sync": 2, "generic_sync": 3, "generic_serial": 4}


def schedule_replay(eps, epns, eons, eonns, start_epoch: int, n_epochs: int):
    """The schedule recurrence alone over epochs [start, start + n) (orc_schedule_replay):
    returns (eons, eonns, positive updates, negative samples); the inputs are not modified."""
    eps = np.ascontiguousarray(eps, np.float64)
    epns = np.ascontiguousarray(epns, np.float64)
    a = np.array(eons, np.float64, copy=True)
    c = np.array(eonns, np.float64, copy=True)
    counts = np.zeros(2, np.uint64)
    lib().orc_schedule_replay(_ptr(eps), _ptr(epns), _ptr(a), _ptr(c), eps.size, start_epoch, n_epochs, _ptr(counts))
    return a, c, int(counts[0]), int(counts[1])


def rng_index(seed: int, epoch: int, edge: int, p: int, n: int) -> int:
    L = lib()
    return int(L.orc_rng_index(L.orc_rng_edge_key(L.orc_rng_epoch_key(seed, epoch), edge), p, n))
