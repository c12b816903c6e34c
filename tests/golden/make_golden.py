"""Writes the golden fixtures of the fit path (tests/golden/*.npz).

The reference (a Rust crate) cannot be built or run in this image and ships
no golden vectors of its own, so the fixtures hold the CPU restatement's
outputs (oracle/umap_oracle.c, glibc libm, -ffp-contract=off) on small seeded
inputs, after that restatement has been pinned by tests/test_oracle_pins.py
(closed forms, brute force, finite differences, library fits).  The device
tests compare against these files without running the oracle; the CPU suite
checks that the oracle still reproduces them bit for bit.

    python tests/golden/make_golden.py      # rewrites the .npz files

Fixtures:
  graph_gmm300.npz   300 points (16-D, 4-component mixture), exact kNN k = 10:
                     the kNN inputs, sigma, rho, symmetric CSR, a, b.
  graph_ring50.npz   the reference's tests.rs generator shape (ring neighbours,
                     unsorted distances, no self slot), k = 10, mix 0.5, local
                     connectivity 2: sigma, rho, CSR.
  sgd_gmm300.npz     Optimizer::new state of graph_gmm300 with a seeded init
                     (200 total epochs), then the embedding after 5 serial
                     epochs and, separately, after 5 epochs of the row-owner
                     synchronous epoch (the device default), with counts.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
from helpers import gmm, knn_exact, ring_test_data  # noqa: E402

F = np.float32


def graph_gmm300():
    x, _ = gmm(300, 16, 4, seed=21)
    idx, d = knn_exact(x, 10)
    m = O.learn_manifold(idx, d)
    return dict(knn_indices=idx, knn_dists=d, sigmas=m.sigmas, rhos=m.rhos, indptr=m.indptr, indices=m.indices,
                data=m.data, a=np.float32(m.a), b=np.float32(m.b)), m


def graph_ring50():
    _, idx, d, _ = ring_test_data(seed=5, n=50, k=10)
    m = O.learn_manifold(idx, d, set_op_mix_ratio=0.5, local_connectivity=2.0)
    return dict(knn_indices=idx, knn_dists=d, sigmas=m.sigmas, rhos=m.rhos, indptr=m.indptr, indices=m.indices,
                data=m.data)


def sgd_gmm300(m):
    init = np.random.default_rng(22).uniform(-10, 10, (300, 2)).astype(F)
    out = dict(init=init, total_epochs=np.uint64(200), seed=np.uint64(7))
    st = O.OptimizerState(m, init, 200, seed=7)
    out.update(head=st.head, tail=st.tail, eps=st.eps, epns=st.epns, normalised=st.embedding.copy())
    st.step_epochs(5, "serial")
    out.update(serial5=st.embedding.copy(), serial5_counts=st.counts.copy())
    st = O.OptimizerState(m, init, 200, seed=7)
    st.step_epochs(5, "sync")
    out.update(sync5=st.embedding.copy(), sync5_counts=st.counts.copy(), sync5_eons=st.eons.copy(),
               sync5_eonns=st.eonns.copy())
    return out


def build_all():
    g, m = graph_gmm300()
    return {"graph_gmm300": g, "graph_ring50": graph_ring50(), "sgd_gmm300": sgd_gmm300(m)}


if __name__ == "__main__":
    for name, arrays in build_all().items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")
