"""Test infrastructure only: a restatement of the reference's checkpoint bytes,
used by tests to check the library's encoder and decoder.  Nothing in the
library imports it.

The reference checkpoints with bincode 1.3 (`bincode::serialize`, its
dev-dependency, umap-rs/Cargo.toml:24; used in tests.rs:132-215 and
README.md:60-126) over serde derives:
  Optimizer        optimizer.rs:27-56
  LearnedManifold  manifold.rs:14-38
  FittedUmap       embedding.rs:438-443
  UmapConfig       config.rs:8-224 (ManifoldParams, GraphParams, OptimizationParams)
None of bincode, serde, ndarray or sprs is under /root/reference, and no
checkpoint bytes are either.  Their published formats, restated:
  * bincode 1.3, default options of the free functions: integers fixed width
    little endian, usize as u64; f32/f64 as IEEE bits little endian; bool one
    byte 0/1; Option a tag byte 0/1 then the value; enum a u32 variant index;
    struct and tuple fields back to back in declaration order; sequences a u64
    length then the elements.
  * ndarray 0.17 (feature "serde"): an array is the struct
    {v: u8 = 1, dim: [usize; D] (serialized as a tuple), data: seq}, data in
    logical (row-major) order.
  * sprs 0.11.3 (feature "serde"): CsMatI<f32, u32, usize> is
    {storage: CompressedStorage (CSR = 0), nrows: usize, ncols: usize,
     indptr: IndPtr {storage: Vec<usize>}, indices: Vec<u32>, data: Vec<f32>}.
Parity of these layouts with real reference output is unpinned (no Rust
toolchain or reference-written checkpoint exists here); tests pin this
restatement against hand-assembled bytes.
"""
import struct

import numpy as np


def _u8(v):
    return struct.pack("<B", v)


def _u32(v):
    return struct.pack("<I", v)


def _u64(v):
    return struct.pack("<Q", v)


def _f32(v):
    return np.float32(v).astype("<f4").tobytes()


def _f64(v):
    return np.float64(v).astype("<f8").tobytes()


def _vec(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return _u64(a.size) + a.astype(np.dtype(dtype).newbyteorder("<")).tobytes()


def _array1(a, dtype):
    a = np.asarray(a, dtype=dtype)
    return _u8(1) + _u64(a.shape[0]) + _vec(a, dtype)


def _array2(a, dtype):
    a = np.asarray(a, dtype=dtype)
    return _u8(1) + _u64(a.shape[0]) + _u64(a.shape[1]) + _vec(a.reshape(-1), dtype)


def _opt(v, enc):
    return _u8(0) if v is None else _u8(1) + enc(v)


def manifold_bytes(m: dict) -> bytes:
    """LearnedManifold {graph, sigmas, rhos, n_vertices, a, b}; m holds those fields
    (graph as indptr/indices/data)."""
    n = int(m["n_vertices"])
    return (_u32(0) + _u64(n) + _u64(n) + _vec(m["indptr"], np.uint64) + _vec(m["indices"], np.uint32) +
            _vec(m["data"], np.float32) + _array1(m["sigmas"], np.float32) + _array1(m["rhos"], np.float32) +
            _u64(n) + _f32(m["a"]) + _f32(m["b"]))


def optimizer_bytes(o: dict) -> bytes:
    """Optimizer fields in declaration order (optimizer.rs:28-56)."""
    return (manifold_bytes(o["manifold"]) + _array1(o["head"], np.uint32) + _array1(o["tail"], np.uint32) +
            _array1(o["epochs_per_sample"], np.float64) + _array2(o["embedding"], np.float32) +
            _array1(o["epoch_of_next_sample"], np.float64) +
            _array1(o["epoch_of_next_negative_sample"], np.float64) +
            _array1(o["epochs_per_negative_sample"], np.float64) + _u64(int(o["current_epoch"])) +
            _u64(int(o["total_epochs"])) + _f32(o["gamma"]) + _f32(o["initial_alpha"]) +
            _f64(o["negative_sample_rate"]) + _u32(int(o["metric_type"])))


def config_bytes(c: dict) -> bytes:
    """UmapConfig {n_components, manifold, graph, optimization} (config.rs:206-224);
    c is nested like the struct, with None for an absent Option."""
    mp, gp, op = c["manifold"], c["graph"], c["optimization"]
    return (_u64(c["n_components"]) + _f32(mp["min_dist"]) + _f32(mp["spread"]) + _opt(mp["a"], _f32) +
            _opt(mp["b"], _f32) + _u64(gp["n_neighbors"]) + _f32(gp["local_connectivity"]) +
            _f32(gp["set_op_mix_ratio"]) + _opt(gp["disconnection_distance"], _f32) +
            _u8(1 if gp["symmetrize"] else 0) + _opt(op["n_epochs"], _u64) + _f32(op["learning_rate"]) +
            _u64(op["negative_sample_rate"]) + _f32(op["repulsion_strength"]))


def fitted_bytes(f: dict) -> bytes:
    """FittedUmap {embedding, manifold, config} (embedding.rs:439-443)."""
    return _array2(f["embedding"], np.float32) + manifold_bytes(f["manifold"]) + config_bytes(f["config"])
