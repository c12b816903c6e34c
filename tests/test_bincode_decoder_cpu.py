"""The checkpoint decoder (umapgpu_*_from_bincode) on malformed input, on the
CPU: it parses and validates on the host before any device work, so every
malformed buffer must come back as UMAPGPU_INVALID_ARGUMENT with a
"bincode: ..." message, never a crash and never an allocation sized by a
corrupt length field.  Valid buffers are built with the restated layout
(oracle/bincode_oracle.py)."""
import ctypes as C
import os
import struct
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import bincode_oracle as B  # noqa: E402

TINY = {"indptr": [0, 1, 2], "indices": [1, 0], "data": [0.5, 0.5], "sigmas": [1.0, 2.0], "rhos": [0.25, 0.0],
        "n_vertices": 2, "a": 1.5, "b": 0.75}
CONFIG = {"n_components": 2, "manifold": {"min_dist": 0.1, "spread": 1.0, "a": None, "b": None},
          "graph": {"n_neighbors": 15, "local_connectivity": 1.0, "set_op_mix_ratio": 1.0,
                    "disconnection_distance": None, "symmetrize": True},
          "optimization": {"n_epochs": None, "learning_rate": 1.0, "negative_sample_rate": 5,
                           "repulsion_strength": 1.0}}
OPT = {"manifold": TINY, "head": [0, 1], "tail": [1, 0], "epochs_per_sample": [1.0, 1.0],
       "embedding": np.array([[1.0, 2.0], [3.0, 4.0]], np.float32), "epoch_of_next_sample": [2.0, 2.0],
       "epoch_of_next_negative_sample": [0.5, 0.5], "epochs_per_negative_sample": [0.2, 0.2], "current_epoch": 1,
       "total_epochs": 10, "gamma": 1.0, "initial_alpha": 1.0, "negative_sample_rate": 5.0, "metric_type": 0}


def _decode(U, kind, data):
    buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data if data else b"\0")
    out = C.c_void_p()
    L = U.lib()
    if kind == "manifold":
        st = L.umapgpu_manifold_from_bincode(buf, len(data), C.byref(out))
    elif kind == "optimizer":
        st = L.umapgpu_optimizer_from_bincode(buf, len(data), 0, C.byref(out))
    else:
        st = L.umapgpu_fitted_from_bincode(buf, len(data), C.byref(out))
    msg = L.umapgpu_last_error().decode()
    if st == U.OK:  # a mutation that stayed a valid checkpoint (only possible with a GPU)
        getattr(L, f"umapgpu_{kind}_destroy")(out)
    return st, msg


def _rejected(U, kind, data, match="bincode: "):
    st, msg = _decode(U, kind, data)
    assert st == U.INVALID_ARGUMENT and msg.startswith(match), (kind, len(data), st, msg)


def test_every_truncation_is_rejected(U):
    cases = {"manifold": B.manifold_bytes(TINY), "optimizer": B.optimizer_bytes(OPT),
             "fitted": B.fitted_bytes({"embedding": OPT["embedding"], "manifold": TINY, "config": CONFIG})}
    for kind, full in cases.items():
        for cut in range(len(full)):
            _rejected(U, kind, full[:cut], "bincode: unexpected end of input")


def test_corrupt_lengths_do_not_allocate(U):
    """Every u64 length prefix of the manifold set to huge values."""
    full = B.manifold_bytes(TINY)
    # offsets of the length prefixes: indptr, indices, data, then the two arrays' dims and lengths
    offs = [20, 20 + 8 + 24, 20 + 8 + 24 + 8 + 8]
    sig = 20 + 8 + 24 + 8 + 8 + 8 + 8
    offs += [sig + 1, sig + 9, sig + 17 + 8 + 1, sig + 17 + 8 + 9]
    for off in offs:
        for big in (2 ** 64 - 1, 2 ** 40, 2 ** 33 + 3):
            bad = full[:off] + struct.pack("<Q", big) + full[off + 8:]
            _rejected(U, "manifold", bad)


def test_tags_and_structure(U):
    full = B.manifold_bytes(TINY)
    for v in (1, 2, 7, 2 ** 31):
        _rejected(U, "manifold", struct.pack("<I", v) + full[4:])  # storage: CSC or unknown
    sig = 20 + 8 + 24 + 8 + 8 + 8 + 8
    for v in (0, 2, 255):
        _rejected(U, "manifold", full[:sig] + bytes([v]) + full[sig + 1:], "bincode: unknown array version")
    # structure: a decreasing indptr, a column out of range, unsorted columns
    _rejected(U, "manifold", B.manifold_bytes({**TINY, "indptr": [0, 2, 1]}))
    _rejected(U, "manifold", B.manifold_bytes({**TINY, "indices": [1, 5]}))
    two = {**TINY, "indptr": [0, 2, 2], "indices": [1, 0], "data": [0.5, 0.5]}
    _rejected(U, "manifold", B.manifold_bytes(two))
    # optimizer: MetricType 2, per-edge arrays of different lengths, embedding rows != n_vertices
    ob = B.optimizer_bytes(OPT)
    _rejected(U, "optimizer", ob[:-4] + struct.pack("<I", 2), "bincode: unknown variant index 2 of MetricType")
    _rejected(U, "optimizer", B.optimizer_bytes({**OPT, "tail": [1]}))
    _rejected(U, "optimizer", B.optimizer_bytes({**OPT, "embedding": np.ones((3, 2), np.float32)}))
    # fitted: bool and Option tags
    fb = B.fitted_bytes({"embedding": OPT["embedding"], "manifold": TINY, "config": CONFIG})
    c0 = len(fb) - len(B.config_bytes(CONFIG))
    sym = c0 + 8 + 4 + 4 + 1 + 1 + 8 + 4 + 4 + 1
    for v in (2, 3, 200):
        _rejected(U, "fitted", fb[:sym] + bytes([v]) + fb[sym + 1:], "bincode: invalid value")
        _rejected(U, "fitted", fb[:c0 + 16] + bytes([v]) + fb[c0 + 17:], "bincode: invalid Option tag")


def test_random_garbage_never_crashes(U):
    """Random buffers (and random overwrites of a valid one) never crash; without
    a GPU none can succeed (a valid one stops at the device)."""
    rng = np.random.default_rng(0)
    full = B.optimizer_bytes(OPT)
    for i in range(3000):
        if i % 2:
            data = rng.integers(0, 256, int(rng.integers(0, 400)), dtype=np.uint8).tobytes()
        else:
            arr = np.frombuffer(full, np.uint8).copy()
            pos = rng.integers(0, arr.size, int(rng.integers(1, 6)))
            arr[pos] = rng.integers(0, 256, pos.size, dtype=np.uint8)
            data = arr[: int(rng.integers(arr.size // 2, arr.size + 1))].tobytes()
        for kind in ("manifold", "optimizer", "fitted"):
            st, _ = _decode(U, kind, data)
            assert st != U.OK or U.device_count() > 0
