"""Pins for the checkpoint-byte restatement (oracle/bincode_oracle.py): byte
strings assembled by hand from the published bincode 1.3 / ndarray 0.17 /
sprs 0.11 serde layouts, written out field by field below, so that a wrong
field order, tag, length prefix or width fails.  CPU only."""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))
import bincode_oracle as B  # noqa: E402


def _h(s: str) -> bytes:
    return bytes.fromhex("".join(s.split()))


def _hexcol(text):
    """The bytes of a dump whose lines are hex, then an optional # comment."""
    return _h("".join(line.split("#")[0] for line in text.splitlines()))


TINY = {"indptr": [0, 1, 2], "indices": [1, 0], "data": [0.5, 0.5], "sigmas": [1.0, 2.0], "rhos": [0.25, 0.0],
        "n_vertices": 2, "a": 1.5, "b": 0.75}
TINY_HEX = """
    00000000                                             # storage: CompressedStorage::CSR
    0200000000000000 0200000000000000                    # nrows, ncols
    0300000000000000 0000000000000000 0100000000000000 0200000000000000   # indptr: len 3, 0 1 2
    0200000000000000 01000000 00000000                   # indices: len 2, 1 0
    0200000000000000 0000003f 0000003f                   # data: 0.5 0.5
    01 0200000000000000 0200000000000000 0000803f 00000040   # sigmas: v=1, dim (2,), len 2, 1.0 2.0
    01 0200000000000000 0200000000000000 0000803e 00000000   # rhos: 0.25 0.0
    0200000000000000                                     # n_vertices
    0000c03f 0000403f                                    # a = 1.5, b = 0.75
"""

DEFAULT_CONFIG = {"n_components": 2,
                  "manifold": {"min_dist": 0.1, "spread": 1.0, "a": None, "b": None},
                  "graph": {"n_neighbors": 15, "local_connectivity": 1.0, "set_op_mix_ratio": 1.0,
                            "disconnection_distance": None, "symmetrize": True},
                  "optimization": {"n_epochs": None, "learning_rate": 1.0, "negative_sample_rate": 5,
                                   "repulsion_strength": 1.0}}
DEFAULT_CONFIG_HEX = """
    0200000000000000            # n_components
    cdcccc3d 0000803f 00 00     # min_dist 0.1, spread 1.0, a None, b None
    0f00000000000000            # n_neighbors 15
    0000803f 0000803f 00 01     # local_connectivity, set_op_mix_ratio, disconnection None, symmetrize true
    00 0000803f                 # n_epochs None, learning_rate 1.0
    0500000000000000 0000803f   # negative_sample_rate 5, repulsion_strength 1.0
"""


def test_manifold_known_bytes():
    assert B.manifold_bytes(TINY) == _hexcol(TINY_HEX)


def test_default_config_known_bytes():
    assert B.config_bytes(DEFAULT_CONFIG) == _hexcol(DEFAULT_CONFIG_HEX)


def test_options_and_bool():
    c = {**DEFAULT_CONFIG,
         "manifold": {"min_dist": 0.0, "spread": 2.0, "a": 1.0, "b": None},
         "graph": {**DEFAULT_CONFIG["graph"], "disconnection_distance": 0.5, "symmetrize": False},
         "optimization": {**DEFAULT_CONFIG["optimization"], "n_epochs": 300}}
    got = B.config_bytes(c)
    assert got == _hexcol("""
        0200000000000000
        00000000 00000040 01 0000803f 00     # min_dist 0, spread 2, a Some(1.0), b None
        0f00000000000000
        0000803f 0000803f 01 0000003f 00     # disconnection Some(0.5), symmetrize false
        01 2c01000000000000 0000803f         # n_epochs Some(300)
        0500000000000000 0000803f
    """)


def test_optimizer_known_bytes():
    emb = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    o = {"manifold": TINY, "head": [0, 1], "tail": [1, 0], "epochs_per_sample": [1.0, 1.0], "embedding": emb,
         "epoch_of_next_sample": [2.0, 2.0], "epoch_of_next_negative_sample": [0.5, 0.5],
         "epochs_per_negative_sample": [0.2, 0.2], "current_epoch": 1, "total_epochs": 10, "gamma": 1.0,
         "initial_alpha": 1.0, "negative_sample_rate": 5.0, "metric_type": 1}
    tail = _hexcol("""
        01 0200000000000000 0200000000000000 00000000 01000000                   # head
        01 0200000000000000 0200000000000000 01000000 00000000                   # tail
        01 0200000000000000 0200000000000000 000000000000f03f 000000000000f03f   # epochs_per_sample 1.0
        01 0200000000000000 0200000000000000 0400000000000000 0000803f 00000040 00004040 00008040  # embedding 2x2
        01 0200000000000000 0200000000000000 0000000000000040 0000000000000040   # epoch_of_next_sample 2.0
        01 0200000000000000 0200000000000000 000000000000e03f 000000000000e03f   # epoch_of_next_negative 0.5
        01 0200000000000000 0200000000000000 9a9999999999c93f 9a9999999999c93f   # epochs_per_negative 0.2
        0100000000000000 0a00000000000000    # current_epoch 1, total_epochs 10
        0000803f 0000803f                    # gamma, initial_alpha
        0000000000001440                     # negative_sample_rate 5.0
        01000000                             # MetricType::Generic
    """)
    assert B.optimizer_bytes(o) == _hexcol(TINY_HEX) + tail


def test_fitted_known_bytes():
    emb = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    head = _hexcol("01 0200000000000000 0200000000000000 0400000000000000 0000803f 00000040 00004040 00008040")
    assert B.fitted_bytes({"embedding": emb, "manifold": TINY, "config": DEFAULT_CONFIG}) == \
        head + _hexcol(TINY_HEX) + _hexcol(DEFAULT_CONFIG_HEX)
