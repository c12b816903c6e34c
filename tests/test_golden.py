"""The committed golden fixtures (tests/golden/*.npz) against the CPU
restatement: it must still reproduce every array bit for bit.  The fixtures
are what the device tests compare against (tests/test_gpu_golden.py)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_oracle_reproduces_golden_fixtures(oracle):
    import make_golden
    for name, arrays in make_golden.build_all().items():
        z = np.load(os.path.join(HERE, "golden", name + ".npz"))
        assert sorted(z.files) == sorted(arrays), name
        for key, val in arrays.items():
            a, b = np.asarray(val), z[key]
            assert a.dtype == b.dtype and a.shape == b.shape, (name, key)
            assert a.tobytes() == b.tobytes(), (name, key)
