"""The device expf restatement (umap-rs_amd/csrc/libm_expf.hpp) equals
glibc's expf -- the function Rust's f32::exp calls in the reference -- bit for
bit.  Compiled as host code with g++ and compared on the CPU: exhaustively on
every float in [-104, 0] (all membership and bisection arguments are <= 0) and
on a dense stride over the remaining range."""
import os
import struct
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("expf") / "libm_expf_check")
    subprocess.run(["g++", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-builtin", "-std=c++17",
                    "-I", os.path.join(REPO, "umap-rs_amd", "csrc"), os.path.join(REPO, "tests", "libm_expf_check.cpp"),
                    "-o", exe, "-lm"], check=True)
    return exe


def bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def run(exe, lo, hi, stride):
    out = subprocess.run([exe, str(lo), str(hi), str(stride)], check=True, capture_output=True, text=True).stdout
    tested, mism = map(int, out.split())
    return tested, mism


@pytest.mark.slow
def test_exhaustive_nonpositive(checker):
    # -0.0 up to -104.0: bit patterns 0x80000000 .. bits(-104.0)
    tested, mism = run(checker, 0x80000000, bits(-104.0), 1)
    assert tested > 1_100_000_000 and mism == 0


def test_strided_everything_else(checker):
    tested, mism = run(checker, 0, 0xffffffff, 4099)  # every sign, exponent, NaN and inf region
    assert tested > 1_000_000 and mism == 0


def test_special_values(checker):
    for x in (0.0, -0.0, float("inf"), float("-inf"), 88.72283, 88.72284, -103.97207, -103.97208, -87.0, 1e-30):
        b = bits(x)
        assert run(checker, b, b, 1)[1] == 0
