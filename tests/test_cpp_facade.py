"""The C++ facade (umap-rs_amd/umap.hpp) compiles against the C ABI, and the
reference's own tests (tests.rs), ported to it in tests/cpp/, pass on the GPU."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "umap-rs_amd")


def _build(out, src="test_reference_api.cpp"):
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(REPO, "include"),
                    "-I", PKG, os.path.join(REPO, "tests", "cpp", src), "-o", out,
                    "-L", PKG, "-lumapgpu", f"-Wl,-rpath,{PKG}"], check=True)


def test_facade_compiles_and_links(tmp_path, U):
    _build(str(tmp_path / "test_reference_api"))
    _build(str(tmp_path / "test_facade_routes"), "test_facade_routes.cpp")


@pytest.mark.gpu
def test_reference_unit_tests_through_cpp_facade(tmp_path, gpu):
    exe = str(tmp_path / "test_reference_api")
    _build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("PASS") == 8


@pytest.mark.gpu
def test_metric_routing_through_cpp_facade(tmp_path, gpu):
    exe = str(tmp_path / "test_facade_routes")
    _build(exe, "test_facade_routes.cpp")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("PASS") == 2
