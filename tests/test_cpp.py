"""Runs the C++ test programs (tests/cpp): the host side on the CPU (plain and under
ASan/UBSan), the engine/runner suite on the GPU, and the compile-only check of the GPU
headers against the reference's own interface headers."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
BUILD = os.path.join(HERE, "_build")
REF = "/root/reference/include/frontend/frontend.hpp"


def binary(name):
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-C", HERE, f"_build/{name}"], check=True, capture_output=True)
    return path


def run(path, timeout=300, env=None):
    p = subprocess.run([path], capture_output=True, text=True, timeout=timeout, env=env)
    print(p.stdout[-4000:])
    print(p.stderr[-4000:])
    return p.returncode


def test_cpp_host():
    assert run(binary("test_host")) == 0


def test_cpp_host_under_asan_ubsan():
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0", UBSAN_OPTIONS="halt_on_error=1")
    assert run(binary("test_host_asan"), env=env) == 0


@pytest.mark.skipif(not os.path.exists(REF), reason="reference tree not present")
def test_gpu_headers_compile_against_reference_headers():
    p = subprocess.run(["make", "-C", HERE, "check-reference"], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


@pytest.mark.gpu
def test_cpp_hip_engine(gpu):
    path = os.path.join(BUILD, "test_hip_engine")
    assert os.path.exists(path), "build the C++ tests first (__graft_entry__.build)"
    assert run(path) == 0
