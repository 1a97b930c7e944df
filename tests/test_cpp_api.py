"""The irlba:: C++ header API: compiles against the C ABI on the CPU, and its
test driver (tests/cpp/test_api.cpp) passes on the GPU."""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "test_api")


def test_headers_compile():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "cppirlba_amd", "include"), SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cpp_api_driver():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
