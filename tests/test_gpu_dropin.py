"""Runs tests/cpp/test_dropin (the reference's C++ test flows against the GPU path through
include/cstfem_hip.hpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "cpp", "_build", "test_dropin")


def test_cpp_dropin():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
