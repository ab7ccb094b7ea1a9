"""Runs the C++ drop-in programs on the GPU:
- tests/cpp/test_dropin: the reference's C++ test flows through include/cstfem_hip.hpp;
- tests/cpp/test_ref_dropin: the reference's own types (its compiled serial core assembles
  the Kirsch L1 system) handed to include/cstfem_hip_reference.hpp -- the adapter solver
  behind cstfem::LinearSolver, the reference's pcg_jacobi loop over the device operator and
  reduction, the device pcg_jacobi, and cst_stiffness bit for bit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "test_dropin")
REF_BIN = os.path.join(HERE, "cpp", "_build", "test_ref_dropin")


def test_cpp_dropin():
    assert os.path.exists(BIN), "build tests/cpp first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout


def test_reference_types_dropin():
    if not os.path.exists(REF_BIN):
        pytest.skip("tests/cpp/_build/test_ref_dropin is built only where the reference's "
                    "headers are mounted")
    r = subprocess.run([REF_BIN, os.path.join(HERE, "golden", "kirsch_L1.msh")],
                       capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout
