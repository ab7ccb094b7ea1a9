import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mpi-cst-fem-solver_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle(oracle_lib.ORACLE_PATH, "oc")


@pytest.fixture(scope="session")
def ref():
    import oracle_lib
    if not os.path.exists(oracle_lib.REF_PATH):
        pytest.skip("oracle/_ref is not built (the reference sources were absent at build time)")
    return oracle_lib.Oracle(oracle_lib.REF_PATH, "ref")


@pytest.fixture(scope="session")
def dev():
    import cstfem_hip
    d = cstfem_hip.Device(0)
    yield d
    d.close()
