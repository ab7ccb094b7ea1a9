"""The C-ABI library loads and exports every function include/cstfem_hip.h declares, and
the Python bindings cover them (no GPU calls)."""
import ctypes
import os
import re

import cstfem_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "cstfem_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cstfem_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(cstfem_hip.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(declared()) <= set(cstfem_hip.SIGNATURES)
    cstfem_hip.lib()  # signatures apply without error


def test_struct_layouts():
    # 8 ints + 8 doubles + 2 long long
    assert ctypes.sizeof(cstfem_hip.SolverStatsC) == 8 * 4 + 8 * 8 + 2 * 8
    # double, 4 ints, double, int (+4 padding)
    assert ctypes.sizeof(cstfem_hip.SolverOptions) == 40
    assert cstfem_hip.SolverOptions.inner_rtol.offset == 24
