"""Sanitizer builds of the host code (CPU only; `make -C tests/cpp sanitize`).

- ThreadSanitizer: the two-thread fine aggregation (csrc/aggregate.cpp, aggregate_pipelined,
  whose leftover pass trails the root pass by twice the graph's bandwidth) equals the
  sequential aggregation on node graphs above its 100,000-node threshold, with no race
  reported;
- AddressSanitizer + UndefinedBehaviorSanitizer: the Gmsh reader (host/io.cpp) on the
  reference's meshes and on every malformed file of test_mesh_io gives the regular build's
  arrays and messages; the case keys on well-formed and malformed texts; every function of
  the C oracle on a small plate, error paths included.
The device path itself cannot run under these sanitizers (no GPU AddressSanitizer here).
"""
import os
import subprocess

import numpy as np
import pytest

import test_mesh_io as tm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BUILD = os.path.join(CPP, "_build")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
REPORTS = ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:",
           "WARNING: ThreadSanitizer")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "sanitize"], cwd=CPP, check=True)
    return BUILD


def run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, **kw)
    for rep in REPORTS:
        assert rep not in r.stderr, r.stderr[-3000:]
    return r


def test_pipelined_aggregation_under_tsan(built):
    r = run([os.path.join(built, "test_aggregate_tsan")])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert r.stdout.count("identical") == 3


def test_case_keys_under_asan_ubsan(built):
    r = run([os.path.join(built, "test_case_keys_asan")])
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert "unknown key 'empty'" in r.stdout and "repeats line 1" in r.stdout


def test_oracle_under_asan_ubsan(built):
    r = run([os.path.join(built, "test_oracle_asan")])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]


def mesh_cases():
    cases = [("L0", os.path.join(tm.GOLD, "kirsch_L0.msh")),
             ("L1", os.path.join(tm.GOLD, "kirsch_L1.msh"))]
    texts = [tm.HEAD + tm.NODES + tm.elements([(1, 2, 3), (1, 4, 3)]),
             tm.HEAD.replace("4.1 0 8", "2.2 0 8") + tm.NODES, tm.HEAD + tm.NODES,
             tm.HEAD + tm.NODES.replace("\n2\n3\n", "\n2\n2\n"),
             tm.HEAD + tm.NODES + tm.elements([(1, 2, 9)]),
             tm.HEAD + tm.NODES + tm.elements([(1, 2, 2)]),
             tm.HEAD + tm.NODES + tm.elements([(1, 2, 3, 4)], etype=3), "$Nodes\n"]
    texts += [t for t, _, _ in tm.MORE]
    return cases, texts


def test_gmsh_reader_under_asan_ubsan(built, tmp_path):
    subprocess.run(["make", "-s", "cstfem_hip_mesh"], cwd=tm.PKG, check=True)
    cases, texts = mesh_cases()
    for k, t in enumerate(texts):
        p = tmp_path / f"c{k}.msh"
        p.write_text(t)
        cases.append((f"c{k}", str(p)))
    for name, path in cases:
        outs = []
        for tool in (os.path.join(built, "mesh_tool_asan"), tm.TOOL):
            prefix = str(tmp_path / f"{name}_{len(outs)}")
            r = run([tool, path, prefix])
            arrays = None
            if r.returncode == 0:
                arrays = [np.loadtxt(prefix + ext, ndmin=2) if os.path.getsize(prefix + ext) else
                          None for ext in (".xy", ".tri")]
            outs.append((r.returncode, r.stdout, r.stderr.strip(), arrays))
        (rc_a, out_a, err_a, arr_a), (rc_b, out_b, err_b, arr_b) = outs
        assert (rc_a, out_a, err_a) == (rc_b, out_b, err_b), name
        if arr_a is not None:
            for a, b in zip(arr_a, arr_b):
                np.testing.assert_array_equal(a, b)
