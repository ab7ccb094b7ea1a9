"""bench.py's command-line contract, checked without a GPU: one GPU only (N > 1 exits with a
message and no JSON line), the option set the driver passes, and the entry points of
__graft_entry__."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    p = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout.strip() == ""  # no JSON line


def test_options_the_driver_passes_are_accepted():
    p = run("--help")
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--config", "--methods", "--no-cpu"):
        assert opt in p.stdout


def test_entry_points():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as g
    assert callable(g.build) and callable(g.smoke)
