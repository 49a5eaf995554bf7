"""bench.py's command-line contract on the host (no device needed)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_help_lists_every_workload():
    r = _run("--help")
    assert r.returncode == 0
    for w in ("disc_stack_bz", "disc_sweep", "random_devices", "twisted", "twisted_fields", "grad_grid",
              "islands", "bands", "beams"):
        assert w in r.stdout
