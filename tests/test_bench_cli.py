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
              "islands", "bands", "beams", "bz_fields"):
        assert w in r.stdout


def test_parity_max_flags_nan_in_either_direction():
    sys.path.insert(0, ROOT)
    import bench
    nan, inf = float("nan"), float("inf")
    assert bench.parity_max(0.0, 0.5, 0.75) == 0.25
    # a NaN from the device against a finite reference value
    assert bench.parity_max(0.0, 0.5, nan) == inf
    # a finite device value where the reference is NaN (evanescent incidence)
    assert bench.parity_max(0.0, nan, 0.5) == inf
    # the same NaN pattern on both sides adds nothing
    assert bench.parity_max(3e-11, nan, nan) == 3e-11
    # NaN and inf do not match each other
    assert bench.parity_max(0.0, nan, inf) == inf
    assert bench.parity_max(2e-11, 0.1, 0.1) == 2e-11
