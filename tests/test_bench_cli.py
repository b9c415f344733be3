"""bench.py measures one GPU: a run asked for several exits at once with a
message that says so (before any GPU or mesh work), and the help text
lists the options the measurement contract uses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2", "--steps", "5", "--warmup", "1")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_contract_options():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--lattice", "--precision", "--batch-cases",
                "--steady-cases"):
        assert opt in r.stdout, opt
