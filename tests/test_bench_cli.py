"""bench.py's command line on the CPU: it measures one GPU, and a request for
more exits with status 2 and says so before anything touches a device."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "one GPU only" in r.stderr and ("--gpus %d" % n) in r.stderr
    assert r.stdout == ""  # no JSON line


def test_unknown_mode_is_rejected():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--mode", "no-such-mode"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2 and "invalid choice" in r.stderr
