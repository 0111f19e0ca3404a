"""bench.py's command line on the CPU: it measures one GPU and refuses
--gpus above 1 with a message before touching a device, and it offers the
workloads its docs name."""
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr


def test_workloads_listed():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0
    for w in ("precision", "batch", "scale", "adjoint", "extras"):
        assert w in p.stdout
