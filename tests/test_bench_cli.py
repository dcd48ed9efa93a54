"""bench.py's command line on the CPU: a request for more than one GPU ends
with a message and exit status 2 before anything touches a device."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_refuses_several_gpus(n):
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", str(n)],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and ("--gpus %d" % n) in r.stderr
    assert r.stdout == ""


def test_bench_help_lists_the_contract_flags():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--help"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--batched", "--no-cpu-baseline"):
        assert flag in r.stdout
