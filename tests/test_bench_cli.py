"""bench.py's command-line contract that needs no GPU: it measures one device
and refuses a multi-GPU request with a message (exit status 2) before it
touches the device or builds any input."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300, cwd=ROOT)
    assert res.returncode == 2, res.stdout + res.stderr
    line = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
