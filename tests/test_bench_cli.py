"""bench.py's command line on the CPU: more than one GPU is refused before any GPU work, with
the reason on the one JSON line it prints."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]
