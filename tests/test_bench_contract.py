"""CPU check of bench.py's command-line contract: more than one GPU is refused
with a JSON message and a nonzero status, before any device is touched."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = [x for x in r.stdout.splitlines() if x.startswith("{")][-1]
    assert "one GPU" in json.loads(line)["error"]
