"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message and exit status 2, before any device work."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU" in (p.stdout + p.stderr)
    assert not p.stdout.lstrip().startswith("{")
