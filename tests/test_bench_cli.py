"""CPU checks of bench.py's command line: a request for more than one GPU
ends at once with a message that says only one GPU is measured (no
multi-GPU path is built), before any GPU work or torch import."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("gpus", [2, 8])
def test_more_than_one_gpu_exits_with_message(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(gpus)],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
