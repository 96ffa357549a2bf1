"""CPU checks of bench.py's command line: a request for more than one GPU
ends at once with a message that says only one GPU is measured (no
multi-GPU path is built), before any GPU work or torch import; the serving
stream's segment records merge into one set of figures without a GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


@pytest.mark.parametrize("gpus", [2, 8])
def test_more_than_one_gpu_exits_with_message(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(gpus)],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_stream_merge_combines_segments(tmp_path):
    # two consecutive segments' records -> percentiles over every request,
    # pair-interactions over the summed busy time, FIFO on the first prefix
    import json

    import numpy as np
    segs = []
    for first, last, lat, busy in ((0, 3, [1.0, 2.0, 3.0], 10.0), (3, 5, [4.0, 5.0], 5.0)):
        k = last - first
        rec = {"first": first, "last": last, "lambda_per_s": 1.2, "tmax_range": [0.2, 2.0],
               "budget_GB": 8.0, "max_active": 64, "grid": "128^3", "policy": "least_work_first",
               "figures": {"busy_s": busy, "makespan_s": busy + 1.0},
               "r": list(range(first, last)), "kind": ["plummer"] * (k - 1) + ["galaxy"],
               "n": [1000] * k, "steps": [97] * k, "arrival_s": list(range(first, last)),
               "latency_s": lat}
        path = tmp_path / f"s{first}.json"
        path.write_text(json.dumps(rec))
        segs.append(str(path))
    # the plain FIFO run of requests 0 and 1 alone (no served figures)
    fifo = {"first": 0, "last": 2, "fifo_only": True, "r": [0, 1], "n": [1000, 1000],
            "steps": [97, 97], "kind": ["plummer", "plummer"], "arrival_s": [0, 1],
            "fifo_service_s": [2.0, 2.0], "fifo_latency_s": [2.0, 3.0]}
    (tmp_path / "fifo.json").write_text(json.dumps(fifo))
    r = subprocess.run([sys.executable, BENCH, "--stream-merge",
                        ",".join(segs[::-1] + [str(tmp_path / "fifo.json")])],
                       capture_output=True, text=True, check=True)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["requests"] == 5 and out["segments"] == [[0, 3], [3, 5]]
    assert out["p50_latency_s"] == 3.0
    assert out["p99_latency_s"] == float(np.percentile([1, 2, 3, 4, 5], 99))
    assert out["pair_interactions"] == 5 * 1000.0 ** 2 * 100
    assert out["Gpairs_per_s_busy"] == 5 * 1e8 / 15.0 / 1e9
    assert out["galaxy_requests"] == 2
    fp = out["fifo_prefix"]
    assert fp["requests"] == [0, 2] and fp["plain_busy_s"] == 4.0
    assert fp["served_p50_latency_s"] == 1.5 and fp["plain_fifo_p50_latency_s"] == 2.5
    # segments that do not join up are refused
    bad = json.loads(open(segs[1]).read())
    bad["first"] = 4
    open(segs[1], "w").write(json.dumps(bad))
    r = subprocess.run([sys.executable, BENCH, "--stream-merge", ",".join(segs)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "not contiguous" in r.stderr
