"""JSON-lines metrics records (empm_amd.metrics) and the bench's
--metrics-jsonl option.  Host only."""
import json
import os
import subprocess
import sys

from empm_amd import metrics


def test_records_append_as_json_lines(tmp_path):
    p = tmp_path / "runs" / "metrics.jsonl"
    a = metrics.append_record(p, {"metric": "particle-substeps/s", "value": 1.6e9}, run="a")
    metrics.append_record(p, {"value": 1.7e9, "roofline": {"frac": 0.26}}, run="b")
    lines = p.read_text().splitlines()
    assert len(lines) == 2 and all(json.loads(x) for x in lines)
    recs = metrics.read_records(p)
    assert recs[0] == a and recs[0]["run"] == "a" and recs[0]["value"] == 1.6e9
    assert recs[1]["roofline"]["frac"] == 0.26 and "time" in recs[1] and "host" in recs[1]


def test_bench_multi_gpu_exit_leaves_no_record(tmp_path):
    # the bench refuses N > 1 before measuring anything (no GPU needed)
    p = tmp_path / "m.jsonl"
    bench = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")
    r = subprocess.run([sys.executable, bench, "--gpus", "2", "--metrics-jsonl", str(p)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and not p.exists()
