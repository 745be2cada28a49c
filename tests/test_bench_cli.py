"""bench.py's host-side contract, on the CPU: more than one GPU is refused
with a message before anything touches a device, and the CPU baseline's
recorded full-size oracle figures come from the committed fixture."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_full_size_recorded_matches_fixture():
    import json

    import bench

    rec = bench.full_size_recorded()
    with open(os.path.join(ROOT, "tests", "golden", "full_size_oracle.json")) as f:
        fix = json.load(f)
    assert set(rec) == {k for k, v in fix.items() if "oracle_seconds" in v}
    assert "grid2d_voronoi_1024x2048_PD" in rec
    for k, v in rec.items():
        assert v["edges"] == fix[k]["edges"] and v["seconds"] == fix[k]["oracle_seconds"]
        assert abs(v["edges_per_s"] * v["seconds"] - v["edges"]) <= 1e-6 * v["edges"]
    # the PD workload of the bench line: 12,469,248 edges
    assert rec["grid2d_voronoi_1024x2048_PD"]["edges"] == 12469248
