"""bench.py --mode bal: the reference's bench_bal contract
(benchmarks/bench_bal.cpp:63-139) on a BAL file written by gba_write_bal --
one JSON object per solver in bench_bal's schema, exit 3 when the anomaly
counter is non-zero."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = {"file", "solver", "loss", "num_cameras", "num_points", "num_observations", "initial_cost",
        "final_cost", "iterations", "total_time_s", "per_iteration_s", "termination", "device"}


def run(path, *extra):
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--mode", "bal", "--file", path, "--pairs", "1",
           "--iterations", "5", *extra]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


def test_bench_bal_schema(gba, tmp_path):
    p = gba.generate(1, 5)
    path = str(tmp_path / "p.txt")
    gba.write_bal(path, p)
    r = run(path, "--json", str(tmp_path / "out.json"))
    assert r.returncode == 0, r.stderr
    objs = [json.loads(l) for l in r.stdout.strip().splitlines()]
    assert [o["solver"] for o in objs] == ["gba-amd-lm-fp64", "gba-amd-lm-mixed"]
    for o in objs:
        assert KEYS <= set(o)
        assert o["num_cameras"] == p.num_cameras and o["iterations"] <= 5
        assert o["final_cost"] < o["initial_cost"]
    assert abs(objs[0]["initial_cost"] - gba.total_squared_error(gba.load_bal(path)) / 2) <= \
        1e-9 * objs[0]["initial_cost"]
    assert os.path.exists(tmp_path / "out.json")


def test_bench_bal_refuses_anomalies(gba, tmp_path):
    p = gba.generate(1, 5)
    p.obs_pixel[3, 0] = np.nan
    path = str(tmp_path / "nan.txt")
    gba.write_bal(path, p)
    r = run(path)
    assert r.returncode == 3, (r.returncode, r.stderr[-500:])
