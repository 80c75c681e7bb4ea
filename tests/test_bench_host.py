"""bench.py's host-side contract, without a GPU: more than one GPU is refused
with a JSON message and exit status 2, and roofline.traffic is read from the
committed rocprofv3 counter summary only for the workload it was taken on."""
import json
import os
import subprocess
import sys
import types

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def _args(**kw):
    a = dict(vertices=4_800_000, edges=69_000_000, beta=2.1, max_degree=20000.0, seed=1)
    a.update(kw)
    return types.SimpleNamespace(**a)


def test_profiled_traffic_matches_the_committed_profile():
    N = 1 << 28
    tri = bench.profiled_traffic("triangle", _args(), N)
    k4 = bench.profiled_traffic("4-clique", _args(), N)
    assert tri is not None and k4 is not None
    # a 2^28-estimator launch moves about three 128-byte lines per estimator
    for t in (tri, k4):
        assert 2.5 * 128 * N < t < 3.5 * 128 * N


def test_profiled_traffic_ignores_other_workloads():
    N = 1 << 28
    assert bench.profiled_traffic("triangle", _args(seed=2), N) is None
    assert bench.profiled_traffic("triangle", _args(edges=1000), N) is None
    assert bench.profiled_traffic("triangle", _args(), N // 2) is None
    assert bench.profiled_traffic("no-such-kernel", _args(), N) is None


def test_time_to_error_summary_reads_every_arm_and_never_raises():
    elp = {"epsilon": 0.01,
           "triangle": {"regression+judge": {"mean_time_s": 0.17, "within_eps": 1.0},
                        "learned+judge": {"mean_time_s": 0.005, "within_eps": 0.8},
                        "pilot+judge": {"mean_time_s": None}},
           "5-house split, completions chosen": {"mean_time_s": 0.02, "within_eps": 1.0}}
    s = bench.tte_summary(elp)
    assert s["triangle"] == {"regression+judge": {"s": 0.17, "within_eps": 1.0},
                             "learned+judge": {"s": 0.005, "within_eps": 0.8}}
    assert s["5-house split, completions chosen"] == {"query": {"s": 0.02, "within_eps": 1.0}}
    assert bench.tte_summary(None) is None
    # records of other shapes are skipped, not fatal
    assert bench.tte_summary({"x": 5, "y": {"learned+judge": 3}, "z": {"queries": []}}) == {}
