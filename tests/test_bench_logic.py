"""CPU: bench.py's host logic (graph chunking, companion records, the one-GPU guard)."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_graph_chunk_divides_timed_steps():
    for K in (1, 7, 20, 97, 100, 500, 5000, 4000):
        for cap in (1, 16, 100):
            c = bench.graph_chunk_for(K, cap)
            assert 1 <= c <= min(cap, K) and K % c == 0
            # the largest such divisor
            assert all(K % d for d in range(c + 1, min(cap, K) + 1))
    assert bench.graph_chunk_for(20, 100) == 20
    assert bench.graph_chunk_for(5000, 100) == 100


def test_companions_are_single_instance_workloads():
    for c in bench.COMPANIONS:
        w = bench.WORKLOADS[c]
        assert w["b_alg"] and w["steps"] > 0 and w["model"] in ("ch", "k93", "kh", "modelb", "k94")


def test_companion_record_keeps_the_measured_fields():
    o = {"config": {"workload": "dendrite: 512x512"}, "value": 1.0, "unit": "cell-updates/s", "steps": 4000,
         "warmup": 200, "ms_per_step": 0.007, "gpu_ms_per_step": 0.0069, "active_fraction": 0.13,
         "full_grid_updates_per_s": 3.0, "finite": True, "cpu_baseline": {"value": 1},
         "roofline": {"achieved": 300.0, "frac": 0.0375, "traffic": None, "b_alg_per_cell": 32,
                      "avg_kernel_us": 8.0, "kernel": "k", "peak": 8000.0}}
    c = bench.companion(o)
    assert c["workload"] == "dendrite: 512x512" and c["steps"] == 4000 and "cpu_baseline" not in c
    assert c["roofline"]["frac"] == 0.0375 and "peak" not in c["roofline"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
