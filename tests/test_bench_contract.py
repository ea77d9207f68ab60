"""CPU: the bench script's contract that needs no GPU.

--gpus above 1 exits with status 2 and one JSON line saying why; every workload and
summariser the default run and the profile recipes call is defined.
"""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_workloads_and_summarisers_are_defined():
    import bench
    for name in ("cpu_baseline", "other_workloads", "bench_batch", "bench_q2", "bench_serve", "bench_upload",
                 "bench_elastic", "summarize_profiles", "summarize_counters", "summarize_q2", "summarize_elastic",
                 "measured_hbm_peaks", "main"):
        assert callable(getattr(bench, name)), name
