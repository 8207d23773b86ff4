"""bench.py's contract that needs no GPU: one GPU only, and the measured HBM
traffic and VALU instruction count of the dominant kernel read from the
committed rocprofv3 summary."""
import argparse
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""          # no JSON line


def test_traffic_of_the_default_workload_comes_from_the_profile():
    b = _bench()
    default = argparse.Namespace(workload="batch", precision="f32")
    fields = {}
    for line in open(b.TRAFFIC_FILE):
        f = line.split()
        if f and not line.startswith("#"):
            fields[f[0]] = f
    traffic, valu, src = b.measured_traffic("source", default)
    f = fields["source"]
    assert traffic == int(float(f[3]) + float(f[4]))   # FETCH + WRITE bytes per launch
    assert valu == int(float(f[5]))                    # SQ_INSTS_VALU per launch
    assert src == os.path.join("profiles", "pmc_traffic_batch64_fp32.txt")
    # every kernel class the bench can name as dominant has a line
    for k in ("source", "trans_theta", "trans_rad", "cfl_force", "trans_vr"):
        assert b.measured_traffic(k, default)[0] > 0
    # other workloads and FP64 were not profiled: null, not a borrowed figure
    assert b.measured_traffic("source", argparse.Namespace(workload="large", precision="f32")) == (None, None, None)
    assert b.measured_traffic("source", argparse.Namespace(workload="batch", precision="f64")) == (None, None, None)
