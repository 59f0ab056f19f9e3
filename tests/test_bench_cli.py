"""bench.py's command-line contract on the CPU: more than one GPU is refused
with a message (the library runs on one GPU), and the CPU-baseline leg returns
the fields the JSON line carries."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_fields(tmp_path):
    code = ("import json, sys; sys.path.insert(0, %r); import bench; "
            "print(json.dumps(bench.cpu_baseline(budget_s=0.2, budget_ref_s=0.1, n=16)))" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["kind"] == "port" and d["cores"] == 2 and d["unit"] == "zone-updates/s"
    assert d["value"] > 0 and d["ref_semantics_value"] > 0
    assert isinstance(d["cpu_model"], str) and d["nproc"] >= 1
