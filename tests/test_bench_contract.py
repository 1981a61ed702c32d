"""bench.py's command-line contract on the CPU: one GPU only (a request for more GPUs exits with
status 2 and says so, before anything touches a GPU), and the CPU baseline's restatement runs
on a small lattice and reports the fields of the JSON line's `cpu_baseline` object."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""


def test_cpu_baseline_fields():
    sys.path.insert(0, str(ROOT))
    import bench
    t, tmin, dinv, prm = bench.real_workload(npts=1024)
    out = bench.cpu_baseline((8, 8, 16), t, tmin, dinv, prm, budget_s=0.2)
    assert out["unit"] == "site-updates/s" and out["kind"] == "port"
    assert out["value"] > 0 and out["cores"] >= 1 and out["ms_per_step"] > 0
    assert "8x8x16" in out["sample"]
