"""bench.py's output line: one JSON object with the keys a reader of the
headline measurement relies on (metric, value, units, step counts, the
roofline of the dominant kernel and the CPU baseline), and its refusal of
more than one GPU."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_bench_line(device):
    sys.path.insert(0, str(ROOT))
    import bench
    r = subprocess.run([sys.executable, "-u", str(ROOT / "bench.py"), "--steps", "3",
                        "--warmup", "1", "--no-batched", "--no-adjoint", "--no-tidal",
                        "--no-ext", "--cpu-seconds", "1"],
                       capture_output=True, text=True, timeout=540, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip().startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    assert d["metric"] == bench.METRIC
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["unit"] and d["dtype"] and d["data"].startswith("synthetic")
    assert "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] in ("hbm", "mfma") and 0 < rf["frac"] < 1
    assert abs(rf["achieved"] / rf["peak"] - rf["frac"]) < 1e-9
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["kind"] in ("port", "reference") and cb["value"] > 0 and cb["cores"] >= 1
