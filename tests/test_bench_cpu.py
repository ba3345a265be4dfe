"""bench.py contract checks that need no GPU."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_baseline_measures_the_reference_rule():
    """The CPU leg counts pairs with r > thr exactly like the oracle and
    reports a TFLOP/s rate with the sample it timed."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "pagani-toolkit_amd"))
    import bench
    from pagani_amd import synth
    x = synth.bold(3, 600, 64, 4)
    out = bench.cpu_baseline(x, 600, 64, 0.3, budget_s=0.5)
    assert out["kind"] == "port" and out["unit"] == "TFLOP/s" and out["value"] > 0
    assert out["cores"] >= 1 and "pairs" in out["sample"]
    assert np.isfinite(out["est_full_subject_s"])
