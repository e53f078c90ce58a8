"""bench.py's contract on the CPU: more than one GPU is refused with a
message before anything touches a device, and the JSON line never carries
NaN or infinity (JSON has neither)."""
import json
import math
import os
import subprocess
import sys

from conftest import ROOT


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_json_line_has_no_nonfinite_numbers():
    sys.path.insert(0, ROOT)
    import bench
    out = bench.json_safe({"value": 1.5, "sigma": math.nan, "legs": [math.inf, {"prec": -math.inf, "n": 3}]})
    line = json.dumps(out, allow_nan=False)
    assert json.loads(line) == {"value": 1.5, "sigma": None, "legs": [None, {"prec": None, "n": 3}]}
