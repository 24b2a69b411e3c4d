"""bench.py pieces that run without a GPU: the refusal of several GPUs and the error helper."""
import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and r.stdout == "" and "one GPU only" in r.stderr


def test_rel_l2_per_variable_momentum_and_ratio():
    rel_l2 = _bench().rel_l2
    b = np.zeros((3, 6, 10))
    b[0] = 2.0
    b[1] = 3.0
    b[2] = 4.0  # ||qy|| / ||qx|| = 4/3
    a = b.copy()
    a[1] += 0.3  # 10 % on qx only
    per, mom, ratio = rel_l2(a, b)
    assert per[0] == 0.0 and abs(per[1] - 0.1) < 1e-15 and per[2] == 0.0
    # momentum vector (qx, qy): 0.3 over sqrt(3^2 + 4^2) = 5 per entry
    assert abs(mom - 0.3 / 5.0) < 1e-15
    assert abs(ratio - 4.0 / 3.0) < 1e-15
