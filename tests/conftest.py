import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C-ABI")


def relerr(a, b):
    """Per-variable relative L2 of a against b ([3, ndof, ne] modal arrays)."""
    import numpy as np
    out = []
    for k in range(3):
        n = np.linalg.norm(b[k])
        d = np.linalg.norm(a[k] - b[k])
        out.append(d / n if n > 0 else d)
    return out
