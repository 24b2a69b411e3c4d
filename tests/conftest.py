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


# The twelve SSP-RK schemes of the reference stepper (explicit_ssp_rk_stepper.cpp:16-257).
SSPRK_SCHEMES = [(1, 1), (2, 2), (3, 3), (4, 3), (5, 3), (6, 3), (7, 3), (8, 3), (5, 4), (6, 4), (7, 4), (8, 4)]


def rk_order_residuals(a, b):
    """Residuals of the Runge-Kutta order conditions (orders 1 to 4, eight conditions) of the
    Shu-Osher scheme y_{i+1} = sum_k a[i,k] y_k + dt b[i,k] F(y_k), u_{n+1} = y_ns, taken to its
    Butcher form (A, weights w, nodes c = A 1)."""
    import numpy as np
    ns = a.shape[0]
    rows = [np.zeros(ns)]
    for i in range(ns):
        r = np.zeros(ns)
        for k in range(i + 1):
            r = r + a[i, k] * rows[k]
            r[k] += b[i, k]
        rows.append(r)
    A, w = np.array(rows[:ns]), rows[ns]
    c = A.sum(axis=1)
    return np.array([w.sum() - 1, w @ c - 1 / 2,
                     w @ c**2 - 1 / 3, w @ (A @ c) - 1 / 6,
                     w @ c**3 - 1 / 4, (w * c) @ (A @ c) - 1 / 8, w @ (A @ c**2) - 1 / 12, w @ (A @ (A @ c)) - 1 / 24])


ORDER_CONDITIONS = {1: 1, 2: 2, 3: 4, 4: 8}  # conditions up to each order


def rk_scheme_ok(a, b, order, tol=1e-9):
    """Order conditions up to `order` hold within tol, rows of a sum to one, and every
    coefficient is non-negative (the strong-stability-preserving form)."""
    import numpy as np
    ns = a.shape[0]
    res = rk_order_residuals(a, b)[:ORDER_CONDITIONS[order]]
    lower = np.tril(np.ones((ns, ns), bool))
    return (bool(np.all(np.abs(res) <= tol)) and bool(np.allclose(a.sum(axis=1), 1.0, atol=tol, rtol=0))
            and bool(np.all(a[lower] >= 0)) and bool(np.all(b[lower] >= 0)))
