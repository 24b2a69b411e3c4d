"""Oracle basis and quadrature against the reference's golden vectors and closed forms.

Pins: Dubiner phi / dphi at p = 10 on 15 points incl. the collapsed vertex
(test/test_basis_dubiner_2d.cpp, tolerances 100 eps / 1000 eps as there); Dunavant and
Gauss-Legendre exactness (test/test_integ_dunavant_2d.cpp:9-24, test_integ_gausslegendre_1d.cpp);
Dubiner orthogonality phi^T W phi = diag(1/minv) (basis_dubiner_2D.cpp:79); and the engine's own
tables (librkdg host helper) equal to the oracle's bit for bit.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle_py import load

GOLD = Path(__file__).resolve().parent / "golden"
EPS = np.finfo(float).eps


def almost_equal(a, b, factor=100.0):  # utilities/almost_equal.hpp:16-22
    m = max(abs(a), abs(b))
    if m < EPS * factor:
        return True
    return abs(a - b) < m * EPS * factor


def dubiner(p, pts):
    lib = load()
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(pts)
    nd = (p + 1) * (p + 2) // 2
    phi, d1, d2 = (np.empty(nd * n) for _ in range(3))
    dp = C.POINTER(C.c_double)
    lib.orc_dubiner(p, n, pts.ctypes.data_as(dp), phi.ctypes.data_as(dp), d1.ctypes.data_as(dp), d2.ctypes.data_as(dp))
    return phi.reshape(nd, n), d1.reshape(nd, n), d2.reshape(nd, n)


def dunavant(p):
    lib = load()
    w, z = np.empty(100), np.empty(200)
    dp = C.POINTER(C.c_double)
    n = lib.orc_dunavant(p, w.ctypes.data_as(dp), z.ctypes.data_as(dp))
    return w[:n], z[:2 * n].reshape(n, 2)


def gauss(p):
    lib = load()
    w, z = np.empty(40), np.empty(40)
    dp = C.POINTER(C.c_double)
    n = lib.orc_gauss_legendre(p, w.ctypes.data_as(dp), z.ctypes.data_as(dp))
    return w[:n], z[:n]


def test_dubiner_p10_golden():
    g = json.load(open(GOLD / "dubiner_p10.json"))
    phi, d1, d2 = dubiner(10, g["points"])
    assert np.all(np.isfinite(phi)) and np.all(np.isfinite(d1)) and np.all(np.isfinite(d2))
    bad = 0
    for dof in range(66):
        for pt in range(15):
            bad += not almost_equal(g["phi"][dof][pt], phi[dof, pt], 100)
            bad += not almost_equal(g["dphi_interleaved"][dof][2 * pt], d1[dof, pt], 1000)
            bad += not almost_equal(g["dphi_interleaved"][dof][2 * pt + 1], d2[dof, pt], 1000)
    assert bad == 0


def test_dubiner_golden_catches_planted_error():
    # A wrong Jacobi weight (2p+1 -> 2p) must fail the golden comparison.
    g = json.load(open(GOLD / "dubiner_p10.json"))
    phi, _, _ = dubiner(10, g["points"])
    phi_bad = phi.copy()
    phi_bad[7] *= 1 + 1e-9
    assert any(not almost_equal(g["phi"][7][pt], phi_bad[7, pt], 100) for pt in range(15) if g["phi"][7][pt] != 0)


@pytest.mark.parametrize("p", range(1, 20))
def test_dunavant_exactness(p):
    w, z = dunavant(p)
    exact = 1 / (p + 1.0) * ((1 - (-1.0) ** p) / (p + 2.0) + 2 * (-1.0) ** p)
    num = sum(z[g, 0] ** p * w[g] for g in range(len(w)))
    assert almost_equal(num, exact, 1e3)


def test_dunavant_point_counts():
    counts = {1: 1, 2: 3, 3: 4, 4: 6, 5: 7, 6: 12, 7: 13, 8: 16, 9: 19, 10: 25, 11: 27, 12: 33, 13: 37, 14: 42,
              15: 48, 16: 52, 17: 61, 18: 70, 19: 73, 20: 79}  # integration_dunavant_2D.cpp:78-121
    for p, n in counts.items():
        assert len(dunavant(p)[0]) == n


@pytest.mark.parametrize("p", range(1, 20))
def test_gauss_legendre_exactness(p):
    w, z = gauss(p)
    exact = 2.0 - (1 - (-1.0) ** (p + 1)) / (p + 1)  # integral of (1 - x^p) over [-1, 1]
    num = sum((1 - z[g] ** p) * w[g] for g in range(len(w)))
    assert almost_equal(num, exact, 1e3)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_dubiner_orthogonality(p):
    w, z = dunavant(2 * p)
    phi, _, _ = dubiner(p, z)
    M = phi @ np.diag(w) @ phi.T
    nd = phi.shape[0]
    minv = []
    for dof in range(nd):
        t = int(np.ceil((-3 + np.sqrt(1 + 8 * (dof + 1))) / 2))
        pp = dof - (t + 1) * t // 2
        qq = t - pp
        minv.append((2 * pp + 1) * (pp + qq + 1) / 2.0)
    assert np.allclose(M, np.diag(1 / np.array(minv)), atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_engine_tables_bitwise_equal_oracle(p):
    from dgswemv2_amd import _native as N
    lib = N.load_library()
    nd, ng = C.c_int32(), C.c_int32()
    lib.rkdg_master_tables(p, C.byref(nd), C.byref(ng), None, None, None, None)
    phi, d1, d2 = (np.empty(nd.value * ng.value) for _ in range(3))
    w = np.empty(ng.value)
    assert lib.rkdg_master_tables(p, None, None, N.dptr(phi), N.dptr(d1), N.dptr(d2), N.dptr(w)) == 0
    wo, zo = dunavant(2 * p)
    po, d1o, d2o = dubiner(p, zo)
    assert np.array_equal(w, wo)
    assert np.array_equal(phi.reshape(nd.value, -1), po)
    assert np.array_equal(d1.reshape(nd.value, -1), d1o)
    assert np.array_equal(d2.reshape(nd.value, -1), d2o)
