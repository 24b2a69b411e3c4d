"""Closed-form pins of the oracle's element and edge geometry, after the reference's
test_boundary_interface.cpp (divergence theorem on one triangle) and test_shape_straight_triangle.cpp.

For a polynomial f of degree p on a triangle and every basis function phi_d of order p,
    int_K (d phi_d / dx) f  +  int_K phi_d (d f / dx)  =  sum over edges int_e phi_d f n_x,
and likewise in y.  The volume rule (Dunavant 2p) and the edge rule (Gauss-Legendre, p + 1 points)
integrate these degrees exactly, so the oracle's volume operators (element.hpp:234-246, 483-487:
J^-1, |det J|, the rule's weights) and its edge operators (boundary.hpp:69-134: normals, surface
Jacobian L/2, the master-edge map of BoundaryToMasterCoordinates, master_triangle.tpp:89-122) must
satisfy the identity to rounding.  It holds on an equilateral, a jittered and a clockwise-listed
triangle for p = 1, 2, 3, and fails with a planted flipped normal, a surface Jacobian of L instead
of L/2, or edge points placed on the neighbouring master edge."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle_py import load

dp = C.POINTER(C.c_double)
PLANT_NORMAL_SIGN, PLANT_SURFJ_FULL, PLANT_EDGE_SHIFT = 9, 10, 11

TRIANGLES = {
    "equilateral": ([-0.5, 0.5, 0.0], [0.0, 0.0, math.sqrt(3.0) / 2.0]),
    "jittered": ([0.13, 2.41, 0.77], [-0.29, 0.35, 1.93]),
    "clockwise": ([0.0, 0.0, 1.5], [0.0, 1.2, 0.1]),  # the reference reorders the coordinates
}


def ptr(a):
    return a.ctypes.data_as(dp)


def poly(p, seed):
    """A random polynomial of total degree p and its x / y derivatives, as callables."""
    rng = np.random.default_rng(seed)
    terms = [(i, j, rng.uniform(-1, 1)) for i in range(p + 1) for j in range(p + 1 - i)]

    def f(x, y):
        return sum(c * x ** i * y ** j for i, j, c in terms)

    def fx(x, y):
        return sum(c * i * x ** (i - 1) * y ** j for i, j, c in terms if i > 0)

    def fy(x, y):
        return sum(c * j * x ** i * y ** (j - 1) for i, j, c in terms if j > 0)
    return f, fx, fy


def residuals(p, vx, vy, seed=1):
    """max |volume - boundary| / scale over d, for the x and y identities."""
    lib = load()
    vx, vy = np.asarray(vx, float), np.asarray(vy, float)
    f, fx, fy = poly(p, seed)
    xy = np.empty(2 * 400)
    n = lib.orc_element_gp_global(p, ptr(vx), ptr(vy), ptr(xy))
    gx, gy = xy[0:2 * n:2].copy(), xy[1:2 * n:2].copy()
    ndof = (p + 1) * (p + 2) // 2

    def vol(fv):
        integ = C.c_double()
        iphi, idx, idy = np.empty(ndof), np.empty(ndof), np.empty(ndof)
        assert lib.orc_element_integrals(p, ptr(vx), ptr(vy), ptr(np.ascontiguousarray(fv)), C.byref(integ),
                                         ptr(iphi), ptr(idx), ptr(idy)) == 0
        return iphi, idx, idy

    _, idx, idy = vol(f(gx, gy))          # int dphi/dx f, int dphi/dy f
    ipx, _, _ = vol(fx(gx, gy))           # int phi df/dx
    ipy, _, _ = vol(fy(gx, gy))           # int phi df/dy
    exy = np.empty(3 * 8 * 2)
    ix, iy = np.empty(ndof), np.empty(ndof)
    ngp = lib.orc_element_edge_integrals(p, ptr(vx), ptr(vy), None, ptr(exy), ptr(ix), ptr(iy))
    assert ngp == p + 1
    ex, ey = exy[0:2 * 3 * ngp:2].copy(), exy[1:2 * 3 * ngp:2].copy()
    fe = np.ascontiguousarray(f(ex, ey))
    assert lib.orc_element_edge_integrals(p, ptr(vx), ptr(vy), ptr(fe), None, ptr(ix), ptr(iy)) == ngp
    scale = max(1.0, np.max(np.abs(idx)), np.max(np.abs(ix)))
    return (float(np.max(np.abs(idx + ipx - ix)) / scale), float(np.max(np.abs(idy + ipy - iy)) / scale))


@pytest.mark.parametrize("shape", sorted(TRIANGLES))
@pytest.mark.parametrize("p", [1, 2, 3])
def test_divergence_theorem_on_one_triangle(shape, p):
    rx, ry = residuals(p, *TRIANGLES[shape])
    assert rx < 1e-12 and ry < 1e-12, (rx, ry)


@pytest.mark.parametrize("plant", [PLANT_NORMAL_SIGN, PLANT_SURFJ_FULL, PLANT_EDGE_SHIFT])
def test_divergence_theorem_planted(plant):
    lib = load()
    lib.orc_set_plant(plant)
    try:
        worst = max(max(residuals(2, *TRIANGLES[s])) for s in TRIANGLES)
    finally:
        lib.orc_set_plant(0)
    assert worst > 1e-3, worst
    # and the unplanted oracle is back to rounding
    assert max(residuals(2, *TRIANGLES["jittered"])) < 1e-12


def test_edge_points_lie_on_their_edges_in_reference_order():
    """Edge k runs from vertex (k+1)%3 to (k+2)%3 (shape_straighttriangle.cpp:14-21): its Gauss
    points are on that segment, ordered from the first vertex to the second, at the GL abscissae."""
    lib = load()
    vx, vy = np.asarray(TRIANGLES["jittered"][0]), np.asarray(TRIANGLES["jittered"][1])
    for p in (1, 2):
        exy = np.empty(3 * 8 * 2)
        ix, iy = np.empty(10), np.empty(10)
        ngp = lib.orc_element_edge_integrals(p, ptr(vx), ptr(vy), None, ptr(exy), ptr(ix), ptr(iy))
        z = np.polynomial.legendre.leggauss(ngp)[0]
        for k in range(3):
            a, b = np.array([vx[(k + 1) % 3], vy[(k + 1) % 3]]), np.array([vx[(k + 2) % 3], vy[(k + 2) % 3]])
            pts = exy[2 * k * ngp:2 * (k + 1) * ngp].reshape(ngp, 2)
            want = a[None, :] + (1 + z)[:, None] / 2 * (b - a)[None, :]
            assert np.allclose(pts, want, atol=1e-13, rtol=0), (p, k)
