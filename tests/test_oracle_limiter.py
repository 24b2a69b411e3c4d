"""Closed-form pins of the oracle's Cockburn-Shu limiting step (swe_CS_slope_limiter.hpp:88-189 with
R / invR of swe_jacobian.hpp:462-480, 531-550).  The reference has no limiter test; these checks
fix the step by what the mathematics and hand-worked arithmetic say:

  (i)   R holds the right eigenvectors of the shallow-water flux Jacobian along a unit direction
        (built here by central differences of the physical flux, not from R), with eigenvalues
        un - c, un, un + c, and invR is its inverse;
  (ii)  a globally linear ze / qx / qy field is returned unchanged, and unflagged, by every element
        whose three neighbours are interior (nu = 1.5, |w_tilda| far above M r^2, so the minmod
        branch decides);
  (iii) a strict local maximum of ze over still water is flattened to its barycentre value, its
        higher modes zeroed and its mean kept;
  (iv)  one element worked by hand: a characteristic clipped to nu |w_delta|, one zeroed on a sign
        mismatch, one kept by the TVB bound, and the positive / negative balance rescaling the
        deltas of every variable;
  (v)   with |w_tilda| <= M r^2 everywhere the limiter keeps w_tilda, which returns a linear
        element unchanged even where its neighbours disagree with it.

Each check has twins with a mistake planted in the oracle (orc_set_plant, CPU only) that must fail
it: a sign in invR, R transposed, alpha's columns exchanged, the vertex matrix T with flipped
signs, theta+ / theta- exchanged, the balance skipped, nu dropped, the TVB branch skipped.
"""
import ctypes as C

import numpy as np
import pytest

from dgswemv2_amd import BC_LAND, _native as N, make_params, quad_mesh
from oracle.oracle_py import OracleSim, load

(PLANT_NONE, PLANT_INVR_SIGN, PLANT_ALPHA_SWAP, PLANT_T_SIGN, PLANT_THETA_SWAP, PLANT_NU_ONE,
 PLANT_R_TRANSPOSE, PLANT_NO_THETA, PLANT_TVB_OFF) = range(9)


@pytest.fixture(autouse=True)
def _no_plant():
    load().orc_set_plant(PLANT_NONE)
    yield
    load().orc_set_plant(PLANT_NONE)


def planted(kind, fn):
    """Runs fn() with mistake `kind` planted in the oracle; True when fn's check fails."""
    lib = load()
    lib.orc_set_plant(kind)
    try:
        fn()
        return False
    except AssertionError:
        return True
    finally:
        lib.orc_set_plant(PLANT_NONE)


def R_of(g, h, u, v, nx, ny):
    out = np.empty(9)
    load().orc_R(g, h, u, v, nx, ny, N.dptr(out))
    return out.reshape(3, 3)


def invR_of(g, h, u, v, nx, ny):
    out = np.empty(9)
    load().orc_invR(g, h, u, v, nx, ny, N.dptr(out))
    return out.reshape(3, 3)


def limit_element(q_lin, q_nb, bath_bar, median, alpha, r_sq, g=1.0, M=1e-8, nu=1.5):
    """The oracle's limiting step on one element; returns (vertex values, deltas, limited)."""
    arrs = [np.ascontiguousarray(a, np.float64).ravel() for a in (q_lin, q_nb, median, alpha, r_sq)]
    qv, dl = np.empty(9), np.empty(9)
    lim = np.empty(3, np.int32)
    load().orc_cs_limit_element(g, M, nu, N.dptr(arrs[0]), N.dptr(arrs[1]), bath_bar, N.dptr(arrs[2]),
                                N.dptr(arrs[3]), N.dptr(arrs[4]), N.dptr(qv), N.dptr(dl),
                                lim.ctypes.data_as(C.POINTER(C.c_int32)))
    return qv.reshape(3, 3), dl.reshape(3, 3), lim


# ---- (i) R and invR ---------------------------------------------------------------------------

def flux_n(q, b, g, nx, ny):
    """F(q) . n of the shallow-water equations in (ze, qx, qy) with h = ze + b and the pressure
    g (ze^2 / 2 + ze b) (swe_flux.hpp:5-25)."""
    ze, qx, qy = q
    h = ze + b
    pe = g * (0.5 * ze * ze + ze * b)
    fx = np.array([qx, qx * qx / h + pe, qx * qy / h])
    fy = np.array([qy, qx * qy / h, qy * qy / h + pe])
    return fx * nx + fy * ny


def jacobian_fd(q, b, g, nx, ny):
    A = np.empty((3, 3))
    for j in range(3):
        s = 1e-6 * max(1.0, abs(q[j]))
        qp, qm = np.array(q, float), np.array(q, float)
        qp[j] += s
        qm[j] -= s
        A[:, j] = (flux_n(qp, b, g, nx, ny) - flux_n(qm, b, g, nx, ny)) / (2 * s)
    return A


def check_eigensystem(seed=3):
    rng = np.random.default_rng(seed)
    for _ in range(20):
        g = 9.81
        b = rng.uniform(1.0, 50.0)
        ze = rng.uniform(-0.5, 2.0)
        h = ze + b
        u, v = rng.uniform(-2, 2, 2)
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        R, iR = R_of(g, h, u, v, nx, ny), invR_of(g, h, u, v, nx, ny)
        assert np.allclose(R @ iR, np.eye(3), atol=1e-13, rtol=0)
        assert np.allclose(iR @ R, np.eye(3), atol=1e-13, rtol=0)
        A = jacobian_fd([ze, u * h, v * h], b, g, nx, ny)
        c, un = np.sqrt(g * h), u * nx + v * ny
        D = iR @ A @ R
        assert np.allclose(D, np.diag([un - c, un, un + c]), atol=1e-6 * (abs(un) + c), rtol=0), D


def test_R_invR_eigensystem():
    check_eigensystem()


@pytest.mark.parametrize("plant", [PLANT_INVR_SIGN, PLANT_R_TRANSPOSE])
def test_R_invR_eigensystem_planted(plant):
    assert planted(plant, check_eigensystem)


# ---- (ii) linear fields ------------------------------------------------------------------------

def linear_modes(mesh, p, coef):
    """Modal state of the linear field q_k(x, y) = c_k0 + c_k1 x + c_k2 y: only the three linear
    Dubiner modes (dofs 0-2, basis_dubiner_2D.cpp:18-23) are nonzero, solved from the vertex values
    (a clockwise element's nodes 1 and 2 swap coordinates, shape_straighttriangle.cpp:9-11)."""
    lib = load()
    nd = (p + 1) * (p + 2) // 2
    pts = np.array([-1.0, -1.0, 1.0, -1.0, -1.0, 1.0])
    phi, d1, d2 = np.empty(nd * 3), np.empty(nd * 3), np.empty(nd * 3)
    assert lib.orc_dubiner(p, 3, N.dptr(pts), N.dptr(phi), N.dptr(d1), N.dptr(d2)) == nd
    Pv = phi.reshape(nd, 3)[:3].T  # [vertex][linear dof]
    xy = mesh.node_xyz[:, :2][mesh.elt_nodes]  # [ne][3][2]
    det = ((xy[:, 1, 0] - xy[:, 0, 0]) * (xy[:, 2, 1] - xy[:, 0, 1]) -
           (xy[:, 2, 0] - xy[:, 0, 0]) * (xy[:, 1, 1] - xy[:, 0, 1]))
    cw = det < 0
    xy[cw, 1], xy[cw, 2] = xy[cw, 2].copy(), xy[cw, 1].copy()
    q = np.zeros((3, nd, mesh.n_elements))
    Pinv = np.linalg.inv(Pv)
    for k in range(3):
        fv = coef[k][0] + coef[k][1] * xy[..., 0] + coef[k][2] * xy[..., 1]  # [ne][3]
        q[k, :3] = (fv @ Pinv.T).T
    return q


def interior_elements(o):
    nb, _ = o.limiter_geometry()
    inner = np.all(nb >= 0, axis=0)
    return np.array([e for e in range(o.ne) if inner[e] and np.all(inner[nb[:, e]])])


def check_linear_field_unchanged(p, jitter):
    mesh = quad_mesh(0, 0, 2000, 2000, 10, 10, bc=(BC_LAND,) * 4, bathy="constant", b0=20.0, jitter=jitter, seed=11)
    prm = make_params(p=p, nstages=2, order=2, dt=1.0, ze0=0.0, slope_limiting=True, sl_M=1e-8, sl_nu=1.5)
    o = OracleSim(mesh, prm)
    coef = [(0.5, 1e-3, -2e-3), (1.0, 2e-3, 5e-4), (-0.5, -1e-3, 3e-3)]
    q0 = linear_modes(mesh, p, coef)
    o.set_state(q0, np.ones(mesh.n_elements, np.uint8))
    assert o.lib.orc_apply_limiter(o.h) == 0
    s = o.state()
    inner = interior_elements(o)
    assert inner.size > 20
    # the minmod branch decides: on every median of these elements the largest characteristic
    # |w_tilda| = |invR (q_mid - q_bar)| is far above M r^2 (invR as pinned by the eigensystem check)
    xy = mesh.node_xyz[:, :2][mesh.elt_nodes]
    med, alp, rsq = np.empty(6), np.empty(6), np.empty(3)
    for e in inner:
        o.lib.orc_limiter_element_geometry(o.h, int(e), N.dptr(med), N.dptr(alp), N.dptr(rsq))
        qv = np.array([[c[0] + c[1] * x + c[2] * y for (x, y) in xy[e]] for c in coef])  # var x vertex
        bar = qv.mean(axis=1)
        h = bar[0] + 20.0
        for b in range(3):
            mid = 0.5 * (qv[:, (b + 1) % 3] + qv[:, (b + 2) % 3])
            wt = invR_of(9.81, h, bar[1] / h, bar[2] / h, med[2 * b], med[2 * b + 1]) @ (mid - bar)
            assert np.max(np.abs(wt)) > 100 * 1e-8 * rsq[b]
    scale = np.max(np.abs(q0), axis=(1, 2))
    for k in range(3):
        assert np.max(np.abs(s["q"][k][:, inner] - q0[k][:, inner])) <= 1e-12 * scale[k], k
        assert s["limited"][k][inner].sum() == 0, k
    o.close()


@pytest.mark.parametrize("p,jitter", [(1, 0.0), (1, 0.2), (2, 0.2)])
def test_linear_field_unchanged(p, jitter):
    check_linear_field_unchanged(p, jitter)


@pytest.mark.parametrize("plant", [PLANT_INVR_SIGN, PLANT_ALPHA_SWAP, PLANT_T_SIGN, PLANT_R_TRANSPOSE])
def test_linear_field_unchanged_planted(plant):
    assert planted(plant, lambda: check_linear_field_unchanged(1, 0.2))


# ---- (iii) a strict local maximum -------------------------------------------------------------

def check_local_max_flattened(p):
    mesh = quad_mesh(0, 0, 1000, 1000, 8, 8, bc=(BC_LAND,) * 4, bathy="constant", b0=10.0, jitter=0.1, seed=5)
    prm = make_params(p=p, nstages=2, order=2, dt=1.0, ze0=0.0, slope_limiting=True, sl_M=1e-8, sl_nu=1.5)
    o = OracleSim(mesh, prm)
    nd = (p + 1) * (p + 2) // 2
    inner = interior_elements(o)
    e = int(inner[len(inner) // 2])
    q = np.zeros((3, nd, mesh.n_elements))
    # a tilted bump on element e: mean 0.2 (mode 0; phi_0 = 1) above its still neighbours, every
    # higher mode set
    q[0, :, e] = np.concatenate([[0.2], 0.006 * np.arange(1, nd)])
    o.set_state(q, np.ones(mesh.n_elements, np.uint8))
    assert o.lib.orc_apply_limiter(o.h) == 0
    s = o.state()
    assert s["limited"][0][e] == 1
    assert s["q"][0, 0, e] == pytest.approx(q[0, 0, e], rel=1e-14)  # the mean is kept
    assert np.max(np.abs(s["q"][0, 1:, e])) <= 1e-15                  # every higher mode zeroed
    assert np.max(np.abs(s["q"][1:, :, e])) == 0.0                    # momentum stays zero
    others = np.ones(mesh.n_elements, bool)
    others[e] = False
    assert np.array_equal(s["q"][:, :, others], q[:, :, others])
    o.close()


@pytest.mark.parametrize("p", [1, 2])
def test_local_max_flattened(p):
    check_local_max_flattened(p)


@pytest.mark.parametrize("plant", [PLANT_THETA_SWAP, PLANT_NO_THETA])
def test_local_max_flattened_planted(plant):
    assert planted(plant, lambda: check_local_max_flattened(1))


# ---- (iv) one element by hand -----------------------------------------------------------------
#
# g = 1, barycentre state (ze, qx, qy) = (1, 0, 0) over bathymetry 3: h = 4, u = v = 0, c = 2.
# With u = v = 0 the characteristic variables along the unit median m are
#   w1 = ze/2 - (m . q)/4,   w2 = m_perp . q = -my qx + mx qy,   w3 = ze/2 + (m . q)/4
# and back, q = w1 r1 + w2 r2 + w3 r3 with r1 = (1, -2 mx, -2 my), r2 = (0, -my, mx), r3 = (1, 2 mx, 2 my).
# Vertex values (var x vertex): ze (0.6, 1.2, 1.2), qx (-0.8, 0.6, 0.2), qy (0, -0.1, 0.1); the edge
# midpoints (edge b opposite vertex b) are ze (1.2, 0.9, 0.9), qx (0.4, -0.3, -0.1), qy (0, 0.05, -0.05).
# Neighbour barycentres: across face 0 (1.1, 0, 0), face 1 (1.3, 0.2, 0), face 2 (0.8, 0, 0).
# Median b uses neighbours b and b+1 (mod 3) with weights alpha_b.  M = 1e-8, r^2 = 1, nu = 1.5.
#
# b = 0, m = (1, 0), alpha = (0.5, 0.5): w(mid) = (0.5, 0, 0.7), w(bar) = (0.5, 0, 0.5),
#   w(nb0) = (0.55, 0, 0.55), w(nb1) = (0.6, 0, 0.7): w_tilda = (0, 0, 0.2),
#   w_delta = 0.5 (0.05, 0, 0.05) + 0.5 (0.1, 0, 0.2) = (0.075, 0, 0.125)
#   -> (0 kept by the TVB bound, 0 kept, min(0.2, 1.5 * 0.125) = 0.1875: clipped to nu |w_delta|)
#   delta_0 = 0.1875 r3 = (0.1875, 0.375, 0)
# b = 1, m = (0, 1), alpha = (0.25, 0.75): w(mid) = (0.4375, 0.3, 0.4625), w(nb1) = (0.65, -0.2, 0.65),
#   w(nb2) = (0.4, 0, 0.4): w_tilda = (-0.0625, 0.3, -0.0375),
#   w_delta = 0.25 (0.15, -0.2, 0.15) + 0.75 (-0.1, 0, -0.1) = (-0.0375, -0.05, -0.0375)
#   -> (-min(0.0625, 0.05625) = -0.05625: clipped; 0: signs differ; -0.0375)
#   delta_1 = -0.05625 (1, 0, -2) - 0.0375 (1, 0, 2) = (-0.09375, 0, 0.0375)
# b = 2, m = (-0.6, -0.8), alpha = (1.0, 0.2): w(mid) = (0.425, -0.05, 0.475), w(nb2) = (0.4, 0, 0.4),
#   w(nb0) = (0.55, 0, 0.55): w_tilda = (-0.075, -0.05, -0.025), w_delta = (-0.09, +0, -0.09)
#   -> (-0.075, 0: signs differ, -0.025)
#   delta_2 = -0.075 (1, 1.2, 1.6) - 0.025 (1, -1.2, -1.6) = (-0.1, -0.06, -0.08)
# Balance per variable (theta+ = min(1, N/P), theta- = min(1, P/N)):
#   ze (0.1875, -0.09375, -0.1): P = 0.1875, N = 0.19375 -> theta- = 30/31: (0.1875, -0.09375*30/31, -0.1*30/31)
#   qx (0.375, 0, -0.06): P = 0.375, N = 0.06 -> theta+ = 0.16: (0.06, 0, -0.06)
#   qy (0, 0.0375, -0.08): P = 0.0375, N = 0.08 -> theta- = 0.46875: (0, 0.0375, -0.0375)
# Vertex v = barycentre + sum_b delta_b T[b][v], T = -1 on the diagonal, +1 off it:
#   ze: v0 = 1 - 0.1875 - 0.1875 = 0.625, v1 = 1 + 0.1875 + (0.09375 - 0.1) * 30/31, v2 = 1 + 0.1875 + (0.1 - 0.09375) * 30/31
#   qx: (-0.12, 0, 0.12); qy: (0, -0.075, 0.075); all three variables change, so all are written back.

HAND = dict(
    q_lin=[[0.6, 1.2, 1.2], [-0.8, 0.6, 0.2], [0.0, -0.1, 0.1]],
    q_nb=[[1.1, 0.0, 0.0], [1.3, 0.2, 0.0], [0.8, 0.0, 0.0]],
    bath_bar=3.0,
    median=[[1.0, 0.0], [0.0, 1.0], [-0.6, -0.8]],
    alpha=[[0.5, 0.5], [0.25, 0.75], [1.0, 0.2]],
    r_sq=[1.0, 1.0, 1.0],
)
HAND_DELTA = np.array([[0.1875, -0.09375 * 30 / 31, -0.1 * 30 / 31], [0.06, 0.0, -0.06], [0.0, 0.0375, -0.0375]])
HAND_VRTX = np.array([[0.625, 1.1875 + (0.09375 - 0.1) * 30 / 31, 1.1875 + (0.1 - 0.09375) * 30 / 31],
                      [-0.12, 0.0, 0.12], [0.0, -0.075, 0.075]])


def check_hand_worked():
    qv, dl, lim = limit_element(**HAND)
    assert np.allclose(dl, HAND_DELTA, atol=1e-14, rtol=0), dl
    assert np.allclose(qv, HAND_VRTX, atol=1e-14, rtol=0), qv
    assert lim.tolist() == [1, 1, 1]


def test_hand_worked_element():
    check_hand_worked()


@pytest.mark.parametrize("plant", [PLANT_INVR_SIGN, PLANT_ALPHA_SWAP, PLANT_T_SIGN, PLANT_THETA_SWAP, PLANT_NU_ONE,
                                   PLANT_R_TRANSPOSE, PLANT_NO_THETA])
def test_hand_worked_element_planted(plant):
    assert planted(plant, check_hand_worked)


# ---- (v) the TVB bound keeps w_tilda ----------------------------------------------------------

def check_tvb_keeps():
    # The hand-worked element with M r^2 = 1 above every |w_tilda| (at most 0.3 here): the limiter
    # keeps w_tilda on every median, delta_b = R invR (mid_b - bar) = mid_b - bar, and the vertex
    # rebuild bar - delta_v + sum of the others returns the linear vertex values: no write-back.
    args = dict(HAND)
    qv, dl, lim = limit_element(M=1.0, **args)
    ql = np.array(HAND["q_lin"])
    mids = np.array([[(r[1] + r[2]) / 2, (r[0] + r[2]) / 2, (r[0] + r[1]) / 2] for r in ql])
    bar = ql.mean(axis=1)
    assert np.allclose(dl, mids - bar[:, None], atol=1e-14, rtol=0), dl
    assert np.allclose(qv, ql, atol=1e-14, rtol=0)
    assert lim.tolist() == [0, 0, 0]


def test_tvb_bound_keeps_w_tilda():
    check_tvb_keeps()


@pytest.mark.parametrize("plant", [PLANT_TVB_OFF, PLANT_INVR_SIGN, PLANT_R_TRANSPOSE, PLANT_T_SIGN])
def test_tvb_bound_keeps_w_tilda_planted(plant):
    assert planted(plant, check_tvb_keeps)
