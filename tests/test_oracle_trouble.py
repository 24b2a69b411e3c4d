"""Closed-form checks of the oracle's trouble check (check_trouble, swe_trouble_check.hpp:5-116).

No reference test covers the `troubled` flag.  What the code fixes, checked on still water with a
step in ze along a grid line of a 0.1 m quad-split mesh (one stage of 1e-7 s, so the state is the
step):

- only interface faces through which "qn" integrates negative count, and the reference forms qn
  from the ze and qx rows of the trace (GlobalCoord::x / y indexing the state, :29-35): in still
  water qn = nx ze, so with ze > 0 the faces with nx < 0 count, with ze < 0 those with nx > 0;
- across the step those faces carry |int (h_in - h_ex)| = |dze| L, elsewhere zero; divided by
  r^((p+1)/2) x (counted perimeter) x (largest vertex depth), a step of 0.5 m is far above 1 and
  a step of 1 mm far below, so the initially troubled elements are exactly the ones on the side
  of the step whose face on it counts;
- six dilation passes then flag everything within five neighbour steps of those (each pass flags
  the current set and moves to its neighbours), and nothing else.
"""
import numpy as np
import pytest

from dgswemv2_amd import make_params, quad_mesh
from oracle.oracle_py import OracleSim
from test_levee import _mode0_scale


def _neighbours(mesh):
    edges = {}
    for e, nodes in enumerate(mesh.elt_nodes.tolist()):
        for k in range(3):
            edges.setdefault(frozenset((nodes[(k + 1) % 3], nodes[(k + 2) % 3])), []).append(e)
    nb = [set() for _ in range(mesh.n_elements)]
    for els in edges.values():
        if len(els) == 2:
            nb[els[0]].add(els[1]), nb[els[1]].add(els[0])
    return nb


def _ball(nb, seeds, radius):
    dist = {int(s): 0 for s in seeds}
    front = list(dist)
    for d in range(1, radius + 1):
        nxt = []
        for e in front:
            for n in nb[e]:
                if n not in dist:
                    dist[n] = d
                    nxt.append(n)
        front = nxt
    return set(dist)


def _troubled_after_step(mesh, p, zl, zr, x0):
    prm = make_params(p=p, nstages=1, order=1, dt=1e-7, g=9.81, slope_limiting=True, sl_M=1e-8, sl_nu=1.5)
    o = OracleSim(mesh, prm)
    cx = mesh.node_xyz[mesh.elt_nodes, 0].mean(axis=1)
    q = np.zeros_like(o.state()["q"])
    q[0, 0] = np.where(cx < x0, zl, zr) * _mode0_scale(mesh, p)
    o.set_state(q)
    o.stage()
    t = o.state()["troubled"].astype(bool)
    o.close()
    return t, cx


def _step_side(mesh, cx, x0, right):
    """Elements on one side of x = x0 with a face on that line."""
    xy = mesh.node_xyz[:, 0]
    out = []
    for e, nodes in enumerate(mesh.elt_nodes.tolist()):
        if (cx[e] > x0) != right:
            continue
        on_line = [abs(xy[v] - x0) < 1e-9 for v in nodes]
        if sum(on_line) == 2:
            out.append(e)
    return out


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("zl,zr,side", [(0.1, 0.6, "right"), (0.6, 0.1, "right"), (-0.2, -0.6, "left"),
                                        (-0.6, -0.2, "left")])
def test_trouble_check_step(p, zl, zr, side):
    mesh = quad_mesh(0.0, 0.0, 1.2, 0.8, 12, 8, bathy="constant", b0=1.0)
    x0 = 0.6
    troubled, cx = _troubled_after_step(mesh, p, zl, zr, x0)
    seeds = _step_side(mesh, cx, x0, right=(side == "right"))
    assert len(seeds) == 8
    want = _ball(_neighbours(mesh), seeds, 5)
    assert set(np.flatnonzero(troubled).tolist()) == want
    # the other side of the step alone would give another set (the side is what the check decides)
    other = _ball(_neighbours(mesh), _step_side(mesh, cx, x0, right=(side != "right")), 5)
    assert other != want


@pytest.mark.parametrize("p", [1, 2])
def test_trouble_check_small_step_and_flat(p):
    mesh = quad_mesh(0.0, 0.0, 1.2, 0.8, 12, 8, bathy="constant", b0=1.0)
    for zl, zr in ((0.3, 0.301), (0.3, 0.3)):
        troubled, _ = _troubled_after_step(mesh, p, zl, zr, 0.6)
        assert not troubled.any(), (zl, zr)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["plain", "fast", "fp32"])
@pytest.mark.parametrize("zl,zr,side", [(0.1, 0.6, "right"), (-0.2, -0.6, "left"), (0.3, 0.301, None)])
def test_gpu_trouble_check_step(precision, zl, zr, side):
    """The engine's trouble check on the same steps (p = 2): the same closed-form sets."""
    from dgswemv2_amd import RKDG_FP32_BULK, RKDG_FP64_FAST, RKDG_FP64_PLAIN, Simulation
    prec = {"plain": RKDG_FP64_PLAIN, "fast": RKDG_FP64_FAST, "fp32": RKDG_FP32_BULK}[precision]
    mesh = quad_mesh(0.0, 0.0, 1.2, 0.8, 12, 8, bathy="constant", b0=1.0)
    prm = make_params(p=2, nstages=1, order=1, dt=1e-7, g=9.81, slope_limiting=True, sl_M=1e-8, sl_nu=1.5)
    cx = mesh.node_xyz[mesh.elt_nodes, 0].mean(axis=1)
    sim = Simulation(mesh, prm, prec)
    q = np.zeros_like(sim.state()["q"])
    q[0, 0] = np.where(cx < 0.6, zl, zr) * _mode0_scale(mesh, 2)
    sim.set_state(q)
    sim.stage()
    got = set(np.flatnonzero(sim.state()["troubled"]).tolist())
    sim.close()
    want = set() if side is None else _ball(_neighbours(mesh), _step_side(mesh, cx, 0.6, side == "right"), 5)
    assert got == want
