"""Ensembles (rkdg_create_batch): every instance of a batch evolves exactly as the same mesh run
alone, under every schedule (one work-group per tile or persistent, eager or graph replay).

Instances of one batch are merged per polynomial order into one element range (each padded to
whole 256-element tiles) with per-instance dt, time, ramp and forcing, so bitwise equality with
the single-mesh engine checks the merge (index offsets, edge/forcing/levee record bases, padding),
the per-instance coefficient tables, the device-side stepper advance, the in-place last stage and
graph replay all at once.
"""
import numpy as np
import pytest

from conftest import relerr
from dgswemv2_amd import (RKDG_FP32_BULK, RKDG_FP64_FAST, RKDG_FP64_PLAIN, Batch, RkdgError, Simulation, make_params,
                          workloads as W)

pytestmark = pytest.mark.gpu

SCHEDULES = [dict(block=128, groups_per_xcd=0, graph=0), dict(block=64, groups_per_xcd=32, graph=0),
             dict(block=64, groups_per_xcd=-1, graph=0), dict(block=256, groups_per_xcd=-1, graph=1),
             dict(block=256, groups_per_xcd=0, graph=1), dict(block=128, groups_per_xcd=64, graph=1)]


def banner(*a):
    print("[gpu-batch]", *a, flush=True)


def tidal_instances():
    out = []
    for scale, p, seed in ((0.03, 1, 3), (0.05, 2, 5), (0.02, 2, 7), (0.04, 1, 9), (0.03, 3, 11)):
        mesh, prm = W.tidal(scale=scale, p=p, seed=seed)
        out.append((mesh, prm))
    return out


def singles(cases, precision, steps):
    res = []
    for mesh, prm in cases:
        s = Simulation(mesh, prm, precision)
        s.step(steps)
        res.append(s.state())
        s.close()
    return res


@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP64_FAST, RKDG_FP32_BULK])
def test_tidal_batch_equals_single_runs(precision):
    cases = tidal_instances()
    steps = 6
    ref = singles(cases, precision, steps)
    for sc in SCHEDULES:
        b = Batch([c[0] for c in cases], [c[1] for c in cases], precision)
        b.set_schedule(**sc)
        b.step(steps)
        for i in range(len(cases)):
            got = b.state(i)
            assert np.array_equal(got["q"], ref[i]["q"]), (sc, i)
            t, st, sg = b.time(i)
            assert st == steps and sg == 0 and t == pytest.approx(steps * cases[i][1].dt, rel=1e-12)
        banner("tidal batch", precision, sc, "bitwise equal to single runs")
        b.close()


@pytest.mark.parametrize("nstages,order", [(2, 2), (3, 3), (5, 4), (8, 4)])
def test_wetting_drying_limiter_batch_equals_single_runs(nstages, order):
    # dam break with wetting-drying and the limiter; 3 and 5 stages exercise the odd-stage wet-flag
    # copy, 8 stages the per-instance coefficient table at its full eight terms
    cases = []
    for scale in (0.04, 0.06, 0.05):
        mesh, prm = W.dambreak(scale=scale)
        prm.nstages, prm.order = nstages, order
        cases.append((mesh, prm))
    steps = 8
    ref = singles(cases, RKDG_FP64_PLAIN, steps)
    for sc in (SCHEDULES[0], SCHEDULES[3]):
        b = Batch([c[0] for c in cases], [c[1] for c in cases], RKDG_FP64_PLAIN)
        b.set_schedule(**sc)
        b.step(steps)
        for i in range(len(cases)):
            got = b.state(i)
            for f in ("q", "wet", "limited", "troubled"):
                assert np.array_equal(got[f], ref[i][f]), (sc, i, f)
        banner("dam-break batch ns", nstages, sc, "bitwise equal incl. flags")
        b.close()


def test_autotune_restores_state_and_picks_a_schedule():
    cases = tidal_instances()
    b = Batch([c[0] for c in cases], [c[1] for c in cases], RKDG_FP32_BULK)
    b.step(2)
    before = [b.state(i)["q"].copy() for i in range(len(cases))]
    best, ms = b.autotune(steps_per_trial=2)
    banner("autotune", best, "ms/step", ms)
    assert best["block"] in (64, 128, 256) and best["graph"] in (0, 1) and ms > 0
    assert b.schedule == best
    for i in range(len(cases)):
        assert np.array_equal(b.state(i)["q"], before[i])
        assert b.time(i)[1] == 2
    b.step(3)
    ref = singles(cases, RKDG_FP32_BULK, 5)
    for i in range(len(cases)):
        assert np.array_equal(b.state(i)["q"], ref[i]["q"])
    b.close()


def test_batch_nan_is_located():
    # a NaN planted in instance 2 is reported with that instance's own element numbering, as the
    # same mesh run alone reports it (the NaN reaches neighbours within the stage, so the lowest
    # index among them is reported)
    cases = tidal_instances()
    b = Batch([c[0] for c in cases], [c[1] for c in cases], RKDG_FP64_PLAIN)
    q = b.state(2)["q"].copy()
    q[1, 0, 17] = np.nan
    b.set_state(2, q)
    with pytest.raises(RkdgError):
        b.step(1)
    inst, elt, step, stage = b.nan_info()
    s = Simulation(cases[2][0], cases[2][1], RKDG_FP64_PLAIN)
    s.set_state(q)
    with pytest.raises(RkdgError):
        s.step(1)
    assert (inst, elt, step, stage) == (2,) + tuple(s.nan_info())
    assert elt >= 0 and step == 0
    b.close()


def test_batch_rejects_mixed_schemes():
    m1, p1 = W.tidal(scale=0.02, p=1)
    m2, p2 = W.dambreak(scale=0.03)
    with pytest.raises(RkdgError):
        Batch([m1, m2], [p1, p2], RKDG_FP64_PLAIN)


def _area(mesh):
    x, y = mesh.node_xyz[:, 0], mesh.node_xyz[:, 1]
    e = mesh.elt_nodes
    return 0.5 * np.abs((x[e[:, 1]] - x[e[:, 0]]) * (y[e[:, 2]] - y[e[:, 0]]) -
                        (x[e[:, 2]] - x[e[:, 0]]) * (y[e[:, 1]] - y[e[:, 0]]))


def test_ensemble_64_instances_full_size():
    """The 64-instance ensemble at full size (triangle counts log-uniform over 20k-400k, 7.63 M in
    all, p in {1, 2}, tide 0.1-1.5 m, SSP-RK3) in FP32 on the schedule the autotuner picks, 5 steps:
    every instance equals its single-mesh run bit for bit and stays finite.  A land-only twin (the
    tide boundary closed, started from the moving state) conserves each instance's mass.  The two
    smallest instances (one of each order) run as an FP64-plain batch against the oracle per step."""
    from dgswemv2_amd import Mesh, BC_LAND, BC_TIDE
    from oracle.oracle_py import OracleSim
    meshes, params, info = W.ensemble(n=64, scale=1.0)
    banner("ensemble 64: triangles", sum(i["triangles"] for i in info))
    b = Batch(meshes, params, RKDG_FP32_BULK)
    best, ms = b.autotune(steps_per_trial=2)
    banner("ensemble 64 autotuned", best, "ms/step", ms)
    steps = 5
    b.step(steps)
    states = [b.state(i) for i in range(64)]
    b.close()
    for i in range(64):
        assert np.all(np.isfinite(states[i]["q"])), i
        s = Simulation(meshes[i], params[i], RKDG_FP32_BULK)
        s.step(steps)
        assert np.array_equal(states[i]["q"], s.state()["q"]), i
        s.close()
    banner("ensemble 64: every instance equals its single run")

    # land-only twin from the moving state: mass (sum of mode 0 x area) kept to FP32 roundoff
    land = [Mesh(m.node_xyz, m.elt_nodes, m.bound_edges, np.where(m.bound_types == BC_TIDE, BC_LAND, m.bound_types),
                 node_manning_n=m.node_manning_n) for m in meshes]
    t = Batch(land, params, RKDG_FP32_BULK)
    t.set_schedule(**best)
    for i in range(64):
        t.set_state(i, states[i]["q"])
    areas = [_area(m) for m in land]
    m0 = [float(np.sum(states[i]["q"][0, 0].astype(np.float64) * areas[i])) for i in range(64)]
    t.step(steps)
    for i in range(64):
        q = t.state(i)["q"]
        m1 = float(np.sum(q[0, 0] * areas[i]))
        scale = float(np.sum(np.abs(q[0, 0]) * areas[i])) + 1e-30
        assert abs(m1 - m0[i]) <= 1e-6 * scale, (i, m0[i], m1)
    t.close()
    banner("ensemble 64: land-only twin conserves mass")

    # the two smallest instances, one per order, in FP64 plain against the oracle per step
    order = sorted(range(64), key=lambda i: info[i]["triangles"])
    pick = [next(i for i in order if info[i]["p"] == 1), next(i for i in order if info[i]["p"] == 2)]
    f = Batch([meshes[i] for i in pick], [params[i] for i in pick], RKDG_FP64_PLAIN)
    f.set_schedule(**best)
    orc = [OracleSim(meshes[i], params[i]) for i in pick]
    for s in range(3):
        f.step(1)
        for k, o in enumerate(orc):
            o.step(1)
            r = max(relerr(f.state(k)["q"], o.state()["q"]))
            assert r <= 1e-12, (pick[k], s, r)
    banner("ensemble 64: smallest instances", [info[i]["triangles"] for i in pick], "match the oracle")
    f.close()
