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

from dgswemv2_amd import (RKDG_FP32_BULK, RKDG_FP64_PLAIN, Batch, RkdgError, Simulation, make_params,
                          workloads as W)

pytestmark = pytest.mark.gpu

SCHEDULES = [dict(block=128, groups_per_xcd=0, graph=0), dict(block=64, groups_per_xcd=32, graph=0),
             dict(block=256, groups_per_xcd=0, graph=1), dict(block=128, groups_per_xcd=64, graph=1)]


def banner(*a):
    print("[gpu-batch]", *a, flush=True)


def tidal_instances():
    out = []
    for scale, p, seed in ((0.03, 1, 3), (0.05, 2, 5), (0.02, 2, 7), (0.04, 1, 9)):
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


@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP32_BULK])
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


@pytest.mark.parametrize("nstages,order", [(2, 2), (3, 3)])
def test_wetting_drying_limiter_batch_equals_single_runs(nstages, order):
    # dam break with wetting-drying and the limiter; 3 stages exercise the odd-stage wet-flag copy
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
