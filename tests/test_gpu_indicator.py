"""Learned troubled-cell indicator on the GPU, against the CPU oracle (oracle/mlp.py).

- features, may-act flags and labels captured on the device vs the oracle's, on the dam break
  run in lockstep with the oracle (the plain FP64 engine is bit-exact with it, so both see the
  same pre-limiter state);
- the gate's BF16-MFMA logits vs the oracle network with the same bf16 rounding points;
- a gate open everywhere reproduces the plain limiter bit for bit (the gated limit pass is the
  plain one where it runs);
- the trainer's forward pass and gradient vs the oracle; training is deterministic and learns;
- what the trained network gains on the dam break: limiter time, misses, L2 against the plain
  limiter;
- retraining during the run (indicator.run_online) keeps the misses low once the flow has moved
  100 steps past the training window, where the fixed network misses many more.
"""
import numpy as np
import pytest

from conftest import relerr
from dgswemv2_amd import RKDG_FP32_BULK, RKDG_FP64_PLAIN, MlpTrainer, RkdgError, Simulation
from dgswemv2_amd import indicator as IND
from dgswemv2_amd import workloads as W
from oracle import mlp as O
from oracle.oracle_py import OracleSim

pytestmark = pytest.mark.gpu


def banner(*a):
    print("[gpu-indicator]", *a, flush=True)


def test_capture_matches_oracle():
    mesh, prm = W.dambreak(scale=0.1)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    o = OracleSim(mesh, prm)  # serial: the plain engine is bit-exact with it
    o.keep_prelimit()
    nb, bb = o.limiter_geometry()
    s.capture(True)
    npos = 0
    for k in range(14):
        s.stage()
        o.stage()
        c = s.get_capture()
        qo = o.prelimit()
        so = o.state()
        Xo, oko = O.features(qo, so["wet"], nb, bb, prm.g)
        yo = O.significant_labels(qo, so["q"], so["limited"], bb, prm.g)
        assert np.array_equal(c["may_act"] != 0, oko), k
        assert np.array_equal(c["label"].astype(float), yo), (k, int(np.sum(c["label"] != yo)))
        assert np.max(np.abs(c["features"] - Xo)) <= 2e-6, (k, np.max(np.abs(c["features"] - Xo)))
        npos += int(yo.sum())
        # the engine state still equals the oracle's (capture does not touch the state)
        assert np.array_equal(s.state()["q"], so["q"])
    assert npos > 100
    banner("capture: features/labels/may-act equal the oracle over 14 stages,", npos, "positive labels")
    s.close()


def _net(seed):
    p = IND.init_weights(seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    p[O.NH * O.NIN:O.NH * O.NIN + O.NH] = rng.normal(0, 0.2, O.NH)  # b1
    return p.astype(np.float32)


def _logit_close(z_gpu, z_ref):
    scale = np.maximum(1.0, np.abs(z_ref))
    d = np.abs(z_gpu - z_ref) / scale
    return np.quantile(d, 0.99), np.max(d)


@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP32_BULK])
def test_gate_logits_match_oracle(precision):
    mesh, prm = W.dambreak(scale=0.1)
    s = Simulation(mesh, prm, precision)
    flat = _net(11)
    thr = -1.5
    s.set_indicator(flat, threshold=thr)
    s.capture(True)
    s.step(5)
    c = s.get_capture(logits=True)
    z_ref, _ = O.forward(O.unflatten(flat.astype(np.float64)), c["features"].astype(np.float64), bf=True)
    q99, qmax = _logit_close(c["logit"], z_ref)
    banner("gate logits precision", precision, "rel diff q99", q99, "max", qmax)
    assert q99 <= 1e-4 and qmax <= 2e-2
    # the gate (reported as the troubled flags) opens where the limiter may act and logit > thr
    gate = s.state()["troubled"] != 0
    assert np.array_equal(gate, (c["may_act"] != 0) & (c["logit"] > np.float32(thr)))
    s.close()


@pytest.mark.parametrize("precision", [RKDG_FP64_PLAIN, RKDG_FP32_BULK])
def test_open_gate_equals_plain_limiter(precision):
    mesh, prm = W.dambreak(scale=0.1)
    a = Simulation(mesh, prm, precision)
    b = Simulation(mesh, prm, precision)
    b.set_indicator(_net(3), threshold=-3.0e38)
    a.step(20)
    b.step(20)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["q"], sb["q"]) and np.array_equal(sa["wet"], sb["wet"])
    assert np.array_equal(sa["limited"], sb["limited"])
    banner("open gate == plain limiter, precision", precision)
    a.close(), b.close()


def _synthetic(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (36, n)).astype(np.float32)
    y = ((X[0] + 0.5 * X[9] - 0.3 * X[20]) > 0.2).astype(np.float32)
    return X, y


def test_trainer_forward_and_gradient_match_oracle():
    X, y = _synthetic(1000, 1)
    flat = _net(7)
    t = MlpTrainer(flat, capacity=4096)
    z = t.forward(X)
    z_ref, _ = O.forward(O.unflatten(flat.astype(np.float64)), X.astype(np.float64), bf=True)
    q99, qmax = _logit_close(z, z_ref)
    assert q99 <= 1e-4 and qmax <= 2e-2, (q99, qmax)
    for n in (1000, 37, 16):  # batch not a multiple of the 16-sample tile / 32-sample K step
        loss, g = t.gradient(X[:, :n], y[:n], pos_weight=3.0)
        lref, gref = O.loss_grad(O.unflatten(flat.astype(np.float64)), X[:, :n].astype(np.float64), y[:n], 3.0, bf=True)
        assert loss == pytest.approx(lref, rel=1e-5)
        gref = O.flatten(gref)
        off = 0
        for name, shp in O.SHAPES:
            m = int(np.prod(shp)) if shp else 1
            a, b = g[off:off + m], gref[off:off + m]
            assert np.linalg.norm(a - b) <= 2e-3 * np.linalg.norm(b) + 1e-9, (n, name, relerr(a[None], b[None]))
            off += m
        banner("trainer gradient n =", n, "loss", loss, "matches the oracle")
    t.close()


def test_trainer_deterministic_and_learns():
    X, y = _synthetic(20000, 2)
    w = []
    for _ in range(2):
        t = MlpTrainer(_net(5), capacity=20000)
        t.set_data(X, y)
        losses = t.train(6, batch=512, lr=3e-3, pos_weight=1.0, seed=9)
        w.append(t.weights())
        z = t.forward(X)
        t.close()
    assert np.array_equal(w[0], w[1])
    acc = np.mean((z > 0) == (y > 0.5))
    banner("synthetic training losses", losses, "accuracy", acc)
    assert losses[-1] < 0.5 * losses[0] and acc > 0.97


def test_indicator_errors():
    mesh, prm = W.tidal(scale=0.02, p=1)  # no limiter
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    with pytest.raises(RkdgError):
        s.set_indicator(_net(1))
    with pytest.raises(RkdgError):
        s.capture(True)
    s.close()
    mesh, prm = W.dambreak(scale=0.05)
    s = Simulation(mesh, prm, RKDG_FP64_PLAIN)
    with pytest.raises(RkdgError):
        s.get_capture()
    s.capture(True)
    s.step(1)
    with pytest.raises(RkdgError):
        s.get_capture(logits=True)  # no indicator set
    s.close()
    t = MlpTrainer(_net(1), capacity=10)
    X, y = _synthetic(11, 3)
    with pytest.raises(RkdgError):
        t.set_data(X, y)
    t.close()


def test_learned_indicator_gain():
    """Dam break (FP32 bulk, 45,000 triangles): the plain limiter labels the first 40 steps, the
    network is trained on the GPU, then the run continues from the same state with the plain
    limiter and with the learned gate.  Over the 10 steps after training the judge (the plain
    limiter's own decision on the learned run's states) counts the misses; the L2 difference to
    the plain run and the limiter time of both are reported.  The network is trained on the
    states it has seen: its misses grow as the flow moves on (documented), so the stated miss
    rate is that of the window right after training."""
    mesh, prm = W.dambreak(scale=0.3)
    a = Simulation(mesh, prm, RKDG_FP32_BULK)  # plain
    b = Simulation(mesh, prm, RKDG_FP32_BULK)  # learned
    X, y, nst = IND.collect(a, 40, every=1, neg_keep=0.25, seed=1)
    assert np.all(np.isfinite(X))
    b.step(40)
    assert np.array_equal(a.state()["q"], b.state()["q"])
    w, thr, info = IND.train(X, y, epochs=6, batch=1024, pos_weight=8.0, recall=0.9999, margin=16.0)
    banner("training:", info)
    b.set_indicator(w, threshold=thr, judge=True)
    a.step(10)
    b.step(10)
    st = b.indicator_stats()
    err = relerr(b.state()["q"], a.state()["q"])
    miss_sig = st["missed_significant"] / max(1, st["significant"])
    banner("judged over 10 steps:", st, "missed significant share", miss_sig, "rel L2 learned vs plain", err)
    # timing without the judge's dry run
    b.set_indicator(w, threshold=thr, judge=False)
    a.reset_timers(), b.reset_timers()
    a.step(10)
    b.step(10)
    ra, rb = a.report(), b.report()
    lim_a = ra["limiter_kernel_ms"] / ra["stages_timed"]
    lim_b = rb["limiter_kernel_ms"] / rb["stages_timed"]
    stage_a = (ra["stage_kernel_ms"] + ra["limiter_kernel_ms"]) / ra["stages_timed"]
    stage_b = (rb["stage_kernel_ms"] + rb["limiter_kernel_ms"]) / rb["stages_timed"]
    banner("limiter ms/stage plain", lim_a, "learned", lim_b, "stage+limiter ms plain", stage_a, "learned", stage_b)
    assert info["val_recall"] >= 0.999
    assert miss_sig <= 0.01 and st["max_missed_change"] <= 1e-2
    assert err[0] <= 1e-6 and err[1] <= 1e-5
    assert lim_b < lim_a
    a.close(), b.close()


def test_online_retraining_500k():
    """500k dam break: trained on 40 warm-up steps, then 100 more steps with the gate fixed or
    retrained every 20 steps on judged data, or checked every 5 and retrained only above 1 % misses;
    the judge over the next 10 steps.  Measured: fixed misses 8.1 % of the significant write-backs
    (largest missed change 7e-3), every 20 steps 0.7 % (2e-4), triggered 0.13 % (5e-5) in 2 rounds."""
    kw = dict(batch=1024, pos_weight=8.0, recall=0.9999, margin=16.0)
    mesh, prm = W.dambreak(scale=1.0)
    banner("online retraining, dam break", mesh.n_elements)
    a = Simulation(mesh, prm, RKDG_FP32_BULK)
    X, y, _ = IND.collect(a, 40, every=1, neg_keep=0.25, seed=1)
    w, thr, _ = IND.train(X, y, epochs=4, **kw)
    st = a.state()
    a.close()
    out = {}
    for mode in ("fixed", "online", "triggered"):
        s = Simulation(mesh, prm, RKDG_FP32_BULK)
        s.set_state(st["q"], st["wet"])
        if mode == "fixed":
            s.set_indicator(w, threshold=thr)
            s.step(100)
            w2, thr2, rounds = w, thr, []
        elif mode == "online":
            w2, thr2, _, _, rounds = IND.run_online(s, w, thr, 100, X=X, y=y, every=20, window=1, epochs=2, **kw)
        else:  # a captured step every 5, retraining only when the gate missed over 1 % of its positives
            w2, thr2, _, _, rounds = IND.run_online(s, w, thr, 100, X=X, y=y, every=5, window=1, epochs=2,
                                                    retrain_above=0.01, **kw)
        s.set_indicator(w2, threshold=thr2, judge=True)
        s.step(10)
        j = s.indicator_stats()
        s.close()
        out[mode] = (j["missed_significant"] / max(1, j["significant"]), j["max_missed_change"], len(rounds))
        banner(mode, out[mode])
    assert out["online"][2] == 4 and out["fixed"][2] == 0
    assert out["online"][0] < 0.25 * out["fixed"][0] and out["online"][0] < 0.02
    assert out["online"][1] < out["fixed"][1]
    # measured: 2 rounds, 0.13 % missed (largest 5e-5)
    assert 1 <= out["triggered"][2] < 4 and out["triggered"][0] < 0.02
