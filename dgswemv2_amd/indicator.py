"""Learned troubled-cell indicator: training-data collection, GPU training, threshold choice.

The network, its features and its kernels are described in include/rkdg.h and
dgswemv2_amd/csrc/indicator.hpp.  The workflow here:

  1. run the simulation with the plain limiter and capture on (Simulation.capture): every
     limiter stage records the features of its state before limiting and the label (the plain
     limiter writes back a change above 1e-8 of the element's physical scale, 1e-6 in FP32);
  2. train the MLP on the GPU (MlpTrainer: BF16-MFMA forward / back-propagation, Adam);
  3. pick the gate threshold on held-out samples so that a target share of the positive
     samples (recall) opens the gate, then lower it by a margin;
  4. Simulation.set_indicator(weights, threshold): the limiter runs only where the gate opens;
  5. optionally one aggregation round (refine): replay the training window under the learned gate
     with capture on (labels then come from the judge's dry run of the plain limit pass on the
     states the gate itself produces), add those samples and retrain from the current weights;
  6. optionally keep retraining during the run (run_online) as the flow leaves the training window.
Only elements where the limiter may act (element and neighbours wet) are training samples;
elsewhere the gate is closed by construction.
"""
import numpy as np

from . import MlpTrainer
from ._native import MLP_HIDDEN, MLP_IN

NPARAM = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + 2 * (MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN) + MLP_HIDDEN + 1


def init_weights(seed=5):
    """He-initialised weights, zero hidden biases, output bias -2 (positives are the minority)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2.0 / MLP_IN), MLP_HIDDEN * MLP_IN), np.zeros(MLP_HIDDEN),
             rng.normal(0, np.sqrt(2.0 / MLP_HIDDEN), MLP_HIDDEN * MLP_HIDDEN), np.zeros(MLP_HIDDEN),
             rng.normal(0, np.sqrt(2.0 / MLP_HIDDEN), MLP_HIDDEN * MLP_HIDDEN), np.zeros(MLP_HIDDEN),
             rng.normal(0, np.sqrt(1.0 / MLP_HIDDEN), MLP_HIDDEN), np.array([-2.0])]
    return np.concatenate(parts).astype(np.float32)


def collect(sim, steps, every=1, neg_keep=1.0, seed=0, gate_misses=None):
    """Steps `sim` (plain limiter) `steps` full steps with capture on and keeps the samples of
    every `every`-th limiter stage where the limiter may act; negatives are kept with
    probability neg_keep.  With an indicator set, `gate_misses` (a threshold) also counts the
    positives whose logit does not exceed it.  Returns (X [36][n] float32, y [n] float32, stages
    sampled), plus (misses, positives) when gate_misses is given."""
    rng = np.random.default_rng(seed)
    ns = sim.params.nstages
    sim.capture(True)
    Xs, ys, k = [], [], 0
    missed = positives = 0
    for _ in range(steps * ns):
        sim.stage()
        k += 1
        if k % every:
            continue
        c = sim.get_capture(logits=gate_misses is not None)
        m = c["may_act"] != 0
        y = c["label"][m].astype(np.float32)
        X = c["features"][:, m]
        if gate_misses is not None:
            pos = y > 0.5
            positives += int(pos.sum())
            missed += int((c["logit"][m][pos] <= gate_misses).sum())
        if neg_keep < 1.0:
            keep = (y > 0) | (rng.uniform(size=y.size) < neg_keep)
            X, y = X[:, keep], y[keep]
        Xs.append(X)
        ys.append(y)
    sim.capture(False)
    out = (np.concatenate(Xs, axis=1), np.concatenate(ys), k // every)
    return out + ((missed, positives),) if gate_misses is not None else out


def choose_threshold(logits, y, recall=0.999):
    """Largest threshold whose gate (logit > threshold) opens for at least `recall` of the positives;
    None when y holds no positive (no threshold is defined by them)."""
    pos = np.sort(np.asarray(logits)[np.asarray(y) > 0.5])
    if pos.size == 0:
        return None
    k = int(np.floor((1.0 - recall) * pos.size))
    return float(np.nextafter(np.float32(pos[k]), np.float32(-np.inf)))


def pick_threshold(z_val, y_val, z_train_pos, recall, margin):
    """Gate threshold from held-out logits z_val / labels y_val, lowered by `margin`.  With no
    held-out positive it falls back to the training positives' logits z_train_pos, and with no
    positive at all to -inf (gate open wherever the limiter may act: the plain limiter's decisions),
    warning in both cases.  Returns (threshold, source) with source "holdout", "training" or "none"."""
    import warnings
    thr = choose_threshold(z_val, y_val, recall)
    if thr is not None:
        return thr - margin, "holdout"
    z_train_pos = np.asarray(z_train_pos)
    if z_train_pos.size:
        warnings.warn("no positive sample in the held-out split: threshold from the training positives")
        return choose_threshold(z_train_pos, np.ones(z_train_pos.size), recall) - margin, "training"
    warnings.warn("no positive sample at all: the gate stays open wherever the limiter may act")
    return float("-inf"), "none"


def train(X, y, epochs=6, batch=1024, lr=2e-3, pos_weight=8.0, recall=0.9999, margin=16.0, holdout=0.1, seed=5,
          init=None):
    """Trains on the GPU and chooses the threshold on a held-out share of the samples: the logit at
    which `recall` of the held-out positives open the gate, lowered by `margin` (logits are
    log-odds).  On the 500k dam break a margin of 16 opens the gate for 3.05 % of the elements
    where the limiter may act, against 2.91 % with none, and halves the learned run's momentum
    error against FP64 plain (rel. L2 of qx 6.2e-7 after 30 steps, 1.4e-6 with no margin) at the
    same speed.  The write-backs it still misses (0.8 % of the significant ones) are elements the
    network scores far below any threshold.  Returns (flat weights, threshold, info)."""
    n = y.size
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    nv = max(1, int(holdout * n))
    # both subsets in sample order: the trainer reshuffles every epoch, and an ordered gather of
    # the 36 feature rows is about twice as fast as a random one
    va, tr = np.sort(perm[:nv]), np.sort(perm[nv:])
    Xt, yt = np.ascontiguousarray(X[:, tr]), np.ascontiguousarray(y[tr])
    t = MlpTrainer(init_weights(seed) if init is None else init, capacity=max(tr.size, nv))
    try:
        t.set_data(Xt, yt)
        losses = t.train(epochs, batch=batch, lr=lr, pos_weight=pos_weight, seed=seed)
        w = t.weights()
        z = t.forward(np.ascontiguousarray(X[:, va]))
        yv = y[va]
        tp = tr[y[tr] > 0.5] if not (yv > 0.5).any() else tr[:0]
        z_tp = t.forward(np.ascontiguousarray(X[:, tp])) if tp.size else np.zeros(0, np.float32)
    finally:
        t.close()
    thr, source = pick_threshold(z, yv, z_tp, recall, margin)
    gate = z > thr
    info = {"samples": int(n), "positives": float(y.mean()), "epoch_loss": [float(v) for v in losses],
            "val_recall": float(gate[yv > 0.5].mean()) if (yv > 0.5).any() else None,
            "val_open": float(gate.mean()), "threshold": thr, "threshold_from": source}
    return w, thr, info


def refine(make_sim, steps, w, thr, X, y, epochs=3, seed=2, **kw):
    """One aggregation round (DAgger): `make_sim()` gives a fresh simulation at the start of the
    training window; it runs `steps` steps under the gate (w, thr) with capture on, so its labels
    are the plain limiter's decisions on the states the gate produces; those samples join (X, y)
    and the network is retrained from w.  On the 500k dam break one round cuts the significant
    write-backs the gate misses after training from 0.8 % to 0.08 % and the learned run's momentum
    error against FP64 plain from 6.2e-7 to 2.8e-7; further rounds were not stable there.
    Returns (weights, threshold, info, X, y)."""
    sim = make_sim()
    try:
        sim.set_indicator(w, threshold=thr)
        Xd, yd, _ = collect(sim, steps, every=1, neg_keep=kw.pop("neg_keep", 0.25), seed=seed)
    finally:
        sim.close()
    X = np.concatenate([X, Xd], axis=1)
    y = np.concatenate([y, yd])
    w, thr, info = train(X, y, epochs=epochs, init=w, **kw)
    info["refine_samples"] = int(yd.size)
    return w, thr, info, X, y



def run_online(sim, w, thr, steps, X=None, y=None, every=20, window=1, epochs=2, keep=4_000_000, neg_keep=0.25,
               seed=3, retrain_above=None, **kw):
    """Steps `sim` `steps` steps under the gate (w, thr) and retrains on judged data as it goes:
    after every `every` gated steps, `window` more steps run with capture on (their labels are the
    judge's dry run of the plain limit pass on the states the gate produces), the samples join a
    buffer that starts from (X, y) and keeps the newest `keep`, and the network is retrained from its
    current weights for `epochs` epochs (train's keyword arguments in **kw).  With `retrain_above`
    set, a round retrains only when the gate missed more than that share of the captured window's
    positives (the samples still join the buffer).  Returns (weights, threshold, X, y, rounds) with
    one info dict per retraining round."""
    rounds = []
    sim.set_indicator(w, threshold=thr)
    done = 0
    while done < steps:
        n = min(every, steps - done)
        sim.step(n)
        done += n
        if done >= steps:
            break
        n = min(window, steps - done)
        if retrain_above is None:
            Xd, yd, _ = collect(sim, n, every=1, neg_keep=neg_keep, seed=seed + done)
        else:
            Xd, yd, _, (missed, positives) = collect(sim, n, every=1, neg_keep=neg_keep, seed=seed + done,
                                                     gate_misses=thr)
        done += n
        X = Xd if X is None else np.concatenate([X, Xd], axis=1)[:, -keep:]
        y = yd if y is None else np.concatenate([y, yd])[-keep:]
        if retrain_above is not None and missed <= retrain_above * max(1, positives):
            continue
        w, thr, info = train(X, y, epochs=epochs, init=w, seed=seed, **kw)
        info["step"] = done
        info["new_samples"] = int(yd.size)
        rounds.append(info)
        sim.set_indicator(w, threshold=thr)
    return w, thr, X, y, rounds
