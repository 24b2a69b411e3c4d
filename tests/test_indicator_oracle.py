"""Pins of the learned indicator's CPU oracle (oracle/mlp.py), independent of its code where the
mathematics allows: bfloat16 rounding against hand-worked bit patterns, the back-propagated
gradient against central differences of the loss, the forward pass against a hand-built
network with a closed-form output, and the features against their defining invariances.
Labels come from the oracle's limiter, whose flags the plain FP64 engine matches bit for bit.
"""
import numpy as np
import pytest

from dgswemv2_amd import workloads as W
from oracle import mlp as O
from oracle.oracle_py import OracleSim


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7  # bfloat16 has 8 significant bits: spacing 2^-7 in [1, 2)
    cases = [(one, one), (one + ulp, one + ulp), (one + ulp / 2, one),            # tie -> even (1.0)
             (one + 3 * ulp / 2, one + 2 * ulp),                                # tie -> even (1 + 2 ulp)
             (one + ulp / 2 + 2.0 ** -20, one + ulp), (-(one + ulp * 0.49), -one),
             (3.0e-39, None), (0.0, 0.0), (65504.0, 65536.0)]
    for x, want in cases:
        got = O.bf16(x)
        if want is None:  # float32 subnormal: keeps the top 8 bits of its pattern
            assert got == np.frombuffer((np.array([x], np.float32).view(np.uint32) + 0x7FFF >> 16 << 16)
                                        .astype(np.uint32).tobytes(), np.float32)[0]
        else:
            assert got == want, (x, got, want)


def _random_net(seed, scale=1.0):
    p = O.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.normal(0, 0.3, O.NH)
    p["b4"] = 0.1
    for k in ("w1", "w2", "w3", "w4"):
        p[k] = p[k] * scale
    return p


def test_exact_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _random_net(1)
    X = rng.uniform(-1, 1, (O.NIN, 40))
    y = (rng.uniform(size=40) < 0.3).astype(float)
    loss, g = O.loss_grad(p, X, y, pos_weight=3.0, bf=False)
    flat, gflat = O.flatten(p), O.flatten(g)
    h = 1e-6
    picks = rng.choice(O.NPARAM, 300, replace=False)
    picks = np.concatenate([picks, [O.NPARAM - 1, O.NPARAM - 2]])  # b4, a w4 entry
    num = np.empty(len(picks))
    for i, k in enumerate(picks):
        a, b = flat.copy(), flat.copy()
        a[k] += h
        b[k] -= h
        la, _ = O.loss_grad(O.unflatten(a), X, y, 3.0, bf=False)
        lb, _ = O.loss_grad(O.unflatten(b), X, y, 3.0, bf=False)
        num[i] = (la - lb) / (2 * h)
    err = np.abs(num - gflat[picks])
    assert np.max(err) <= 1e-7 + 1e-5 * np.max(np.abs(gflat[picks])), np.max(err)


def test_forward_hand_built_network():
    """Layer 1 copies features 0..35 into neurons 0..35 (and their negatives into 36..63 for the
    first 28); layers 2 and 3 are identities; w4 = 1 on neurons 0..35 and -1 on 36..63.  Then
    z = sum_i relu(x_i) - sum_{i<28} relu(-x_i) + b4 exactly (all inputs are bf16 numbers)."""
    p = {k: np.zeros(s) if s else 0.0 for k, s in O.SHAPES}
    for i in range(36):
        p["w1"][i, i] = 1.0
    for i in range(28):
        p["w1"][36 + i, i] = -1.0
    p["w2"] = np.eye(64)
    p["w3"] = np.eye(64)
    p["w4"] = np.concatenate([np.ones(36), -np.ones(28)])
    p["b4"] = 0.25
    rng = np.random.default_rng(3)
    X = O.bf16(rng.uniform(-1, 1, (36, 50)))
    want = np.maximum(X, 0).sum(axis=0) - np.maximum(-X[:28], 0).sum(axis=0) + 0.25
    for bf in (True, False):
        z, _ = O.forward(p, X, bf=bf)
        assert np.allclose(z, want, rtol=0, atol=1e-12)


def test_bf16_gradient_close_to_exact():
    # bf16 rounding of activations and back-propagated values (relative 2^-9 each) changes the
    # gradient by a few per cent where contributions cancel; the bf16 gradient is still the gradient
    rng = np.random.default_rng(2)
    p = _random_net(2)
    X = rng.uniform(-1, 1, (O.NIN, 256))
    y = (rng.uniform(size=256) < 0.4).astype(float)
    _, ge = O.loss_grad(p, X, y, 2.0, bf=False)
    _, gb = O.loss_grad(p, X, y, 2.0, bf=True)
    for k in ("w1", "w2", "w3", "w4", "b1", "b2", "b3"):
        a, b = np.ravel(ge[k]), np.ravel(gb[k])
        assert np.linalg.norm(a - b) <= 0.1 * np.linalg.norm(a), k


def _dambreak_stage(scale=0.06, stages=7):
    mesh, prm = W.dambreak(scale=scale)
    o = OracleSim(mesh, prm)
    o.keep_prelimit()
    for _ in range(stages):
        o.stage()
    nb, bb = o.limiter_geometry()
    return mesh, prm, o, nb, bb


def test_feature_invariances():
    mesh, prm, o, nb, bb = _dambreak_stage()
    q = o.prelimit()
    wet = o.state()["wet"]
    X, ok = O.features(q, wet, nb, bb, prm.g)
    assert X.shape == (36, mesh.n_elements) and np.all(np.abs(X) <= 1.0)
    # the shape entries of each variable: max |.| is 1 where the variable varies
    for v in range(3):
        shape_idx = [9 * k + 3 * v + i for k in range(4) for i in range(3) if (k, i) != (0, 0)]
        m = np.abs(X[shape_idx]).max(axis=0)
        assert np.all((np.abs(m - 1.0) < 1e-12) | (m == 0.0))
    # negating the momenta negates their shape entries and keeps their size entries
    q2 = q.copy()
    q2[1:] *= -1
    X2, _ = O.features(q2, wet, nb, bb, prm.g)
    for v in (1, 2):
        for k in range(4):
            for i in range(3):
                f = 9 * k + 3 * v + i
                if (k, i) == (0, 0):
                    assert np.array_equal(X2[f], X[f])
                else:
                    assert np.allclose(X2[f], -X[f], atol=0)
    # adding a constant to ze everywhere leaves the shape entries of ze unchanged
    q3 = q.copy()
    q3[0, 0] += 0.37
    X3, _ = O.features(q3, wet, nb, bb, prm.g)
    sidx = [9 * k + i for k in range(4) for i in range(3) if (k, i) != (0, 0)]
    assert np.allclose(X3[sidx], X[sidx], atol=1e-9)
    # a boundary face sees the element itself: its mode-0 entry is zero there
    bnd = np.where(nb[0] < 0)[0]
    assert len(bnd) and np.all(X[9 * 1 + 0][bnd] == 0.0)
    # the limiter may act only where the element and its neighbours are wet
    w = wet != 0
    expect = w & np.all([np.where(nb[k] >= 0, w[np.maximum(nb[k], 0)], w) for k in range(3)], axis=0)
    assert np.array_equal(ok, expect)


def test_feature_hand_example():
    """One interior element with hand-chosen modes: ze varies by 0.2 at most (at neighbour 2's
    mode 0), qx not at all, qy only in the element's own slope mode."""
    ne = 4
    q = np.zeros((3, 3, ne))
    q[0, 0] = [1.0, 1.1, 0.9, 1.2]      # ze mode 0 of e, n0, n1, n2
    q[0, 1, 0] = 0.05
    q[2, 2, 0] = -0.3
    nb = np.array([[1, -1, -1, -1], [2, -1, -1, -1], [3, -1, -1, -1]])
    bath_bc = np.array([3.0, 3.0, 3.0, 3.0])
    X, ok = O.features(q, np.ones(ne, np.uint8), nb, bath_bc, 9.81)
    x = X[:, 0]
    h = 4.0
    assert x[0] == pytest.approx(1 + np.log10(0.2 / h) / 8, abs=1e-12)   # ze size (m = |1.2 - 1.0|)
    assert x[1] == pytest.approx(0.05 / 0.2) and x[9] == pytest.approx(0.1 / 0.2)
    assert x[18] == pytest.approx(-0.1 / 0.2) and x[27] == pytest.approx(1.0)
    assert x[3] == -1.0 and np.all(x[[4, 5, 12, 13, 14, 21, 22, 23, 30, 31, 32]] == 0.0)   # qx
    sq = h * np.sqrt(9.81 * h)
    assert x[6] == pytest.approx(1 + np.log10(0.3 / sq) / 8) and x[8] == pytest.approx(-1.0)
    assert ok[0]


def test_rounding_level_variation_has_no_shape():
    """A variable that varies by 1e-300 (or by 1e-17 of its scale) carries no shape: its shape
    entries are 0 and its size -1 -- features stay finite where 1/m would overflow."""
    ne = 4
    q = np.zeros((3, 3, ne))
    q[0, 0] = 2.0
    q[1, 1, 0] = 1e-300
    q[2, 0, 1] = 1e-17 * 1.0 * np.sqrt(9.81)   # qy of neighbour 0: 1e-17 of s_q (h_s = 1)
    nb = np.array([[1, -1, -1, -1], [2, -1, -1, -1], [3, -1, -1, -1]])
    X, _ = O.features(q, np.ones(ne, np.uint8), nb, np.full(ne, -1.0), 9.81)
    assert np.all(np.isfinite(X))
    assert np.all(X[[4, 5, 13, 14, 22, 23, 31, 32, 7, 8, 15, 16, 17]][:, 0] == 0.0)
    assert X[3, 0] == -1.0 and X[6, 0] == -1.0


def test_labels_are_significant_write_backs():
    mesh, prm, o, nb, bb = _dambreak_stage(stages=9)
    q = o.prelimit()
    st = o.state()
    y = O.significant_labels(q, st["q"], st["limited"], bb, prm.g)
    raw = O.labels(st["limited"])
    assert y.sum() > 0 and raw.sum() > y.sum()
    assert np.all(raw[y > 0] > 0)          # a significant write-back is a write-back
    # where nothing was written back the state is unchanged
    untouched = raw == 0
    assert np.array_equal(st["q"][:, :, untouched], q[:, :, untouched])


def test_threshold_choice_and_fallbacks():
    """The gate threshold: `recall` of the held-out positives open the gate; with no held-out
    positive the training positives decide; with none at all the gate opens everywhere (-inf), and
    never silently closes (a +inf threshold would switch the limiter off)."""
    import warnings
    from dgswemv2_amd.indicator import choose_threshold, pick_threshold
    z = np.array([-3.0, 0.5, 2.0, 4.0, -1.0], np.float32)
    y = np.array([0, 1, 1, 1, 0], np.float32)
    t = choose_threshold(z, y, recall=1.0)
    assert t < 0.5 and np.all((z > t)[y > 0.5])
    assert choose_threshold(z, np.zeros(5), recall=0.99) is None
    thr, src = pick_threshold(z, y, np.zeros(0), 1.0, 2.0)
    assert src == "holdout" and thr == pytest.approx(t - 2.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        thr, src = pick_threshold(z, np.zeros(5), np.array([1.5, 3.0], np.float32), 1.0, 2.0)
        assert src == "training" and thr < 1.5 - 2.0 + 1e-6 and len(w) == 1
        thr, src = pick_threshold(z, np.zeros(5), np.zeros(0), 1.0, 2.0)
        assert src == "none" and thr == float("-inf") and np.all(z > thr)
