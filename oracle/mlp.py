"""CPU oracle of the learned troubled-cell indicator -- TEST INFRASTRUCTURE ONLY.

Used by tests/ (and bench.py's reporting) as the checker of the engine's indicator kernels; the
engine never imports it.  The indicator has no counterpart in the reference: what it predicts is
the reference limiter's write-back decision (slope_limiting_kernel, swe_CS_slope_limiter.hpp:181-187,
OR over ze/qx/qy), taken from the oracle's own limiter, whose flags the plain FP64 engine matches
bit for bit.

Features (36 per element e).  Entities k = 0..3: e itself, then its neighbours across local faces
0, 1, 2 (e itself across a boundary face, as the limiter substitutes its own barycentre there).
For variable v (ze, qx, qy), modes i = 0, 1, 2:
    c[v][k][i] = q[v][i][E_k] - (i == 0) q[v][0][e]      (mode 0 relative to the element's own)
    m_v = max over k, i of |c[v][k][i]|
    feature[9 k + 3 v + i] = c[v][k][i] / m_v               the shape
and the three entries that are zero by construction (k = 0, i = 0) carry the size of the
variation relative to the physical scale of the element:
    feature[3 v] = clamp(1 + log10(m_v / s_v) / 8, -1, 1),
    s_ze = h_s,  s_qx = s_qy = h_s sqrt(g h_s),  h_s = max(q[ze][0][e] + b_bc[e], 0.01).
A variable with m_v <= 1e-16 s_v varies only at rounding level: its shape entries are 0 and
its size entry -1.
The limiter's decision depends on the shape (it compares slopes with neighbour differences); the
size separates write-backs that change the state from those at the level of rounding.
The gate may open only where the limiter can act: the element and all its neighbours wet.

Network: z = w4 . relu(Z3) + b4, Z3 = W3 H2 + b3, H2 = relu(W2 H1 + b2), H1 = relu(W1 x + b1),
hidden width 64.  On the GPU the three hidden layers run as BF16 MFMA products with FP32
accumulation: x, W1..W3, H1, H2 and the back-propagated dZ1..dZ3 are rounded to bfloat16
(round to nearest even); biases, Z3, w4, b4 and the logit stay FP32.  `bf=True` reproduces those
rounding points (in float64 otherwise); `bf=False` is the exact network, whose gradient the tests
pin by central differences.

Loss: weighted binary cross-entropy on the logit, mean over the batch,
    L = (1/B) sum_n w_n [softplus(z_n) - y_n z_n],   w_n = pos_weight if y_n else 1.
"""
import numpy as np

NIN, NH = 36, 64
H_FLOOR = 0.01
VAR_FLOOR = 1.0e-16
# parameter layout (the C struct rkdg_mlp_weights, minus its threshold)
SHAPES = (("w1", (NH, NIN)), ("b1", (NH,)), ("w2", (NH, NH)), ("b2", (NH,)), ("w3", (NH, NH)), ("b3", (NH,)),
          ("w4", (NH,)), ("b4", ()))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), through float32 as the GPU converts."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(flat):
    flat = np.asarray(flat, np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp).copy() if shp else float(flat[o])
        o += n
    return out


def flatten(prm):
    return np.concatenate([np.atleast_1d(np.asarray(prm[name], np.float64)).reshape(-1) for name, _ in SHAPES])


def init_params(seed=5):
    """He-initialised weights, zero biases, b4 = -2 (rare positives)."""
    rng = np.random.default_rng(seed)
    p = {"w1": rng.normal(0, np.sqrt(2.0 / NIN), (NH, NIN)), "b1": np.zeros(NH),
         "w2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
         "w3": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b3": np.zeros(NH),
         "w4": rng.normal(0, np.sqrt(1.0 / NH), NH), "b4": -2.0}
    return p


def features(q, wet, neighbor, bath_bc, g):
    """q [3][ndof][ne] modal state, wet [ne], neighbor [3][ne], bath_bc [ne] ->
    (X [36][ne] float64, gate_ok [ne] bool)."""
    q = np.asarray(q, np.float64)
    ne = q.shape[2]
    e = np.arange(ne)
    ent = [e] + [np.where(neighbor[k] >= 0, neighbor[k], e) for k in range(3)]
    hs = np.maximum(q[0, 0] + bath_bc, H_FLOOR)
    sq = hs * np.sqrt(g * hs)
    scale = (hs, sq, sq)
    X = np.empty((NIN, ne))
    for v in range(3):
        c = np.stack([q[v, i, ent[k]] - (q[v, 0, e] if i == 0 else 0.0) for k in range(4) for i in range(3)])
        m = np.abs(c).max(axis=0)
        r = m / scale[v]
        varies = r > VAR_FLOOR
        inv = np.where(varies, 1.0 / np.where(varies, m, 1.0), 0.0)
        for k in range(4):
            for i in range(3):
                X[9 * k + 3 * v + i] = c[3 * k + i] * inv
        X[3 * v] = np.where(varies, np.clip(1.0 + np.log10(np.where(varies, r, 1.0)) / 8.0, -1.0, 1.0), -1.0)
    w = np.asarray(wet) != 0
    ok = w.copy()
    for k in range(1, 4):
        ok &= w[ent[k]]
    return X, ok


def labels(limited):
    """Per-element label: the limiter wrote back any of ze/qx/qy."""
    return (np.asarray(limited).reshape(3, -1) != 0).any(axis=0).astype(np.float64)


def _r(x, bf):
    return bf16(x) if bf else np.asarray(x, np.float64)


def forward(prm, X, bf=True):
    """X [36][B] -> (logits [B], cache)."""
    xb = _r(X, bf)
    W1, W2, W3 = _r(prm["w1"], bf), _r(prm["w2"], bf), _r(prm["w3"], bf)
    Z1 = W1 @ xb + np.asarray(prm["b1"])[:, None]
    H1 = _r(np.maximum(Z1, 0.0), bf)
    Z2 = W2 @ H1 + np.asarray(prm["b2"])[:, None]
    H2 = _r(np.maximum(Z2, 0.0), bf)
    Z3 = W3 @ H2 + np.asarray(prm["b3"])[:, None]
    h3 = np.maximum(Z3, 0.0)
    z = np.asarray(prm["w4"]) @ h3 + prm["b4"]
    return z, dict(xb=xb, W1=W1, W2=W2, W3=W3, H1=H1, H2=H2, h3=h3)


def loss_grad(prm, X, y, pos_weight=1.0, bf=True):
    """Loss and gradient (dict of the parameter shapes) of the weighted BCE mean over the batch."""
    z, c = forward(prm, X, bf)
    y = np.asarray(y, np.float64)
    B = z.shape[0]
    w = np.where(y > 0.5, pos_weight, 1.0)
    loss = float(np.sum(w * (np.logaddexp(0.0, z) - y * z)) / B)
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / B
    g = {"w4": c["h3"] @ dz, "b4": float(np.sum(dz))}
    dZ3 = _r(np.asarray(prm["w4"])[:, None] * dz[None, :] * (c["h3"] > 0), bf)
    g["w3"] = dZ3 @ c["H2"].T
    g["b3"] = dZ3.sum(axis=1)
    dZ2 = _r((c["W3"].T @ dZ3) * (c["H2"] > 0), bf)
    g["w2"] = dZ2 @ c["H1"].T
    g["b2"] = dZ2.sum(axis=1)
    dZ1 = _r((c["W2"].T @ dZ2) * (c["H1"] > 0), bf)
    g["w1"] = dZ1 @ c["xb"].T
    g["b1"] = dZ1.sum(axis=1)
    return loss, g


SIG_TOL = 1.0e-8


def vertex_values(q):
    """Linear part of a modal state [3][ndof][ne] at the three vertices -> [3][3][ne]
    (q_lin = q T_bl, basis_dubiner_2D.cpp:214-234)."""
    q = np.asarray(q, np.float64)
    return np.stack([q[:, 0] - q[:, 1] - q[:, 2], q[:, 0] - q[:, 1] + q[:, 2], q[:, 0] + 2.0 * q[:, 1]], axis=1)


def significant_labels(q_pre, q_post, limited, bath_bc, g, tol=SIG_TOL):
    """Per-element label: the limiter wrote back a variable whose vertex values changed by more
    than tol relative to that variable's feature scale (2-norm over the three vertices)."""
    h = np.maximum(np.asarray(q_pre)[0, 0] + bath_bc, H_FLOOR)
    sq = h * np.sqrt(g * h)
    scale = np.stack([h, sq, sq])
    d = np.sqrt(((vertex_values(q_post) - vertex_values(q_pre)) ** 2).sum(axis=1)) / scale
    lim = np.asarray(limited).reshape(3, -1) != 0
    return (lim & (d > tol)).any(axis=0).astype(np.float64)
