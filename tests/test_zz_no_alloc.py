"""rkdg_mlp_forward_noalloc against rkdg_mlp_forward (the trainer's forward on the same weights):
two calls in a row through the C-ABI, the second with fewer samples, then one at the size limit;
the logits must be equal bit for bit (same kernel, same inputs)."""
import ctypes as C

import numpy as np
import pytest

from dgswemv2_amd import MlpTrainer
from dgswemv2_amd import _native as N
from dgswemv2_amd import indicator as IND

pytestmark = pytest.mark.gpu

NMAX = 16384  # RKDG_MLP_FORWARD_NOALLOC_MAX


def noalloc_forward(lib, flat, X):
    X = np.ascontiguousarray(X, np.float32)
    w = N.weights_from_flat(flat)
    out = np.full(X.shape[1], np.nan, np.float32)
    st = lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), X.shape[1], N.fptr(out))
    assert st == N.RKDG_OK, st
    return out


def test_noalloc_forward_equals_trainer_forward():
    lib = N.load_library()
    rng = np.random.default_rng(17)
    flat = IND.init_weights(23).astype(np.float32)
    tr = MlpTrainer(flat, capacity=NMAX)
    try:
        for n in (5000, 777, NMAX):  # a call, a smaller one after it, then the limit
            X = rng.standard_normal((N.MLP_IN, n)).astype(np.float32)
            print(f"noalloc forward n={n}", flush=True)
            ref = tr.forward(X)
            got = noalloc_forward(lib, flat, X)
            assert got.shape == ref.shape
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
                (n, float(np.max(np.abs(got - ref))))
            assert np.all(np.isfinite(got))
        # the fixed array keeps the last call's contents: a second set of weights must not see them
        flat2 = IND.init_weights(29).astype(np.float32)
        X = rng.standard_normal((N.MLP_IN, 64)).astype(np.float32)
        tr2 = MlpTrainer(flat2, capacity=64)
        try:
            assert np.array_equal(noalloc_forward(lib, flat2, X).view(np.uint32), tr2.forward(X).view(np.uint32))
        finally:
            tr2.close()
    finally:
        tr.close()


def test_noalloc_forward_rejects_sizes_beyond_the_limit():
    lib = N.load_library()
    w = N.weights_from_flat(IND.init_weights(1).astype(np.float32))
    X = np.zeros((N.MLP_IN, NMAX + 1), np.float32)
    out = np.zeros(NMAX + 1, np.float32)
    assert lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), NMAX + 1, N.fptr(out)) == N.RKDG_ERR_ARG
    assert lib.rkdg_mlp_forward_noalloc(C.byref(w), N.fptr(X), 0, N.fptr(out)) == N.RKDG_ERR_ARG
