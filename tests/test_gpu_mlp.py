"""The learned guess's network kernels (calchem_amd/csrc/mlp.hip) against the
float64 oracle (tests/oracle_mlp.py, pinned by finite differences): forward
pass, loss and every parameter gradient at the production sizes (64 -> 128
-> 128 -> 225) on a batch that splits into uneven reduction chunks; training
lowers the loss and is bit-reproducible for a seed."""
import numpy as np
import pytest

import oracle_mlp
from calchem_amd import learned_guess as LG

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, LG.D_IN)).astype(np.float32)
    t = rng.normal(size=(n, LG.D_OUT)).astype(np.float32)
    m = (rng.uniform(size=(n, LG.D_OUT)) > 0.4).astype(np.float32)
    return x, t, m


def test_forward_and_gradients_match_the_oracle():
    p = LG.init_params(9)
    net = LG.DeviceMlp(p)
    x, t, m = _batch(2500, 1)  # 2500 rows: reduction chunks of 1024, 1024, 452
    y = net.forward(x)
    y_ref, _ = oracle_mlp.forward(p, x, LG.D_IN, LG.HIDDEN, LG.D_OUT)
    assert np.abs(y - y_ref).max() < 1e-5 * max(1.0, np.abs(y_ref).max())
    loss, g = net.loss_grad(x, t, m)
    loss_ref, g_ref = oracle_mlp.loss_grad(p, x, t, m, LG.D_IN, LG.HIDDEN, LG.D_OUT)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"loss {loss:.7f} vs {loss_ref:.7f}; gradient relative error {rel:.2e}, "
          f"max abs {np.abs(g - g_ref).max():.2e} of {np.abs(g_ref).max():.2e}")
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert rel < 1e-5
    net.close()


def test_training_lowers_the_loss_and_is_reproducible():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(6000, LG.D_IN)).astype(np.float32)
    A = rng.normal(size=(LG.D_IN, LG.D_OUT)).astype(np.float32) / 8.0
    t = np.tanh(x @ A).astype(np.float32)
    m = np.ones_like(t)
    out = []
    for _ in range(2):
        net = LG.DeviceMlp(LG.init_params(3))
        hist = net.train(x, t, m, epochs=16, batch=1000, lr=3e-3, seed=11)
        out.append(net.params())
        net.close()
        print("loss per epoch", np.round(hist, 5))
        assert hist[-1] < 0.5 * hist[0] and np.all(np.diff(hist) < 0)
    assert np.array_equal(out[0], out[1])
