"""The library's BF16-MFMA network kernels against the FP64 CPU restatement
(oracle/mlp_ref.py) with BF16 rounding emulated at the device's rounding
points: forward outputs, loss and every parameter block's gradient.  The
remaining difference is FP32 accumulation (and rare one-ulp BF16 flips of an
activation), far below what a wrong operand layout, transpose or missing term
produces."""
import numpy as np
import pytest

from netflux_amd.mlp import Mlp
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _net(seed=3):
    net = Mlp(seed)
    p = M.init_params(seed).astype(np.float32)
    net.params = p
    shift = np.array([0.5, 0.05, 1.5, 0.5, 0.25, 0.1], np.float32)
    scale = np.array([4.0, 20.0, 0.6, 2.0, 4.0, 8.0], np.float32)
    oshift = np.array([1.0, 0.5, 0.0, 0.0, 0.7], np.float32)
    oscale = np.array([0.1, 0.2, 0.05, 0.01, 0.08], np.float32)
    net.set_norm(shift, scale, oshift, oscale)
    return net, p.astype(np.float64), (shift, scale, oshift, oscale)


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = np.column_stack([rng.uniform(0.3, 0.8, n), rng.uniform(0.02, 0.1, n), rng.uniform(0, 3, n),
                         rng.uniform(0, 1, n), rng.uniform(0, 0.5, n), rng.uniform(0, 0.25, n)]).astype(np.float32)
    y = np.column_stack([1 + 0.1 * np.sin(3 * x[:, 2]), x[:, 0] * (1 + 0.2 * x[:, 3]), 0.05 * np.cos(x[:, 2]),
                         0.01 * x[:, 4], 0.7 + 0.05 * x[:, 5]]).astype(np.float32)
    return x, y


def test_predict_matches_oracle():
    print("[case] mlp predict n=1000", flush=True)
    net, p, (sh, sc, osh, osc) = _net()
    x, _ = _data(1000)
    y = net.predict(x)
    yn_ref, _ = M.forward(p.astype(np.float32), M.normalise_inputs(x, sh, sc), emulate=True)
    y_ref = yn_ref * osc + osh
    err = np.abs(y - y_ref) / osc
    assert err.max() < 2e-3, err.max()


def test_predict_independent_of_batching():
    print("[case] mlp predict chunking n=300001", flush=True)
    net, _, _ = _net()
    x, _ = _data(300001, seed=5)
    y = net.predict(x)
    pick = np.r_[0:7, 262140:262150, 300000]
    np.testing.assert_array_equal(net.predict(x[pick]), y[pick])


@pytest.mark.parametrize("n", [256, 3000])
def test_loss_and_gradient_match_oracle(n):
    print(f"[case] mlp loss/grad n={n}", flush=True)
    net, p, (sh, sc, osh, osc) = _net()
    x, y = _data(n, seed=1)
    loss, g = net.loss_grad(x, y)
    xn = M.normalise_inputs(x, sh, sc)
    tn = (y.astype(np.float32) - osh) * (np.float32(1) / osc)
    l_ref, g_ref = M.loss_grad(p.astype(np.float32), xn, tn.astype(np.float64), emulate=True)
    assert abs(loss - l_ref) <= 1e-4 * l_ref
    off = 0
    for shape in M.SHAPES:
        k = int(np.prod(shape))
        a, b = g[off:off + k], g_ref[off:off + k]
        assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b), (shape, np.linalg.norm(a - b) / np.linalg.norm(b))
        off += k


def test_training_reduces_loss():
    print("[case] mlp train 20k samples x 15 epochs", flush=True)
    net = Mlp(7)
    x, y = _data(20480, seed=2)
    st = net.train(x, y, epochs=15, batch=1024, lr=2e-3, seed=1)
    assert st["steps"] == 15 * 20
    assert st["loss_last"] < 0.01 * st["loss_first"], st
    xv, yv = _data(4096, seed=9)
    pred = net.predict(xv)
    _, _, osh, osc = net.get_norm()
    rel = np.sqrt(np.mean(((pred - yv) / osc) ** 2))
    assert rel < 0.1, rel
