"""Pins the CPU restatement of the warm-start network (oracle/mlp_ref.py):
its gradient against central finite differences, GELU against closed forms,
and the BF16 rounding against known bit patterns."""
import numpy as np

from oracle import mlp_ref as M


def test_gelu_closed_forms():
    assert M.gelu(0.0) == 0.0
    assert M.gelu_d(0.0) == 0.5
    x = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(M.gelu(x) - M.gelu(-x), x, rtol=0, atol=1e-14)
    assert abs(M.gelu(10.0) - 10.0) < 1e-12 and abs(M.gelu(-10.0)) < 1e-12
    # GELU(1) = Phi(1) = 0.841344746068543 (standard normal CDF at 1)
    assert abs(M.gelu(1.0) - 0.8413447460685429) < 1e-15
    h = 1e-6
    np.testing.assert_allclose((M.gelu(x + h) - M.gelu(x - h)) / (2 * h), M.gelu_d(x), atol=1e-8)


def test_bf16_rounding():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert M.bf16(-3.0) == -3.0 and M.bf16(0.0) == 0.0
    x = np.random.default_rng(0).normal(size=1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3)
    x = rng.normal(size=(64, 6))
    t = rng.normal(size=(64, 5))
    loss, g = M.loss_grad(p, x, t)
    assert g.shape == (M.N_PARAMS,)
    off = 0
    h = 1e-6
    for shape in M.SHAPES:
        n = int(np.prod(shape))
        for i in rng.choice(n, size=min(n, 6), replace=False):
            q = p.copy()
            q[off + i] += h
            lp, _ = M.loss_grad(q, x, t)
            q[off + i] -= 2 * h
            lm, _ = M.loss_grad(q, x, t)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[off + i]) <= 1e-7 + 1e-5 * abs(fd), (shape, i, fd, g[off + i])
        off += n


def test_emulated_rounding_changes_little():
    rng = np.random.default_rng(2)
    p = M.init_params(4)
    x = rng.normal(size=(256, 6))
    t = rng.normal(size=(256, 5))
    l0, g0 = M.loss_grad(p, x, t)
    l1, g1 = M.loss_grad(p, x, t, emulate=True)
    assert abs(l1 - l0) < 1e-2 * l0
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)
