"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the warm-start network (MLP 6 -> 128 -> 128 -> 128
-> 5, exact-erf GELU, mean-squared-error loss over the normalised outputs),
used by tests/ to check the library's BF16-MFMA forward and backward kernels.
The reference has no counterpart (its only initial condition is the
freestream, src/SolverData.cpp:27-34); this module states the mathematics.

Parameters are flattened W1 [128][6], b1, W2 [128][128], b2, W3, b3,
W4 [5][128], b4, each W row-major [out][in].  Samples are rows.

emulate=True rounds to BF16 exactly where the device does -- weights, the
normalised inputs, each layer's GELU output, and the output and hidden-layer
error signals that enter products -- and keeps everything else in FP64, so
the remaining device/oracle difference is FP32 accumulation alone.

Pinned by: central finite differences of loss_grad (FP64, emulate=False),
GELU closed forms and the identity GELU(x) - GELU(-x) = x, and the BF16
rounding of known bit patterns (tests/test_mlp_oracle.py).
"""
import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT = 6, 128, 5
SHAPES = [(N_HID, N_IN), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,), (N_OUT, N_HID), (N_OUT,)]
N_PARAMS = int(sum(np.prod(s) for s in SHAPES))


def unpack(p):
    out, off = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(np.asarray(p[off:off + n], np.float64).reshape(s))
        off += n
    return out


def pack(blocks):
    return np.concatenate([np.asarray(b, np.float64).ravel() for b in blocks])


def bf16(a):
    """Round to the nearest BF16 value (ties to even) through FP32."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def normalise_inputs(x, in_shift, in_scale):
    """(x - shift) * scale in FP32, as the device computes it."""
    x = np.asarray(x, np.float32)
    return ((x - np.asarray(in_shift, np.float32)) * np.asarray(in_scale, np.float32)).astype(np.float64)


def forward(p, xn, emulate=False):
    """Normalised inputs xn [n][6] -> normalised outputs [n][5] and the cache."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    a = r(xn)
    cache = {"a0": a, "z": [], "a": []}
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        z = a @ r(W).T + b
        a = r(gelu(z))
        cache["z"].append(z)
        cache["a"].append(a)
    y = a @ r(W4).T + b4
    return y, cache


def loss_grad(p, xn, tn, emulate=False):
    """Loss = mean over samples and outputs of (y - t)^2 and its gradient."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, c = forward(p, xn, emulate)
    n = len(xn)
    d = y - tn
    loss = float(np.sum(d * d) / (N_OUT * n))
    dy = 2.0 * d / (N_OUT * n)
    dyb = r(dy)
    a0, (a1, a2, a3), (z1, z2, z3) = c["a0"], c["a"], c["z"]
    g = [None] * 8
    g[6] = dyb.T @ a3
    g[7] = dy.sum(0)
    da = dyb @ r(W4)
    for l, (W, z, a_in) in zip((2, 1, 0), ((W3, z3, a2), (W2, z2, a1), (W1, z1, a0))):
        dz = da * gelu_d(z)
        dzb = r(dz)
        g[2 * l] = dzb.T @ a_in
        g[2 * l + 1] = dz.sum(0)
        if l > 0:
            da = dzb @ r(W)
    return loss, pack(g)


def init_params(seed=0):
    """uniform(+-1/sqrt(fan_in)) weights and biases (any fixed draw serves a test)."""
    rng = np.random.default_rng(seed)
    blocks = []
    for s, fan in zip(SHAPES, (N_IN, N_IN, N_HID, N_HID, N_HID, N_HID, N_HID, N_HID)):
        blocks.append(rng.uniform(-1, 1, s) / np.sqrt(fan))
    return pack(blocks)
