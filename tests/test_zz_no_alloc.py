"""nf_mlp_predict_noalloc -- inference with the activations in a fixed
device array, no allocation -- against nf_mlp_predict on the same network and
inputs: bit for bit, on two calls in a row (the second smaller, so stale
contents of the shared array from the first call would show), on a stream
of the caller's, and the size limit rejected with an error.  The test's own
device buffers come from the HIP runtime the library is linked to."""
import ctypes as C

import numpy as np
import pytest

from netflux_amd import _abi as A
from netflux_amd.mlp import Mlp

pytestmark = pytest.mark.gpu

MAX_SAMPLES = 16384
H2D, D2H = 1, 2


def _hip():
    A.lib()  # loads libnetflux.so and with it libamdhip64.so.7
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _net(seed=5):
    net = Mlp(seed)
    net.set_norm(np.array([0.5, 0.05, 1.5, 0.5, 0.25, 0.1], np.float32),
                 np.array([4.0, 20.0, 0.6, 2.0, 4.0, 8.0], np.float32),
                 np.array([1.0, 0.5, 0.0, 0.0, 0.7], np.float32),
                 np.array([0.1, 0.2, 0.05, 0.01, 0.08], np.float32))
    return net


def _features(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0.3, 0.8, n), rng.uniform(0.02, 0.1, n), rng.uniform(0, 3, n),
                            rng.uniform(0, 1, n), rng.uniform(0, 0.5, n),
                            rng.uniform(0, 0.25, n)]).astype(np.float32)


def _noalloc(net, hip, x, stream=None):
    """Runs the entry point on device copies of x; output buffer prefilled
    with NaN so unwritten entries show."""
    xd, yd = DevBuf(hip, max(x.nbytes, 4)), DevBuf(hip, max(len(x) * 5 * 4, 4))
    try:
        xd.put(x)
        yd.put(np.full((len(x), 5), np.nan, np.float32))
        rc = A.lib().nf_mlp_predict_noalloc(net._h, xd.p, len(x), yd.p, stream)
        msg = A.lib().nf_mlp_last_error(net._h).decode()
        if rc == A.NF_OK:
            # a null stream means the network's own (non-blocking) stream
            assert (hip.hipStreamSynchronize(stream) if stream else hip.hipDeviceSynchronize()) == 0
        return rc, msg, yd.get((len(x), 5))
    finally:
        xd.free()
        yd.free()


def test_noalloc_matches_predict_twice():
    print("[case] noalloc predict: n=16384 then n=1000", flush=True)
    hip = _hip()
    net = _net()
    for n, seed in ((MAX_SAMPLES, 1), (1000, 2)):
        x = _features(n, seed)
        ref = net.predict(x)
        rc, msg, y = _noalloc(net, hip, x)
        assert rc == A.NF_OK, msg
        assert np.isfinite(y).all()
        np.testing.assert_array_equal(y, ref)


def test_noalloc_on_a_caller_stream_and_limit():
    print("[case] noalloc predict on a caller stream; n over the limit", flush=True)
    hip = _hip()
    net = _net(seed=6)
    x = _features(777, 3)
    ref = net.predict(x)
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0  # default priority
    try:
        rc, msg, y = _noalloc(net, hip, x, s)
        assert rc == A.NF_OK, msg
        np.testing.assert_array_equal(y, ref)
    finally:
        hip.hipStreamDestroy(s)
    rc, msg, _ = _noalloc(net, hip, _features(MAX_SAMPLES + 1, 4))
    assert rc == A.NF_E_ARG
    assert "exceed the maximum of 16384" in msg
