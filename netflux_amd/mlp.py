"""Warm-start network over the C ABI: MLP 6 -> 128 -> 128 -> 128 -> 5 (GELU),
trained and evaluated by the library's BF16-MFMA kernels.

Features per cell: (Mach, bump height, centroid x, y, z, wall distance);
outputs: primitive variables (rho, u, v, w, p)."""
import ctypes as C

import numpy as np

from . import _abi as A

N_IN, N_HID, N_OUT = 6, 128, 5


def _f32(a):
    return a.ctypes.data_as(A.f32p)


def param_slices():
    """Offsets of W1, b1, W2, b2, W3, b3, W4, b4 in the flat parameter vector."""
    shapes = [(N_HID, N_IN), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,),
              (N_OUT, N_HID), (N_OUT,)]
    out, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append((slice(off, off + n), s))
        off += n
    return out


class Mlp:
    def __init__(self, seed=0):
        h = C.c_void_p()
        rc = A.lib().nf_mlp_create(seed, C.byref(h))
        if rc != A.NF_OK:
            raise RuntimeError(f"nf_mlp_create failed with status {rc}")
        self._h = h
        self.n_params = A.lib().nf_mlp_n_params()

    def __del__(self):
        if getattr(self, "_h", None):
            A.lib().nf_mlp_destroy(self._h)
            self._h = None

    def _check(self, rc, what):
        if rc != A.NF_OK:
            raise RuntimeError(f"{what} failed (status {rc}): {A.lib().nf_mlp_last_error(self._h).decode()}")

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._check(A.lib().nf_mlp_get_params(self._h, _f32(p)), "nf_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.n_params,)
        self._check(A.lib().nf_mlp_set_params(self._h, _f32(p)), "nf_mlp_set_params")

    def set_norm(self, in_shift, in_scale, out_shift, out_scale):
        a = [np.ascontiguousarray(v, np.float32) for v in (in_shift, in_scale, out_shift, out_scale)]
        self._check(A.lib().nf_mlp_set_norm(self._h, *map(_f32, a)), "nf_mlp_set_norm")

    def get_norm(self):
        a = [np.zeros(6, np.float32), np.zeros(6, np.float32), np.zeros(5, np.float32), np.zeros(5, np.float32)]
        self._check(A.lib().nf_mlp_get_norm(self._h, *map(_f32, a)), "nf_mlp_get_norm")
        return a

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), N_OUT), np.float32)
        self._check(A.lib().nf_mlp_predict(self._h, _f32(x), len(x), _f32(y)), "nf_mlp_predict")
        return y

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32)
        self._check(A.lib().nf_mlp_loss_grad(self._h, _f32(x), _f32(y), len(x), C.byref(loss), _f32(g)),
                    "nf_mlp_loss_grad")
        return loss.value, g

    def train(self, x, y, epochs=10, batch=4096, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, lr_decay=1.0,
              seed=0):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        o = A.TrainOpts(epochs, batch, lr, beta1, beta2, eps, lr_decay, seed)
        st = A.TrainStats()
        self._check(A.lib().nf_mlp_train(self._h, _f32(x), _f32(y), len(x), C.byref(o), C.byref(st)),
                    "nf_mlp_train")
        return {f: getattr(st, f) for f, _ in A.TrainStats._fields_}
