"""Learned surrogate of the impurity solver (no counterpart in the reference).

``MLP`` is the device network of the C-ABI (``ed_surrogate_*``): a GELU MLP
trained and evaluated on the GPU with bf16 MFMA GEMMs, fp32 accumulation,
fp32 master weights and Adam.  The cluster-model family it learns, the
training corpus and the VCA drivers that use it as a proposer live in
``pyqcm_amd.vca_surrogate``.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class MLP:
    """n_in -> `depth` hidden GELU layers of `width` -> n_out, on the device."""

    def __init__(self, n_in, n_out, width=256, depth=4, seed=0):
        h = C.c_int64()
        check(lib().ed_surrogate_create(int(n_in), int(n_out), int(width), int(depth), int(seed), C.byref(h)))
        self.h = h.value
        self.n_in, self.n_out, self.width, self.depth = int(n_in), int(n_out), int(width), int(depth)
        n = C.c_int64()
        check(lib().ed_surrogate_n_params(self.h, C.byref(n)))
        self.n_params = n.value

    def __del__(self):
        try:
            lib().ed_surrogate_destroy(self.h)
        except Exception:
            pass

    def set_scaling(self, x_mean=None, x_std=None, y_mean=None, y_std=None, loss_weight=None):
        args = [None if v is None else _f64(v) for v in (x_mean, x_std, y_mean, y_std, loss_weight)]
        for v, n in zip(args, (self.n_in, self.n_in, self.n_out, self.n_out, self.n_out)):
            if v is not None and v.size != n:
                raise ValueError("scaling vector of length %d, expected %d" % (v.size, n))
        check(lib().ed_surrogate_set_scaling(self.h, *[_ptr(v) for v in args]))

    def set_optimizer(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().ed_surrogate_set_optimizer(self.h, float(lr), float(beta1), float(beta2), float(eps)))

    def get_weights(self):
        w = np.zeros(self.n_params)
        check(lib().ed_surrogate_get_weights(self.h, _ptr(w)))
        return w

    def set_weights(self, w):
        w = _f64(w)
        if w.size != self.n_params:
            raise ValueError("expected %d weights" % self.n_params)
        check(lib().ed_surrogate_set_weights(self.h, _ptr(w)))

    def predict(self, x):
        x = _f64(x).reshape(-1, self.n_in)
        y = np.zeros((x.shape[0], self.n_out))
        if x.shape[0]:
            check(lib().ed_surrogate_predict(self.h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def loss_grad(self, x, y):
        x = _f64(x).reshape(-1, self.n_in)
        y = _f64(y).reshape(-1, self.n_out)
        g = np.zeros(self.n_params)
        loss = C.c_double()
        check(lib().ed_surrogate_loss_grad(self.h, x.shape[0], _ptr(x), _ptr(y), C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, x, y, epochs, batch, seed=0):
        """Returns (mean training loss per epoch, device ms per step)."""
        x = _f64(x).reshape(-1, self.n_in)
        y = _f64(y).reshape(-1, self.n_out)
        hist = np.zeros(max(int(epochs), 1))
        ms = C.c_double()
        check(lib().ed_surrogate_train(self.h, x.shape[0], _ptr(x), _ptr(y), int(epochs), int(batch), int(seed),
                                       _ptr(hist), C.byref(ms)))
        return hist[:int(epochs)], ms.value

    def time(self, batch, nrep=10):
        """(ms per forward, ms per forward + backward) on `batch` resident rows."""
        a, b = C.c_double(), C.c_double()
        check(lib().ed_surrogate_time(self.h, int(batch), int(nrep), C.byref(a), C.byref(b)))
        return a.value, b.value

    def flops(self, batch):
        """(forward, forward + backward) GEMM flops for `batch` rows."""
        dims = [(self.n_in if l == 0 else self.width, self.n_out if l == self.depth else self.width)
                for l in range(self.depth + 1)]
        fwd = sum(2.0 * batch * i * o for (i, o) in dims)
        # backward: dW for every layer, dA for all but the first
        bwd = sum(2.0 * batch * i * o for (i, o) in dims) + sum(2.0 * batch * i * o for (i, o) in dims[1:])
        return fwd, fwd + bwd
