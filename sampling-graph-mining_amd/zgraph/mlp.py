"""Device MLP of the learned error-latency profile (BF16 MFMA kernels)."""
import ctypes as C

import numpy as np

from ._capi import check, lib

IN_DIM = 16


class MLP:
    """16-64-64-64-1 ReLU network; parameters live in HBM (fp32 master)."""

    def __init__(self, seed=1):
        self.h = C.c_void_p()
        check(lib().zg_mlp_create(seed, C.byref(self.h)))
        self.nparam = lib().zg_mlp_num_params()
        assert lib().zg_mlp_input_dim() == IN_DIM

    def params(self):
        p = np.empty(self.nparam, np.float32)
        check(lib().zg_mlp_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.nparam,)
        check(lib().zg_mlp_set_params(self.h, p.ctypes.data))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        assert X.ndim == 2 and X.shape[1] == IN_DIM
        y = np.empty(X.shape[0], np.float32)
        check(lib().zg_mlp_forward(self.h, X.ctypes.data, X.shape[0], y.ctypes.data))
        return y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(self.nparam, np.float32)
        loss = C.c_float()
        check(lib().zg_mlp_gradient(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], g.ctypes.data,
                                    C.byref(loss)))
        return g, loss.value

    def train(self, X, T, steps, batch=256, lr=1e-3, seed=7):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = C.c_float()
        check(lib().zg_mlp_train(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], steps, batch, lr, seed,
                                 C.byref(loss)))
        return loss.value

    def __del__(self):
        try:
            if self.h:
                lib().zg_mlp_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass
