"""Python view of the warm-start network in libofx.so (ofx_mlp_*).

All evaluation and training run in the library's MFMA kernels on the GPU;
this module only moves arrays across the C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .solver import OfxError, load_library

NIN, NOUT = 9, 5


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _lib():
    L = load_library()
    if not getattr(L, "_mlp_ready", False):
        P = ctypes.c_void_p
        fp = ctypes.POINTER(ctypes.c_float)
        L.ofx_mlp_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_ulonglong]
        L.ofx_mlp_destroy.argtypes = [P]
        L.ofx_mlp_param_count.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.ofx_mlp_get_params.argtypes = [P, fp]
        L.ofx_mlp_set_params.argtypes = [P, fp]
        L.ofx_mlp_forward.argtypes = [P, fp, ctypes.c_int, fp]
        L.ofx_mlp_forward_noalloc.argtypes = [P, P, ctypes.c_int, P, P]
        L.ofx_mlp_set_forward_variant.argtypes = [P, ctypes.c_int]
        L.ofx_mlp_gradients.argtypes = [P, fp, fp, ctypes.c_int, fp, fp]
        L.ofx_mlp_train.argtypes = [P, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_ulonglong, fp]
        L._mlp_ready = True
    return L


class Mlp:
    """BF16 MLP 9 -> 256 x 5 (SiLU) -> 5 on one GPU."""

    def __init__(self, device: int = 0, seed: int = 1):
        self._L = _lib()
        self.h = ctypes.c_void_p()
        self._check(self._L.ofx_mlp_create(ctypes.byref(self.h), device, seed))
        n = ctypes.c_int()
        self._check(self._L.ofx_mlp_param_count(self.h, ctypes.byref(n)))
        self.n_params = n.value

    def _check(self, rc):
        if rc != 0:
            raise OfxError(f"ofx error {rc}: {self._L.ofx_last_error().decode()}")

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._check(self._L.ofx_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.n_params,)
        self._check(self._L.ofx_mlp_set_params(self.h, _fp(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        y = np.zeros((n, NOUT), np.float32)
        step = 1 << 20
        for i in range(0, n, step):
            xs = np.ascontiguousarray(x[i:i + step])
            ys = np.zeros((xs.shape[0], NOUT), np.float32)
            self._check(self._L.ofx_mlp_forward(self.h, _fp(xs), xs.shape[0], _fp(ys)))
            y[i:i + step] = ys
        return y

    def set_forward_variant(self, variant) -> None:
        """1 (default): all layers in one kernel; 0: one GEMM kernel per layer. Same results."""
        self._check(self._L.ofx_mlp_set_forward_variant(self.h, int(variant)))

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self._check(self._L.ofx_mlp_gradients(self.h, _fp(x), _fp(t), x.shape[0], _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, batch=8192, steps=1000, lr=1e-3, seed=7) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(steps, np.float32)
        self._check(self._L.ofx_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], batch, steps, lr, seed, _fp(losses)))
        return losses

    def close(self):
        if self.h:
            self._L.ofx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
