"""CPU oracle of the warm-start network (test infrastructure only).

FP64 restatement of the GPU MLP in oneflow_amd/csrc/ofx_mlp.hip: 9 inputs
(padded to 16) -> 5 hidden layers of 256 with SiLU -> 5 outputs (padded to
64), mean-squared-error loss over the 5 outputs.  With `bf16=True` the values
are rounded to BF16 at exactly the points the GPU rounds them (inputs, weights,
hidden activations, the output gradient and the pre-activation gradients fed
to the next GEMM), and everything else is exact FP64; the GPU differs from it
only by FP32 accumulation order and its exp.  With `bf16=False` it is the
plain FP64 network, whose gradient tests/test_mlp_oracle.py pins against
central finite differences.

Parameter layout (flat vector): for each layer W[kin][kout] then b[kout],
kin = 16, 256, 256, 256, 256, 256 and kout = 256, 256, 256, 256, 256, 64.
Only tests/, __graft_entry__.smoke() and bench.py's baseline leg use this
module.
"""
from __future__ import annotations

import numpy as np

NIN, NIN_P, NHID, NOUT, NOUT_P = 9, 16, 256, 5, 64
KIN = [NIN_P, NHID, NHID, NHID, NHID, NHID]
KOUT = [NHID, NHID, NHID, NHID, NHID, NOUT_P]


def param_count() -> int:
    return sum(i * o + o for i, o in zip(KIN, KOUT))


def to_bf16(a) -> np.ndarray:
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for i, k in zip(KIN, KOUT):
        w = p[o:o + i * k].reshape(i, k)
        o += i * k
        b = p[o:o + k]
        o += k
        out.append((w, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


class MlpOracle:
    def __init__(self, params, bf16: bool = True):
        self.layers = unpack(params)
        self.bf16 = bf16

    def _r(self, a):
        return to_bf16(a) if self.bf16 else np.asarray(a, np.float64)

    def _forward(self, x):
        n = x.shape[0]
        h = np.zeros((n, NIN_P))
        h[:, :NIN] = x
        h = self._r(h)
        hs, zs = [h], []
        for w, b in self.layers[:-1]:
            z = h @ self._r(w) + b
            zs.append(z)
            h = self._r(silu(z))
            hs.append(h)
        w, b = self.layers[-1]
        y = h @ self._r(w) + b
        return y, hs, zs

    def forward(self, x) -> np.ndarray:
        y, _, _ = self._forward(np.asarray(x, np.float64))
        return y[:, :NOUT]

    def gradients(self, x, t):
        """(loss, flat gradient) of mean((y - t)^2) over n samples x 5 outputs."""
        x = np.asarray(x, np.float64)
        t = np.asarray(t, np.float64)
        n = x.shape[0]
        y, hs, zs = self._forward(x)
        d = y[:, :NOUT] - t
        loss = float(np.mean(d * d))
        gy = np.zeros_like(y)
        gy[:, :NOUT] = 2.0 * d / (n * NOUT)
        grads = [None] * len(self.layers)
        g_gemm = self._r(gy)  # what the next GEMMs see
        g_bias = gy
        for l in range(len(self.layers) - 1, -1, -1):
            w, _ = self.layers[l]
            grads[l] = (hs[l].T @ g_gemm, g_bias.sum(axis=0))
            if l == 0:
                break
            g = (g_gemm @ self._r(w).T) * dsilu(zs[l - 1])
            g_bias = g
            g_gemm = self._r(g)
        flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
        return loss, flat
