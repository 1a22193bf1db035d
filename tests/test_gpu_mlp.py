"""The warm-start network's MFMA kernels against its CPU oracle.

oracle/mlp_oracle.py rounds to BF16 at the same points as the GPU and is
exact FP64 elsewhere, so the GPU differs from it only by FP32 accumulation
order, its exp and reciprocal, and the rare BF16 rounding flips these cause.  Stated bounds:
outputs within 2e-3 normwise relative, loss within 1e-4 relative, every
layer's weight and bias gradient within 5e-3 normwise relative.  A short
training run must cut the loss on a smooth regression target tenfold.
"""
import numpy as np
import pytest

from oracle.mlp_oracle import KIN, KOUT, NIN, NOUT, MlpOracle
from tests.test_mlp_oracle import random_params

pytestmark = pytest.mark.gpu


def net(params=None):
    from oneflow_amd.mlp import Mlp
    m = Mlp(seed=3)
    if params is not None:
        m.set_params(params.astype(np.float32))
    return m


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    p = random_params(rng, scale=1.0).astype(np.float32)
    x = rng.normal(0, 1, (1000, NIN)).astype(np.float32)  # not a multiple of 64: padded rows
    m = net(p)
    y = m.forward(x)
    yo = MlpOracle(p, bf16=True).forward(x)
    e = float(np.linalg.norm(y - yo) / np.linalg.norm(yo))
    print(f"\n[mlp] forward rel err {e:.3e}", flush=True)
    assert e <= 2e-3, e


@pytest.mark.parametrize("n", [1050, 70_000, 70_050])
def test_one_kernel_forward_equals_layered(n):
    """The one-kernel forward pass (activations in LDS) against one GEMM
    kernel per layer: the same bits.  1050 and 70,050 samples leave the last
    128-sample workgroup half filled; 70,000 takes the layered pass's large
    tiles."""
    rng = np.random.default_rng(7)
    p = random_params(rng, scale=1.0).astype(np.float32)
    x = rng.normal(0, 1, (n, NIN)).astype(np.float32)
    m = net(p)
    m.set_forward_variant(0)
    y0 = m.forward(x)
    m.set_forward_variant(1)
    y1 = m.forward(x)
    print(f"\n[mlp] one-kernel forward, n={n}: max |diff| {float(np.max(np.abs(y1 - y0))):.3e}", flush=True)
    assert np.array_equal(y0, y1)


def test_gradients_match_oracle():
    rng = np.random.default_rng(2)
    p = random_params(rng, scale=1.0).astype(np.float32)
    x = rng.normal(0, 1, (777, NIN)).astype(np.float32)
    t = rng.normal(0, 1, (777, NOUT)).astype(np.float32)
    m = net(p)
    loss, g = m.gradients(x, t)
    lo, go = MlpOracle(p, bf16=True).gradients(x, t)
    assert abs(loss - lo) <= 1e-4 * abs(lo), (loss, lo)
    o = 0
    worst = 0.0
    for l, (i, k) in enumerate(zip(KIN, KOUT)):
        for name, cnt in (("W", i * k), ("b", k)):
            a, b = g[o:o + cnt], go[o:o + cnt]
            e = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
            worst = max(worst, e)
            assert e <= 5e-3, (l, name, e)
            o += cnt
    print(f"\n[mlp] loss {loss:.6e} vs {lo:.6e}; worst gradient block rel err {worst:.3e}", flush=True)


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, (20000, NIN)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], np.cos(x[:, 2] + x[:, 3]), x[:, 4] ** 2 - x[:, 5],
                  np.tanh(3 * x[:, 6]), x[:, 7] * x[:, 8]], axis=1).astype(np.float32)
    m = net()
    losses = m.train(x, t, batch=4096, steps=400, lr=2e-3)
    first, last = float(losses[:5].mean()), float(losses[-20:].mean())
    print(f"\n[mlp] training loss {first:.4e} -> {last:.4e}", flush=True)
    assert np.all(np.isfinite(losses)) and last < 0.1 * first, (first, last)
