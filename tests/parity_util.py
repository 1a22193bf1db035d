"""Comparison helpers for GPU-vs-oracle parity tests."""
import numpy as np


def rel_err(gpu, ref):
    """Normwise relative error max|gpu-ref| / max|ref| (0 if both are all zero)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(gpu - ref)) if ref.size else 0.0
    if den == 0.0:
        return float(num)
    return float(num / den)


def bitwise_frac(gpu, ref):
    gpu = np.asarray(gpu, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    return float(np.mean(gpu == ref)) if ref.size else 1.0


def report(name, gpu, ref):
    e = rel_err(gpu, ref)
    b = bitwise_frac(gpu, ref)
    print(f"  {name:14s} rel_err={e:.3e} bitwise={b*100:.2f}%", flush=True)
    return e
