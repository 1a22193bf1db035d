"""The allocation-free forward pass (ofx_mlp_forward_noalloc) against
ofx_mlp_forward through the C ABI: caller-owned device buffers, the fixed
per-GPU work array, two calls in a row (the second smaller, so the array
holds the first call's larger contents; the first on the network's own
stream, the second on a caller's stream), and the size limit.

The device buffers come from the HIP runtime that libofx.so is linked
against, called through ctypes, so the test needs no other GPU library in
the process."""
import ctypes

import numpy as np
import pytest

from oracle.mlp_oracle import NIN
from tests.test_mlp_oracle import random_params

pytestmark = pytest.mark.gpu

MAXN = 8192
H2D, D2H = 1, 2


def hip_runtime():
    hip = ctypes.CDLL("libamdhip64.so.7")  # already loaded by libofx.so
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamDestroy.argtypes = [vp]
    return hip


def test_noalloc_forward_equals_forward():
    from oneflow_amd.mlp import Mlp, _lib
    L = _lib()
    net = Mlp(seed=11)
    hip = hip_runtime()
    rng = np.random.default_rng(5)
    net.set_params(random_params(rng, scale=1.0).astype(np.float32))
    xd, yd, st = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(xd), MAXN * NIN * 4) == 0
    assert hip.hipMalloc(ctypes.byref(yd), MAXN * 5 * 4) == 0
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    for n, stream in ((MAXN, None), (300, st)):
        x = rng.normal(0, 1, (n, NIN)).astype(np.float32)
        want = net.forward(x)
        y = np.full((n, 5), np.nan, np.float32)
        assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
        assert hip.hipMemcpy(yd, y.ctypes.data, y.nbytes, H2D) == 0
        print(f"[noalloc] n={n} stream={'caller' if stream else 'network'}", flush=True)
        rc = L.ofx_mlp_forward_noalloc(net.h, xd, n, yd, stream)
        assert rc == 0, L.ofx_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(y.ctypes.data, yd, y.nbytes, D2H) == 0
        assert np.array_equal(y, want), float(np.nanmax(np.abs(y - want)))
    # beyond the fixed size: rejected before anything is launched
    rc = L.ofx_mlp_forward_noalloc(net.h, xd, MAXN + 1, yd, None)
    assert rc == -1 and b"8192" in L.ofx_last_error()
    assert hip.hipStreamDestroy(st) == 0
    assert hip.hipFree(xd) == 0 and hip.hipFree(yd) == 0
    net.close()
