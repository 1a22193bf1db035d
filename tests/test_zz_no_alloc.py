"""The allocation-free forward pass (ofx_mlp_forward_noalloc) against
ofx_mlp_forward through the C ABI: caller-owned device buffers, the fixed
per-GPU work array, two calls in a row (the second smaller, so the array
holds the first call's larger contents), and the size limit."""
import numpy as np
import pytest

from oracle.mlp_oracle import NIN
from tests.test_mlp_oracle import random_params

pytestmark = pytest.mark.gpu

MAXN = 8192


def test_noalloc_forward_equals_forward():
    import torch
    from oneflow_amd.mlp import Mlp, _lib
    L = _lib()
    net = Mlp(seed=11)
    rng = np.random.default_rng(5)
    net.set_params(random_params(rng, scale=1.0).astype(np.float32))
    stream = torch.cuda.current_stream().cuda_stream
    for n in (MAXN, 300):
        x = rng.normal(0, 1, (n, NIN)).astype(np.float32)
        want = net.forward(x)
        xd = torch.from_numpy(x).to("cuda")
        yd = torch.full((n, 5), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        print(f"[noalloc] n={n}", flush=True)
        rc = L.ofx_mlp_forward_noalloc(net.h, xd.data_ptr(), n, yd.data_ptr(), stream)
        assert rc == 0, L.ofx_last_error().decode()
        torch.cuda.synchronize()
        got = yd.cpu().numpy()
        assert np.array_equal(got, want), float(np.nanmax(np.abs(got - want)))
    # beyond the fixed size: rejected before anything is launched
    rc = L.ofx_mlp_forward_noalloc(net.h, xd.data_ptr(), MAXN + 1, yd.data_ptr(), stream)
    assert rc == -1 and b"8192" in L.ofx_last_error()
    net.close()
