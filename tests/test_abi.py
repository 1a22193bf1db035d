"""The C-ABI library loads and exports every entry point include/ofx.h declares (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "ofx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ofx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "ofx_residual_ns" in names and "ofx_lusgs_ns" in names and len(names) >= 20


def test_library_exports_all_declared_symbols():
    lib = ROOT / "oneflow_amd" / "libofx.so"
    if not lib.exists():
        pytest.skip("libofx.so not built in this checkout")
    L = ctypes.CDLL(str(lib))
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_oracle_exports_mirror():
    from oracle.pyoracle import lib
    L = lib()
    for n in declared():
        if n in ("ofx_ctx_create", "ofx_ctx_destroy", "ofx_mesh_upload", "ofx_sync",
                 "ofx_time_kernel", "ofx_copy_bench", "ofx_set_face_tile", "ofx_set_kernel_variant", "ofx_get_kernel_variant",
                 "ofx_autotune", "ofx_set_storage_order") or \
                n.startswith("ofx_batch_") or n.startswith("ofx_mlp_") or n.startswith("ofx_zones_"):
            continue  # context plumbing, GPU launch tuning, batching, the predictor: no CPU counterpart
        assert hasattr(L, n.replace("ofx_", "ofo_")), n
