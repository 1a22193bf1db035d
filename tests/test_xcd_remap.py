"""The XCD block remaps are bijections (host build of ofx_common.h, no GPU).

Every residual kernel maps its hardware block index through xcd_remap or
xcd_remap_seg; a map that missed or repeated a block would skip or redo
cells.  tests/xcd_remap_check.hip checks 20,000 random launch shapes and
colour-segment cuts exhaustively.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
HIPCC = "/opt/rocm/bin/hipcc"


def test_xcd_remap_bijective(tmp_path):
    if not Path(HIPCC).exists() and not shutil.which("hipcc"):
        pytest.skip("hipcc not available")
    exe = tmp_path / "xcd_remap_check"
    subprocess.run([HIPCC if Path(HIPCC).exists() else "hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    str(HERE / "xcd_remap_check.hip"), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.startswith("ok "), out
