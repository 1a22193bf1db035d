"""Build libofx.so (HIP, gfx950) in-tree.

hipcc compiles every .hip source of oneflow_amd/csrc for gfx950 with
-ffp-contract=off (FP64 parity with the reference's x86-64 arithmetic) and
packed FP32 instructions disabled, and links one shared library next to this
file.  The build always recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
OUT = HERE / "libofx.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False) -> Path:
    srcs = sorted(SRC.glob("*.hip"))
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    procs = []
    for s in srcs:
        o = objdir / (s.stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), s))
        objs.append(o)
    failed = False
    for p, s in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l]
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- hipcc failed on {s.name}\n" + "\n".join(lines) + "\n")
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed")
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
