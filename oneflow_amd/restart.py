"""Restart files and residual histories in the reference's formats.

flow.dat (one zone, NS solver, as Restart::Dump writes it through the zone's
data book, codes/restart/src/Restart.cpp:99-194, codes/database/src/DataBaseIO.cpp:57-74):

    int64  zone payload length (bytes that follow)
    int32  outer step count           (Iteration::outerSteps)
    float64 physical time             (ctrl.currTime)
    float64 q1[5][nTCell], q2[5][nTCell]     primitive state, two time levels
    float64 res1[5][nTCell], res2[5][nTCell] residuals of those levels

A steady restart takes q = q1 (Restart::ReadUnsteady) and continues from the
stored step count.  res.dat / turbres.dat rows are written as
ResidualTask::DumpResidual does (codes/residual/src/ResidualTask.cpp:140-188):
"iter sub-iter " then every norm left-aligned in 13 columns, scientific with 5
digits, each followed by a space.  wallaero.dat (DUMP_PRESSURE_COEFF, the
idump_pres branch of CalcAeroForce, codes/force/src/AeroForceTask.cpp:242-260,
353-362): a title block, "Zone  i = N ", then per wall face x y z -cp cf
left-aligned in 20 columns, scientific with 10 digits.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class FlowRestart:
    outer_steps: int
    time: float
    q1: np.ndarray    # [5][nTCell]
    q2: np.ndarray
    res1: np.ndarray
    res2: np.ndarray


def read_flow_dat(path, n_tcell: int, n_equ: int = 5) -> FlowRestart:
    b = Path(path).read_bytes()
    (length,) = struct.unpack("<q", b[:8])
    if length != len(b) - 8:
        raise ValueError(f"flow.dat: zone length {length} does not match the file ({len(b) - 8} bytes)")
    (steps,) = struct.unpack("<i", b[8:12])
    (time,) = struct.unpack("<d", b[12:20])
    d = np.frombuffer(b, dtype="<f8", offset=20)
    n = n_equ * n_tcell
    if d.size != 4 * n:
        raise ValueError(f"flow.dat: {d.size} values, expected 4 x {n_equ} x {n_tcell}")
    f = [d[k * n:(k + 1) * n].reshape(n_equ, n_tcell).copy() for k in range(4)]
    return FlowRestart(steps, time, *f)


def write_flow_dat(path, r: FlowRestart) -> None:
    body = struct.pack("<i", r.outer_steps) + struct.pack("<d", r.time)
    for a in (r.q1, r.q2, r.res1, r.res2):
        body += np.ascontiguousarray(a, dtype="<f8").tobytes()
    Path(path).write_bytes(struct.pack("<q", len(body)) + body)


def _sci5(v: float) -> str:
    return f"{v:.5e}".ljust(13)


def residual_lines(first_step: int, norms, n_var: int) -> list[str]:
    """Header + one row per step, as res.dat (n_var = 5) or turbres.dat (1)."""
    head = ['Title="THE RESIDUAL OF ONEFLOW"', "Variables=", '"iter"', '"sub-iter"']
    head += [f'"res{k + 1}"' for k in range(n_var)]
    rows = []
    for i, row in enumerate(np.asarray(norms)):
        rows.append(f"{first_step + i} 1 " + "".join(_sci5(float(v)) + " " for v in row[:n_var]))
    return head + rows


def write_residuals(path, first_step: int, norms, n_var: int) -> None:
    Path(path).write_text("\n".join(residual_lines(first_step, norms, n_var)) + "\n")


def wallaero_lines(rows) -> list[str]:
    """wallaero.dat text for wall rows [n][5] (x, y, z, -cp, cf)."""
    head = ['title="THE FLOW FIELD OF ONEFLOW"', "variables=", '"x"', '"y"', '"z"', '"-cp"', '"cf"',
            f"Zone  i = {len(rows)} "]
    return head + ["".join(f"{float(v):.10e}".ljust(20) for v in r) for r in np.asarray(rows)]


def write_wallaero(path, rows) -> None:
    Path(path).write_text("\n".join(wallaero_lines(rows)) + "\n")


AERO_COLUMNS = ["CL", "CD", "CD_PR", "CD_SF", "CDL2", "Xcp", "Fx", "Fy", "Fz", "Cmx", "Cmy", "Cmz"]


def aero_lines(first_step: int, rows, dt: float = 1.0e-3, time0: float = 0.0) -> list[str]:
    """aero.dat text (AerodynamicForceTask::Dump, codes/force/src/AeroForceTask.cpp:96-150):
    the header, then per step `iter  sub-iter  time` and the 12 value columns
    of `rows` [n][12] (ofx_aero_row), four spaces after every field.  The time
    column is the reference's ctrl.currTime, advanced by global_dt each outer
    step (codes/multigrid/src/Multigrid.cpp:137-155) from time0."""
    head = ['Title="Aerodynamic Force"', "Variables=", '"iter"', '"sub-iter"', '"time"']
    head += [f'"{c}"' for c in AERO_COLUMNS]
    out, t = [], time0
    for i, row in enumerate(np.asarray(rows)):
        t += dt
        fields = [str(first_step + i), "1", f"{t:.6e}"] + [f"{float(v):.4e}" for v in row[:12]]
        out.append("".join(f + "    " for f in fields))
    return head + out


def write_aero(path, first_step: int, rows, dt: float = 1.0e-3, time0: float = 0.0) -> None:
    Path(path).write_text("\n".join(aero_lines(first_step, rows, dt, time0)) + "\n")
