"""Helpers shared by the golden-file tests."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference"


def load_res(path, nvar):
    """Rows of a res.dat / turbres.dat file: [iter, sub-iter, res1..resN] -> res columns."""
    rows = [l.split() for l in Path(path).read_text().splitlines() if l[:1].isdigit()]
    return np.array([[float(v) for v in r[2:2 + nvar]] for r in rows])


def harness_equal(a: str, b: str, tol: float = 1.0e-8) -> bool:
    """The reference harness rule (test/test.py:38-79): equal text, else |a-b| <= tol."""
    if a == b:
        return True
    return abs(float(a) - float(b)) <= tol
