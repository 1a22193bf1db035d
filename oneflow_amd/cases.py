"""Seeded case generators for the batched airfoil sweep.

Each case is a NACA 4-digit airfoil (camber m in [0, 4] %, camber position
p = 0.4, thickness t in [8, 18] %) on the hybrid quad/triangle O-grid of
`meshgen.naca_hybrid` with 40 wall layers sized for y+ ~ 1 at the case's
Reynolds number, a target cell count drawn in [20k, 200k], Mach in
[0.3, 0.85], AoA in [-2, 6] deg and Re in [3e6, 1e7]; Roe + SA +
Venkatakrishnan, FP64, multicolour numbering.  The same seed always gives the
same cases.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import meshgen
from .config import CaseConfig

# cells of the reference layout: n_surf * (n_quad + 2 * n_outer) = 512 * (40 + 196)
_REF_CELLS = 512 * (40 + 2 * 98)


@dataclass
class AirfoilCase:
    m: float
    p: float
    t: float
    mach: float
    aoa: float
    reynolds: float
    target_cells: int

    def config(self, **kw) -> CaseConfig:
        base = dict(ischeme=1, ilim=2, ivischeme=0, vismodel=3, mach_ref=self.mach, aoa_degree=self.aoa,
                    reynolds=self.reynolds, centropy1=0.001, centropy2=0.001, vencat_coef=0.001)
        base.update(kw)
        return CaseConfig(**base)

    def grid(self, numbering: str = "multicolor", face_order: str = "cell"):
        """The case's grid; with the multicolour numbering, face_order "spatial"
        sorts the interior faces by their cells' ids before it."""
        s = math.sqrt(self.target_cells / _REF_CELLS)
        n_surf = max(64, int(round(512 * s / 8.0)) * 8)
        n_outer = max(16, int(round(98 * s)))
        g = meshgen.naca_hybrid(n_surf=n_surf, n_quad=40, ratio=1.15, reynolds=self.reynolds, n_outer=n_outer,
                                radius=30.0, m=self.m, p=self.p, t=self.t)
        if numbering == "multicolor":
            order = meshgen.multicolor_order(g)
            g = meshgen.renumber(g, order)
            if face_order == "spatial":
                g = meshgen.sort_interior_faces(g, order)
        return g

    def features(self) -> np.ndarray:
        """The case-level inputs of the warm-start network: m, p, t, log10 Re, Mach, AoA (rad)."""
        return np.array([self.m, self.p, self.t, math.log10(self.reynolds), self.mach, math.radians(self.aoa)])


def airfoil_sweep(n: int = 64, seed: int = 2026, min_cells: int = 20_000, max_cells: int = 200_000,
                  mach=(0.3, 0.85), aoa=(-2.0, 6.0), reynolds=(3.0e6, 1.0e7)) -> list[AirfoilCase]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(AirfoilCase(m=float(rng.uniform(0.0, 0.04)), p=0.4, t=float(rng.uniform(0.08, 0.18)),
                               mach=float(rng.uniform(*mach)), aoa=float(rng.uniform(*aoa)),
                               reynolds=float(rng.uniform(*reynolds)),
                               target_cells=int(rng.uniform(min_cells, max_cells))))
    return out
