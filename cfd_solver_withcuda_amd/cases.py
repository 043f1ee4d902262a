"""Seeded case families for batched steady solves (the reference's Test_Cases/ defines one case
per run, e.g. Test_Cases/Flow_Over_Bump.cpp:141-143 for the bump's freestream rho = 1.4, p = 1):

  ramp_batch   compression ramps, angle ~ U[5, 20] deg, Mach ~ U[1.5, 4], cell counts
               log-uniform in [min_cells, max_cells], AUSM + MUSCL/Van Leer, LTS TVD-RK3;
  bump_family  circular-arc bumps, (Mach, thickness) by Latin hypercube over [1.2, 3] x
               [0.04, 0.12], HLLC + MUSCL/Van Leer + WLSQ, LTS TVD-RK3.

Random draws use the package's SplitMix64 stream (meshgen.uniform01) so that a seed fixes every
instance.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import meshgen
from .solver import freestream

RAMP_BCS = ("freestream", "exit", "exit", "wall", "wall")
BUMP_BCS = ("freestream", "exit", "wall", "wall", "wall")
TRI_FRACTION = 0.3  # quads split into two triangles


@dataclass
class Case:
    mesh: meshgen.Mesh
    fs: tuple
    mach: float
    param: float  # ramp angle (deg) or bump thickness


def _u(seed, stream, n):
    return meshgen.uniform01(seed, stream, np.arange(n, dtype=np.uint64))


def ramp_batch(n: int, seed: int = 3, min_cells: int = 20000, max_cells: int = 500000,
               length: float = 1.5):
    ang = 5.0 + 15.0 * _u(seed, 101, n)
    mach = 1.5 + 2.5 * _u(seed, 102, n)
    cells = np.exp(np.log(min_cells) + (np.log(max_cells) - np.log(min_cells)) * _u(seed, 103, n))
    out = []
    for i in range(n):
        quads = cells[i] / (1.0 + TRI_FRACTION)
        ny = max(4, int(round(math.sqrt(quads / length))))
        nx = max(4, int(round(quads / ny)))
        m = meshgen.ramp_mesh(nx, ny, float(ang[i]), length=length, seed=seed * 1000003 + i)
        out.append(Case(m, tuple(freestream(float(mach[i]))), float(mach[i]), float(ang[i])))
    return out


def latin_hypercube(n: int, seed: int, stream: int):
    """n points in [0,1)^2, one per row and column stratum."""
    u = _u(seed, stream, 2 * n).reshape(n, 2)
    out = np.empty((n, 2))
    for d in range(2):
        perm = np.argsort(_u(seed, stream + 1 + d, n), kind="stable")
        out[:, d] = (perm + u[:, d]) / n
    return out


def bump_family(n: int, seed: int = 5, cells: int = 60000, stream: int = 200,
                mach_range=(1.2, 3.0), thick_range=(0.04, 0.12)):
    """n bump cases on meshes of about `cells` cells (3:1 aspect grid)."""
    lh = latin_hypercube(n, seed, stream)
    mach = mach_range[0] + (mach_range[1] - mach_range[0]) * lh[:, 0]
    thick = thick_range[0] + (thick_range[1] - thick_range[0]) * lh[:, 1]
    quads = cells / (1.0 + TRI_FRACTION)
    ny = max(4, int(round(math.sqrt(quads / 3.0))))
    nx = 3 * ny
    out = []
    for i in range(n):
        m = meshgen.bump_mesh(nx, ny, float(thick[i]), seed=seed * 1000003 + i)
        out.append(Case(m, tuple(freestream(float(mach[i]))), float(mach[i]), float(thick[i])))
    return out


__all__ = ["Case", "ramp_batch", "bump_family", "latin_hypercube", "RAMP_BCS", "BUMP_BCS"]
