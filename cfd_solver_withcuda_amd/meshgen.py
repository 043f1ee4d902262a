"""Seeded mixed triangle/quadrilateral mesh generator and legacy-VTK writer.

The generator produces the meshes of the solver's workloads: a structured grid fitted to the
geometry (shock tube, circular-arc bump channel, forward-facing step, compression ramp), whose
interior nodes are jittered uniformly by +-`jitter` of the local spacing and whose quads are split
into two triangles with probability `tri_frac`. All randomness comes from SplitMix64 hashes of
(seed, stream, index), so a mesh is a pure function of its parameters.

Meshes are returned as (points[n,3] float64, cells[m,4] int32 with -1 padding for triangles). The
VTK writer emits the legacy ASCII format read by the reference's Read_VTK_Grid
(src/Read_Gmsh_File.cpp:11-144): POINTS, CELLS (3/4-node cells only), CELL_TYPES (5/9).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on a uint64 array."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Deterministic U[0,1) doubles for indices `idx` in stream `stream` of `seed`."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.array([(seed * 0x100000001B3 + stream) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
        h = splitmix64(idx.astype(np.uint64) ^ base)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Mesh:
    points: np.ndarray  # [n,3] float64
    cells: np.ndarray  # [m,4] int32, -1 padded (triangles)
    name: str = ""

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def n_tri(self) -> int:
        return int((self.cells[:, 3] < 0).sum())


def structured_to_mixed(X: np.ndarray, Y: np.ndarray, active: np.ndarray | None = None,
                        seed: int = 0, jitter: float = 0.1, tri_frac: float = 0.3,
                        name: str = "") -> Mesh:
    """Turn a structured node grid X,Y [ny+1, nx+1] into a jittered mixed mesh.

    active[j, i] selects which quads exist (None = all). Nodes touching a missing quad or the
    outer frame stay fixed, so every boundary edge keeps its exact coordinates.
    """
    nyp, nxp = X.shape
    ny, nx = nyp - 1, nxp - 1
    if active is None:
        active = np.ones((ny, nx), dtype=bool)
    # A node is interior when all four surrounding quads exist.
    pad = np.zeros((ny + 2, nx + 2), dtype=bool)
    pad[1:-1, 1:-1] = active
    interior = pad[:-1, :-1] & pad[:-1, 1:] & pad[1:, :-1] & pad[1:, 1:]
    # local spacing (central, one-sided at the frame)
    dxl = np.empty_like(X)
    dyl = np.empty_like(Y)
    dxl[:, 1:-1] = 0.5 * (X[:, 2:] - X[:, :-2])
    dxl[:, 0] = X[:, 1] - X[:, 0]
    dxl[:, -1] = X[:, -1] - X[:, -2]
    dyl[1:-1, :] = 0.5 * (Y[2:, :] - Y[:-2, :])
    dyl[0, :] = Y[1, :] - Y[0, :]
    dyl[-1, :] = Y[-1, :] - Y[-2, :]
    nid = np.arange(nyp * nxp, dtype=np.int64).reshape(nyp, nxp)
    Xj = X.copy()
    Yj = Y.copy()
    if jitter > 0.0:
        rx = uniform01(seed, 1, nid) * 2.0 - 1.0
        ry = uniform01(seed, 2, nid) * 2.0 - 1.0
        Xj = np.where(interior, X + jitter * rx * dxl, X)
        Yj = np.where(interior, Y + jitter * ry * dyl, Y)
    # Keep only nodes used by an active quad; renumber them densely in row-major order.
    used = np.zeros((nyp, nxp), dtype=bool)
    used[:-1, :-1] |= active
    used[:-1, 1:] |= active
    used[1:, :-1] |= active
    used[1:, 1:] |= active
    new_id = -np.ones(nyp * nxp, dtype=np.int64)
    flat_used = used.ravel()
    new_id[flat_used] = np.arange(int(flat_used.sum()))
    pts = np.zeros((int(flat_used.sum()), 3), dtype=np.float64)
    pts[:, 0] = Xj.ravel()[flat_used]
    pts[:, 1] = Yj.ravel()[flat_used]
    # quads in row-major order, corners (BL, BR, TR, TL)
    jj, ii = np.nonzero(active)
    bl = new_id[nid[jj, ii]]
    br = new_id[nid[jj, ii + 1]]
    tr = new_id[nid[jj + 1, ii + 1]]
    tl = new_id[nid[jj + 1, ii]]
    qid = (jj.astype(np.int64) * nx + ii)
    split = uniform01(seed, 3, qid) < tri_frac
    diag = uniform01(seed, 4, qid) < 0.5
    nq = len(qid)
    counts = np.where(split, 2, 1)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    cells = -np.ones((int(counts.sum()), 4), dtype=np.int32)
    # unsplit quads
    q = ~split
    cells[start[q]] = np.stack([bl[q], br[q], tr[q], tl[q]], axis=1)
    # split along BL-TR
    s0 = split & ~diag
    cells[start[s0], :3] = np.stack([bl[s0], br[s0], tr[s0]], axis=1)
    cells[start[s0] + 1, :3] = np.stack([bl[s0], tr[s0], tl[s0]], axis=1)
    # split along BR-TL
    s1 = split & diag
    cells[start[s1], :3] = np.stack([bl[s1], br[s1], tl[s1]], axis=1)
    cells[start[s1] + 1, :3] = np.stack([br[s1], tr[s1], tl[s1]], axis=1)
    del nq
    return Mesh(pts, cells, name)


def _grid(x0, x1, nx, y0, y1, ny):
    x = x0 + (x1 - x0) * (np.arange(nx + 1, dtype=np.float64) / nx)
    y = y0 + (y1 - y0) * (np.arange(ny + 1, dtype=np.float64) / ny)
    X, Y = np.meshgrid(x, y)
    return X, Y


def shock_tube_mesh(nx: int = 400, ny: int = 40, length: float = 1.0, height: float = 0.1,
                    seed: int = 0, jitter: float = 0.0, tri_frac: float = 0.0) -> Mesh:
    """Sod shock tube [0,length]x[0,height] (uniform quads by default)."""
    X, Y = _grid(0.0, length, nx, 0.0, height, ny)
    return structured_to_mixed(X, Y, None, seed, jitter, tri_frac, f"sod_{nx}x{ny}")


def bump_mesh(nx: int = 1400, ny: int = 700, thickness: float = 0.10, seed: int = 1,
              jitter: float = 0.1, tri_frac: float = 0.3) -> Mesh:
    """Channel [-1,2]x[0,1] with a circular-arc bump of chord 1 on x in [0,1].

    The arc through (0,0), (0.5,t), (1,0) has radius R = (0.25 + t^2)/(2t); grid lines are
    fitted so the bottom row follows the bump and the top row stays at y = 1.
    """
    x = -1.0 + 3.0 * (np.arange(nx + 1, dtype=np.float64) / nx)
    eta = np.arange(ny + 1, dtype=np.float64) / ny
    yb = np.zeros_like(x)
    if thickness > 0.0:
        R = (0.25 + thickness * thickness) / (2.0 * thickness)
        on = (x > 0.0) & (x < 1.0)
        yb[on] = np.sqrt(R * R - (x[on] - 0.5) ** 2) - (R - thickness)
    X = np.broadcast_to(x[None, :], (ny + 1, nx + 1)).copy()
    Y = yb[None, :] + (1.0 - yb[None, :]) * eta[:, None]
    return structured_to_mixed(X, Y, None, seed, jitter, tri_frac, f"bump_{nx}x{ny}_t{thickness:g}")


def ffs_mesh(h: float = 1.0 / 1560, seed: int = 2, jitter: float = 0.1,
             tri_frac: float = 0.3) -> Mesh:
    """Forward-facing step: [0,3]x[0,1] minus the step [0.6,3]x[0,0.2], uniform spacing h."""
    nx = int(round(3.0 / h))
    ny = int(round(1.0 / h))
    X, Y = _grid(0.0, 3.0, nx, 0.0, 1.0, ny)
    xc = 0.5 * (X[:-1, :-1] + X[:-1, 1:])
    yc = 0.5 * (Y[:-1, :-1] + Y[1:, :-1])
    active = ~((xc > 0.6) & (yc < 0.2))
    return structured_to_mixed(X, Y, active, seed, jitter, tri_frac, f"ffs_{nx}x{ny}")


def ramp_mesh(nx: int, ny: int, angle_deg: float, x_corner: float = 0.5, length: float = 1.5,
              height: float = 1.0, seed: int = 3, jitter: float = 0.1,
              tri_frac: float = 0.3) -> Mesh:
    """Compression ramp: flat wall up to x_corner, then a ramp at `angle_deg`; top at y=height."""
    x = length * (np.arange(nx + 1, dtype=np.float64) / nx)
    eta = np.arange(ny + 1, dtype=np.float64) / ny
    yb = np.where(x > x_corner, (x - x_corner) * math.tan(math.radians(angle_deg)), 0.0)
    X = np.broadcast_to(x[None, :], (ny + 1, nx + 1)).copy()
    Y = yb[None, :] + (height - yb[None, :]) * eta[:, None]
    return structured_to_mixed(X, Y, None, seed, jitter, tri_frac, f"ramp_{nx}x{ny}_{angle_deg:g}")


def write_vtk(path: str, mesh: Mesh) -> None:
    """Legacy ASCII VTK unstructured grid with full double precision coordinates."""
    pts = mesh.points
    cells = mesh.cells
    nn = np.where(cells[:, 3] < 0, 3, 4)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\n")
        f.write(f"{mesh.name or 'mesh'}\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {len(pts)} double\n")
        np.savetxt(f, pts, fmt="%.17g")
        f.write(f"CELLS {len(cells)} {int((nn + 1).sum())}\n")
        for row, k in zip(cells, nn):
            f.write(" ".join(str(int(v)) for v in ([k] + list(row[:k]))) + "\n")
        f.write(f"CELL_TYPES {len(cells)}\n")
        f.write("\n".join("5" if k == 3 else "9" for k in nn) + "\n")


def read_vtk(path: str) -> Mesh:
    """Reads the POINTS/CELLS sections of a legacy ASCII VTK file (3/4-node cells only)."""
    with open(path) as f:
        tokens = f.read().split()
    pts = None
    cells = []
    i = 0
    while i < len(tokens):
        t = tokens[i]
        if t == "POINTS":
            n = int(tokens[i + 1])
            vals = np.array(tokens[i + 3:i + 3 + 3 * n], dtype=np.float64)
            pts = vals.reshape(n, 3)
            i += 3 + 3 * n
        elif t == "CELLS":
            m = int(tokens[i + 1])
            i += 3
            for _ in range(m):
                k = int(tokens[i])
                ids = [int(v) for v in tokens[i + 1:i + 1 + k]]
                i += 1 + k
                if k in (3, 4):
                    cells.append(ids + [-1] * (4 - k))
        else:
            i += 1
    return Mesh(pts, np.array(cells, dtype=np.int32).reshape(-1, 4), path)
