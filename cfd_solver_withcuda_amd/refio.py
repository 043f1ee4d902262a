"""Reference-compatible files, so runs of this library can be compared with the reference's and
opened in the same tools:

  write_solution / read_solution   Write_Solution type 1 (src/output_files.cpp:492-556): header
                                   (nx ny nz), No_Physical_Cells, iterations, then per cell
                                   "id dt rho P T u v M P0"; read back as the restart of
                                   Initialize(const string&) (src/Initialize.cpp:126-179);
  append_error                     Write_Error_File (src/output_files.cpp:558-575): one row of
                                   Error[0..3] per call;
  write_vtk_solution               the grid (Read_Write_Grid) plus Append_Solution's CELL_DATA
                                   blocks (src/Create_Vtk_File.cpp:5-158);
  config_from_json                 the reference's solver JSON (src/Configuration_Read.cpp:23-192,
                                   json_Files/Solver_Config.json) mapped to make_config()
                                   arguments.

Numbers are written as C++ streams write doubles by default (6 significant digits, %g).
Derived quantities follow the reference's primitive recovery (src/Basic_Functions.cpp:85-142):
velocities below 1e-10 in magnitude flushed to zero, T = P / (rho R) with R = 287, M = |V| / a,
P0 = P (1 + (gamma-1)/2 M^2)^(gamma/(gamma-1)). They are output post-processing on the host; the
solution itself comes from the device.
"""
from __future__ import annotations

import json
import math
import re

import numpy as np

GAMMA = 1.4
R_GC = 287.0


def _g(x) -> str:
    return f"{float(x):g}"


def primitives(U: np.ndarray) -> dict:
    U = np.asarray(U, np.float64)
    rho = U[:, 0]
    u = U[:, 1] / rho
    v = U[:, 2] / rho
    u = np.where(np.abs(u) <= 1e-10, 0.0, u)
    v = np.where(np.abs(v) <= 1e-10, 0.0, v)
    q2 = u * u + v * v
    p = (GAMMA - 1.0) * (U[:, 3] - 0.5 * rho * q2)
    T = p / rho / R_GC
    a = np.sqrt(GAMMA * p / rho)
    M = np.sqrt(q2) / a
    P0 = p * (1.0 + 0.5 * (GAMMA - 1.0) * M * M) ** (GAMMA / (GAMMA - 1.0))
    return dict(rho=rho, u=u, v=v, p=p, T=T, a=a, M=M, P0=P0)


def write_solution(path: str, U: np.ndarray, dt: np.ndarray, iterations: int, header=(0, 0, 0)):
    """Write_Solution(file, 1): cell ids are the caller's (input-file) order."""
    w = primitives(U)
    with open(path, "w") as f:
        f.write("\t".join(str(int(h)) for h in header) + "\n")
        f.write(f"{len(U)}\n{int(iterations)}\n")
        for i in range(len(U)):
            f.write(f"{i}\t\t{_g(dt[i])}\t{_g(w['rho'][i])}\t{_g(w['p'][i])}\t{_g(w['T'][i])}\t\t\t\t"
                    f"{_g(w['u'][i])}\t\t\t\t{_g(w['v'][i])}\t\t\t\t{_g(w['M'][i])}\t\t\t\t{_g(w['P0'][i])}\n")


def read_solution(path: str, header_len: int = 3):
    """Reads a Write_Solution type-1 file; returns (U [N,4] rebuilt from rho, P, u, v as the
    restart does, dt [N], iterations, header)."""
    tok = open(path).read().split()
    header = [int(t) for t in tok[:header_len]]
    n = int(tok[header_len])
    it = int(tok[header_len + 1])
    vals = np.array(tok[header_len + 2: header_len + 2 + 9 * n], dtype=np.float64).reshape(n, 9)
    dt, rho, p, u, v = vals[:, 1], vals[:, 2], vals[:, 3], vals[:, 5], vals[:, 6]
    U = np.column_stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)])
    return U, dt, it, header


def append_error(path: str, err4):
    with open(path, "a") as f:
        f.write("".join(f"{_g(e)}\t" for e in err4) + "\n")


def write_vtk_solution(path: str, points: np.ndarray, cells: np.ndarray, U: np.ndarray):
    """Legacy VTK unstructured grid with the reference's cell-data scalars and velocity vectors."""
    pts = np.asarray(points, np.float64)
    cells = np.asarray(cells)
    nn = np.where(cells[:, 3] < 0, 3, 4)
    w = primitives(U)
    n = len(cells)
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nsolution\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS\t{len(pts)}\tdouble\n")
        for x, y, z in pts:
            f.write(f"{_g(x)}\t{_g(y)}\t{_g(z)}\n")
        f.write(f"CELLS\t{n}\t{int((nn + 1).sum())}\n")
        for row, k in zip(cells, nn):
            f.write("\t".join(str(int(x)) for x in [k] + list(row[:k])) + "\n")
        f.write(f"CELL_TYPES\t{n}\n")
        f.write("\n".join("5" if k == 3 else "9" for k in nn) + "\n")
        f.write(f"CELL_DATA\t{n}\n")
        for name, vals in (("Pressure", w["p"]), ("Temperature", w["T"]), ("Density", w["rho"]),
                           ("U_Velocity", w["u"]), ("V_Velocity", w["v"]), ("Total_Pressure", w["P0"]),
                           ("Entropy", np.log(w["p"] / w["rho"] ** GAMMA)), ("Mach_Number", w["M"])):
            f.write(f"SCALARS\t {name}\t double\nLOOKUP_TABLE default\n")
            f.write("\n".join(_g(x) for x in vals) + "\n")
        f.write("Vectors\t Velocity\t double \n")
        for a, b in zip(w["u"], w["v"]):
            f.write(f"{_g(a)}\t{_g(b)}\t0\n")


# The reference's Dissipation_Type (src/Net_Flux.cpp:74-91): 1 LLF, 3 ROE are on this path; 101
# HLLC and 102 AUSM are the values libcfdhost's driver accepts for the other two fluxes.
_DISSIPATION = {1: "llf", 3: "roe", 101: "hllc", 102: "ausm"}


def config_from_json(path: str) -> dict:
    """make_config() keyword arguments from the reference's solver JSON (comments allowed, as
    the reference's jsoncpp reader allows them), with the reference's meaning of each key
    (src/Configuration_Read.cpp:23-87, src/Solver.cpp:93-118,190-205):

    - Simulation.CFL; Solver.Local_Time_Stepping;
    - Solver.Dissipation_Type: 1 LLF, 3 ROE (101 HLLC and 102 AUSM are this library's values);
    - Solver.Time_Accurate: Runge_Kutta_Method (the reference's TVD-RK3 in delta form, rk 3) when
      true, else Explicit_Method (forward Euler, rk 1), both as the reference's iteration
      (ref_mode 1);
    - Solver.Is_Second_Order: Evaluate_Cell_Net_Flux_2O; with ROE that is ROE_2O (Roe dissipation
      on MinMod-limited differences) with LimiterCoefficients.Limiter_Zeta / Limiter_Zeta1; other
      fluxes use MUSCL with minmod and Green-Gauss gradients;
    - Solver.Is_WENO: WENO5 faces (order 3) with the Rusanov (LLF) flux whatever Dissipation_Type,
      as the reference's Calculate_Face_WENO_Flux (src/WENO2D.cpp:458-555);
    - Solver.Is_Viscous: laminar terms with Sutherland's law (Pr 0.72, R 287) at order 2 or 3; they
      exist in the native iteration only (ref_mode 0, TVD-RK3); the caller gives the walls ghost
      type "noslip_wall", the reference's Is_Viscous_Wall.

    A key the file does not hold reads as false or 0, as jsoncpp's asBool() / asDouble() return
    for a missing member; a CFL that is missing or not positive is rejected."""
    text = open(path).read()
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    d = json.loads(text)
    sim, sol, lim = d.get("Simulation", {}), d.get("Solver", {}), d.get("LimiterCoefficients", {})
    diss = int(sol.get("Dissipation_Type", 0))
    if diss not in _DISSIPATION:
        raise ValueError(f"Dissipation_Type {diss} is not on this path (1 LLF or 3 ROE)")
    cfl = float(sim.get("CFL", 0.0))
    if not cfl > 0.0:
        raise ValueError("Simulation.CFL must be positive")
    order = 3 if sol.get("Is_WENO", False) else (2 if sol.get("Is_Second_Order", False) else 1)
    kw = dict(flux="llf" if order == 3 else _DISSIPATION[diss], order=order, cfl=cfl,
              lts=1 if sol.get("Local_Time_Stepping", False) else 0,
              rk=3 if sol.get("Time_Accurate", False) else 1, ref_mode=1)
    if order >= 2:
        kw.update(limiter="minmod", gradient="gg")
    if order == 2 and kw["flux"] == "roe":
        kw.update(zeta=float(lim.get("Limiter_Zeta", 0.0)), zeta1=float(lim.get("Limiter_Zeta1", 0.0)))
    if sol.get("Is_Viscous", False):
        if order < 2:
            raise ValueError("Is_Viscous needs Is_Second_Order or Is_WENO (cell gradients)")
        kw.update(viscous="sutherland", prandtl=0.72, r_gas=287.0, ref_mode=0, rk=3)
        kw.pop("zeta", None)
        kw.pop("zeta1", None)
    return kw


__all__ = ["primitives", "write_solution", "read_solution", "append_error", "write_vtk_solution",
           "config_from_json"]
