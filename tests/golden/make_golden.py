"""Writes the golden fixtures of tests/golden from the reference solver.

Runs only where the reference build exists (`make -f oracle/Makefile.ref`, which compiles the
reference's src/*.cpp from /root/reference). Every fixture is data: the inputs (mesh generator
parameters or a mesh file the reference ships, initial states) and the reference's outputs.

    python tests/golden/make_golden.py

Fixtures:
  mesh_<name>.npz   per-cell connectivity/geometry arrays of Read_VTK_Grid + Form_Cells
  prim.npz          Calculate_Primitive_Variables on edge-case states (velocity flush, etc.)
  sod_roe.npz       Sod tube 100x10: Roe-1O and LLF-1O residuals at t=0, 25 TVD-RK3 steps
  bump_mixed.npz    jittered mixed bump mesh: ghost BC arrays, Roe-1O/LLF-1O/Roe-2O/AUSM
                    residuals, Green-Gauss gradients, LTS RK3 and forward-Euler steps
  sod_solution_ref.txt, sod_error_ref.txt
                    the reference's Write_Solution (type 1) of the Sod initial state after one
                    Roe-1O residual, and its Write_Error_File rows of 3 TVD-RK3 steps

    python tests/golden/make_golden.py io    # only the two text files
    python tests/golden/make_golden.py weno  # only weno_ref.npz

  weno_ref.npz      WENO_Reconstruction (src/WENO2D.cpp:9-79) of 1203 five-cell stencils
                    (smooth, stepped, constant, random), both shifts
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cfd_solver_withcuda_amd import meshgen  # noqa: E402
from oracle.oracle import Reference, cons_from_prim  # noqa: E402

# generator parameters of the generated meshes (recreated by the tests)
GEN_MESHES = {
    "sod_40x4": ("shock_tube_mesh", dict(nx=40, ny=4)),
    "sod_100x10": ("shock_tube_mesh", dict(nx=100, ny=10)),
    "bump_40x20": ("bump_mesh", dict(nx=40, ny=20, thickness=0.1, seed=7)),
    "ffs_h20": ("ffs_mesh", dict(h=1.0 / 20, seed=3)),
    "ramp_30x10": ("ramp_mesh", dict(nx=30, ny=10, angle_deg=15.0, seed=4)),
}
FILE_MESHES = ["Ramp_15o_52_18.vtk", "Test_Domain_Triangle.vtk"]

MESH_FIELDS = ["nf", "nodes", "nbr", "nrm", "length", "ctr", "area", "inv_area", "dist", "kind",
               "inlet", "exit", "wall"]


def gen_mesh(name):
    fn, kw = GEN_MESHES[name]
    return getattr(meshgen, fn)(**kw)


def mesh_arrays_dict(a):
    return {f: getattr(a, f) for f in MESH_FIELDS}


def load(ref, mesh):
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m.vtk")
        meshgen.write_vtk(p, mesh)
        ref.load_vtk(p)


def smooth_state(ctr, nc, rho0=1.4, u0=1.65, p0=1.0):
    x, y = ctr[:nc, 0], ctr[:nc, 1]
    rho = rho0 * (1 + 0.1 * np.sin(3 * x) * np.cos(2 * y))
    u = u0 * (1 + 0.05 * np.cos(4 * y))
    v = 0.1 * np.sin(2 * x)
    p = p0 + 0.1 * np.cos(5 * x + y)
    return np.ascontiguousarray(cons_from_prim(rho, u, v, p))


def io_files(ref):
    m = gen_mesh("sod_100x10")
    load(ref, m)
    a = ref.arrays()
    nc = a.nphys
    xc = a.ctr[:nc, 0]
    left = xc < 0.5
    U0 = np.ascontiguousarray(cons_from_prim(np.where(left, 1.0, 0.125), 0 * xc, 0 * xc,
                                             np.where(left, 1.0, 0.1)))
    ref.configure(dissipation=3, second_order=0, time_accurate=1, lts=0, cfl=0.5, rho_in=1.0, p_in=1.0)
    ref.init_state(U0)
    ref.apply_bc()
    ref.residual()
    ref.write_solution(os.path.join(HERE, "sod_solution_ref.txt"), 0, (101, 11, 1))
    err_path = os.path.join(HERE, "sod_error_ref.txt")
    if os.path.exists(err_path):
        os.remove(err_path)
    ref.init_state(U0)
    for _ in range(3):
        ref.step()
        ref.write_error(err_path)


def weno_vectors(ref):
    """WENO_Reconstruction (src/WENO2D.cpp:9-79) on smooth, stepped, constant and random stencils,
    both shifts: weno_ref.npz (inputs [n, 5], outputs [n, 2])."""
    rng = np.random.default_rng(2024)
    x = np.arange(5, dtype=np.float64)
    rows = [np.full(5, 1.3), np.array([1.0, 1.0, 1.0, 0.125, 0.125]), np.array([0.1, 0.1, 2.5, 2.5, 2.5])]
    for _ in range(400):
        c = rng.normal(size=4) * 10 ** rng.uniform(-4, 2)
        rows.append(c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3)          # smooth
        k = rng.integers(1, 5)
        rows.append(np.where(np.arange(5) < k, rng.uniform(0.1, 3), rng.uniform(0.1, 3)))  # step
        rows.append(rng.normal(size=5) * 10 ** rng.uniform(-6, 4))             # random
    inp = np.array(rows)
    out = np.array([[ref.weno(*r, 0), ref.weno(*r, 1)] for r in inp])
    np.savez_compressed(os.path.join(HERE, "weno_ref.npz"), stencil=inp, value=out)
    print("wrote weno_ref", inp.shape)


def main():
    ref = Reference()
    if sys.argv[1:] == ["io"]:
        io_files(ref)
        return
    if sys.argv[1:] == ["weno"]:
        weno_vectors(ref)
        return
    out = {}
    for name in GEN_MESHES:
        load(ref, gen_mesh(name))
        out[f"mesh_{name}"] = mesh_arrays_dict(ref.arrays())
    for fname in FILE_MESHES:
        ref.load_vtk(os.path.join(HERE, fname))
        out[f"mesh_{fname[:-4]}"] = mesh_arrays_dict(ref.arrays())

    # primitive recovery on edge cases
    states = np.array([[1.0, 0.0, 0.0, 2.5], [1.3, 1e-11, -0.7, 2.9], [0.125, -1e-10, 1e-10, 0.25],
                       [2.0, 3.0, -4.0, 30.0], [1.4, 2.31, 0.0, 4.4057], [0.7, 5e-11, 2e-10, 1.1]])
    prims = np.array([ref.prim(s) for s in states])
    out["prim"] = dict(U=states, prim=prims)

    # Sod shock tube, 100x10 quads
    m = gen_mesh("sod_100x10")
    load(ref, m)
    a = ref.arrays()
    nc = a.nphys
    xc = a.ctr[:nc, 0]
    left = xc < 0.5
    U0 = np.ascontiguousarray(cons_from_prim(np.where(left, 1.0, 0.125), 0 * xc, 0 * xc,
                                             np.where(left, 1.0, 0.1)))
    sod = dict(U0=U0)
    for diss, tag in [(3, "roe"), (1, "llf")]:
        ref.configure(dissipation=diss, second_order=0, time_accurate=1, lts=0, cfl=0.5,
                      rho_in=1.0, p_in=1.0)
        ref.init_state(U0)
        ref.apply_bc()
        R, dt = ref.residual()
        sod[f"R_{tag}"] = R
        sod[f"dt_{tag}"] = dt
    ref.configure(dissipation=3, second_order=0, time_accurate=1, lts=0, cfl=0.5, rho_in=1.0, p_in=1.0)
    ref.init_state(U0)
    errs, dts, times = [], [], []
    for s in range(25):
        e, mdt, tt = ref.step()
        errs.append(e)
        dts.append(mdt)
        times.append(tt)
        if s == 4:
            sod["U_step5"] = ref.state()[0][:nc].copy()
    sod["err_hist"] = np.array(errs)
    sod["dt_hist"] = np.array(dts)
    sod["time_hist"] = np.array(times)
    sod["U_step25"] = ref.state()[0][:nc].copy()
    out["sod_roe"] = sod

    # jittered mixed bump mesh
    m = gen_mesh("bump_40x20")
    load(ref, m)
    a = ref.arrays()
    nc = a.nphys
    U0 = smooth_state(a.ctr, nc)
    bump = dict(U0=U0)
    ref.configure(dissipation=3, second_order=0, time_accurate=1, cfl=0.5, rho_in=1.4, p_in=1.0)
    ref.init_state(U0)
    ref.apply_bc()
    Ua, Pa = ref.state()
    bump["U_all_bc"] = Ua
    bump["prim_all_bc"] = Pa
    bump["R_roe1"], bump["dt"] = ref.residual()
    ref.configure(dissipation=1, second_order=0, time_accurate=1, cfl=0.5, rho_in=1.4, p_in=1.0)
    bump["R_llf1"], _ = ref.residual()
    ref.configure(dissipation=3, second_order=1, time_accurate=1, cfl=0.5, zeta=0.5, zeta1=0.5,
                  rho_in=1.4, p_in=1.0)
    bump["R_roe2o"], _ = ref.residual()
    bump["R_ausm"] = ref.ausm()
    bump["gg"] = np.stack([ref.gg_gradient(v) for v in (0, 1, 2, 4)], axis=1)
    for tag, kw, n in [("lts", dict(time_accurate=1, lts=1), 10), ("euler", dict(time_accurate=0, lts=0), 10)]:
        ref.configure(dissipation=3, second_order=0, cfl=0.5, rho_in=1.4, p_in=1.0, **kw)
        ref.init_state(U0)
        errs = []
        for _ in range(n):
            errs.append(ref.step()[0])
        bump[f"err_{tag}"] = np.array(errs)
        bump[f"U_{tag}"] = ref.state()[0][:nc].copy()
    out["bump_mixed"] = bump

    for key, d in out.items():
        np.savez_compressed(os.path.join(HERE, f"{key}.npz"), **d)
        print("wrote", key)


if __name__ == "__main__":
    main()
