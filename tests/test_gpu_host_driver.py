"""GPU test of the reference-named C++ host driver (include/cfd_host.hpp, libcfdhost.so): a C++
program (tests/host/sod_driver.cpp) runs the reference's Inviscid_Solver loop -- Read_VTK_Grid,
Apply_Boundary_Conditions, Runge_Kutta_Method, Estimate_Error, Update -- on the Sod tube and
must reproduce the reference's own 25-step outputs (tests/golden/sod_roe.npz)."""
import os
import subprocess

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "host", "_build", "sod_driver")


def test_host_driver_sod_vs_reference_golden(tmp_path):
    assert os.path.exists(EXE), "build first: python -m cfd_solver_withcuda_amd.build"
    g = np.load(os.path.join(ROOT, "tests", "golden", "sod_roe.npz"), allow_pickle=False)
    m = meshgen.shock_tube_mesh(100, 10)
    vtk = tmp_path / "sod.vtk"
    meshgen.write_vtk(str(vtk), m)
    u0 = tmp_path / "U0.bin"
    np.ascontiguousarray(g["U0"], np.float64).tofile(u0)
    out = tmp_path / "out.bin"
    sol, errf = tmp_path / "sol.txt", tmp_path / "err.txt"
    r = subprocess.run([EXE, str(vtk), str(u0), "25", str(out), str(sol), str(errf)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    n = len(m.cells)
    d = np.fromfile(out, np.float64)
    U = d[:4 * n].reshape(n, 4)
    hist = d[4 * n:4 * n + 25 * 5].reshape(25, 5)
    R0 = d[4 * n + 125:8 * n + 125].reshape(n, 4)
    dt0 = d[8 * n + 125:]
    assert np.allclose(hist[:, :4], g["err_hist"], rtol=1e-10, atol=1e-13)
    assert np.allclose(hist[:, 4], g["dt_hist"], rtol=1e-13, atol=0)
    assert np.abs(U - g["U_step25"]).max() <= 1e-12 * np.abs(g["U_step25"]).max()
    # Evaluate_Cell_Net_Flux_1O of the initial state against the reference's net flux and del_t
    assert np.allclose(R0, g["R_roe"], rtol=0, atol=1e-13 * np.abs(g["R_roe"]).max())
    assert np.allclose(dt0, g["dt_roe"], rtol=1e-14, atol=0)
    # Write_Solution / Write_Error_File byte for byte as the reference writes them
    assert sol.read_text() == open(os.path.join(ROOT, "tests", "golden", "sod_solution_ref.txt")).read()
    assert errf.read_text() == open(os.path.join(ROOT, "tests", "golden", "sod_error_ref.txt")).read()


def test_host_driver_weno_vs_oracle(tmp_path):
    """Is_WENO set: Evaluate_Cell_Net_Flux_WENO of the Sod state and 10 reference iterations
    (Runge_Kutta_Method, Rusanov flux on WENO5 states, MUSCL/Van Leer + WLSQ on the faces without a
    quadrilateral line) equal the oracle's bit for bit."""
    from oracle.oracle import Oracle, make_cfg

    assert os.path.exists(EXE), "build first: python -m cfd_solver_withcuda_amd.build"
    g = np.load(os.path.join(ROOT, "tests", "golden", "sod_roe.npz"), allow_pickle=False)
    m = meshgen.shock_tube_mesh(100, 10)
    vtk = tmp_path / "sod.vtk"
    meshgen.write_vtk(str(vtk), m)
    u0 = tmp_path / "U0.bin"
    U0 = np.ascontiguousarray(g["U0"], np.float64)
    U0.tofile(u0)
    out = tmp_path / "out.bin"
    nsteps = 10
    r = subprocess.run([EXE, str(vtk), str(u0), str(nsteps), str(out), "weno"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    n = len(m.cells)
    d = np.fromfile(out, np.float64)
    U = d[:4 * n].reshape(n, 4)
    R0 = d[4 * n + 5 * nsteps:8 * n + 5 * nsteps].reshape(n, 4)
    dt0 = d[8 * n + 5 * nsteps:]
    om = Oracle().mesh(m.points, m.cells)
    cfg = make_cfg(flux="llf", order=3, limiter="vanleer", gradient="wlsq", roe_bdiss=0, cfl=0.5,
                   bc=("inlet_ref", "exit", "wall", "wall", "wall"), fs=(1.0, 0.0, 0.0, 1.0))
    Ro, dto, _ = om.residual(cfg, U0)
    assert np.array_equal(R0, Ro) and np.array_equal(dt0, dto)
    Uo = U0.copy()
    for _ in range(nsteps):
        om.step(cfg, Uo, rk=3, lts=0, ref=True)
    assert np.array_equal(U, Uo), float(np.abs(U - Uo).max())


PLUGIN_EXE = os.path.join(ROOT, "tests", "host", "_build", "plugin_driver")


def test_dissipation_plugin_point_sweeps(tmp_path):
    """The reference's per-cell loop over Calculate_Flux_For_All_Faces(cell, fn) and
    Evaluate_Time_Step(cell) (src/Net_Flux.cpp:65-94) through the host driver's plugin point:
    every scheme's sweep (ROE, LLF, HLLC, AUSM, ROE_2O, a registered caller function) equals the bulk
    Evaluate_Cell_Net_Flux_1O/2O bit for bit, a state change between sweeps is seen, and unknown
    functions, direct stand-in calls and bad cells throw (checked inside the program). The ROE sweep
    of the Sod state equals the reference's own net flux and del_t."""
    assert os.path.exists(PLUGIN_EXE), "build first: python -m cfd_solver_withcuda_amd.build"
    g = np.load(os.path.join(ROOT, "tests", "golden", "sod_roe.npz"), allow_pickle=False)
    m = meshgen.shock_tube_mesh(100, 10)
    vtk = tmp_path / "sod.vtk"
    meshgen.write_vtk(str(vtk), m)
    u0 = tmp_path / "U0.bin"
    np.ascontiguousarray(g["U0"], np.float64).tofile(u0)
    out = tmp_path / "out.bin"
    r = subprocess.run([PLUGIN_EXE, str(vtk), str(u0), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "0 failures" in r.stdout
    n = len(m.cells)
    d = np.fromfile(out, np.float64)
    R, dt = d[:4 * n].reshape(n, 4), d[4 * n:]
    assert np.allclose(R, g["R_roe"], rtol=0, atol=1e-13 * np.abs(g["R_roe"]).max())
    assert np.allclose(dt, g["dt_roe"], rtol=1e-14, atol=0)
