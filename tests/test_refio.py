"""Reference-compatible files (cfd_solver_withcuda_amd/refio.py) against the reference's own
writers: Write_Solution type 1 and Write_Error_File outputs on the Sod tube
(tests/golden/sod_solution_ref.txt, sod_error_ref.txt), the restart round trip, the VTK solution
layout and the JSON configuration mapping. CPU only."""
import os

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen, refio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sod():
    return np.load(os.path.join(GOLD, "sod_roe.npz"), allow_pickle=False)


def test_write_solution_matches_reference_text(tmp_path):
    g = sod()
    p = tmp_path / "sol.txt"
    refio.write_solution(str(p), g["U0"], g["dt_roe"], 0, (101, 11, 1))
    assert p.read_text() == open(os.path.join(GOLD, "sod_solution_ref.txt")).read()


def test_error_rows_match_reference_text(tmp_path):
    g = sod()
    p = tmp_path / "err.txt"
    for k in range(3):
        refio.append_error(str(p), g["err_hist"][k])
    assert p.read_text() == open(os.path.join(GOLD, "sod_error_ref.txt")).read()


def test_restart_round_trip(tmp_path):
    g = sod()
    p = tmp_path / "sol.txt"
    U = g["U_step25"]
    refio.write_solution(str(p), U, g["dt_roe"], 25, (101, 11, 1))
    U2, dt, it, hdr = refio.read_solution(str(p))
    assert it == 25 and hdr == [101, 11, 1] and U2.shape == U.shape
    assert np.allclose(U2, U, rtol=2e-5, atol=1e-6)


def test_vtk_solution_layout(tmp_path):
    m = meshgen.bump_mesh(12, 6, 0.1, seed=2)
    U = np.tile([1.4, 1.4 * 1.65, 0.0, 1.0 / 0.4 + 0.5 * 1.4 * 1.65 ** 2], (len(m.cells), 1))
    p = tmp_path / "sol.vtk"
    refio.write_vtk_solution(str(p), m.points, m.cells, U)
    text = p.read_text()
    back = meshgen.read_vtk(str(p))
    assert np.array_equal(back.cells, m.cells)
    for name in ("Pressure", "Temperature", "Density", "U_Velocity", "V_Velocity", "Total_Pressure",
                 "Entropy", "Mach_Number"):
        assert f"SCALARS\t {name}\t double" in text
    assert f"CELL_DATA\t{len(m.cells)}" in text and "Vectors\t Velocity" in text


def test_config_from_reference_json(tmp_path):
    """Keys keep the reference's meaning (src/Configuration_Read.cpp:23-87, src/Solver.cpp:93-118)."""
    p = tmp_path / "cfg.json"
    p.write_text('{ "Simulation": {"CFL": 0.4, // comment\n "Total_Iterations": 10},'
                 ' "Solver": {"Dissipation_Type": 3, "Is_Second_Order": false, "Local_Time_Stepping": true,'
                 ' "Time_Accurate": true} }')
    assert refio.config_from_json(str(p)) == dict(flux="roe", order=1, cfl=0.4, lts=1, rk=3, ref_mode=1)
    # Time_Accurate false (or missing) is Explicit_Method: forward Euler
    p.write_text('{ "Simulation": {"CFL": 0.4}, "Solver": {"Dissipation_Type": 3} }')
    assert refio.config_from_json(str(p)) == dict(flux="roe", order=1, cfl=0.4, lts=0, rk=1, ref_mode=1)
    # Is_Second_Order with ROE is ROE_2O with the limiter coefficients
    p.write_text('{ "Simulation": {"CFL": 0.5}, "Solver": {"Dissipation_Type": 3, "Is_Second_Order": true,'
                 ' "Time_Accurate": true}, "LimiterCoefficients": {"Limiter_Zeta": 0.5, "Limiter_Zeta1": 0.25} }')
    assert refio.config_from_json(str(p)) == dict(flux="roe", order=2, cfl=0.5, lts=0, rk=3, ref_mode=1,
                                                  limiter="minmod", gradient="gg", zeta=0.5, zeta1=0.25)
    # WENO faces use the Rusanov flux whatever Dissipation_Type
    p.write_text('{ "Simulation": {"CFL": 0.6}, "Solver": {"Dissipation_Type": 3, "Is_Second_Order": true,'
                 ' "Is_WENO": true, "Time_Accurate": true} }')
    assert refio.config_from_json(str(p)) == dict(flux="llf", order=3, cfl=0.6, lts=0, rk=3, ref_mode=1,
                                                  limiter="minmod", gradient="gg")
    # viscous terms run in the native iteration, at order 2 or 3
    p.write_text('{ "Simulation": {"CFL": 0.5}, "Solver": {"Dissipation_Type": 101, "Is_Second_Order": true,'
                 ' "Is_Viscous": true} }')
    assert refio.config_from_json(str(p)) == dict(flux="hllc", order=2, cfl=0.5, lts=0, rk=3, ref_mode=0,
                                                  limiter="minmod", gradient="gg", viscous="sutherland",
                                                  prandtl=0.72, r_gas=287.0)
    p.write_text('{ "Simulation": {"CFL": 0.5}, "Solver": {"Dissipation_Type": 3, "Is_WENO": true,'
                 ' "Is_Viscous": true} }')
    assert refio.config_from_json(str(p)) == dict(flux="llf", order=3, cfl=0.5, lts=0, rk=3, ref_mode=0,
                                                  limiter="minmod", gradient="gg", viscous="sutherland",
                                                  prandtl=0.72, r_gas=287.0)
    for bad in ('{ "Simulation": {"CFL": 0.5}, "Solver": {"Dissipation_Type": 3, "Is_Viscous": true} }',
                '{ "Solver": {"Dissipation_Type": 3} }',
                '{ "Simulation": {"CFL": 0.5}, "Solver": {"Dissipation_Type": 2} }'):
        p.write_text(bad)
        with pytest.raises(ValueError):
            refio.config_from_json(str(p))
