"""Pins the CPU oracle to the reference solver's golden vectors (tests/golden/*.npz).

The fixtures were written by tests/golden/make_golden.py from the reference build; these
tests need neither the reference nor a GPU.
"""
import os

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from oracle.oracle import Oracle, cons_from_prim, make_cfg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GEN = {
    "sod_40x4": lambda: meshgen.shock_tube_mesh(nx=40, ny=4),
    "sod_100x10": lambda: meshgen.shock_tube_mesh(nx=100, ny=10),
    "bump_40x20": lambda: meshgen.bump_mesh(nx=40, ny=20, thickness=0.1, seed=7),
    "ffs_h20": lambda: meshgen.ffs_mesh(h=1.0 / 20, seed=3),
    "ramp_30x10": lambda: meshgen.ramp_mesh(nx=30, ny=10, angle_deg=15.0, seed=4),
    "Ramp_15o_52_18": lambda: meshgen.read_vtk(os.path.join(GOLD, "Ramp_15o_52_18.vtk")),
    "Test_Domain_Triangle": lambda: meshgen.read_vtk(os.path.join(GOLD, "Test_Domain_Triangle.vtk")),
}


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def gold(name):
    return np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False)


@pytest.mark.parametrize("name", sorted(GEN))
def test_connectivity_bit_exact(orc, name):
    """Read_VTK_Grid + Form_Cells arrays reproduced bit for bit (src/Read_Gmsh_File.cpp:11-621)."""
    g = gold(f"mesh_{name}")
    m = GEN[name]()
    a = orc.mesh(m.points, m.cells).arrays()
    msk = np.arange(4)[None, :] < g["nf"][:, None]
    assert np.array_equal(a.nf, g["nf"])
    for f in ("nodes", "nbr", "kind", "length", "dist"):
        assert np.array_equal(getattr(a, f)[msk], g[f][msk]), f
    assert np.array_equal(a.nrm[np.repeat(msk, 2, axis=1)], g["nrm"][np.repeat(msk, 2, axis=1)])
    for f in ("ctr", "area", "inv_area", "inlet", "exit", "wall"):
        assert np.array_equal(getattr(a, f), g[f]), f


def test_ramp_ghost_count_is_one_per_boundary_face(orc):
    """The reference's own 52x18 ramp grid: 2*(51+17) boundary faces -> 132 ghosts."""
    m = GEN["Ramp_15o_52_18"]()
    om = orc.mesh(m.points, m.cells)
    assert om.counts == (800, 132, 16, 16, 100)


def test_prim_bit_exact(orc):
    g = gold("prim")
    for U, P in zip(g["U"], g["prim"]):
        assert np.array_equal(orc.prim(U), P)


def _sod(orc):
    m = GEN["sod_100x10"]()
    return orc.mesh(m.points, m.cells)


def test_sod_residuals(orc):
    g = gold("sod_roe")
    om = _sod(orc)
    for tag, flux in (("roe", "roe"), ("llf", "llf")):
        R, dt, _ = om.residual(make_cfg(flux=flux, order=1, cfl=0.5, fs=(1.0, 0, 0, 1.0)), g["U0"])
        Rg = g[f"R_{tag}"]
        assert np.abs(R - Rg).max() <= 1e-13 * np.abs(Rg).max(), tag
        assert np.array_equal(dt, g[f"dt_{tag}"])


def test_sod_rk3_steps(orc):
    """25 iterations of Inviscid_Solver (TVD-RK3 delta form, Roe-1O, CFL 0.5)."""
    g = gold("sod_roe")
    om = _sod(orc)
    cfg = make_cfg(flux="roe", order=1, cfl=0.5, fs=(1.0, 0, 0, 1.0))
    U = np.ascontiguousarray(g["U0"].copy())
    t = 0.0
    for s in range(25):
        err, dt = om.step(cfg, U, rk=3, lts=0, ref=True)
        t += dt
        assert np.allclose(err, g["err_hist"][s], rtol=1e-10, atol=1e-13), s
        assert abs(dt - g["dt_hist"][s]) <= 1e-13 * dt
        if s == 4:
            assert np.abs(U - g["U_step5"]).max() <= 1e-12 * np.abs(U).max()
    assert np.abs(U - g["U_step25"]).max() <= 1e-12 * np.abs(U).max()
    assert abs(t - g["time_hist"][-1]) <= 1e-12 * t


def _bump(orc):
    m = GEN["bump_40x20"]()
    return orc.mesh(m.points, m.cells)


def test_bump_ghost_bcs_bit_exact(orc):
    """Apply_Boundary_Conditions: supersonic inlet, supersonic exit, inviscid wall."""
    g = gold("bump_mixed")
    om = _bump(orc)
    nc = om.nphys
    U = g["U_all_bc"].copy()
    P = g["prim_all_bc"].copy()
    U[nc:] = 0.0
    P[nc:] = 0.0
    om.apply_bc_ref(U, P, 1.4, 1.0)
    assert np.array_equal(U[nc:], g["U_all_bc"][nc:])
    assert np.array_equal(P[nc:], g["prim_all_bc"][nc:])


def test_bump_residuals(orc):
    g = gold("bump_mixed")
    om = _bump(orc)
    fs = (1.4, 0.0, 0.0, 1.0)
    R, dt, _ = om.residual(make_cfg(flux="roe", order=1, cfl=0.5, fs=fs), g["U0"])
    assert np.abs(R - g["R_roe1"]).max() <= 1e-13 * np.abs(g["R_roe1"]).max()
    assert np.array_equal(dt, g["dt"])
    R, _, _ = om.residual(make_cfg(flux="llf", order=1, cfl=0.5, fs=fs), g["U0"])
    assert np.abs(R - g["R_llf1"]).max() <= 1e-13 * np.abs(g["R_llf1"]).max()
    R, dt2 = om.roe2o_ref(g["U_all_bc"], g["prim_all_bc"], 0.5, 0.5, 0.5)
    assert np.array_equal(R, g["R_roe2o"])
    assert np.array_equal(om.ausm_ref(g["U_all_bc"]), g["R_ausm"])
    for i, var in enumerate((0, 1, 2, 4)):
        assert np.array_equal(om.gg_gradient_ref(g["prim_all_bc"], var), g["gg"][:, i])


def test_bump_green_gauss_face_loop_matches_reference(orc):
    """The face-loop GG gradient (ghost faces use the cell value) equals the reference's."""
    g = gold("bump_mixed")
    om = _bump(orc)
    nc = om.nphys
    P = g["prim_all_bc"][:nc]
    W5 = np.ascontiguousarray(np.stack([P[:, 0], P[:, 1], P[:, 2], P[:, 4], P[:, 5]], axis=1))
    grad = om.gradients(make_cfg(gradient="gg"), W5)
    for i in range(4):
        assert np.array_equal(grad[:, 2 * i:2 * i + 2], g["gg"][:, i])


@pytest.mark.parametrize("tag,rk,lts", [("lts", 3, 1), ("euler", 1, 0)])
def test_bump_steps(orc, tag, rk, lts):
    g = gold("bump_mixed")
    om = _bump(orc)
    cfg = make_cfg(flux="roe", order=1, cfl=0.5, fs=(1.4, 0, 0, 1.0))
    U = np.ascontiguousarray(g["U0"].copy())
    for s in range(10):
        err, _ = om.step(cfg, U, rk=rk, lts=lts, ref=True)
        assert np.allclose(err, g[f"err_{tag}"][s], rtol=1e-10, atol=1e-13)
    assert np.abs(U - g[f"U_{tag}"]).max() <= 1e-12 * np.abs(U).max()


def test_cons_from_prim_matches_reference_expression():
    U = cons_from_prim(np.array([1.0]), np.array([0.3]), np.array([-0.2]), np.array([2.0]))
    assert U[0, 3] == (2.0 / (1.4 - 1.0)) + 1.0 * (0.5 * (0.3 * 0.3 + -0.2 * -0.2))
