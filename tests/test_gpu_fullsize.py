"""Parity at the benchmark configurations' full sizes: the 1400 x 700 jittered mixed bump mesh
(1.27M cells) and the h = 1/1560 forward-facing-step mesh (7.97M cells). The oracle evaluates a
whole residual of either in seconds, so the device is checked against it bit for bit at full
size, not only through size-independent properties: residual, local time step and gradients,
one full time step of each configuration's scheme, and independence of the tiling."""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config
from oracle.oracle import Oracle, make_cfg

pytestmark = pytest.mark.gpu

BCS = ("freestream", "exit", "wall", "wall", "wall")


def smooth_state(om, mach):
    """A non-uniform admissible state, so every flux branch and limiter sees real data."""
    c = om.arrays().ctr[: om.nphys]
    x, y = c[:, 0], c[:, 1]
    return np.ascontiguousarray(cons_from_prim(1.4 * (1 + 0.08 * np.sin(3 * x) * np.cos(2 * y)),
                                               mach * (1 + 0.05 * np.cos(4 * y)), 0.1 * np.sin(2 * x),
                                               1.0 + 0.1 * np.cos(5 * x + y)))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def test_bump_full_size_bit_exact(orc):
    """Config 1's mesh and scheme: HLLC + MUSCL/Van Leer + WLSQ, local time stepping, TVD-RK3."""
    m = meshgen.bump_mesh(1400, 700, 0.10, seed=1)
    om = orc.mesh(m.points, m.cells)
    U = smooth_state(om, 1.65)
    kw = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, bc=BCS, fs=freestream(1.65))
    ocfg = make_cfg(roe_bdiss=1, **kw)
    Ro, dto, go = om.residual(ocfg, U)
    print(f"case bump 1400x700 full size ({len(m.cells)} cells): residual", flush=True)
    s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(roe_bdiss=1, rk=3, lts=1, **kw))
    s.set_state(U)
    R, dt, g = s.residual()
    assert np.array_equal(R, Ro)
    assert np.array_equal(dt, dto)
    assert np.array_equal(g, go)
    print("case bump full size: LTS TVD-RK3 step", flush=True)
    Uo = U.copy()
    om.step(ocfg, Uo, rk=3, lts=1)
    s.step(1)
    assert np.array_equal(s.get_state(), Uo)
    print("case bump full size: tiling 64 / 128 faces budget", flush=True)
    s2 = Solver(Mesh(m.points, m.cells, tile=64, face_budget=128), make_config(roe_bdiss=1, rk=3, lts=1, block=64, **kw))
    s2.set_state(U)
    R2, dt2, _ = s2.residual()
    assert np.array_equal(R2, Ro) and np.array_equal(dt2, dto)


def test_step_mesh_full_size_bit_exact(orc):
    """Config 2's mesh and scheme: second-order Roe (MUSCL with minmod, Green-Gauss gradients),
    classical RK4 with the global time step at CFL 0.4."""
    m = meshgen.ffs_mesh(1.0 / 1560, seed=2)
    om = orc.mesh(m.points, m.cells)
    U = smooth_state(om, 3.0)
    kw = dict(flux="roe", order=2, limiter="minmod", gradient="gg", cfl=0.4, bc=BCS, fs=freestream(3.0))
    ocfg = make_cfg(roe_bdiss=1, **kw)
    Ro, dto, go = om.residual(ocfg, U)
    print(f"case forward step h=1/1560 full size ({len(m.cells)} cells): residual", flush=True)
    s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(roe_bdiss=1, rk=4, lts=0, **kw))
    s.set_state(U)
    R, dt, g = s.residual()
    assert np.array_equal(R, Ro)
    assert np.array_equal(dt, dto)
    assert np.array_equal(g, go)
    print("case forward step full size: RK4 step, global dt", flush=True)
    Uo = U.copy()
    _, dto_step = om.step(ocfg, Uo, rk=4, lts=0)
    err, dts = s.step(1, history=True)
    assert dts[0] == dto_step
    assert np.array_equal(s.get_state(), Uo)
