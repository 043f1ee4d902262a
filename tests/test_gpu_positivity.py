"""GPU parity of the positivity fallback of the native MUSCL stages against the oracle.

The detecting flux pass flags cells whose stage state fails the admissibility check and marks the
tiles holding their faces; the recomputing pass redoes those tiles with first-order states on the
flagged cells' faces. Both are compared bit for bit with the oracle's restatement (on_step with
positivity), on cases where the fallback fires: the Mach 3 forward-facing step from an impulsive
start (which turns non-physical within 25 steps without it) and a strong blast in a closed box.
"""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config
from oracle.oracle import Oracle, make_cfg

pytestmark = pytest.mark.gpu

STEP_BCS = ("freestream", "exit", "wall", "wall", "wall")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def step_mesh():
    m = meshgen.ffs_mesh(0.025, seed=2)
    fs = freestream(3.0)
    U0 = np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (len(m.cells), 1))
    return m, fs, np.ascontiguousarray(U0)


def blast_box():
    X, Y = np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 1, 41))
    m = meshgen.structured_to_mixed(X, Y, None, 5, 0.1, 0.5, "box")
    return m


@pytest.mark.parametrize("scheme,rk,lts,nsteps,mode", [
    (dict(flux="roe", limiter="minmod", gradient="gg"), 4, 0, 40, 1),
    (dict(flux="hllc", limiter="vanleer", gradient="wlsq"), 3, 1, 30, 1),
    (dict(flux="ausm", limiter="superbee", gradient="wlsq"), 3, 0, 30, 1),
    (dict(flux="roe", limiter="minmod", gradient="gg"), 4, 0, 40, 2),
])
def test_step_mesh_with_fallback_matches_oracle(orc, scheme, rk, lts, nsteps, mode):
    print(f"case positivity {mode} step mesh {scheme} rk{rk} lts{lts}", flush=True)
    m, fs, U0 = step_mesh()
    kw = dict(order=2, roe_bdiss=1, cfl=0.4, bc=STEP_BCS, fs=fs, **scheme)
    om = orc.mesh(m.points, m.cells)
    s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256),
               make_config(rk=rk, lts=lts, block=128, positivity=mode, **kw))
    s.set_state(U0)
    Uo = U0.copy()
    flagged = 0
    errs = []
    for n in range(nsteps):
        e, _ = om.step(make_cfg(positivity=mode, **kw), Uo, rk=rk, lts=lts)
        errs.append(e)
        flagged += orc.last_fallback_cells()
    err, _ = s.step(nsteps, history=True)
    assert flagged > 0
    U = s.get_state()
    assert np.array_equal(U, Uo), rel(U, Uo)
    errs = np.array(errs)
    fin = np.isfinite(errs)
    assert np.array_equal(fin, np.isfinite(err))
    assert np.allclose(err[fin], errs[fin], rtol=1e-12, atol=0)


def test_blast_box_with_fallback_matches_oracle(orc):
    print("case positivity blast box", flush=True)
    m = blast_box()
    om = orc.mesh(m.points, m.cells)
    ctr = om.arrays().ctr[: om.nphys]
    r = np.hypot(ctr[:, 0] - 0.5, ctr[:, 1] - 0.5)
    U0 = np.ascontiguousarray(cons_from_prim(np.ones(len(r)), 0 * r, 0 * r, np.where(r < 0.1, 1000.0, 0.01)))
    kw = dict(flux="hllc", order=2, limiter="vanleer", gradient="gg", cfl=0.4, bc=("wall",) * 5)
    for tile, budget, block in ((64, 128, 64), (128, 256, 128)):
        s = Solver(Mesh(m.points, m.cells, tile=tile, face_budget=budget),
                   make_config(rk=3, lts=0, block=block, positivity=1, **kw))
        s.set_state(U0)
        s.step(12)
        Uo = U0.copy()
        flagged = 0
        for _ in range(12):
            om.step(make_cfg(positivity=1, **kw), Uo, rk=3, lts=0)
            flagged += orc.last_fallback_cells()
        assert flagged > 0
        assert np.array_equal(s.get_state(), Uo), (tile, rel(s.get_state(), Uo))


def test_fallback_off_and_on_agree_on_smooth_flow(orc):
    """On a smooth entropy wave nothing is flagged: the two-pass stages equal the plain ones."""
    print("case positivity smooth flow", flush=True)
    X, Y = np.meshgrid(np.linspace(0, 1, 81), np.linspace(0, 0.5, 41))
    m = meshgen.structured_to_mixed(X, Y, None, 9, 0.1, 0.3, "channel")
    mesh = Mesh(m.points, m.cells, tile=128, face_budget=256)
    ctr = mesh.arrays().ctr[: mesh.nphys]
    x, y = ctr[:, 0], ctr[:, 1]
    fs = freestream(1.65)
    rho = fs[0] * (1 + 0.05 * np.exp(-((x - 0.4) ** 2 + (y - 0.25) ** 2) / 0.02))
    U0 = np.ascontiguousarray(cons_from_prim(rho, fs[1] + 0 * x, 0 * x, fs[3] + 0 * x))
    kw = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.6,
              bc=("freestream",) * 5, fs=fs, rk=4, lts=0, block=128)
    a = Solver(mesh, make_config(**kw))
    b = Solver(mesh, make_config(positivity=1, **kw))
    a.set_state(U0)
    b.set_state(U0)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.get_state(), b.get_state())


def test_fallback_configuration_is_checked():
    m, fs, U0 = step_mesh()
    mesh = Mesh(m.points, m.cells, tile=128, face_budget=256)
    for bad in (dict(order=1), dict(order=3), dict(ref_mode=1, rk=3),
                dict(viscous="constant", mu=1e-3), dict(positivity=3)):
        kw = dict(dict(flux="roe", order=2, limiter="minmod", gradient="gg", cfl=0.4, bc=STEP_BCS, fs=fs,
                       rk=4, lts=0, block=128, positivity=1), **bad)
        with pytest.raises(Exception):
            Solver(mesh, make_config(**kw))


def test_fallback_with_limiter_freezing_matches_oracle(orc):
    """The fallback's passes with a freezing state set (the kernel variants that carry the frozen
    limiter factors) against the oracle: 4 plain, 1 recording and 5 frozen LTS RK3 steps of the
    step mesh from the impulsive start, HLLC + MUSCL / Van Leer + WLSQ."""
    print("case positivity + freezing step mesh", flush=True)
    m, fs, U0 = step_mesh()
    kw = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.4, bc=STEP_BCS, fs=fs)
    om = orc.mesh(m.points, m.cells)
    s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256),
               make_config(rk=3, lts=1, block=128, positivity=1, **kw))
    s.set_state(U0)
    Uo = U0.copy()
    ocfg = make_cfg(positivity=1, **kw)
    phi = np.zeros(8 * om.nfaces, np.float32)
    ocfg.phi = phi.ctypes.data
    flagged = 0
    for mode, n in ((0, 4), (1, 1), (2, 5)):
        s.set_freeze(mode)
        s.step(n)
        ocfg.freeze = mode
        for _ in range(n):
            om.step(ocfg, Uo, rk=3, lts=1)
            flagged += orc.last_fallback_cells()
        assert np.array_equal(s.get_state(), Uo), (mode, rel(s.get_state(), Uo))
    assert flagged > 0
