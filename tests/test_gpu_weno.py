"""GPU parity of WENO5 (order 3) against the CPU oracle: residual, local time step and gradients
bit for bit on uniform quads, on a jittered mixed bump mesh (faces without a complete line fall
back to MUSCL) and on a triangle-heavy mesh, for every flux; independence of the tiling; fused
RK3/RK4 steps and a Sod march bit for bit; batched WENO solves bit for bit as single solvers;
the reference's iteration with WENO stages; unsupported combinations rejected."""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.batch import Batch
from cfd_solver_withcuda_amd.solver import CfdError, Mesh, Solver, cons_from_prim, freestream, make_config
from oracle.oracle import Oracle, make_cfg

pytestmark = pytest.mark.gpu

FS = (1.4, 1.65, 0.0, 1.0)
BCS = ("freestream", "exit", "wall", "wall", "wall")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def smooth_state(ctr, n):
    x, y = ctr[:n, 0], ctr[:n, 1]
    rho = 1.4 * (1 + 0.1 * np.sin(3 * x) * np.cos(2 * y))
    u = 1.65 * (1 + 0.05 * np.cos(4 * y))
    v = 0.1 * np.sin(2 * x)
    p = 1.0 + 0.1 * np.cos(5 * x + y)
    return cons_from_prim(rho, u, v, p)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


MESHES = {
    "quads": lambda: meshgen.shock_tube_mesh(48, 24, length=2.0, height=1.0),
    "bump_mixed": lambda: meshgen.bump_mesh(40, 20, 0.1, seed=7),
    "tri_heavy": lambda: meshgen.shock_tube_mesh(36, 18, length=2.0, height=1.0, seed=13, jitter=0.1, tri_frac=0.6),
}


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("flux,lim,grad", [("hllc", "vanleer", "wlsq"), ("roe", "minmod", "gg"),
                                           ("llf", "superbee", "wlsq"), ("ausm", "none", "gg")])
def test_weno_residual_bit_exact_vs_oracle(orc, mesh, flux, lim, grad):
    print(f"case weno mesh={mesh} flux={flux} lim={lim} grad={grad}", flush=True)
    m = MESHES[mesh]()
    om = orc.mesh(m.points, m.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux=flux, order=3, limiter=lim, gradient=grad, roe_bdiss=1, cfl=0.7, bc=BCS, fs=FS)
    Ro, dto, go = om.residual(make_cfg(**kw), U)
    s = Solver(Mesh(m.points, m.cells, tile=64, renumber=1), make_config(**kw))
    s.set_state(U)
    R, dt, g = s.residual()
    assert np.array_equal(g, go) and np.array_equal(dt, dto)
    assert np.array_equal(R, Ro), rel(R, Ro)


@pytest.mark.parametrize("tile,budget,block", [(128, 256, 128), (32, 0, 64), (64, 128, 64)])
def test_weno_residual_independent_of_tiling(orc, tile, budget, block):
    print(f"case weno tiling tile={tile} budget={budget} block={block}", flush=True)
    m = MESHES["bump_mixed"]()
    om = orc.mesh(m.points, m.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.7, bc=BCS, fs=FS)
    Ro, _, _ = om.residual(make_cfg(**kw), U)
    s = Solver(Mesh(m.points, m.cells, tile=tile, face_budget=budget), make_config(block=block, **kw))
    s.set_state(U)
    R, _, _ = s.residual()
    assert np.array_equal(R, Ro), rel(R, Ro)


@pytest.mark.parametrize("mesh,flux,rk,lts", [("bump_mixed", "hllc", 3, 1), ("bump_mixed", "roe", 4, 0),
                                              ("bump_mixed", "ausm", 3, 0), ("quads", "hllc", 3, 1),
                                              ("quads", "llf", 4, 0)])
def test_weno_steps_vs_oracle(orc, mesh, flux, rk, lts):
    """Fused stages: the state bit-exact after 6 steps. On the quadrilateral mesh most tiles have
    no MUSCL face, so the steps' gradient pass runs over the boundary tiles only."""
    print(f"case weno steps mesh={mesh} flux={flux} rk={rk} lts={lts}", flush=True)
    m = MESHES[mesh]()
    om = orc.mesh(m.points, m.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux=flux, order=3, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.6, bc=BCS, fs=FS)
    ocfg = make_cfg(**kw)
    Uo = np.ascontiguousarray(U0.copy())
    errs = []
    for _ in range(6):
        e, _ = om.step(ocfg, Uo, rk=rk, lts=lts)
        errs.append(e)
    s = Solver(Mesh(m.points, m.cells, tile=64), make_config(rk=rk, lts=lts, **kw))
    s.set_state(U0)
    err, _ = s.step(6, history=True)
    U = s.get_state()
    assert np.array_equal(U, Uo), rel(U, Uo)
    assert np.allclose(err, np.array(errs), rtol=1e-12, atol=0)


def test_weno_sod_march_vs_oracle(orc):
    """Sod on 200 x 4 quads, 150 global-dt TVD-RK3 steps with HLLC: bit for bit."""
    print("case weno sod 200x4, 150 steps", flush=True)
    nx = 200
    m = meshgen.shock_tube_mesh(nx, 4, length=1.0, height=4.0 / nx)
    om = orc.mesh(m.points, m.cells)
    x = om.arrays().ctr[: om.nphys, 0]
    left = x < 0.5
    U0 = np.ascontiguousarray(cons_from_prim(np.where(left, 1.0, 0.125), 0 * x, 0 * x, np.where(left, 1.0, 0.1)))
    kw = dict(flux="hllc", order=3, limiter="minmod", gradient="wlsq", cfl=0.5,
              bc=("exit", "exit", "symmetry", "symmetry", "wall"))
    Uo = U0.copy()
    for _ in range(150):
        om.step(make_cfg(**kw), Uo, rk=3, lts=0)
    s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(rk=3, lts=0, block=128, **kw))
    s.set_state(U0)
    s.step(150)
    assert np.array_equal(s.get_state(), Uo)


@pytest.mark.parametrize("rk", [1, 3])
def test_weno_reference_iteration_vs_oracle(orc, rk):
    """The reference's iteration (ref_mode: per-stage global dt, delta form, exit ghosts from U^n)
    with WENO5 stages, forward Euler or RK3: 8 steps bit for bit."""
    print(f"case weno reference iteration rk={rk}", flush=True)
    m = MESHES["quads"]()
    om = orc.mesh(m.points, m.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux="llf", order=3, limiter="vanleer", gradient="wlsq", cfl=0.5, bc=BCS, fs=FS)
    Uo = np.ascontiguousarray(U0.copy())
    for _ in range(8):
        om.step(make_cfg(**kw), Uo, rk=rk, lts=0, ref=True)
    s = Solver(Mesh(m.points, m.cells, tile=64), make_config(rk=rk, lts=0, ref_mode=1, **kw))
    s.set_state(U0)
    s.step(8)
    assert np.array_equal(s.get_state(), Uo), rel(s.get_state(), Uo)


@pytest.mark.parametrize("mesh", ["quads", "bump_mixed"])
def test_weno_with_viscous_terms_vs_oracle(orc, mesh):
    """WENO5 faces with laminar viscous terms and no-slip walls: residual, time step and gradients
    bit for bit, then 4 LTS RK3 steps bit for bit."""
    print(f"case weno viscous mesh={mesh}", flush=True)
    m = MESHES[mesh]()
    om = orc.mesh(m.points, m.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.6, fs=FS,
              bc=("freestream", "exit", "symmetry", "noslip_wall", "noslip_wall"), viscous="constant", mu=2e-3)
    Ro, dto, go = om.residual(make_cfg(**kw), U0)
    s = Solver(Mesh(m.points, m.cells, tile=64), make_config(rk=3, lts=1, **kw))
    s.set_state(U0)
    R, dt, g = s.residual()
    assert np.array_equal(g, go) and np.array_equal(dt, dto)
    assert np.array_equal(R, Ro), rel(R, Ro)
    Uo = np.ascontiguousarray(U0.copy())
    for _ in range(4):
        om.step(make_cfg(**kw), Uo, rk=3, lts=1)
    s.step(4)
    assert np.array_equal(s.get_state(), Uo), rel(s.get_state(), Uo)


@pytest.mark.parametrize("bad", [dict(freeze_tol=1e-3), dict(freeze_stall=50)])
def test_weno_unsupported_combinations_rejected(bad):
    m = meshgen.shock_tube_mesh(8, 4)
    kw = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", cfl=0.5, fs=FS, bc=BCS)
    kw.update(bad)
    with pytest.raises(CfdError):
        Solver(Mesh(m.points, m.cells), make_config(**kw))


def test_weno_batch_bit_exact_vs_single_solvers():
    """Three ramps of uneven size (quadrilateral lines and triangles) in one batch: 5 WENO steps
    equal each mesh's own solver bit for bit."""
    print("case weno batch vs single, 3 ramps, 5 steps", flush=True)
    specs = [(24, 12, 8.0, 2.0, 41), (40, 18, 15.0, 3.0, 42), (16, 10, 5.0, 1.6, 43)]
    meshes = [meshgen.ramp_mesh(nx, ny, th, seed=sd) for nx, ny, th, _, sd in specs]
    fs = np.array([freestream(M) for *_, M, _ in specs])
    kw = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
              bc=("freestream", "exit", "exit", "wall", "wall"), block=128)
    b = Batch(meshes, fs, make_config(**kw))
    b.solve(tol=-1.0, max_steps=5, check_every=5)
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(fs=fs[i], **kw))
        s.set_state(np.tile(cons_from_prim(np.array([fs[i][0]]), fs[i][1], fs[i][2], fs[i][3]), (len(m.cells), 1)))
        s.step(5)
        assert np.array_equal(Ub[i], s.get_state()), i


def test_weno_batch_rejects_freezing():
    m = meshgen.ramp_mesh(16, 8, 10.0, seed=1)
    with pytest.raises(CfdError, match="freezing"):
        Batch([m], np.array([freestream(2.0)]), make_config(flux="hllc", order=3, rk=3, lts=1, freeze_stall=20,
                                                               bc=("freestream", "exit", "exit", "wall", "wall")))
