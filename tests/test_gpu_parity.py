"""GPU parity of the native HIP path against the CPU oracle and the reference's golden vectors.

All comparisons are FP64. The kernels are compiled without contraction and keep the oracle's
operation order, so most checks are bit-exact; the few reductions whose summation order differs
(error-norm partial sums) use a 1e-12 relative tolerance, stated per test.
"""
import os

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, make_config
from oracle.oracle import Oracle, make_cfg

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FS = (1.4, 1.65 * np.sqrt(1.0), 0.0, 1.0)  # rho, u (Mach 1.65 at a = 1), v, p


@pytest.fixture(scope="module")
def orc():
    return Oracle()


@pytest.fixture(scope="module")
def bump():
    m = meshgen.bump_mesh(40, 20, 0.1, seed=7)
    return m


def smooth_state(ctr, n):
    x, y = ctr[:n, 0], ctr[:n, 1]
    rho = 1.4 * (1 + 0.1 * np.sin(3 * x) * np.cos(2 * y))
    u = 1.65 * (1 + 0.05 * np.cos(4 * y))
    v = 0.1 * np.sin(2 * x)
    p = 1.0 + 0.1 * np.cos(5 * x + y)
    return cons_from_prim(rho, u, v, p)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


BCS_NATIVE = ("freestream", "exit", "wall", "wall", "wall")
CASES = []
for flux in ("roe", "hllc", "llf", "ausm"):
    CASES.append((flux, 1, "none", "gg"))
    for lim in ("minmod", "vanleer", "superbee", "vanalbada", "none"):
        for grad in ("gg", "wlsq"):
            CASES.append((flux, 2, lim, grad))


@pytest.mark.parametrize("flux,order,lim,grad", CASES)
def test_residual_bit_exact_vs_oracle(orc, bump, flux, order, lim, grad):
    """R, local dt and gradients of one residual evaluation equal the oracle's bit for bit."""
    print(f"case flux={flux} order={order} lim={lim} grad={grad}", flush=True)
    om = orc.mesh(bump.points, bump.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    ocfg = make_cfg(flux=flux, order=order, limiter=lim, gradient=grad, roe_bdiss=1, cfl=0.7,
                    bc=BCS_NATIVE, fs=FS)
    Ro, dto, go = om.residual(ocfg, U)
    mesh = Mesh(bump.points, bump.cells, tile=64, renumber=1)
    s = Solver(mesh, make_config(flux=flux, order=order, limiter=lim, gradient=grad, roe_bdiss=1,
                                 cfl=0.7, bc=BCS_NATIVE, fs=FS))
    s.set_state(U)
    R, dt, g = s.residual()
    assert np.array_equal(dt, dto)
    if order == 2:
        assert np.array_equal(g, go), rel(g, go)
    assert np.array_equal(R, Ro), rel(R, Ro)


@pytest.mark.parametrize("tri_frac,jitter", [(0.0, 0.0), (0.0, 0.15), (1.0, 0.1), (0.5, 0.1)])
@pytest.mark.parametrize("flux,grad", [("hllc", "wlsq"), ("roe", "gg"), ("ausm", "wlsq"), ("llf", "gg")])
def test_residual_on_quad_triangle_and_mixed_meshes(orc, tri_frac, jitter, flux, grad):
    """Uniform and jittered quads, all triangles, and half split: the second-order residual,
    local dt and gradients equal the oracle's bit for bit whatever the cell shapes."""
    print(f"case tri_frac={tri_frac} jitter={jitter} flux={flux} grad={grad}", flush=True)
    m = meshgen.shock_tube_mesh(36, 18, length=2.0, height=1.0, seed=13, jitter=jitter, tri_frac=tri_frac)
    om = orc.mesh(m.points, m.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux=flux, order=2, limiter="vanleer", gradient=grad, cfl=0.7, bc=BCS_NATIVE, fs=FS)
    Ro, dto, go = om.residual(make_cfg(roe_bdiss=1, **kw), U)
    s = Solver(Mesh(m.points, m.cells, tile=64, renumber=1), make_config(roe_bdiss=1, **kw))
    s.set_state(U)
    R, dt, g = s.residual()
    assert np.array_equal(dt, dto) and np.array_equal(g, go) and np.array_equal(R, Ro), rel(R, Ro)


@pytest.mark.parametrize("tile,renumber,block,remap", [(32, 0, 64, 0), (128, 1, 128, 1), (256, 1, 256, 1),
                                                       (32, 1, 64, 3), (64, 1, 128, 16), (32, 1, 64, 4096)])
def test_residual_independent_of_tiling(orc, bump, tile, renumber, block, remap):
    """Tile size, renumbering, block and the tile order over the XCDs (xcd_remap: block order, one
    run per XCD, runs of k tiles dealt in turn) change nothing: each order visits every tile once."""
    om = orc.mesh(bump.points, bump.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.7, bc=BCS_NATIVE, fs=FS)
    Ro, _, _ = om.residual(make_cfg(roe_bdiss=1, **cfg), U)
    s = Solver(Mesh(bump.points, bump.cells, tile=tile, renumber=renumber),
               make_config(block=block, xcd_remap=remap, **cfg))
    s.set_state(U)
    R, _, _ = s.residual()
    assert np.array_equal(R, Ro)


@pytest.mark.parametrize("remap", [0, 1, 5, 16])
def test_steps_independent_of_xcd_order(bump, remap):
    """Fused RK3 stages with local time stepping: the state after 3 steps is the same bit for bit
    whatever order the tiles are dealt to the XCDs in."""
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.7, bc=BCS_NATIVE, fs=FS,
               rk=3, lts=1)
    mesh = Mesh(bump.points, bump.cells, tile=32)
    U0 = smooth_state(mesh.arrays().ctr, mesh.nphys)
    out = []
    for r in (2, remap):
        s = Solver(mesh, make_config(block=64, xcd_remap=r, **cfg))
        s.set_state(U0)
        s.step(3)
        out.append(s.get_state())
    assert np.array_equal(out[0], out[1])


def test_residual_deterministic(bump):
    mesh = Mesh(bump.points, bump.cells, tile=64)
    s = Solver(mesh, make_config(flux="ausm", order=2, limiter="vanleer", gradient="wlsq", bc=BCS_NATIVE, fs=FS))
    U = smooth_state(mesh.arrays().ctr, mesh.nphys)
    s.set_state(U)
    R1, _, _ = s.residual()
    R2, _, _ = s.residual()
    assert np.array_equal(R1, R2)


@pytest.mark.parametrize("bc", ["farfield", "symmetry", "inlet_ref", "subsonic_inlet_ref"])
def test_boundary_types_vs_oracle(orc, bump, bc):
    om = orc.mesh(bump.points, bump.cells)
    U = smooth_state(om.arrays().ctr, om.nphys)
    bcs = (bc, "exit", "wall", bc, "symmetry")
    Ro, _, _ = om.residual(make_cfg(flux="roe", order=2, limiter="minmod", gradient="wlsq",
                                    roe_bdiss=1, bc=bcs, fs=FS), U)
    s = Solver(Mesh(bump.points, bump.cells, tile=64),
               make_config(flux="roe", order=2, limiter="minmod", gradient="wlsq", roe_bdiss=1, bc=bcs, fs=FS))
    s.set_state(U)
    R, _, _ = s.residual()
    if bc == "farfield":
        # the far-field ghost state calls pow(); device pow and glibc pow are not both
        # correctly rounded, so this case is compared to 1e-13 relative, not bitwise
        assert rel(R, Ro) <= 1e-13
    else:
        assert np.array_equal(R, Ro)


@pytest.mark.parametrize("bcs", [BCS_NATIVE, ("farfield", "exit", "symmetry", "wall", "wall"),
                                 ("inlet_ref", "subsonic_inlet_ref", "freestream", "symmetry", "wall")])
def test_ghost_states_vs_oracle(orc, bump, bcs):
    """cfd_apply_bc: every ghost (reference ghost order) equals the oracle's ghost of its owner
    cell; bit-exact except the far field (pow, 1e-13 relative)."""
    print(f"case bcs={bcs}", flush=True)
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(bc=bcs, fs=FS))
    a = s.mesh.arrays()
    U = smooth_state(a.ctr, s.mesh.nphys)
    s.set_state(U)
    Ug = s.ghost_states()
    trip = np.concatenate([a.inlet.reshape(-1, 3), a.exit.reshape(-1, 3), a.wall.reshape(-1, 3)])
    assert len(trip) == s.mesh.nghost == len(Ug)
    for c, f, g in trip:
        kind = a.kind[c, f]
        want = orc.ghost_cons(bcs[kind - 1], U[c], a.nrm[c, 2 * f], a.nrm[c, 2 * f + 1], FS)
        got = Ug[g - s.mesh.nphys]
        if bcs[kind - 1] == "farfield":
            assert np.allclose(got, want, rtol=1e-13, atol=0), (c, f, got, want)
        else:
            assert np.array_equal(got, want), (c, f, bcs[kind - 1], got, want)


FS_SUB = (1.4, 0.5, 0.05, 1.0)  # a Mach 0.5 prescribed state


def subsonic_state(ctr, n):
    x, y = ctr[:n, 0], ctr[:n, 1]
    return cons_from_prim(1.4 * (1 + 0.05 * np.sin(3 * x)), 0.5 * (1 + 0.1 * np.cos(4 * y)),
                          0.05 * np.sin(2 * x), 1.0 + 0.08 * np.cos(5 * x + y))


@pytest.mark.parametrize("p_exit", [0.0, 0.93])
def test_subsonic_boundaries_vs_oracle(orc, bump, p_exit):
    """The back-pressure exit (9) and the total-condition inlet (10): ghost states, residuals at
    first and second order and LTS-RK3 steps against the oracle. The inlet calls pow(), so those
    comparisons are to 1e-13 relative; the exit ghosts are bit-exact."""
    print(f"case subsonic boundaries p_exit={p_exit}", flush=True)
    bcs = ("subsonic_inlet", "subsonic_exit", "wall", "wall", "wall")
    om = orc.mesh(bump.points, bump.cells)
    U = subsonic_state(om.arrays().ctr, om.nphys)
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(bc=bcs, fs=FS_SUB, p_exit=p_exit))
    a = s.mesh.arrays()
    s.set_state(U)
    Ug = s.ghost_states()
    trip = np.concatenate([a.inlet.reshape(-1, 3), a.exit.reshape(-1, 3), a.wall.reshape(-1, 3)])
    for c, f, g in trip:
        kind = a.kind[c, f]
        want = orc.ghost_cons(bcs[kind - 1], U[c], a.nrm[c, 2 * f], a.nrm[c, 2 * f + 1], FS_SUB, p_exit=p_exit)
        got = Ug[g - s.mesh.nphys]
        if bcs[kind - 1] == "subsonic_inlet":
            assert np.allclose(got, want, rtol=1e-13, atol=0), (c, f, got, want)
        else:
            assert np.array_equal(got, want), (c, f, bcs[kind - 1], got, want)
    for kw in (dict(flux="roe", order=1), dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq")):
        Ro, dto, _ = om.residual(make_cfg(roe_bdiss=1, bc=bcs, fs=FS_SUB, p_exit=p_exit, cfl=0.7, **kw), U)
        s.configure(make_config(roe_bdiss=1, bc=bcs, fs=FS_SUB, p_exit=p_exit, cfl=0.7, rk=3, lts=1, **kw))
        s.set_state(U)
        R, dt, _ = s.residual()
        assert rel(R, Ro) <= 1e-13 and np.array_equal(dt, dto), (kw, rel(R, Ro))
        s.step(5)
        Uo = np.ascontiguousarray(U.copy())
        for _ in range(5):
            om.step(make_cfg(roe_bdiss=1, bc=bcs, fs=FS_SUB, p_exit=p_exit, cfl=0.7, **kw), Uo, rk=3, lts=1)
        assert rel(s.get_state(), Uo) <= 1e-12, (kw, rel(s.get_state(), Uo))


@pytest.mark.parametrize("flux,order,rk,lts", [("roe", 1, 3, 0), ("hllc", 2, 3, 1), ("ausm", 2, 4, 0),
                                              ("llf", 2, 4, 1), ("hllc", 2, 3, 0)])
def test_native_steps_vs_oracle(orc, bump, flux, order, rk, lts):
    """Fused RK stages: state bit-exact after 6 steps; error norm (tree-summed) to 1e-12."""
    om = orc.mesh(bump.points, bump.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux=flux, order=order, limiter="vanleer", gradient="wlsq", cfl=0.6, bc=BCS_NATIVE, fs=FS)
    ocfg = make_cfg(roe_bdiss=1, **kw)
    Uo = U0.copy()
    errs = []
    for _ in range(6):
        e, _ = om.step(ocfg, Uo, rk=rk, lts=lts)
        errs.append(e)
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(roe_bdiss=1, rk=rk, lts=lts, **kw))
    s.set_state(U0)
    err, _ = s.step(6, history=True)
    U = s.get_state()
    assert np.array_equal(U, Uo), rel(U, Uo)
    assert np.allclose(err, np.array(errs), rtol=1e-12, atol=0)


def _sod_setup():
    m = meshgen.shock_tube_mesh(100, 10)
    g = np.load(os.path.join(GOLD, "sod_roe.npz"), allow_pickle=False)
    return m, g


def test_sod_reference_iteration_vs_golden():
    """The reference's Inviscid_Solver iteration (Roe-1O, TVD-RK3 delta form, CFL 0.5) on the
    Sod tube: 25 steps against the reference's own outputs."""
    m, g = _sod_setup()
    s = Solver(Mesh(m.points, m.cells, tile=128),
               make_config(flux="roe", order=1, cfl=0.5, fs=(1.0, 0.0, 0.0, 1.0), rk=3, ref_mode=1))
    s.set_state(g["U0"])
    err, dt = s.step(25, history=True)
    U = s.get_state()
    assert np.allclose(err, g["err_hist"], rtol=1e-10, atol=1e-13)
    assert np.allclose(dt, g["dt_hist"], rtol=1e-13, atol=0)
    assert rel(U, g["U_step25"]) <= 1e-12
    assert abs(dt.sum() - g["time_hist"][-1]) <= 1e-12 * g["time_hist"][-1]


SOD_JSON = """{
    "TestCase": {"Test_Case": 0, "Test_Case_Name": "Sod_Shock_Tube"},
    "Simulation": {"Initialize_Type": 0, "Is_Implicit_Method": false, "Total_Iterations": 25,
                   "CFL": 0.5, "Terminating_Time": 0.2, "Is_Time_Dependent": true},
    "Solver": {"Solver_Type": 0, "Solver_Name": "Euler", "Is_Conservative": true, "Is_Viscous": false,
               "Limiter_Case": 0, "Area_Weighted_Average": 1, "Flux_Type": 2, "NUM_FLUX_COMPONENTS": 4,
               "Is_Second_Order": false, "Time_Accurate": true, "Local_Time_Stepping": false,
               "Non_Dimensional_Form": false, "Is_WENO": false, "Dissipation_Type": 3,
               "Is_MOVERS_1": false, "Enable_Entropy_Fix": true},
    "LimiterCoefficients": {"Limiter_Zeta": 0.5, "Limiter_Zeta1": 0.5
                            // "Epsilon": 1.0e-6
                           }
}"""


def test_reference_json_drives_sod_golden(tmp_path):
    """A solver file in the reference's schema (json_Files/Solver_Config.json) selects the
    reference's iteration (Time_Accurate: TVD-RK3 in delta form; Dissipation_Type 3: Roe), which
    reproduces the reference's own 25-step Sod histories."""
    from cfd_solver_withcuda_amd import refio

    p = tmp_path / "Solver_Config.json"
    p.write_text(SOD_JSON)
    kw = refio.config_from_json(str(p))
    assert kw["ref_mode"] == 1 and kw["rk"] == 3 and kw["flux"] == "roe" and kw["order"] == 1
    m, g = _sod_setup()
    s = Solver(Mesh(m.points, m.cells, tile=128), make_config(fs=(1.0, 0.0, 0.0, 1.0), **kw))
    s.set_state(g["U0"])
    err, dt = s.step(25, history=True)
    assert np.allclose(err, g["err_hist"], rtol=1e-10, atol=1e-13)
    assert np.allclose(dt, g["dt_hist"], rtol=1e-13, atol=0)
    assert rel(s.get_state(), g["U_step25"]) <= 1e-12
    # Time_Accurate false is the reference's Explicit_Method: forward Euler with the global dt
    p.write_text(SOD_JSON.replace('"Time_Accurate": true', '"Time_Accurate": false'))
    kw = refio.config_from_json(str(p))
    assert kw["rk"] == 1
    s.configure(make_config(fs=(1.0, 0.0, 0.0, 1.0), **kw))
    s.set_state(g["U0"])
    s.step(3)
    om = Oracle().mesh(m.points, m.cells)
    Uo = np.ascontiguousarray(g["U0"].copy())
    for _ in range(3):
        om.step(make_cfg(flux="roe", order=1, cfl=0.5, fs=(1.0, 0, 0, 1.0)), Uo, rk=1, lts=0, ref=True)
    assert np.array_equal(s.get_state(), Uo)


def test_sod_reference_iteration_vs_oracle_bit_exact(orc):
    m, g = _sod_setup()
    om = orc.mesh(m.points, m.cells)
    cfg = make_cfg(flux="roe", order=1, cfl=0.5, fs=(1.0, 0, 0, 1.0))
    Uo = np.ascontiguousarray(g["U0"].copy())
    for _ in range(10):
        om.step(cfg, Uo, rk=3, lts=0, ref=True)
    s = Solver(Mesh(m.points, m.cells, tile=128),
               make_config(flux="roe", order=1, cfl=0.5, fs=(1.0, 0.0, 0.0, 1.0), rk=3, ref_mode=1))
    s.set_state(g["U0"])
    s.step(10)
    assert np.array_equal(s.get_state(), Uo)


@pytest.mark.parametrize("tag,rk,lts", [("lts", 3, 1), ("euler", 1, 0)])
def test_bump_reference_iteration_vs_golden(tag, rk, lts):
    g = np.load(os.path.join(GOLD, "bump_mixed.npz"), allow_pickle=False)
    m = meshgen.bump_mesh(40, 20, 0.1, seed=7)
    s = Solver(Mesh(m.points, m.cells, tile=64),
               make_config(flux="roe", order=1, cfl=0.5, fs=(1.4, 0.0, 0.0, 1.0), rk=rk, lts=lts, ref_mode=1))
    s.set_state(g["U0"])
    err, _ = s.step(10, history=True)
    assert np.allclose(err, g[f"err_{tag}"], rtol=1e-10, atol=1e-13)
    assert rel(s.get_state(), g[f"U_{tag}"]) <= 1e-12


def test_freestream_preserved(bump):
    """Uniform freestream with freestream/far-field ghosts stays uniform (residual at roundoff)."""
    mesh = Mesh(bump.points, bump.cells, tile=64)
    s = Solver(mesh, make_config(flux="hllc", order=2, limiter="vanleer", gradient="wlsq",
                                 bc=("freestream",) * 5, fs=FS))
    U = np.tile(cons_from_prim(np.array([FS[0]]), FS[1], FS[2], FS[3]), (mesh.nphys, 1))
    s.set_state(U)
    R, _, _ = s.residual()
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("flux,lim,grad", [("hllc", "vanleer", "wlsq"), ("ausm", "minmod", "gg"),
                                           ("hllc", "vanalbada", "wlsq")])
def test_limiter_freezing_vs_oracle(orc, bump, flux, lim, grad):
    """Record-then-frozen limiter factors (cfd_set_freeze): states bit-exact with the oracle's
    freezing over 3 plain, 1 recording and 4 frozen LTS RK3 steps."""
    print(f"case freezing {flux} {lim} {grad}", flush=True)
    om = orc.mesh(bump.points, bump.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    kw = dict(flux=flux, order=2, limiter=lim, gradient=grad, cfl=0.7, bc=BCS_NATIVE, fs=FS)
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(rk=3, lts=1, **kw))
    s.set_state(U0)
    U = np.ascontiguousarray(U0.copy())
    ocfg = make_cfg(**kw)
    phi = np.zeros(8 * om.nfaces, np.float32)
    ocfg.phi = phi.ctypes.data
    for mode, n in ((0, 3), (1, 1), (2, 4)):
        s.set_freeze(mode)
        s.step(n)
        ocfg.freeze = mode
        for _ in range(n):
            om.step(ocfg, U, rk=3, lts=1)
        assert np.array_equal(s.get_state(), U), mode


def test_reference_second_order_roe_vs_golden():
    """ref_mode at order 2 with Roe: the reference's Evaluate_Cell_Net_Flux_2O (Roe dissipation on
    MinMod-limited differences, zeta = zeta1 = 0.5) against the reference's own residual of the
    jittered mixed bump (triangles fall back to first order), bit for bit; dt as the reference's."""
    g = np.load(os.path.join(GOLD, "bump_mixed.npz"), allow_pickle=False)
    m = meshgen.bump_mesh(40, 20, 0.1, seed=7)
    s = Solver(Mesh(m.points, m.cells, tile=64),
               make_config(flux="roe", order=2, cfl=0.5, fs=(1.4, 0.0, 0.0, 1.0), rk=3, ref_mode=1))
    s.set_state(g["U0"])
    R, dt, _ = s.residual()
    assert np.array_equal(R, g["R_roe2o"]), rel(R, g["R_roe2o"])
    assert np.array_equal(dt, g["dt"])
    s.step(3)
    assert np.all(np.isfinite(s.get_state()))


def test_steps_follow_reconfiguration_and_freezing(orc, bump):
    """Changing the Courant number, the flux and the time stepping between steps of one solver,
    then switching limiter freezing on (which allocates its buffers), gives the oracle's states bit
    for bit after every segment."""
    print("case steps across reconfiguration", flush=True)
    om = orc.mesh(bump.points, bump.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    base = dict(order=2, limiter="vanleer", gradient="wlsq", bc=BCS_NATIVE, fs=FS, roe_bdiss=1)
    segs = [dict(flux="hllc", cfl=0.6, rk=3, lts=1), dict(flux="hllc", cfl=0.4, rk=3, lts=1),
            dict(flux="roe", cfl=0.5, rk=4, lts=0), dict(flux="roe", cfl=0.5, rk=3, lts=0)]
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(**base, **segs[0]))
    s.set_state(U0)
    Uo = np.ascontiguousarray(U0.copy())
    for i, sg in enumerate(segs):
        s.configure(make_config(**base, **sg))
        s.step(3)
        ocfg = make_cfg(flux=sg["flux"], order=2, limiter="vanleer", gradient="wlsq", roe_bdiss=1,
                        cfl=sg["cfl"], bc=BCS_NATIVE, fs=FS)
        for _ in range(3):
            om.step(ocfg, Uo, rk=sg["rk"], lts=sg["lts"])
        assert np.array_equal(s.get_state(), Uo), (i, rel(s.get_state(), Uo))
    # freezing: record on the next step, then reuse (buffers allocated between captured steps)
    s.configure(make_config(**base, **segs[0]))
    s.set_freeze(1)
    s.step(1)
    s.set_freeze(2)
    s.step(2)
    phi = np.zeros(8 * om.nfaces, np.float32)
    ocfg = make_cfg(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", roe_bdiss=1, cfl=0.6,
                    bc=BCS_NATIVE, fs=FS)
    ocfg.phi = phi.ctypes.data
    ocfg.freeze = 1
    om.step(ocfg, Uo, rk=3, lts=1)
    ocfg.freeze = 2
    for _ in range(2):
        om.step(ocfg, Uo, rk=3, lts=1)
    assert np.array_equal(s.get_state(), Uo), rel(s.get_state(), Uo)


def test_reconfiguration_after_freezing_evaluates_the_limiter(orc, bump):
    """Recorded limiter factors belong to the scheme that recorded them: after freezing (mode 2)
    a change of limiter makes the next steps evaluate the new limiter, as the oracle does without
    freezing; the explicit mode survives a reconfiguration that keeps the scheme."""
    print("case reconfiguration after freezing", flush=True)
    om = orc.mesh(bump.points, bump.cells)
    U0 = smooth_state(om.arrays().ctr, om.nphys)
    base = dict(order=2, gradient="wlsq", bc=BCS_NATIVE, fs=FS, roe_bdiss=1, flux="hllc", cfl=0.6, rk=3, lts=1)
    s = Solver(Mesh(bump.points, bump.cells, tile=64), make_config(limiter="vanleer", **base))
    s.set_state(U0)
    s.set_freeze(1)
    s.step(1)
    s.set_freeze(2)
    s.step(1)
    U1 = s.get_state()
    s.configure(make_config(limiter="minmod", **base))
    s.step(2)
    Uo = np.ascontiguousarray(U1.copy())
    ocfg = make_cfg(flux="hllc", order=2, limiter="minmod", gradient="wlsq", roe_bdiss=1, cfl=0.6,
                    bc=BCS_NATIVE, fs=FS)
    for _ in range(2):
        om.step(ocfg, Uo, rk=3, lts=1)
    assert np.array_equal(s.get_state(), Uo), rel(s.get_state(), Uo)
    # same scheme, new Courant number: the recorded factors stay in use
    s.configure(make_config(limiter="minmod", **base))
    s.set_state(U0)
    s.set_freeze(1)
    s.step(1)
    s.set_freeze(2)
    s.configure(make_config(limiter="minmod", **dict(base, cfl=0.5)))
    s.step(1)
    phi = np.zeros(8 * om.nfaces, np.float32)
    Uo = np.ascontiguousarray(U0.copy())
    ocfg = make_cfg(flux="hllc", order=2, limiter="minmod", gradient="wlsq", roe_bdiss=1, cfl=0.6,
                    bc=BCS_NATIVE, fs=FS)
    ocfg.phi = phi.ctypes.data
    ocfg.freeze = 1
    om.step(ocfg, Uo, rk=3, lts=1)
    ocfg.freeze = 2
    ocfg.cfl = 0.5
    om.step(ocfg, Uo, rk=3, lts=1)
    assert np.array_equal(s.get_state(), Uo), rel(s.get_state(), Uo)
