"""The oracle's laminar viscous terms pinned by closed forms (CPU only).

The reference has no test or golden vector for its viscous fluxes, so each check here is a field
for which the discrete viscous residual is known exactly, chosen so that a plausible slip (a
transposed or mis-weighted stress component, a wrong heat-flux sign or conductivity, a dropped
time-step factor) fails one of them:
  - mu = 0 leaves residual and time step bit-identical to the Euler path;
  - a linear velocity field has constant stress: no net viscous momentum on cells whose gradients
    are exact (weighted least squares, two layers inside), and on uniform quads the energy term
    is minus the dissipation tau : grad u times the area;
  - u = 4 U y (1 - y): the x-momentum term is 8 mu U A (Poiseuille), with Sutherland's mu(273.15 K)
    = 1.716e-5 (air);
  - p quadratic in y at rest: the energy term is -k T_yy A with k = mu gamma R / ((gamma - 1) Pr);
  - the time step of a uniform quad at rest: CFL h^2 / (4 a h + 16 gamma mu / (rho Pr));
  - a decaying shear wave u = U cos(pi y) between symmetry planes follows exp(-nu pi^2 t).
The residual's viscous part is the difference of two residuals (viscous on minus off): the
inviscid part is computed identically in both.
"""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from oracle.oracle import Oracle, cons_from_prim, make_cfg

GAMMA = 1.4
BCS = ("exit", "exit", "symmetry", "symmetry", "wall")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def deep_cells(a, layers=2):
    """Cells whose neighbourhoods `layers` deep contain no ghost."""
    n = a.nphys
    ok = np.ones(n, bool)
    for c in range(n):
        if np.any(a.nbr[c, : a.nf[c]] >= n):
            ok[c] = False
    for _ in range(layers):
        nxt = ok.copy()
        for c in range(n):
            if ok[c] and not all(ok[j] for j in a.nbr[c, : a.nf[c]]):
                nxt[c] = False
        ok = nxt
    return ok


def viscous_part(om, U, **kw):
    base = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.5, bc=BCS,
                fs=(1.0, 0.0, 0.0, 1.0 / GAMMA))
    base.update({k: v for k, v in kw.items() if k in ("flux", "gradient", "limiter", "bc")})
    visc = {k: v for k, v in kw.items() if k in ("viscous", "mu", "prandtl", "r_gas")}
    R1, dt1, _ = om.residual(make_cfg(**base, **visc), U)
    R0, dt0, _ = om.residual(make_cfg(**base), U)
    return R1 - R0, dt1, dt0


@pytest.mark.parametrize("flux", ["roe", "hllc", "llf", "ausm"])
@pytest.mark.parametrize("grad", ["gg", "wlsq"])
def test_zero_viscosity_is_the_euler_path(orc, flux, grad):
    m = meshgen.shock_tube_mesh(16, 8, length=1.0, height=0.5, seed=3, jitter=0.15, tri_frac=0.4)
    om = orc.mesh(m.points, m.cells)
    x, y = om.arrays().ctr[: om.nphys, 0], om.arrays().ctr[: om.nphys, 1]
    U = cons_from_prim(1 + 0.1 * np.sin(3 * x), 0.3 + 0.1 * y, 0.05 * np.cos(2 * x), 1 + 0.05 * x * y)
    cfg = dict(flux=flux, order=2, limiter="vanleer", gradient=grad, bc=("exit", "exit", "noslip_wall",
                                                                        "noslip_wall", "wall"))
    R1, dt1, g1 = om.residual(make_cfg(**cfg, viscous="constant", mu=0.0), U)
    R0, dt0, g0 = om.residual(make_cfg(**cfg), U)
    assert np.array_equal(R1, R0) and np.array_equal(dt1, dt0) and np.array_equal(g1, g0)


def test_linear_velocity_has_no_net_viscous_momentum(orc):
    m = meshgen.shock_tube_mesh(20, 20, length=1.0, height=1.0, seed=5, jitter=0.2, tri_frac=0.3)
    om = orc.mesh(m.points, m.cells)
    a = om.arrays()
    x, y = a.ctr[: om.nphys, 0], a.ctr[: om.nphys, 1]
    U = cons_from_prim(np.ones_like(x), 0.1 + 0.3 * x - 0.2 * y, -0.1 + 0.25 * x + 0.15 * y,
                       np.full_like(x, 1.0 / GAMMA))
    mu = 0.05
    d, _, _ = viscous_part(om, U, viscous="constant", mu=mu)
    deep = deep_cells(a)
    assert deep.sum() > 50
    # a face flux is about mu |grad u| h = 1e-3; rounding leaves ~1e-18
    assert np.abs(d[deep, 0]).max() == 0.0
    assert np.abs(d[deep, 1:3]).max() < 1e-15


def test_linear_velocity_energy_is_minus_the_dissipation(orc):
    n = 16
    m = meshgen.shock_tube_mesh(n, n, length=1.0, height=1.0)
    om = orc.mesh(m.points, m.cells)
    a = om.arrays()
    x, y = a.ctr[: om.nphys, 0], a.ctr[: om.nphys, 1]
    ux, uy, vx, vy = 0.3, -0.2, 0.25, 0.15
    U = cons_from_prim(np.ones_like(x), 0.1 + ux * x + uy * y, -0.1 + vx * x + vy * y,
                       np.full_like(x, 1.0 / GAMMA))
    mu = 0.05
    d, _, _ = viscous_part(om, U, viscous="constant", mu=mu)
    txx = 2.0 / 3.0 * mu * (2 * ux - vy)
    tyy = 2.0 / 3.0 * mu * (2 * vy - ux)
    txy = mu * (uy + vx)
    phi = txx * ux + txy * (uy + vx) + tyy * vy
    deep = deep_cells(a)
    A = a.area[: om.nphys]
    np.testing.assert_allclose(d[deep, 3], -phi * A[deep], rtol=1e-10, atol=0)
    assert np.abs(d[deep, 1:3]).max() < 1e-15


@pytest.mark.parametrize("viscous,mu,rho,p", [("constant", 0.02, 1.0, 1.0 / GAMMA),
                                              ("sutherland", 0.0, 1.0, 287.0 * 273.15)])
def test_poiseuille_profile_momentum(orc, viscous, mu, rho, p):
    n = 16
    m = meshgen.shock_tube_mesh(n, n, length=1.0, height=1.0)
    om = orc.mesh(m.points, m.cells)
    a = om.arrays()
    x, y = a.ctr[: om.nphys, 0], a.ctr[: om.nphys, 1]
    Uw = 0.05 * np.sqrt(GAMMA * p / rho)
    U = cons_from_prim(np.full_like(x, rho), 4 * Uw * y * (1 - y), np.zeros_like(x), np.full_like(x, p))
    d, _, _ = viscous_part(om, U, viscous=viscous, mu=mu, r_gas=287.0)
    deep = deep_cells(a)
    A = a.area[: om.nphys]
    mu_eff = mu if viscous == "constant" else 1.716e-5  # air at 273.15 K (Sutherland's mu_0)
    rtol = 1e-10 if viscous == "constant" else 2e-4
    np.testing.assert_allclose(d[deep, 1], 8 * mu_eff * Uw * A[deep], rtol=rtol, atol=0)
    # the pressure flux (p h ~ 5e3 at 273 K) leaves rounding of ~1e-12 in the difference
    assert np.abs(d[deep, 2]).max() < 1e-6 * 8 * mu_eff * Uw * A.max()


def test_heat_conduction_sign_and_conductivity(orc):
    n = 16
    m = meshgen.shock_tube_mesh(n, n, length=1.0, height=1.0)
    om = orc.mesh(m.points, m.cells)
    a = om.arrays()
    x, y = a.ctr[: om.nphys, 0], a.ctr[: om.nphys, 1]
    p0, beta, rho, mu, pr, R = 1.0 / GAMMA, 0.3, 1.2, 0.01, 0.72, 287.0
    U = cons_from_prim(np.full_like(x, rho), np.zeros_like(x), np.zeros_like(x), p0 * (1 + beta * y * y))
    d, _, _ = viscous_part(om, U, viscous="constant", mu=mu, prandtl=pr, r_gas=R)
    deep = deep_cells(a)
    A = a.area[: om.nphys]
    k = mu * GAMMA * R / ((GAMMA - 1) * pr)
    T_yy = 2 * p0 * beta / (rho * R)
    np.testing.assert_allclose(d[deep, 3], -k * T_yy * A[deep], rtol=1e-9, atol=0)
    assert np.abs(d[deep, :3]).max() == 0.0


def test_viscous_time_step_closed_form(orc):
    n, L = 8, 2.0
    h = L / n
    m = meshgen.shock_tube_mesh(n, n, length=L, height=L)
    om = orc.mesh(m.points, m.cells)
    rho, p, mu, pr, cfl = 1.3, 0.9, 0.07, 0.7, 0.6
    U = cons_from_prim(np.full(om.nphys, rho), np.zeros(om.nphys), np.zeros(om.nphys), np.full(om.nphys, p))
    cfg = make_cfg(flux="hllc", order=2, gradient="wlsq", cfl=cfl, bc=BCS, viscous="constant", mu=mu,
                   prandtl=pr)
    _, dt, _ = om.residual(cfg, U)
    a_s = np.sqrt(GAMMA * p / rho)
    want = cfl * h * h / (4 * a_s * h + 16 * GAMMA * mu / (rho * pr))
    np.testing.assert_allclose(dt, want, rtol=1e-13)


def test_noslip_ghost_state(orc):
    W = np.array([1.2, 0.3, -0.4, 0.9, np.sqrt(GAMMA * 0.9 / 1.2)])
    g = orc.ghost_state("noslip_wall", W, 0.6, 0.8, (1.0, 0.0, 0.0, 1.0))
    np.testing.assert_allclose(g, [1.2, -0.3, 0.4, 0.9, W[4]], rtol=1e-15)


def test_shear_wave_decays_at_the_viscous_rate(orc):
    """u = U cos(pi y) on [0, 1/4] x [0, 1] between symmetry planes, Mach 0.05: the amplitude
    decays as exp(-nu pi^2 t) (HLLC and Roe add no dissipation to a pure shear wave)."""
    nx, ny = 2, 32
    m = meshgen.shock_tube_mesh(nx, ny, length=nx / ny, height=1.0)
    om = orc.mesh(m.points, m.cells)
    y = om.arrays().ctr[: om.nphys, 1]
    rho, p, mu = 1.0, 1.0 / GAMMA, 0.01
    Uw = 0.05
    shape = np.cos(np.pi * y)
    U = cons_from_prim(np.full_like(y, rho), Uw * shape, np.zeros_like(y), np.full_like(y, p))
    cfg = make_cfg(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, bc=BCS,
                   viscous="constant", mu=mu)
    t = 0.0
    while t < 0.5:
        _, dt = om.step(cfg, U, rk=3, lts=0)
        t += dt
    amp = float(np.sum(U[:, 1] / U[:, 0] * shape) / np.sum(shape * shape))
    want = Uw * np.exp(-(mu / rho) * np.pi ** 2 * t)
    assert abs(amp / want - 1) < 2e-3, (amp, want, t)


def test_viscous_terms_on_weno_faces(orc):
    """WENO5 (order 3) with viscous terms: mu = 0 is the inviscid WENO residual bit for bit, and the
    shear wave decays at the viscous rate with WENO faces in the interior of the line mesh."""
    m = meshgen.shock_tube_mesh(16, 8, length=1.0, height=0.5, seed=3, jitter=0.1)
    om = orc.mesh(m.points, m.cells)
    x, y = om.arrays().ctr[: om.nphys, 0], om.arrays().ctr[: om.nphys, 1]
    U = cons_from_prim(1 + 0.1 * np.sin(3 * x), 0.3 + 0.1 * y, 0.05 * np.cos(2 * x), 1 + 0.05 * x * y)
    cfg = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", bc=BCS)
    R1, dt1, _ = om.residual(make_cfg(**cfg, viscous="constant", mu=0.0), U)
    R0, dt0, _ = om.residual(make_cfg(**cfg), U)
    assert np.array_equal(R1, R0) and np.array_equal(dt1, dt0)

    nx, ny = 2, 32
    m = meshgen.shock_tube_mesh(nx, ny, length=nx / ny, height=1.0)
    om = orc.mesh(m.points, m.cells)
    y = om.arrays().ctr[: om.nphys, 1]
    rho, p, mu, Uw = 1.0, 1.0 / GAMMA, 0.01, 0.05
    shape = np.cos(np.pi * y)
    U = np.ascontiguousarray(cons_from_prim(np.full_like(y, rho), Uw * shape, np.zeros_like(y), np.full_like(y, p)))
    c = make_cfg(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", cfl=0.8, bc=BCS, viscous="constant", mu=mu)
    t = 0.0
    while t < 0.5:
        _, dt = om.step(c, U, rk=3, lts=0)
        t += dt
    amp = float(np.sum(U[:, 1] / U[:, 0] * shape) / np.sum(shape * shape))
    assert abs(amp / (Uw * np.exp(-(mu / rho) * np.pi ** 2 * t)) - 1) < 2e-3

