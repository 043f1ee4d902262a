"""The positivity fallback of the native MUSCL stages, pinned on the oracle (CPU).

The fallback has no counterpart in the reference. A stage state that has a non-positive density or
pressure, or leaves the relaxed range of its neighbourhood's stage inputs, is recomputed with
first-order states on the faces of the flagged cells. These tests pin what it must do:
- keep the step mesh's second-order march (Mach 3, Roe + minmod + Green-Gauss, RK4, CFL 0.4) and a
  strong blast physical, where both turn non-physical without it;
- conserve mass and energy exactly in a closed box (the recomputed faces change both cells' fluxes);
- leave smooth flows bit for bit unchanged.
"""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import cons_from_prim, freestream
from oracle.oracle import Oracle, make_cfg

STEP_BCS = ("freestream", "exit", "wall", "wall", "wall")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def rho_p(U):
    rho = U[:, 0]
    return rho, 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)


def step_case(orc):
    m = meshgen.ffs_mesh(0.025, seed=2)
    om = orc.mesh(m.points, m.cells)
    fs = freestream(3.0)
    U = np.ascontiguousarray(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (om.nphys, 1)))
    return om, fs, U


def blast_case(orc):
    X, Y = np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 1, 41))
    m = meshgen.structured_to_mixed(X, Y, None, 5, 0.1, 0.5, "box")
    om = orc.mesh(m.points, m.cells)
    ctr = om.arrays().ctr[: om.nphys]
    r = np.hypot(ctr[:, 0] - 0.5, ctr[:, 1] - 0.5)
    U = np.ascontiguousarray(cons_from_prim(np.ones(len(r)), 0 * r, 0 * r, np.where(r < 0.1, 1000.0, 0.01)))
    return om, U


def test_step_mesh_second_order_needs_the_fallback(orc):
    om, fs, U0 = step_case(orc)
    kw = dict(flux="roe", order=2, limiter="minmod", gradient="gg", roe_bdiss=1, cfl=0.4, bc=STEP_BCS, fs=fs)
    U = U0.copy()
    bad_at = None
    for n in range(60):
        om.step(make_cfg(**kw), U, rk=4, lts=0)
        rho, p = rho_p(U)
        if not (np.all(rho > 0) and np.all(p > 0)):
            bad_at = n
            break
    assert bad_at is not None and bad_at < 40  # the corner cells go non-physical
    U = U0.copy()
    flagged = 0
    for n in range(100):
        err, _ = om.step(make_cfg(positivity=1, **kw), U, rk=4, lts=0)
        flagged += orc.last_fallback_cells()
        assert np.isfinite(err[0]) and np.all(np.isfinite(U))  # Error[2] divides by rho v = 0 (reference norm)
    rho, p = rho_p(U)
    assert flagged > 0 and rho.min() > 0.5 and p.min() > 0.5


def test_blast_in_a_closed_box_stays_physical_and_conservative(orc):
    om, U0 = blast_case(orc)
    area = om.arrays().area[: om.nphys]
    kw = dict(flux="hllc", order=2, limiter="vanleer", gradient="gg", cfl=0.4, bc=("wall",) * 5)
    U = U0.copy()
    err, _ = om.step(make_cfg(**kw), U, rk=3, lts=0)
    assert not np.all(np.isfinite(U))  # without the fallback the first step breaks down
    U = U0.copy()
    tot0 = (U * area[:, None]).sum(0)
    flagged = 0
    for _ in range(30):
        om.step(make_cfg(positivity=1, **kw), U, rk=3, lts=0)
        flagged += orc.last_fallback_cells()
    rho, p = rho_p(U)
    tot = (U * area[:, None]).sum(0)
    assert flagged > 0 and rho.min() > 0 and p.min() > 0
    # slip walls pass no mass or energy; recomputed faces change both of their cells' fluxes
    assert abs(tot[0] - tot0[0]) <= 1e-12 * tot0[0]
    assert abs(tot[3] - tot0[3]) <= 1e-12 * tot0[3]


@pytest.mark.parametrize("flux,limiter,gradient", [("hllc", "vanleer", "wlsq"), ("roe", "minmod", "gg")])
def test_fallback_leaves_smooth_flow_unchanged(orc, flux, limiter, gradient):
    """A smooth entropy wave carried by a Mach 1.65 stream through a jittered mixed mesh (far-field
    ghosts all round): no stage state leaves its neighbourhood's range, so the fallback flags
    nothing and the march is the plain one bit for bit."""
    X, Y = np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 0.5, 21))
    m = meshgen.structured_to_mixed(X, Y, None, 9, 0.1, 0.3, "channel")
    om = orc.mesh(m.points, m.cells)
    ctr = om.arrays().ctr[: om.nphys]
    x, y = ctr[:, 0], ctr[:, 1]
    fs = freestream(1.65)
    rho = fs[0] * (1 + 0.05 * np.exp(-((x - 0.4) ** 2 + (y - 0.25) ** 2) / 0.02))
    U0 = np.ascontiguousarray(cons_from_prim(rho, fs[1] + 0 * x, 0 * x, fs[3] + 0 * x))
    kw = dict(flux=flux, order=2, limiter=limiter, gradient=gradient, roe_bdiss=1, cfl=0.6,
              bc=("freestream",) * 5, fs=fs)
    Ua, Ub = U0.copy(), U0.copy()
    for _ in range(10):
        om.step(make_cfg(**kw), Ua, rk=3, lts=0)
        om.step(make_cfg(positivity=1, **kw), Ub, rk=3, lts=0)
        assert orc.last_fallback_cells() == 0
    assert np.array_equal(Ua, Ub)
