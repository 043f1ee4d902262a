"""Pins of the back-pressure exit and the total-condition subsonic inlet (oracle, CPU).

Neither exists in working form in the reference: its subsonic exit function is empty
(src/Exit_Boundary_Conditions.cpp:83-85; the characteristic formula at :8-66 is never called) and
its subsonic inlet has an energy error (src/Inlet_Boundary_Conditions.cpp:32). They are pinned by
what the physics fixes: a uniform subsonic stream is preserved, the exit's face state carries the
back pressure exactly, supersonic outflow is copied, the inlet is isentropic with the prescribed
total conditions, and a channel relaxes to the back pressure with the isentropic Mach number.
"""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import cons_from_prim
from oracle.oracle import Oracle, make_cfg

G = 1.4
BCS = ("subsonic_inlet", "subsonic_exit", "wall", "wall", "wall")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def w5(r, u, v, p):
    return np.array([r, u, v, p, np.sqrt(G * p / r)])


def p_of(U):
    return (G - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / U[0])


def test_exit_face_state_carries_the_back_pressure(orc):
    W = w5(1.2, 0.4, 0.1, 1.1)
    n = (0.8, 0.6)
    # interior pressure equal to the back pressure: the ghost is the interior state
    g = orc.ghost_state("subsonic_exit", W, *n, (1.4, 0.5, 0.0, 1.0), p_exit=1.1)
    assert np.allclose(g, W, rtol=1e-15, atol=1e-15)
    # otherwise the face state (U_cell + U_ghost) / 2 has exactly the back pressure, and the jump
    # across the face is the outgoing characteristic: dp = rho c du_n, drho = dp / c^2
    Uc = cons_from_prim(np.array([W[0]]), W[1], W[2], W[3])[0]
    Ug = orc.ghost_cons("subsonic_exit", Uc, *n, (1.4, 0.5, 0.0, 1.0), p_exit=0.9)
    Uf = 0.5 * (Uc + Ug)
    assert abs(p_of(Uf) - 0.9) <= 1e-14
    c = W[4]
    dp = 0.9 - W[3]
    assert abs(Uf[0] - (W[0] + dp / c ** 2)) <= 1e-14
    un_f = (Uf[1] * n[0] + Uf[2] * n[1]) / Uf[0]
    assert abs(un_f - (W[1] * n[0] + W[2] * n[1] - dp / (W[0] * c))) <= 1e-14
    # tangential velocity is kept
    ut_f = (-Uf[1] * n[1] + Uf[2] * n[0]) / Uf[0]
    assert abs(ut_f - (-W[1] * n[1] + W[2] * n[0])) <= 1e-14
    # supersonic outflow copies the interior state
    Wsup = w5(1.0, 2.0, 0.0, 1.0)
    assert np.array_equal(orc.ghost_state("subsonic_exit", Wsup, 1.0, 0.0, (1, 0, 0, 1), p_exit=0.5), Wsup)


def test_inlet_is_isentropic_with_the_prescribed_total_conditions(orc):
    fs = (1.4, 0.6, 0.2, 1.0)
    q2 = fs[1] ** 2 + fs[2] ** 2
    f = 1 + 0.5 * (G - 1) * q2 / (G * fs[3] / fs[0])
    p0, r0 = fs[3] * f ** (G / (G - 1)), fs[0] * f ** (1 / (G - 1))
    h0 = G / (G - 1) * p0 / r0
    # interior pressure equal to the prescribed one: the prescribed state itself
    g = orc.ghost_state("subsonic_inlet", w5(1.3, 0.1, 0.0, fs[3]), -1.0, 0.0, fs)
    assert np.allclose(g[:4], fs, rtol=1e-14, atol=1e-15)
    for p in (0.8, 0.95, 1.05, 1.2):
        g = orc.ghost_state("subsonic_inlet", w5(1.3, 0.1, 0.0, p), -1.0, 0.0, fs)
        r, u, v, pg = g[:4]
        assert pg == pytest.approx(min(p, p0), rel=1e-15)
        assert pg / r ** G == pytest.approx(p0 / r0 ** G, rel=1e-13)           # isentropic
        assert G / (G - 1) * pg / r + 0.5 * (u * u + v * v) == pytest.approx(h0, rel=1e-13)  # total enthalpy
        assert u * fs[2] - v * fs[1] == pytest.approx(0.0, abs=1e-14)           # the prescribed direction
    # at or above the total pressure the inflow stagnates
    g = orc.ghost_state("subsonic_inlet", w5(1.3, 0.1, 0.0, 2.0), -1.0, 0.0, fs)
    assert g[1] == 0.0 and g[2] == 0.0 and g[3] == pytest.approx(p0, rel=1e-15)


@pytest.mark.parametrize("order", [1, 2])
def test_uniform_subsonic_stream_is_preserved(orc, order):
    X, Y = np.meshgrid(np.linspace(0, 2, 41), np.linspace(0, 0.5, 11))
    m = meshgen.structured_to_mixed(X, Y, None, 4, 0.1, 0.3, "channel")
    om = orc.mesh(m.points, m.cells)
    fs = (1.4, 0.5, 0.0, 1.0)
    U = np.ascontiguousarray(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (om.nphys, 1)))
    R, _, _ = om.residual(make_cfg(flux="hllc", order=order, limiter="vanleer", gradient="wlsq", roe_bdiss=1,
                                   bc=BCS, fs=fs), U)
    assert np.abs(R).max() <= 1e-12


def test_channel_relaxes_to_the_back_pressure(orc):
    """A straight channel fed by the total-condition inlet (Mach 0.4 prescribed state) and closed by
    a back pressure above the prescribed static pressure decelerates to the uniform isentropic
    state of that pressure: exit pressure = back pressure, Mach = M(p_exit / p0)."""
    X, Y = np.meshgrid(np.linspace(0, 1, 31), np.linspace(0, 0.2, 7))
    m = meshgen.structured_to_mixed(X, Y, None, 4, 0.0, 0.0, "channel")
    om = orc.mesh(m.points, m.cells)
    fs = (1.4, 0.4, 0.0, 1.0)
    pb = 1.03
    U = np.ascontiguousarray(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (om.nphys, 1)))
    cfg = make_cfg(flux="roe", order=1, roe_bdiss=1, cfl=0.8, bc=BCS, fs=fs, p_exit=pb)
    for _ in range(6000):
        err, _ = om.step(cfg, U, rk=3, lts=1)
    assert err[0] < 1e-8
    r, u, p = U[:, 0], U[:, 1] / U[:, 0], p_of(U.T)
    p0 = fs[3] * (1 + 0.2 * 0.4 ** 2) ** 3.5
    m_exp = np.sqrt(5 * ((p0 / pb) ** (1 / 3.5) - 1))
    assert np.allclose(p, pb, rtol=2e-3)
    assert np.allclose(u / np.sqrt(G * p / r), m_exp, rtol=1e-2)
