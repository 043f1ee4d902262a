"""Pins of the reference-layout restatement (oracle/ref_layout.c: the launch_* kernels' formulas
from CUDA_KERNELS/, which the reference never ran on a CPU) by what the mathematics fixes: the
first-order fluxes equal the native face fluxes (themselves pinned by the reference and by exact
Riemann solutions); limiter functions at known r; limiters, MUSCL and Barth-Jespersen on linear
data; Venkatakrishnan's limits; the boundary states' invariants; the RK stage wrappers' order of
accuracy; the second-order Roe flux's first-order and uniform-state limits. CPU only."""
import ctypes
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, make_cfg

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
G = 1.4


def d(a):
    return a.ctypes.data_as(_dp)


def i(a):
    return a.ctypes.data_as(_ip)


@pytest.fixture(scope="module")
def L():
    return Oracle().lib


def row_mesh(n, h=0.1):
    """n quads in a row, faces (bottom, right, top, left); neighbours across right/left."""
    nb = -np.ones((n, 4), np.int32)
    for c in range(n):
        if c + 1 < n:
            nb[c, 1] = c + 1
        if c > 0:
            nb[c, 3] = c - 1
    nrm = np.tile(np.array([0, -1, 1, 0, 0, 1, -1, 0], np.float64), (n, 1))
    dist = np.full((n, 4), h)
    return np.ascontiguousarray(nb), np.ascontiguousarray(nrm), np.ascontiguousarray(dist)


def cons(rho, u, v, p):
    return np.array([rho, rho * u, rho * v, p / (G - 1) + 0.5 * rho * (u * u + v * v)])


def p6(W):
    rho, u, v, p = W
    return [rho, u, v, p, p / rho / 287.0, math.sqrt(G * p / rho)]


@pytest.mark.parametrize("scheme,flux", [(0, "roe"), (1, "hllc"), (2, "llf")])
def test_first_order_fluxes_equal_native_face_fluxes(L, scheme, flux):
    """Two cells across one face: the reference-layout HLLC and LLF equal the native face fluxes
    to rounding. The reference's Roe kernel text forms 0.5 (F_L + F_R - D) with
    D = 0.5 sum |lambda| alpha R (Roe_Flux_Cuda_Kernels.cu:197-238), i.e. half the dissipation of
    its CPU Roe: it equals the mean of the native Roe flux and the central average."""
    orc = Oracle()
    cfg = make_cfg(flux=flux, roe_bdiss=1)
    rng = np.random.default_rng(scheme)
    for _ in range(30):
        WL = [rng.uniform(0.3, 2.5), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.3, 2.5)]
        WR = [rng.uniform(0.3, 2.5), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.3, 2.5)]
        t = rng.uniform(0, 2 * math.pi)
        nx, ny = math.cos(t), math.sin(t)
        U = np.ascontiguousarray(np.stack([cons(*WL), cons(*WR)]))
        P = np.ascontiguousarray(np.array([p6(WL), p6(WR)]))
        nb = np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32)
        nrm = np.zeros((2, 8))
        nrm[0, 0:2] = nx, ny
        nrm[1, 0:2] = -nx, -ny
        ar = np.array([[0.7, 0, 0, 0], [0.7, 0, 0, 0]])
        out = np.zeros((2, 4, 4))
        L.rl_flux(scheme, d(U), d(P), i(nb), d(nrm), d(ar), d(out), 2, G)
        w = lambda W: np.array(list(W) + [math.sqrt(G * W[3] / W[0])])
        ref = orc.face_flux(cfg, w(WL), w(WR), nx, ny, 0.7)
        if flux == "roe":
            central = orc.face_flux(make_cfg(flux="roe", roe_bdiss=0), w(WL), w(WR), nx, ny, 0.7, boundary=1)
            ref = 0.5 * (ref + central)
        scale = np.abs(ref).max() + 1.0
        assert np.allclose(out[0, 0], ref, rtol=0, atol=1e-12 * scale), (flux, out[0, 0], ref)
        assert np.allclose(out[1, 0], -ref, rtol=0, atol=1e-12 * scale)


PHI = {0: lambda r: max(0.0, min(1.0, r)), 1: lambda r: 0.0 if r <= 0 else 2 * r / (1 + r),
       2: lambda r: max(0.0, min(2 * r, 1.0), min(r, 2.0)), 3: lambda r: 0.0 if r <= 0 else (r * r + r) / (r * r + 1)}


@pytest.mark.parametrize("typ", sorted(PHI))
def test_limiter_functions_at_known_ratios(L, typ):
    L.rl_phi.restype = ctypes.c_double
    for r in (-1.0, 0.0, 0.5, 1.0, 2.0, 3.0):
        for b in (1.0, -0.3):
            assert L.rl_phi(typ, r * b, b) == pytest.approx(PHI[typ](r), rel=1e-14, abs=1e-15), (typ, r)
    assert L.rl_phi(typ, 1.0, 0.0) == 0.0  # no forward difference: no slope


@pytest.mark.parametrize("typ", [0, 1, 2, 3])
def test_limiters_on_linear_data(L, typ):
    """Linear data on a row of equal quads: r = 1 on interior faces toward a neighbour, so every
    limiter returns phi(1) = 1; faces without a backward neighbour get 0."""
    n = 8
    nb, nrm, dist = row_mesh(n)
    x = np.arange(n) * 0.1
    U = np.ascontiguousarray(np.stack([1 + 0.5 * x, 0.3 + x, -0.2 * x, 2.5 + 0.1 * x], 1))
    phi = np.zeros((n, 4, 4))
    L.rl_limiter(d(U), i(nb), d(dist), d(phi), n, 1.0, typ, 0)
    assert np.allclose(phi[1:-1, 1], 1.0) and np.allclose(phi[1:-1, 3], 1.0)
    assert np.all(phi[0, 1] == 0.0) and np.all(phi[:, 0] == 0.0)


def test_venkatakrishnan_limits(L):
    """No backward difference gives 0; a very large K (epsilon) returns the backward difference;
    the expression is odd in the data."""
    n = 3
    nb, nrm, dist = row_mesh(n)
    vol = np.full(n, 0.01)
    for K, check in ((1.0, "zero"), (1e6, "eps")):
        U = np.ascontiguousarray(np.array([[1.0] * 4, [1.0] * 4, [1.3] * 4]))
        phi = np.zeros((n, 4, 4))
        L.rl_venkat(d(U), i(nb), d(vol), d(phi), n, K)
        if check == "zero":
            assert np.allclose(phi[1, 1], 0.0)       # dm = U1 - U0 = 0
        U2 = np.ascontiguousarray(np.array([[1.0] * 4, [1.2] * 4, [1.3] * 4]))
        phi2 = np.zeros((n, 4, 4))
        L.rl_venkat(d(U2), i(nb), d(vol), d(phi2), n, K)
        if check == "eps":
            assert np.allclose(phi2[1, 1], 0.2, rtol=1e-6)  # -> dm
        U3 = np.ascontiguousarray(-U2)
        phi3 = np.zeros((n, 4, 4))
        L.rl_venkat(d(U3), i(nb), d(vol), d(phi3), n, K)
        assert np.allclose(phi3[1, 1], -phi2[1, 1])


def test_barth_jespersen_linear_and_extremum(L):
    n = 6
    nb, nrm, dist = row_mesh(n)
    x = np.arange(n) * 0.1
    U = np.ascontiguousarray(np.stack([1 + x, 1 + x, 1 + x, 1 + x], 1))
    unlim = np.zeros((n, 4, 4))
    unlim[:, 1, :] = 0.5 / 0.1 * 0.1  # half the neighbour difference toward +x: within bounds
    unlim[:, 3, :] = -0.5 / 0.1 * 0.1
    phi = np.zeros((n, 4, 4))
    L.rl_bj(d(U), i(nb), d(dist), d(unlim), d(phi), n)
    assert np.allclose(phi[1:-1, 1], 1.0) and np.allclose(phi[1:-1, 3], 1.0)
    Um = np.ascontiguousarray(np.array([[1.0] * 4, [2.0] * 4, [1.0] * 4]))  # cell 1 is a maximum
    nb3, _, dist3 = row_mesh(3)
    un3 = np.zeros((3, 4, 4))
    un3[:, 1, :] = 1.0
    phi3 = np.zeros((3, 4, 4))
    L.rl_bj(d(Um), i(nb3), d(dist3), d(un3), d(phi3), 3)
    assert np.all(phi3[1, 1] == 0.0)  # an increase beyond the local maximum is cut to zero


def test_muscl_linear_data_face_values_agree(L):
    """phi = 1, kappa = 1 on linear data: both sides reconstruct the face midpoint value."""
    n = 6
    nb, nrm, dist = row_mesh(n)
    x = np.arange(n) * 0.1
    U = np.ascontiguousarray(np.stack([1 + x, 0.5 - x, 0.2 * x, 2 + x], 1))
    phi = np.ones((n, 4, 4))
    QL = np.zeros((n, 4, 4))
    QR = np.zeros((n, 4, 4))
    L.rl_muscl(d(U), i(nb), d(dist), d(phi), d(QL), d(QR), n, 1.0)
    mid = 0.5 * (U[:-1] + U[1:])
    assert np.allclose(QL[:-1, 1], mid) and np.allclose(QR[:-1, 1], mid)
    assert np.allclose(QL[:, 0], U) and np.allclose(QR[:, 0], U)  # boundary faces copy the cell


def test_boundary_states(L):
    rng = np.random.default_rng(4)
    W = [1.2, 0.7, -0.4, 0.9]
    bc = np.array([1.4, 2.0, 0.1, 1.0])
    for kind in (5, 6):  # inviscid wall, symmetry: V.n flips, rho and |V| kept
        t = rng.uniform(0, 2 * math.pi)
        nrm = np.zeros((1, 8))
        nrm[0, :2] = math.cos(t), math.sin(t)
        U = np.ascontiguousarray(cons(*W)[None, :])
        L.rl_bc(kind, d(U), i(np.array([0], np.int32)), d(nrm), d(bc), 1)
        u, v = U[0, 1] / U[0, 0], U[0, 2] / U[0, 0]
        n = nrm[0, :2]
        assert U[0, 0] == W[0]
        assert u * n[0] + v * n[1] == pytest.approx(-(W[1] * n[0] + W[2] * n[1]), abs=1e-14)
        assert -u * n[1] + v * n[0] == pytest.approx(-W[1] * n[1] + W[2] * n[0], abs=1e-14)
    U = np.ascontiguousarray(cons(*W)[None, :])
    L.rl_bc(1, d(U), i(np.array([0], np.int32)), d(np.zeros((1, 8))), d(bc), 1)
    assert np.allclose(U[0], cons(*bc))  # supersonic inlet: all prescribed
    U = np.ascontiguousarray(cons(*W)[None, :])
    U0 = U.copy()
    L.rl_bc(3, d(U), i(np.array([0], np.int32)), d(np.zeros((1, 8))), d(bc), 1)
    assert np.array_equal(U, U0)  # supersonic exit: unchanged
    nrm = np.zeros((1, 8))
    nrm[0, :2] = 1.0, 0.0
    U = np.ascontiguousarray(cons(*bc)[None, :])
    L.rl_bc(7, d(U), i(np.array([0], np.int32)), d(nrm), d(bc), 1)
    assert np.allclose(U[0], cons(*bc), rtol=1e-12)  # far field: freestream is a fixed point


@pytest.mark.parametrize("rk,order", [(3, 3), (4, 4)])
def test_rk_stage_wrappers_order(L, rk, order):
    """The reference-layout stages on du/dt = -R(u) (A u with a rotation and a decay): RK3 in the
    reference's delta form (U + dU) and classical RK4 converge at orders 3 and 4."""
    w = 2.0

    def R(u):
        return np.ascontiguousarray(-np.array([[-w * u[0, 1], w * u[0, 0], -0.5 * u[0, 2] + u[0, 3],
                                                -u[0, 2] - 0.5 * u[0, 3]]]))

    def exact(t):
        e, c, s = math.exp(-0.5 * t), math.cos(t), math.sin(t)
        return np.array([math.cos(w * t), math.sin(w * t), e * (0.3 * c - 0.2 * s), e * (-0.2 * c - 0.3 * s)])

    ia = np.ones(1)
    errs = []
    for nsteps in (40, 80, 160):
        h = 2.0 / nsteps
        U = np.array([[1.0, 0.0, 0.3, -0.2]])
        for _ in range(nsteps):
            if rk == 3:
                Urk = np.zeros_like(U)
                L.rl_rk3_stage(d(U), d(R(U)), None, d(Urk), d(ia), h, 1, 1)
                L.rl_rk3_stage(d(U), d(R(Urk)), None, d(Urk), d(ia), h, 1, 2)
                L.rl_rk3_stage(d(U), d(R(Urk)), None, d(Urk), d(ia), h, 1, 3)
                U = U + Urk
            else:
                Us = np.zeros_like(U)
                K = np.zeros_like(U)
                L.rl_rk4_stage(d(U), d(R(U)), d(Us), d(K), d(ia), h, 1, 1)
                for st in (2, 3, 4):
                    L.rl_rk4_stage(d(U), d(R(Us.copy())), d(Us), d(K), d(ia), h, 1, st)
                U = Us.copy()
        errs.append(np.abs(U[0] - exact(2.0)).max())
    rates = [math.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert all(abs(r - order) < 0.2 for r in rates), (rk, errs, rates)


def test_second_order_roe_limits(L):
    """rl_roe2: with zero limiter values it is the first-order Roe flux; on a uniform state it is
    the physical flux whatever the limiter values."""
    rng = np.random.default_rng(2)
    WL, WR = [1.1, 0.6, 0.2, 1.3], [0.8, 0.9, -0.1, 0.9]
    U = np.ascontiguousarray(np.stack([cons(*WL), cons(*WR)]))
    P = np.ascontiguousarray(np.array([p6(WL), p6(WR)]))
    nb = np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32)
    nrm = np.zeros((2, 8))
    nrm[0, :2] = 0.6, 0.8
    nrm[1, :2] = -0.6, -0.8
    ar = np.array([[0.5, 0, 0, 0], [0.5, 0, 0, 0]])
    ctr = np.array([[0.0, 0.0], [0.6, 0.8]])
    fctr = np.zeros((2, 4, 2))
    fctr[:, 0] = [0.3, 0.4]
    out1 = np.zeros((2, 4, 4))
    L.rl_flux(0, d(U), d(P), i(nb), d(nrm), d(ar), d(out1), 2, G)
    out2 = np.zeros((2, 4, 4))
    L.rl_roe2(d(U), i(nb), d(nrm), d(ar), d(np.zeros((2, 4, 4))), d(ctr), d(fctr), d(out2), 2, G)
    assert np.allclose(out2[0, 0], out1[0, 0], rtol=1e-12, atol=1e-12)
    Uu = np.ascontiguousarray(np.stack([cons(*WL), cons(*WL)]))
    out3 = np.zeros((2, 4, 4))
    L.rl_roe2(d(Uu), i(nb), d(nrm), d(ar), d(rng.uniform(0, 1, (2, 4, 4))), d(ctr), d(fctr), d(out3), 2, G)
    rho, u, v, p = WL
    vn = u * 0.6 + v * 0.8
    F = np.array([rho * vn, rho * u * vn + p * 0.6, rho * v * vn + p * 0.8, (cons(*WL)[3] + p) * vn]) * 0.5
    assert np.allclose(out3[0, 0], F, rtol=1e-12, atol=1e-13)
