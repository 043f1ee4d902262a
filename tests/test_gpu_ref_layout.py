"""GPU parity of the reference-layout entry points (include/cfd_hip_ref.h) against the oracle's
restatement (oracle/ref_layout.c), on a jittered mixed mesh in the reference's 4-slot layout.
FP64, bit-exact except the far-field BC (pow(): 1e-13 relative)."""
import ctypes

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd._lib import DeviceArray, lib
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def d(a):
    return a.ctypes.data_as(_dp)


def i(a):
    return a.ctypes.data_as(_ip)


@pytest.fixture(scope="module")
def setup():
    orc = Oracle()
    m = meshgen.bump_mesh(30, 12, 0.1, seed=21)
    om = orc.mesh(m.points, m.cells)
    a = om.arrays()
    N = a.nphys
    nb = a.nbr.copy()
    nb[nb >= N] = -1
    x, y = a.ctr[:N, 0], a.ctr[:N, 1]
    rho = 1.4 * (1 + 0.2 * np.tanh(8 * (x - 0.4)) + 0.05 * np.cos(3 * y))
    u = 1.2 + 0.3 * np.sin(2 * x)
    v = 0.2 * np.cos(3 * x + y)
    p = 1.0 + 0.3 * np.tanh(6 * (0.6 - x))
    vm = 0.5 * (u * u + v * v)
    U = np.ascontiguousarray(np.stack([rho, rho * u, rho * v, p / (1.4 - 1.0) + rho * vm], 1))
    pr = orc.prims(U)
    P6 = np.ascontiguousarray(np.stack([pr[:, 0], pr[:, 1], pr[:, 2], pr[:, 4], pr[:, 3], pr[:, 5]], 1))
    fctr = np.zeros((N, 4, 2))
    for c in range(N):
        for f in range(a.nf[c]):
            p0 = m.points[a.nodes[c, f], :2]
            p1 = m.points[a.nodes[c, (f + 1) % a.nf[c]], :2]
            fctr[c, f] = 0.5 * (p0 + p1)
    return dict(orc=orc, N=N, nb=np.ascontiguousarray(nb, np.int32), nrm=np.ascontiguousarray(a.nrm),
                ar=np.ascontiguousarray(a.length), dist=np.ascontiguousarray(np.where(a.dist > 0, a.dist, 1.0)),
                U=U, P=P6, ctr=np.ascontiguousarray(a.ctr[:N, :2]), fctr=np.ascontiguousarray(fctr),
                vol=np.ascontiguousarray(a.area[:N]), ia=np.ascontiguousarray(a.inv_area[:N]))


def dev(s, *names):
    return [DeviceArray(s[n]) for n in names]


@pytest.mark.parametrize("scheme,name", [(0, "launch_roe_flux"), (1, "launch_hllc_flux"), (2, "launch_llf_flux")])
def test_first_order_fluxes(setup, scheme, name):
    s = setup
    N = s["N"]
    dU, dP, dnb, dn, da = dev(s, "U", "P", "nb", "nrm", "ar")
    out0 = np.full((N, 4, 4), 7.0)
    dout = DeviceArray(out0)
    rc = getattr(lib(), name)(dU.ptr, dP.ptr, dnb.ptr, dn.ptr, da.ptr, dout.ptr, N, ctypes.c_double(1.4), 256, None)
    assert rc == 0
    ref = out0.copy()
    s["orc"].lib.rl_flux(scheme, d(s["U"]), d(s["P"]), i(s["nb"]), d(s["nrm"]), d(s["ar"]), d(ref), N, 1.4)
    assert np.array_equal(dout.get(), ref)
    assert (ref == 7.0).any()  # boundary faces are skipped (left untouched), as in the reference


@pytest.mark.parametrize("name,typ,three", [("launch_minmod_limiter", 0, 0), ("launch_minmod_limiter", 0, 1),
                                            ("launch_vanleer_limiter", 1, 0), ("launch_superbee_limiter", 2, 0),
                                            ("launch_vanalbada_limiter", 3, 0), ("launch_generic_limiter", 1, 0),
                                            ("launch_generic_limiter", 5, 0)])
def test_limiters(setup, name, typ, three):
    s = setup
    N = s["N"]
    dU, dnb, dd = dev(s, "U", "nb", "dist")
    dphi = DeviceArray(np.zeros((N, 4, 4)))
    L = lib()
    zeta = ctypes.c_double(0.5)
    if name == "launch_minmod_limiter":
        rc = L.launch_minmod_limiter(dU.ptr, dnb.ptr, dd.ptr, dphi.ptr, N, zeta, three, 128, None)
    elif name == "launch_generic_limiter":
        rc = L.launch_generic_limiter(dU.ptr, dnb.ptr, dd.ptr, dphi.ptr, N, typ, zeta, 128, None)
        typ = typ if typ <= 3 else 0
    else:
        rc = getattr(L, name)(dU.ptr, dnb.ptr, dd.ptr, dphi.ptr, N, zeta, 128, None)
    assert rc == 0
    ref = np.zeros((N, 4, 4))
    s["orc"].lib.rl_limiter(d(s["U"]), i(s["nb"]), d(s["dist"]), d(ref), N, 0.5, typ, three)
    got = dphi.get()
    assert np.array_equal(got, ref)
    assert (ref > 0).any() and (ref < 1).any()  # not the text's constant zero


def test_venkatakrishnan_and_barth_jespersen(setup):
    s = setup
    N = s["N"]
    L = lib()
    dU, dnb, dd, dv = dev(s, "U", "nb", "dist", "vol")
    dphi = DeviceArray(np.zeros((N, 4, 4)))
    assert L.launch_venkatakrishnan_limiter(dU.ptr, dnb.ptr, dd.ptr, dv.ptr, dphi.ptr, N, ctypes.c_double(0.3), 256, None) == 0
    ref = np.zeros((N, 4, 4))
    s["orc"].lib.rl_venkat(d(s["U"]), i(s["nb"]), d(s["vol"]), d(ref), N, 0.3)
    assert np.array_equal(dphi.get(), ref)
    unlim = np.ascontiguousarray(np.sin(np.arange(N * 16).reshape(N, 4, 4)) * 0.3)
    dun = DeviceArray(unlim)
    assert L.launch_barth_jespersen_limiter(dU.ptr, dnb.ptr, dd.ptr, dun.ptr, dphi.ptr, N, 256, None) == 0
    s["orc"].lib.rl_bj(d(s["U"]), i(s["nb"]), d(s["dist"]), d(unlim), d(ref), N)
    assert np.array_equal(dphi.get(), ref)


def test_muscl_and_second_order_roe(setup):
    s = setup
    N = s["N"]
    L = lib()
    phi = np.zeros((N, 4, 4))
    s["orc"].lib.rl_limiter(d(s["U"]), i(s["nb"]), d(s["dist"]), d(phi), N, 1.0, 1, 0)
    dU, dP, dnb, dn, da, dd, dc, dfc = dev(s, "U", "P", "nb", "nrm", "ar", "dist", "ctr", "fctr")
    dphi = DeviceArray(phi)
    dQL, dQR = DeviceArray(np.zeros((N, 4, 4))), DeviceArray(np.zeros((N, 4, 4)))
    assert L.launch_muscl_reconstruction(dU.ptr, dnb.ptr, dd.ptr, dphi.ptr, dQL.ptr, dQR.ptr, N, ctypes.c_double(0.5), 256, None) == 0
    QL, QR = np.zeros((N, 4, 4)), np.zeros((N, 4, 4))
    s["orc"].lib.rl_muscl(d(s["U"]), i(s["nb"]), d(s["dist"]), d(phi), d(QL), d(QR), N, 0.5)
    assert np.array_equal(dQL.get(), QL) and np.array_equal(dQR.get(), QR)
    dout = DeviceArray(np.zeros((N, 4, 4)))
    assert L.launch_roe_flux_2nd_order(dU.ptr, dP.ptr, dnb.ptr, dn.ptr, da.ptr, dphi.ptr, dc.ptr, dfc.ptr, dout.ptr, N, ctypes.c_double(1.4), 256, None) == 0
    ref = np.zeros((N, 4, 4))
    s["orc"].lib.rl_roe2(d(s["U"]), i(s["nb"]), d(s["nrm"]), d(s["ar"]), d(phi), d(s["ctr"]), d(s["fctr"]), d(ref), N, 1.4)
    assert np.array_equal(dout.get(), ref)


BCS = [("launch_subsonic_inlet_bc", 0), ("launch_supersonic_inlet_bc", 1), ("launch_subsonic_exit_bc", 2),
       ("launch_supersonic_exit_bc", 3), ("launch_viscous_wall_bc", 4), ("launch_inviscid_wall_bc", 5),
       ("launch_symmetry_bc", 6), ("launch_farfield_bc", 7)]


class Inlet(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("Rho", "u", "v", "P", "T")]


class Exit(ctypes.Structure):
    _fields_ = [("P", ctypes.c_double)]


class Far(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("Rho_inf", "u_inf", "v_inf", "P_inf", "Mach")]


@pytest.mark.parametrize("name,kind", BCS)
def test_boundary_conditions_in_place(setup, name, kind):
    s = setup
    N = s["N"]
    L = lib()
    cells = np.ascontiguousarray(np.arange(0, N, 7, dtype=np.int32))
    dU = DeviceArray(s["U"])
    dl, dn = DeviceArray(cells), DeviceArray(s["nrm"])
    bcv = np.array([1.3, 1.1, 0.05, 0.9])
    g = ctypes.c_double(1.4)
    n = len(cells)
    if kind == 0:
        rc = L.launch_subsonic_inlet_bc(dU.ptr, dl.ptr, dn.ptr, ctypes.byref(Inlet(1.3, 1.1, 0.05, 0.9, 1.0)), n, g, 256, None)
    elif kind == 1:
        rc = L.launch_supersonic_inlet_bc(dU.ptr, dl.ptr, ctypes.byref(Inlet(1.3, 1.1, 0.05, 0.9, 1.0)), n, g, 256, None)
    elif kind == 2:
        rc = L.launch_subsonic_exit_bc(dU.ptr, dl.ptr, dn.ptr, ctypes.byref(Exit(0.9)), n, g, 256, None)
    elif kind == 3:
        rc = L.launch_supersonic_exit_bc(dU.ptr, dl.ptr, n, 256, None)
    elif kind == 4:
        rc = L.launch_viscous_wall_bc(dU.ptr, dl.ptr, n, 256, None)
    elif kind in (5, 6):
        rc = getattr(L, name)(dU.ptr, dl.ptr, dn.ptr, n, 256, None)
    else:
        rc = L.launch_farfield_bc(dU.ptr, dl.ptr, dn.ptr, ctypes.byref(Far(1.3, 1.1, 0.05, 0.9, 1.2)), n, g, 256, None)
    assert rc == 0
    ref = s["U"].copy()
    s["orc"].lib.rl_bc(kind, d(ref), i(cells), d(s["nrm"]), d(np.ascontiguousarray(bcv)), n)
    got = dU.get()
    if kind == 7:
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    else:
        assert np.array_equal(got, ref)


def test_rk_stages(setup):
    s = setup
    N = s["N"]
    L = lib()
    R = np.ascontiguousarray(np.cos(np.arange(N * 4).reshape(N, 4)) * 0.01)
    dU, dR, dia = DeviceArray(s["U"]), DeviceArray(R), DeviceArray(s["ia"])
    Urk = DeviceArray(np.zeros((N, 4)))
    ref = np.zeros((N, 4))
    for st in (1, 2, 3):
        assert L.launch_tvd_rk3_stage(dU.ptr, dR.ptr, None, Urk.ptr, dia.ptr, ctypes.c_double(1e-3), N, st, 128, None) == 0
        s["orc"].lib.rl_rk3_stage(d(s["U"]), d(R), None, d(ref), d(s["ia"]), 1e-3, N, st)
        assert np.array_equal(Urk.get(), ref)
    Us, K = DeviceArray(np.zeros((N, 4))), DeviceArray(np.zeros((N, 4)))
    rUs, rK = np.zeros((N, 4)), np.zeros((N, 4))
    for st in (1, 2, 3, 4):
        assert L.launch_rk4_stage(dU.ptr, dR.ptr, Us.ptr, K.ptr, dia.ptr, ctypes.c_double(1e-3), N, st, 64, None) == 0
        s["orc"].lib.rl_rk4_stage(d(s["U"]), d(R), d(rUs), d(rK), d(s["ia"]), 1e-3, N, st)
        assert np.array_equal(Us.get(), rUs) and np.array_equal(K.get(), rK)


def test_argument_errors_and_empty_launches(setup):
    L = lib()
    assert L.launch_roe_flux(None, None, None, None, None, None, 0, ctypes.c_double(1.4), 256, None) == 0
    assert L.launch_roe_flux(None, None, None, None, None, None, 5, ctypes.c_double(1.4), 256, None) == -1
    assert L.launch_tvd_rk3_stage(None, None, None, None, None, ctypes.c_double(1.0), 3, 7, 256, None) == -1
    L.get_limiter_name.restype = ctypes.c_char_p
    assert L.get_limiter_name(2) == b"Superbee"
