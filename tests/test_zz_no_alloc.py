"""cfd_residual_noalloc (the allocation-free residual) against cfd_residual, through the C-ABI.

Two calls in a row on one GPU, the second on a smaller mesh, each compared with a solver's
cfd_residual of the same state (same kernels: bit for bit). The state and outputs are device
buffers from cfd_device_alloc, in the library's renumbered cell order.
"""
import ctypes

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd._lib import CfdError, check, dptr, lib
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config

pytestmark = pytest.mark.gpu

BCS = ("freestream", "exit", "wall", "wall", "wall")


def _dev(nbytes):
    p = ctypes.c_void_p()
    check(lib().cfd_device_alloc(ctypes.c_size_t(nbytes), ctypes.byref(p)), "cfd_device_alloc")
    return p


def _state(mesh, mach, seed):
    ctr = mesh.arrays().ctr[: mesh.nphys]
    rng = np.random.default_rng(seed)
    fs = freestream(mach)
    x, y = ctr[:, 0], ctr[:, 1]
    rho = fs[0] * (1 + 0.1 * np.sin(3 * x) * np.cos(2 * y)) * (1 + 0.01 * rng.random(len(x)))
    u = fs[1] * (1 + 0.05 * np.cos(4 * y))
    v = 0.1 * np.sin(2 * x)
    p = fs[3] * (1 + 0.1 * np.cos(5 * x + y))
    return np.ascontiguousarray(cons_from_prim(rho, u, v, p)), fs


def _noalloc(mesh, cfg, U, with_dt=True):
    """cfd_residual_noalloc on U (caller order): R, dt (caller order) and the minimum dt."""
    n = mesh.nphys
    perm = mesh.permutation()  # new -> old
    Ur = np.ascontiguousarray(U[perm])
    dU, dR, ddt, dmin = _dev(Ur.nbytes), _dev(32 * n), _dev(8 * n), _dev(8)
    L = lib()
    try:
        check(L.cfd_memcpy_h2d(dU, dptr(Ur), ctypes.c_size_t(Ur.nbytes)), "h2d")
        check(L.cfd_residual_noalloc(mesh.h, ctypes.byref(cfg), dU, dR, ddt if with_dt else None, dmin, None),
              "cfd_residual_noalloc")
        check(L.cfd_synchronize(), "sync")
        Rr, dtr, m = np.zeros((n, 4)), np.zeros(n), np.zeros(1)
        check(L.cfd_memcpy_d2h(dptr(Rr), dR, ctypes.c_size_t(Rr.nbytes)), "d2h")
        check(L.cfd_memcpy_d2h(dptr(dtr), ddt, ctypes.c_size_t(dtr.nbytes)), "d2h")
        check(L.cfd_memcpy_d2h(dptr(m), dmin, ctypes.c_size_t(8)), "d2h")
    finally:
        for p in (dU, dR, ddt, dmin):
            L.cfd_device_free(p)
    R, dt = np.zeros_like(Rr), np.zeros_like(dtr)
    R[perm] = Rr
    dt[perm] = dtr
    return R, dt, float(m[0])


@pytest.mark.parametrize("scheme", [dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq"),
                                    dict(flux="roe", order=1)])
def test_noalloc_residual_matches_cfd_residual_twice(scheme):
    for nx, ny, seed in ((160, 80, 1), (60, 24, 2)):  # the second call on a smaller mesh
        print(f"case noalloc residual {scheme['flux']} order {scheme['order']} mesh {nx}x{ny}", flush=True)
        m = meshgen.bump_mesh(nx, ny, 0.06, seed=seed)
        mesh = Mesh(m.points, m.cells, tile=128, face_budget=256)
        U, fs = _state(mesh, 1.8, seed)
        cfg = make_config(cfl=0.7, rk=3, lts=1, roe_bdiss=1, bc=BCS, fs=fs, block=128, **scheme)
        s = Solver(mesh, cfg)
        s.set_state(U)
        R0, dt0, _ = s.residual()
        R1, dt1, dmin = _noalloc(mesh, cfg, U)
        assert np.array_equal(R1, R0)
        assert np.array_equal(dt1, dt0)
        assert dmin == dt0.min()


def test_noalloc_rejects_what_it_cannot_hold():
    m = meshgen.bump_mesh(24, 12, 0.06, seed=3)
    mesh = Mesh(m.points, m.cells, tile=64)
    U, fs = _state(mesh, 1.8, 3)
    for bad in (dict(order=3), dict(ref_mode=1, order=1)):
        cfg = make_config(**dict(dict(flux="roe", order=2, cfl=0.5, bc=BCS, fs=fs, block=128), **bad))
        with pytest.raises(CfdError):
            _noalloc(mesh, cfg, U)
    # more cells than the scratch array holds (CFD_NOALLOC_MAX_CELLS = 2^20): rejected before any launch
    big = meshgen.bump_mesh(1200, 700, 0.06, seed=4)
    assert len(big.cells) > (1 << 20)
    mesh = Mesh(big.points, big.cells, tile=128, face_budget=256)
    U = np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (mesh.nphys, 1))
    with pytest.raises(CfdError, match="CFD_NOALLOC_MAX_CELLS"):
        _noalloc(mesh, make_config(flux="roe", order=1, cfl=0.5, bc=BCS, fs=fs, block=128), U)
