"""ctypes bindings of the CPU oracle (oracle/cfd_oracle.c) and of the reference build.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg. The product package never imports this module.

`Oracle` wraps oracle/_build/liboracle.so (built by build_oracle() below).
`Reference` wraps oracle/_ref/libcfdref.so, the reference solver compiled from its own sources
by oracle/Makefile.ref; it exists only where that build ran.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcfdref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

OBC = dict(inlet_ref=1, exit=2, wall=3, symmetry=4, freestream=5, farfield=6, subsonic_inlet_ref=7,
           noslip_wall=8, subsonic_exit=9, subsonic_inlet=10)
OVISC = dict(off=0, constant=1, sutherland=2)
OFLUX = dict(roe=0, hllc=1, llf=2, ausm=3)
OLIM = dict(none=0, minmod=1, vanleer=2, superbee=3, vanalbada=4)
OGRAD = dict(gg=0, wlsq=1)


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class ONatCfg(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_int), ("order", ctypes.c_int), ("limiter", ctypes.c_int),
                ("gradient", ctypes.c_int), ("roe_bdiss", ctypes.c_int), ("cfl", ctypes.c_double),
                ("bc", ctypes.c_int * 6), ("fs", ctypes.c_double * 4), ("freeze", ctypes.c_int),
                ("phi", ctypes.c_void_p), ("viscous", ctypes.c_int), ("mu", ctypes.c_double),
                ("prandtl", ctypes.c_double), ("r_gas", ctypes.c_double), ("positivity", ctypes.c_int), ("p_exit", ctypes.c_double)]


def make_cfg(flux="roe", order=1, limiter="vanleer", gradient="gg", roe_bdiss=0, cfl=0.5,
             bc=("inlet_ref", "exit", "wall", "wall", "wall"), fs=(1.0, 0.0, 0.0, 1.0), viscous="off",
             mu=0.0, prandtl=0.72, r_gas=287.0, positivity=0, p_exit=0.0) -> ONatCfg:
    c = ONatCfg()
    c.flux = OFLUX[flux]
    c.order = order
    c.limiter = OLIM[limiter]
    c.gradient = OGRAD[gradient]
    c.roe_bdiss = roe_bdiss
    c.cfl = cfl
    c.bc[0] = 0
    for k, name in enumerate(bc):
        c.bc[k + 1] = OBC[name]
    for k in range(4):
        c.fs[k] = fs[k]
    c.viscous = OVISC[viscous] if isinstance(viscous, str) else int(viscous)
    c.mu = mu
    c.prandtl = prandtl
    c.r_gas = r_gas
    c.positivity = int(positivity)
    c.p_exit = float(p_exit)
    return c


ORACLE_SOURCES = ("cfd_oracle.c", "ref_layout.c")


def build_oracle(force: bool = False) -> str:
    """Compiles the oracle's C sources with gcc (contraction off, like the reference build)."""
    srcs = [os.path.join(HERE, f) for f in ORACLE_SOURCES]
    os.makedirs(os.path.dirname(ORACLE_SO), exist_ok=True)
    newest = max(os.path.getmtime(f) for f in srcs + [os.path.join(HERE, "cfd_oracle.h")])
    if force or not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < newest:
        subprocess.check_call(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                               "-fopenmp", "-Wall", "-o", ORACLE_SO] + srcs + ["-lm"])
    return ORACLE_SO


@dataclass
class MeshArrays:
    """Per-cell mesh arrays in the reference's layout (4 face slots per cell)."""
    nphys: int
    nghost: int
    nf: np.ndarray
    nodes: np.ndarray
    nbr: np.ndarray
    nrm: np.ndarray
    length: np.ndarray
    ctr: np.ndarray
    area: np.ndarray
    inv_area: np.ndarray
    dist: np.ndarray
    kind: np.ndarray
    inlet: np.ndarray
    exit: np.ndarray
    wall: np.ndarray
    extra: dict = field(default_factory=dict)

    def slot_mask(self) -> np.ndarray:
        return np.arange(4)[None, :] < self.nf[:, None]


def _alloc_mesh_arrays(counts):
    nphys, nghost, nin, nex, nwall = counts
    nt = nphys + nghost
    return MeshArrays(
        nphys, nghost,
        np.zeros(nphys, np.int32), -np.ones((nphys, 4), np.int32), -np.ones((nphys, 4), np.int32),
        np.zeros((nphys, 8)), np.zeros((nphys, 4)), np.zeros((nt, 3)), np.zeros(nt),
        np.zeros(nt), np.zeros((nphys, 4)), np.zeros((nphys, 4), np.int32),
        np.zeros((nin, 3), np.int32), np.zeros((nex, 3), np.int32), np.zeros((nwall, 3), np.int32))


def _export_args(a: MeshArrays):
    return (_i(a.nf), _i(a.nodes), _i(a.nbr), _d(a.nrm), _d(a.length), _d(a.ctr), _d(a.area),
            _d(a.inv_area), _d(a.dist), _i(a.kind), _i(a.inlet), _i(a.exit), _i(a.wall))


class Oracle:
    def __init__(self):
        self.lib = ctypes.CDLL(build_oracle())
        L = self.lib
        L.om_build.restype = ctypes.c_void_p
        L.om_build.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip]
        L.om_free.argtypes = [ctypes.c_void_p]
        L.om_counts.argtypes = [ctypes.c_void_p] + [_ip] * 5
        L.om_export.argtypes = [ctypes.c_void_p] + [_ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip]
        L.or_prim.argtypes = [_dp, _dp]
        L.on_ghost_state.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.om_nfaces.restype = ctypes.c_int
        L.om_nfaces.argtypes = [ctypes.c_void_p]
        L.on_ghost_cons.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.or_weno5.restype = ctypes.c_double
        L.or_weno5.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int]
        L.or_limiter_slope.restype = ctypes.c_double
        L.or_limiter_slope.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        cfgp = ctypes.POINTER(ONatCfg)
        L.on_residual.argtypes = [ctypes.c_void_p, cfgp, _dp, _dp, _dp, _dp, _dp]
        L.on_gradients.argtypes = [ctypes.c_void_p, cfgp, _dp, _dp]
        L.on_face_flux.argtypes = [cfgp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.on_step.argtypes = [ctypes.c_void_p, cfgp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.on_refstep.argtypes = [ctypes.c_void_p, cfgp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.on_last_fallback_cells.restype = ctypes.c_int
        L.on_last_fallback_cells.argtypes = []
        L.or_threads.restype = ctypes.c_int
        L.or_threads.argtypes = []
        L.or_error.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_rk_stage.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.or_gg_gradient.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp]
        L.or_ausm_cell.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_roe2o_residual.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.or_apply_bc_ref.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        # reference-layout restatements (oracle/ref_layout.c)
        L.rl_flux.argtypes = [ctypes.c_int, _dp, _dp, _ip, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.rl_roe2.argtypes = [_dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.rl_phi.restype = ctypes.c_double
        L.rl_phi.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.rl_limiter.argtypes = [_dp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.rl_venkat.argtypes = [_dp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.rl_bj.argtypes = [_dp, _ip, _dp, _dp, _dp, ctypes.c_int]
        L.rl_muscl.argtypes = [_dp, _ip, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double]
        L.rl_bc.argtypes = [ctypes.c_int, _dp, _ip, _dp, _dp, ctypes.c_int]
        L.rl_rk3_stage.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.rl_rk4_stage.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]

    # ---- mesh ----
    def mesh(self, points: np.ndarray, cells: np.ndarray) -> "OracleMesh":
        return OracleMesh(self, points, cells)

    def threads(self) -> int:
        """Threads of the residual's parallel loops (OMP_NUM_THREADS, else all host threads)."""
        return int(self.lib.or_threads())

    def last_fallback_cells(self) -> int:
        """Cells the positivity fallback recomputed at first order during the last step."""
        return int(self.lib.on_last_fallback_cells())

    def prim(self, U4) -> np.ndarray:
        U = np.ascontiguousarray(U4, dtype=np.float64)
        P = np.zeros(11)
        self.lib.or_prim(_d(U), _d(P))
        return P

    def prims(self, U: np.ndarray) -> np.ndarray:
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 4)
        out = np.zeros((len(U), 11))
        for i in range(len(U)):
            self.lib.or_prim(_d(U[i]), _d(out[i]))
        return out

    def weno5(self, a, b, c, d, e, shift: int) -> float:
        return self.lib.or_weno5(a, b, c, d, e, shift)

    def limiter_slope(self, lim: str, a: float, b: float) -> float:
        return self.lib.or_limiter_slope(OLIM[lim], a, b)

    def face_flux(self, cfg: ONatCfg, WL, WR, nx, ny, dl, boundary=0) -> np.ndarray:
        WL = np.ascontiguousarray(WL, np.float64)
        WR = np.ascontiguousarray(WR, np.float64)
        F = np.zeros(4)
        self.lib.on_face_flux(ctypes.byref(cfg), boundary, _d(WL), _d(WR), nx, ny, dl, _d(F))
        return F

    @staticmethod
    def _fs5(fs, p_exit):
        """(rho, u, v, p) and the exits' static pressure (p_exit, else p)."""
        f = [float(x) for x in fs][:4]
        return np.ascontiguousarray(f + [p_exit if p_exit else f[3]], np.float64)

    def ghost_state(self, bc: str, W5, nx, ny, fs, p_exit=0.0) -> np.ndarray:
        W = np.ascontiguousarray(W5, np.float64)
        f = self._fs5(fs, p_exit)
        out = np.zeros(5)
        self.lib.on_ghost_state(OBC[bc], _d(W), nx, ny, _d(f), _d(out))
        return out

    def ghost_cons(self, bc: str, U4, nx, ny, fs, p_exit=0.0) -> np.ndarray:
        U = np.ascontiguousarray(U4, np.float64)
        f = self._fs5(fs, p_exit)
        out = np.zeros(4)
        self.lib.on_ghost_cons(OBC[bc], _d(U), nx, ny, _d(f), _d(out))
        return out

    def rk_stage(self, rk, stage, Un, Us, acc, R, h, out):
        """One integrator stage over n cells (arrays [n,4], h [n]); acc and out updated in place."""
        self.lib.or_rk_stage(rk, stage, len(h), _d(Un), _d(Us), _d(acc), _d(R), _d(h), _d(out))

    def error(self, dU: np.ndarray, U: np.ndarray) -> np.ndarray:
        dU = np.ascontiguousarray(dU, np.float64)
        U = np.ascontiguousarray(U, np.float64)
        e = np.zeros(4)
        self.lib.or_error(len(U) // 4 if U.ndim == 1 else U.shape[0], _d(dU), _d(U), _d(e))
        return e


class OracleMesh:
    def __init__(self, orc: Oracle, points: np.ndarray, cells: np.ndarray):
        self.orc = orc
        self.points = np.ascontiguousarray(points, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int32)
        self.h = orc.lib.om_build(len(self.points), _d(self.points), len(self.cells), _i(self.cells))
        c = [ctypes.c_int() for _ in range(5)]
        orc.lib.om_counts(self.h, *[ctypes.byref(x) for x in c])
        self.counts = tuple(x.value for x in c)
        self.nphys, self.nghost = self.counts[0], self.counts[1]
        self._arrays = None

    def __del__(self):
        try:
            self.orc.lib.om_free(self.h)
        except Exception:
            pass

    @property
    def nfaces(self) -> int:
        return self.orc.lib.om_nfaces(self.h)

    def arrays(self) -> MeshArrays:
        if self._arrays is None:
            a = _alloc_mesh_arrays(self.counts)
            self.orc.lib.om_export(self.h, *_export_args(a))
            self._arrays = a
        return self._arrays

    def residual(self, cfg: ONatCfg, U: np.ndarray, U_exit_src: np.ndarray | None = None):
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros((self.nphys, 4))
        dt = np.zeros(self.nphys)
        grad = np.zeros((self.nphys, 8))
        src = None if U_exit_src is None else np.ascontiguousarray(U_exit_src, np.float64)
        self.orc.lib.on_residual(self.h, ctypes.byref(cfg), _d(U), None if src is None else _d(src),
                                 _d(R), _d(dt), _d(grad))
        return R, dt, grad

    def gradients(self, cfg: ONatCfg, W5: np.ndarray) -> np.ndarray:
        W5 = np.ascontiguousarray(W5, np.float64)
        g = np.zeros((self.nphys, 8))
        self.orc.lib.on_gradients(self.h, ctypes.byref(cfg), _d(W5), _d(g))
        return g

    def step(self, cfg: ONatCfg, U: np.ndarray, rk=3, lts=0, ref=False):
        """Advances U in place; returns (err4, dt)."""
        assert U.flags.c_contiguous and U.dtype == np.float64
        err = np.zeros(4)
        dt = ctypes.c_double()
        fn = self.orc.lib.on_refstep if ref else self.orc.lib.on_step
        fn(self.h, ctypes.byref(cfg), rk, lts, _d(U), _d(err), ctypes.byref(dt))
        return err, dt.value

    def gg_gradient_ref(self, prim_all: np.ndarray, var: int) -> np.ndarray:
        prim_all = np.ascontiguousarray(prim_all, np.float64)
        g = np.zeros((self.nphys, 2))
        self.orc.lib.or_gg_gradient(self.h, _d(prim_all), var, _d(g))
        return g

    def ausm_ref(self, U_all: np.ndarray) -> np.ndarray:
        U_all = np.ascontiguousarray(U_all, np.float64)
        R = np.zeros((self.nphys, 4))
        self.orc.lib.or_ausm_cell(self.h, _d(U_all), _d(R))
        return R

    def roe2o_ref(self, U_all, prim_all, zeta, zeta1, cfl):
        U_all = np.ascontiguousarray(U_all, np.float64)
        prim_all = np.ascontiguousarray(prim_all, np.float64)
        R = np.zeros((self.nphys, 4))
        dt = np.zeros(self.nphys)
        self.orc.lib.or_roe2o_residual(self.h, _d(U_all), _d(prim_all), zeta, zeta1, cfl, _d(R), _d(dt))
        return R, dt

    def apply_bc_ref(self, U_all, prim_all, rho_in, p_in, inlet="inlet_ref", U_exit_src=None):
        src = None if U_exit_src is None else np.ascontiguousarray(U_exit_src, np.float64)
        self.orc.lib.or_apply_bc_ref(self.h, OBC[inlet], rho_in, p_in, _d(U_all), _d(prim_all),
                                     None if src is None else _d(src))


class Reference:
    """The reference solver built from /root/reference/src (oracle/Makefile.ref)."""

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def __init__(self):
        self.lib = ctypes.CDLL(REF_SO)
        L = self.lib
        L.ref_load_vtk.argtypes = [ctypes.c_char_p]
        L.ref_counts.argtypes = [_ip] * 5
        L.ref_export_mesh.argtypes = [_ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip]
        L.ref_configure.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int]
        L.ref_init_state.argtypes = [_dp]
        L.ref_set_state_U.argtypes = [_dp]
        L.ref_residual.argtypes = [_dp, _dp]
        L.ref_ausm.argtypes = [_dp]
        L.ref_weno.restype = ctypes.c_double
        L.ref_weno.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int]
        L.ref_step.argtypes = [_dp, _dp, _dp]
        L.ref_get_state.argtypes = [_dp, _dp]
        L.ref_get_delU.argtypes = [_dp]
        L.ref_gg_gradient.argtypes = [ctypes.c_int, _dp]
        L.ref_prim.argtypes = [_dp, _dp]
        L.ref_write_solution.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ref_write_error.argtypes = [ctypes.c_char_p]
        self.counts = None

    def load_vtk(self, path: str):
        rc = self.lib.ref_load_vtk(path.encode())
        if rc != 0:
            raise RuntimeError(f"reference failed to load {path}")
        c = [ctypes.c_int() for _ in range(5)]
        self.lib.ref_counts(*[ctypes.byref(x) for x in c])
        self.counts = tuple(x.value for x in c)
        self.nphys, self.nghost = self.counts[0], self.counts[1]

    def arrays(self) -> MeshArrays:
        a = _alloc_mesh_arrays(self.counts)
        self.lib.ref_export_mesh(*_export_args(a))
        return a

    def configure(self, dissipation=3, second_order=0, time_accurate=1, lts=0, cfl=0.5,
                  zeta=0.5, zeta1=0.5, rho_in=1.0, p_in=1.0, inlet_subsonic=0, exit_subsonic=0):
        self.lib.ref_configure(dissipation, second_order, time_accurate, lts, cfl, zeta, zeta1,
                               rho_in, p_in, inlet_subsonic, exit_subsonic)

    def init_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, np.float64)
        self.lib.ref_init_state(_d(U))

    def set_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, np.float64)
        self.lib.ref_set_state_U(_d(U))

    def apply_bc(self):
        self.lib.ref_apply_bc()

    def residual(self):
        R = np.zeros((self.nphys, 4))
        dt = np.zeros(self.nphys)
        self.lib.ref_residual(_d(R), _d(dt))
        return R, dt

    def ausm(self):
        R = np.zeros((self.nphys, 4))
        self.lib.ref_ausm(_d(R))
        return R

    def step(self):
        err = np.zeros(4)
        mdt = ctypes.c_double()
        tt = ctypes.c_double()
        self.lib.ref_step(_d(err), ctypes.byref(mdt), ctypes.byref(tt))
        return err, mdt.value, tt.value

    def write_solution(self, path: str, iterations: int, header=(0, 0, 0)):
        """Write_Solution(path, 1) of the reference's current state."""
        self.lib.ref_write_solution(path.encode(), iterations, *[int(h) for h in header])

    def write_error(self, path: str):
        self.lib.ref_write_error(path.encode())

    def state(self):
        nt = self.nphys + self.nghost
        U = np.zeros((nt, 4))
        P = np.zeros((nt, 11))
        self.lib.ref_get_state(_d(U), _d(P))
        return U, P

    def weno(self, a, b, c, d, e, shift: int) -> float:
        """The reference's WENO_Reconstruction of one variable."""
        return self.lib.ref_weno(a, b, c, d, e, shift)

    def gg_gradient(self, var: int) -> np.ndarray:
        g = np.zeros((self.nphys, 2))
        self.lib.ref_gg_gradient(var, _d(g))
        return g

    def prim(self, U4) -> np.ndarray:
        U = np.ascontiguousarray(U4, np.float64)
        P = np.zeros(11)
        self.lib.ref_prim(_d(U), _d(P))
        return P


def cons_from_prim(rho, u, v, p):
    """Calculate_Computational_Variables case 2 (src/Basic_Functions.cpp:171-206), vectorised."""
    rho = np.asarray(rho, np.float64)
    vmag = 0.5 * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, (p / (1.4 - 1.0)) + rho * vmag], axis=-1)
