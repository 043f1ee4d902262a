"""Host-side interface of the finite-volume path over libcfdhip.so.

It mirrors the reference's driver (rameshkolluru43/CFD_Solver_withCUDA):
  Mesh(points, cells)          Form_Cells (src/Grid_Computations.cpp:319-392)
  Solver.residual()            Evaluate_Cell_Net_Flux_1O/2O (src/Net_Flux.cpp:46-128)
  Solver.step()                one Inviscid_Solver iteration (src/Solver.cpp:93-118)
  Solver.run_steady(tol)       the iteration loop until Error[0] <= tol
with the reference's flags made explicit (flux, order, limiter, gradient, rk, lts, cfl).
All arithmetic runs in the HIP kernels; this module only moves arrays.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import CfdError, Config, check, dptr, iptr, lib, uptr

FLUX = dict(roe=0, hllc=1, llf=2, ausm=3)
LIMITER = dict(none=0, minmod=1, vanleer=2, superbee=3, vanalbada=4)
GRADIENT = dict(gg=0, wlsq=1)
BC = dict(inlet_ref=1, exit=2, wall=3, symmetry=4, freestream=5, farfield=6, subsonic_inlet_ref=7,
          noslip_wall=8, subsonic_exit=9, subsonic_inlet=10)
VISCOUS = dict(off=0, constant=1, sutherland=2)

# ghost types of the reference's face kinds (LEFT->Inlet, RIGHT->Exit, rest->Wall,
# src/Read_Gmsh_File.cpp:579-621)
REF_BCS = ("inlet_ref", "exit", "wall", "wall", "wall")


def make_config(flux="roe", order=1, limiter="vanleer", gradient="gg", roe_bdiss=0, cfl=0.5,
                bc=REF_BCS, fs=(1.0, 0.0, 0.0, 1.0), rk=3, lts=0, ref_mode=0, block=256,
                xcd_remap=16, grad_lds=1, freeze_tol=0.0, freeze_stall=0, zeta=0.5, zeta1=0.5,
                viscous="off", mu=0.0, prandtl=0.72, r_gas=287.0, positivity=0, p_exit=0.0) -> Config:
    """Scheme settings (cfd_config_t). viscous: "off" (Euler), "constant" (viscosity mu) or
    "sutherland" (Sutherland's law in T = p / (rho r_gas)); laminar viscous terms need order 2 or 3."""
    c = Config()
    c.flux = FLUX[flux]
    c.order = int(order)
    c.limiter = LIMITER[limiter]
    c.gradient = GRADIENT[gradient]
    c.roe_bdiss = int(roe_bdiss)
    c.cfl = float(cfl)
    c.bc[0] = 0
    if len(bc) != 5:
        raise ValueError("bc needs one ghost type per face kind LEFT, RIGHT, TOP, BOTTOM, WALL")
    for k, name in enumerate(bc):
        c.bc[k + 1] = BC[name]
    for k in range(4):
        c.fs[k] = float(fs[k])
    c.rk = int(rk)
    c.lts = int(lts)
    c.ref_mode = int(ref_mode)
    c.block = int(block)
    c.xcd_remap = int(xcd_remap)
    c.grad_lds = int(grad_lds)
    c.freeze_tol = float(freeze_tol)
    c.freeze_stall = int(freeze_stall)
    c.zeta = float(zeta)
    c.zeta1 = float(zeta1)
    c.viscous = VISCOUS[viscous] if isinstance(viscous, str) else int(viscous)
    c.mu = float(mu)
    c.prandtl = float(prandtl)
    c.r_gas = float(r_gas)
    c.positivity = int(positivity)
    c.p_exit = float(p_exit)
    return c


def freestream(mach: float, rho: float = 1.4, p: float = 1.0, angle_deg: float = 0.0):
    """Primitive freestream (rho, u, v, p) with sound speed sqrt(1.4 p / rho)."""
    a = np.sqrt(1.4 * p / rho)
    th = np.radians(angle_deg)
    return (rho, mach * a * np.cos(th), mach * a * np.sin(th), p)


def cons_from_prim(rho, u, v, p) -> np.ndarray:
    """Conserved variables, in the reference's expression order (src/Basic_Functions.cpp:171-206)."""
    rho = np.asarray(rho, np.float64)
    u = np.broadcast_to(np.asarray(u, np.float64), rho.shape)
    v = np.broadcast_to(np.asarray(v, np.float64), rho.shape)
    p = np.broadcast_to(np.asarray(p, np.float64), rho.shape)
    vmag = 0.5 * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, (p / (1.4 - 1.0)) + rho * vmag], axis=-1))


@dataclass
class MeshArrays:
    """The reference's per-cell mesh arrays (4 face slots per cell)."""
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


class Mesh:
    """A mixed triangle/quad mesh with the reference's connectivity, tiled for the GPU."""

    def __init__(self, points: np.ndarray, cells: np.ndarray, tile: int = 256, renumber: int = 1,
                 device: bool = True, face_budget: int = 0):
        L = lib()
        self.points = np.ascontiguousarray(points, np.float64)
        self.cells = np.ascontiguousarray(cells, np.int32)
        if self.points.ndim != 2 or self.points.shape[1] != 3:
            raise ValueError("points must be [n, 3]")
        if self.cells.ndim != 2 or self.cells.shape[1] != 4:
            raise ValueError("cells must be [m, 4] (triangles padded with -1)")
        h = ctypes.c_void_p()
        fn = L.cfd_mesh_create if device else L.cfd_mesh_create_host
        check(fn(len(self.points), dptr(self.points), len(self.cells), iptr(self.cells), tile,
                 face_budget, renumber, ctypes.byref(h)), "cfd_mesh_create")
        self.h = h
        self.device = device
        c = np.zeros(9, np.int32)
        check(L.cfd_mesh_counts(self.h, iptr(c)), "cfd_mesh_counts")
        (self.nphys, self.nghost, self.n_inlet, self.n_exit, self.n_wall, self.nfaces,
         self.ntiles, self.nface_entries, self.max_tile_faces) = (int(x) for x in c)
        self.tile = tile

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().cfd_mesh_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def counts(self):
        return (self.nphys, self.nghost, self.n_inlet, self.n_exit, self.n_wall)

    def arrays(self) -> MeshArrays:
        n, g = self.nphys, self.nghost
        a = MeshArrays(n, g, np.zeros(n, np.int32), -np.ones((n, 4), np.int32),
                       -np.ones((n, 4), np.int32), np.zeros((n, 8)), np.zeros((n, 4)),
                       np.zeros((n + g, 3)), np.zeros(n + g), np.zeros(n + g), np.zeros((n, 4)),
                       np.zeros((n, 4), np.int32), np.zeros((self.n_inlet, 3), np.int32),
                       np.zeros((self.n_exit, 3), np.int32), np.zeros((self.n_wall, 3), np.int32))
        check(lib().cfd_mesh_export(self.h, iptr(a.nf), iptr(a.nodes), iptr(a.nbr), dptr(a.nrm),
                                    dptr(a.length), dptr(a.ctr), dptr(a.area), dptr(a.inv_area),
                                    dptr(a.dist), iptr(a.kind), iptr(a.inlet), iptr(a.exit),
                                    iptr(a.wall)), "cfd_mesh_export")
        return a

    def permutation(self) -> np.ndarray:
        p = np.zeros(self.nphys, np.int32)
        check(lib().cfd_mesh_permutation(self.h, iptr(p)), "cfd_mesh_permutation")
        return p

    def layout(self):
        tc = np.zeros(self.ntiles + 1, np.int32)
        tf = np.zeros(self.ntiles + 1, np.int32)
        fL = np.zeros(self.nface_entries, np.int32)
        fR = np.zeros(self.nface_entries, np.int32)
        sl = np.zeros(4 * self.nphys, np.uint16)
        check(lib().cfd_mesh_layout_export(self.h, iptr(tc), iptr(tf), iptr(fL), iptr(fR), uptr(sl)),
              "cfd_mesh_layout_export")
        return tc, tf, fL, fR, sl.reshape(-1, 4)


class Solver:
    """Solver state resident in HBM plus the scheme configuration."""

    def __init__(self, mesh: Mesh, config: Config | None = None, **kw):
        self.mesh = mesh
        self.config = config if config is not None else make_config(**kw)
        h = ctypes.c_void_p()
        check(lib().cfd_solver_create(mesh.h, ctypes.byref(self.config), ctypes.byref(h)),
              "cfd_solver_create")
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().cfd_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def configure(self, config: Config | None = None, **kw):
        self.config = config if config is not None else make_config(**kw)
        check(lib().cfd_solver_configure(self.h, ctypes.byref(self.config)), "cfd_solver_configure")

    def set_state(self, U: np.ndarray):
        U = np.ascontiguousarray(U, np.float64)
        if U.shape != (self.mesh.nphys, 4):
            raise ValueError(f"state must be [{self.mesh.nphys}, 4]")
        check(lib().cfd_set_state(self.h, dptr(U)), "cfd_set_state")

    def get_state(self) -> np.ndarray:
        U = np.zeros((self.mesh.nphys, 4))
        check(lib().cfd_get_state(self.h, dptr(U)), "cfd_get_state")
        return U

    def state_device_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(lib().cfd_state_device(self.h, ctypes.byref(p)), "cfd_state_device")
        return int(p.value)

    def residual(self, stream: int = 0):
        n = self.mesh.nphys
        R = np.zeros((n, 4))
        dt = np.zeros(n)
        grad = np.zeros((n, 8))
        check(lib().cfd_residual(self.h, dptr(R), dptr(dt), dptr(grad), ctypes.c_void_p(stream)),
              "cfd_residual")
        return R, dt, grad

    def set_freeze(self, mode: int):
        """Limiter freezing state: 0 off, 1 record on the next step, 2 use the recorded factors."""
        check(lib().cfd_set_freeze(self.h, int(mode)), "cfd_set_freeze")

    def ghost_states(self, stream: int = 0) -> np.ndarray:
        """Conserved ghost states [nghost, 4] of the current state, ghost k = cell nphys + k."""
        Ug = np.zeros((self.mesh.nghost, 4))
        check(lib().cfd_apply_bc(self.h, dptr(Ug), ctypes.c_void_p(stream)), "cfd_apply_bc")
        return Ug

    def step(self, nsteps: int = 1, history: bool = False, stream: int = 0):
        """Advances nsteps; with history returns (err[n,4], dt[n])."""
        if history:
            err = np.zeros((nsteps, 4))
            dt = np.zeros(nsteps)
            check(lib().cfd_step(self.h, nsteps, dptr(err), dptr(dt), ctypes.c_void_p(stream)), "cfd_step")
            return err, dt
        check(lib().cfd_step(self.h, nsteps, None, None, ctypes.c_void_p(stream)), "cfd_step")
        return None

    def time_pass(self, which: str, reps: int, stream: int = 0) -> float:
        """Milliseconds for `reps` launches of 'grad', 'flux' (fused stage), 'step' or 'resid' (the
        flux pass in residual mode) (HIP events)."""
        ms = ctypes.c_float()
        check(lib().cfd_time_pass(self.h, dict(grad=0, flux=1, step=2, resid=3)[which], reps,
                                  ctypes.byref(ms), ctypes.c_void_p(stream)), "cfd_time_pass")
        return float(ms.value)

    def run_steady(self, tol: float = 1e-8, max_steps: int = 100000, check_every: int = 50,
                   stream: int = 0):
        """Marches to Error[0] <= tol; returns (converged, steps, err4)."""
        steps = ctypes.c_int()
        err = np.zeros(4)
        rc = lib().cfd_run_steady(self.h, tol, max_steps, check_every, ctypes.byref(steps),
                                  dptr(err), ctypes.c_void_p(stream))
        if rc not in (0, -4):
            check(rc, "cfd_run_steady")
        return rc == 0, steps.value, err


__all__ = ["Mesh", "Solver", "make_config", "freestream", "cons_from_prim", "CfdError", "MeshArrays"]
