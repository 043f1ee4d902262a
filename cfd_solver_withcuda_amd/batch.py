"""Batched steady solves over libcfdhip.so (cfd_batch_*).

Many independent meshes of uneven size (the reference solves one Inviscid_Solver case per run,
src/Solver.cpp:93-118) are concatenated into one device mesh and marched together with local
time stepping until each instance's Error[0] (Estimate_Error, src/Error_Estimate_Update.cpp:10-30)
reaches the tolerance. One RK step of all instances is one hipGraph launch; converged instances
drop out of the tile list on the device.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import PROGRESS_FN, Config, check, dptr, iptr, lib


class Batch:
    def __init__(self, meshes, fs, config: Config, nthreads: int | None = None):
        """meshes: sequence of objects with .points [n,3] and .cells [m,4] (-1 pads triangles);
        fs: [ninst, 4] prescribed (rho, u, v, p) per instance; config.block = 0 autotunes."""
        self.ninst = len(meshes)
        fs = np.ascontiguousarray(fs, np.float64).reshape(self.ninst, 4)
        pt_off = np.zeros(self.ninst + 1, np.int32)
        cell_off = np.zeros(self.ninst + 1, np.int32)
        for i, m in enumerate(meshes):
            pt_off[i + 1] = pt_off[i] + len(m.points)
            cell_off[i + 1] = cell_off[i] + len(m.cells)
        pts = np.ascontiguousarray(np.concatenate([np.asarray(m.points, np.float64) for m in meshes]))
        cells = np.ascontiguousarray(np.concatenate([np.asarray(m.cells, np.int32) for m in meshes]))
        self.cell_off = cell_off
        self.config = config
        nthreads = nthreads or min(16, os.cpu_count() or 1)
        h = ctypes.c_void_p()
        check(lib().cfd_batch_create(self.ninst, iptr(pt_off), dptr(pts), iptr(cell_off), iptr(cells),
                                     dptr(fs), ctypes.byref(config), nthreads, ctypes.byref(h)),
              "cfd_batch_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().cfd_batch_destroy(h)
            except Exception:
                pass
            self.h = None

    def info(self) -> dict:
        o = np.zeros(8, np.int32)
        secs = np.zeros(2)
        check(lib().cfd_batch_info(self.h, iptr(o), dptr(secs)), "cfd_batch_info")
        keys = ("ninst", "cells", "tiles", "face_entries", "tile", "face_budget", "block", "max_tile_faces")
        d = {k: int(v) for k, v in zip(keys, o)}
        d["build_s"], d["autotune_s"] = float(secs[0]), float(secs[1])
        src = ctypes.c_int()
        check(lib().cfd_batch_tuning(self.h, ctypes.byref(src)), "cfd_batch_tuning")
        d["shape_source"] = ("given", "autotuned", "autotune cache")[src.value]
        return d

    @property
    def ncells(self) -> int:
        return int(self.cell_off[-1])

    def set_state(self, U):
        """U: [total cells, 4] or a list of per-instance [n_i, 4] arrays."""
        if isinstance(U, (list, tuple)):
            U = np.concatenate(U)
        U = np.ascontiguousarray(U, np.float64)
        if U.shape != (self.ncells, 4):
            raise ValueError(f"state must be [{self.ncells}, 4]")
        check(lib().cfd_batch_set_state(self.h, dptr(U)), "cfd_batch_set_state")

    def get_state(self) -> list:
        U = np.zeros((self.ncells, 4))
        check(lib().cfd_batch_get_state(self.h, dptr(U)), "cfd_batch_get_state")
        return [U[self.cell_off[i]:self.cell_off[i + 1]] for i in range(self.ninst)]

    def solve(self, tol: float = 1e-8, max_steps: int = 100000, check_every: int = 100,
              stream: int = 0, progress=None, progress_every: int = 10) -> dict:
        """progress(steps_done, marching) is called every `progress_every` status checks."""
        n = self.ninst
        cb = None
        if progress is not None:
            cb = PROGRESS_FN(lambda done, marching, user: progress(int(done), int(marching)))
            check(lib().cfd_batch_set_progress(self.h, cb, None, int(progress_every)), "cfd_batch_set_progress")
        steps = np.zeros(n, np.int32)
        err = np.zeros((n, 4))
        status = np.zeros(n, np.int32)
        conv = np.zeros(n)
        wall = ctypes.c_double()
        try:
            rc = lib().cfd_batch_solve(self.h, tol, max_steps, check_every, iptr(steps), dptr(err),
                                       iptr(status), dptr(conv), ctypes.byref(wall), ctypes.c_void_p(stream))
        finally:
            if cb is not None:
                lib().cfd_batch_set_progress(self.h, PROGRESS_FN(), None, 0)
        if rc not in (0, -4):
            check(rc, "cfd_batch_solve")
        return dict(converged=rc == 0, steps=steps, err=err, status=status, conv_s=conv,
                    wall_s=float(wall.value))

    def set_freeze(self, mode=None):
        """Per-instance limiter-freezing mode of later solves: 0 evaluated limiter until the
        freezing trigger, 1 record on the first step and march frozen (None: all 0)."""
        if mode is None:
            check(lib().cfd_batch_set_freeze(self.h, None), "cfd_batch_set_freeze")
            return
        m = np.ascontiguousarray(mode, np.int32)
        if m.shape != (self.ninst,):
            raise ValueError(f"mode must have {self.ninst} entries")
        check(lib().cfd_batch_set_freeze(self.h, iptr(m)), "cfd_batch_set_freeze")

    def set_cfl(self, cfl: float):
        """Courant number of later solves."""
        check(lib().cfd_batch_set_cfl(self.h, float(cfl)), "cfd_batch_set_cfl")

    def set_launch(self, mode):
        """Launch of the gradient and flux/stage passes in later solves: "tiles" (one workgroup per
        tile) or "persistent" (workgroups that walk the tiles still marching; order 1 or 2,
        inviscid). Results are identical."""
        m = {"tiles": 0, "persistent": 1}[mode] if isinstance(mode, str) else int(mode)
        check(lib().cfd_batch_set_launch(self.h, m), "cfd_batch_set_launch")

    def history(self) -> np.ndarray:
        """Error[0] per check of the last solve, [checks, ninst]."""
        n = ctypes.c_int()
        check(lib().cfd_batch_history(self.h, ctypes.byref(n), None), "cfd_batch_history")
        h = np.zeros((n.value, self.ninst))
        check(lib().cfd_batch_history(self.h, ctypes.byref(n), dptr(h)), "cfd_batch_history")
        return h

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_float()
        check(lib().cfd_batch_time_steps(self.h, nsteps, ctypes.byref(ms)), "cfd_batch_time_steps")
        return float(ms.value)


__all__ = ["Batch"]
