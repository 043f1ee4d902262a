"""Loader of libcfdhip.so (the HIP implementation) with ctypes signatures of its C-ABI.

There is no CPU fallback: if the library is missing the import of any solver object raises.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
# CFD_HIP_LIB selects a separately built (e.g. differently tuned) library
LIB_PATH = os.environ.get("CFD_HIP_LIB") or os.path.join(PKG, "libcfdhip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ushort)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

_lib = None


class CfdError(RuntimeError):
    pass


class Config(ctypes.Structure):
    """cfd_config_t (include/cfd_hip.h)."""
    _fields_ = [("flux", _i), ("order", _i), ("limiter", _i), ("gradient", _i), ("roe_bdiss", _i),
                ("cfl", _d), ("bc", _i * 6), ("fs", _d * 4), ("rk", _i), ("lts", _i),
                ("ref_mode", _i), ("block", _i), ("xcd_remap", _i), ("grad_lds", _i),
                ("freeze_tol", _d), ("freeze_stall", _i), ("zeta", _d), ("zeta1", _d),
                ("viscous", _i), ("mu", _d), ("prandtl", _d), ("r_gas", _d), ("positivity", _i), ("p_exit", _d)]


# cfd_progress_fn: void (*)(int steps_done, int marching, void *user)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


def lib() -> ctypes.CDLL:
    """The loaded library; raises CfdError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CfdError(f"{LIB_PATH} is missing: build it with `python -m cfd_solver_withcuda_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    cfgp = ctypes.POINTER(Config)
    _sig(L, "cfd_last_error", ctypes.c_char_p, [])
    _sig(L, "cfd_version", _i, [])
    for nm in ("cfd_mesh_create", "cfd_mesh_create_host"):
        _sig(L, nm, _i, [_i, _dp, _i, _ip, _i, _i, _i, ctypes.POINTER(_vp)])
    _sig(L, "cfd_mesh_destroy", None, [_vp])
    _sig(L, "cfd_mesh_counts", _i, [_vp, _ip])
    _sig(L, "cfd_mesh_export", _i, [_vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip])
    _sig(L, "cfd_mesh_permutation", _i, [_vp, _ip])
    _sig(L, "cfd_mesh_layout_export", _i, [_vp, _ip, _ip, _ip, _ip, _up])
    _sig(L, "cfd_solver_create", _i, [_vp, cfgp, ctypes.POINTER(_vp)])
    _sig(L, "cfd_solver_destroy", None, [_vp])
    _sig(L, "cfd_solver_configure", _i, [_vp, cfgp])
    _sig(L, "cfd_set_state", _i, [_vp, _dp])
    _sig(L, "cfd_get_state", _i, [_vp, _dp])
    _sig(L, "cfd_state_device", _i, [_vp, ctypes.POINTER(_vp)])
    _sig(L, "cfd_residual", _i, [_vp, _dp, _dp, _dp, _vp])
    _sig(L, "cfd_residual_noalloc", _i, [_vp, cfgp, _vp, _vp, _vp, _vp, _vp])
    _sig(L, "cfd_step", _i, [_vp, _i, _dp, _dp, _vp])
    _sig(L, "cfd_apply_bc", _i, [_vp, _dp, _vp])
    _sig(L, "cfd_set_freeze", _i, [_vp, _i])
    _sig(L, "cfd_batch_create", _i, [_i, _ip, _dp, _ip, _ip, _dp, cfgp, _i, ctypes.POINTER(_vp)])
    _sig(L, "cfd_batch_destroy", None, [_vp])
    _sig(L, "cfd_batch_info", _i, [_vp, _ip, _dp])
    _sig(L, "cfd_batch_tuning", _i, [_vp, _ip])
    _sig(L, "cfd_batch_set_state", _i, [_vp, _dp])
    _sig(L, "cfd_batch_get_state", _i, [_vp, _dp])
    _sig(L, "cfd_batch_solve", _i, [_vp, _d, _i, _i, _ip, _dp, _ip, _dp, _dp, _vp])
    _sig(L, "cfd_batch_time_steps", _i, [_vp, _i, ctypes.POINTER(ctypes.c_float)])
    _sig(L, "cfd_batch_set_freeze", _i, [_vp, _ip])
    _sig(L, "cfd_batch_set_cfl", _i, [_vp, _d])
    _sig(L, "cfd_batch_set_launch", _i, [_vp, _i])
    _sig(L, "cfd_batch_set_progress", _i, [_vp, PROGRESS_FN, _vp, _i])
    _sig(L, "cfd_batch_history", _i, [_vp, _ip, _dp])
    _fp = ctypes.POINTER(ctypes.c_float)
    _sig(L, "cfd_mlp_create", _i, [_i, _i, _i, _i, _i, ctypes.c_ulonglong, ctypes.POINTER(_vp)])
    _sig(L, "cfd_mlp_destroy", None, [_vp])
    _sig(L, "cfd_mlp_num_params", _i, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)])
    _sig(L, "cfd_mlp_set_params", _i, [_vp, _fp, _fp])
    _sig(L, "cfd_mlp_get_params", _i, [_vp, _fp, _fp])
    _sig(L, "cfd_mlp_set_normalization", _i, [_vp, _fp, _fp, _fp, _fp])
    _sig(L, "cfd_mlp_gradients", _i, [_vp, _fp, _fp, _i, _fp, _fp, _dp])
    _sig(L, "cfd_mlp_train", _i, [_vp, _fp, _fp, _i, _i, _i, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_ulonglong, _fp, _i])
    _sig(L, "cfd_mlp_predict", _i, [_vp, _fp, _i, _fp])
    _sig(L, "cfd_mlp_set_launch", _i, [_vp, _i])
    _sig(L, "cfd_mlp_time_step", _i, [_vp, _i, _i, _fp])
    _sig(L, "cfd_run_steady", _i, [_vp, _d, _i, _i, _ip, _dp, _vp])
    _sig(L, "cfd_time_pass", _i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_float), _vp])
    _sig(L, "cfd_device_alloc", _i, [ctypes.c_size_t, ctypes.POINTER(_vp)])
    _sig(L, "cfd_device_free", _i, [_vp])
    _sig(L, "cfd_memcpy_h2d", _i, [_vp, _vp, ctypes.c_size_t])
    _sig(L, "cfd_memcpy_d2h", _i, [_vp, _vp, ctypes.c_size_t])
    _sig(L, "cfd_synchronize", _i, [])
    _lib = L
    return L


def check(rc: int, what: str = "") -> int:
    if rc != 0:
        msg = lib().cfd_last_error().decode(errors="replace")
        raise CfdError(f"{what}: error {rc}: {msg}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def iptr(a):
    return a.ctypes.data_as(_ip)


def uptr(a):
    return a.ctypes.data_as(_up)


class DeviceArray:
    """A device copy of a numpy array (allocated and copied through the library)."""

    def __init__(self, host):
        import numpy as np
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        p = _vp()
        check(lib().cfd_device_alloc(self.nbytes, ctypes.byref(p)), "cfd_device_alloc")
        self.ptr = p
        check(lib().cfd_memcpy_h2d(self.ptr, self.host.ctypes.data_as(_vp), self.nbytes), "cfd_memcpy_h2d")

    def get(self):
        import numpy as np
        out = np.empty_like(self.host)
        check(lib().cfd_synchronize(), "cfd_synchronize")
        check(lib().cfd_memcpy_d2h(out.ctypes.data_as(_vp), self.ptr, self.nbytes), "cfd_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().cfd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
