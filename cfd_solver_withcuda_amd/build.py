"""Builds libcfdhip.so in-tree with hipcc for gfx950.

    python -m cfd_solver_withcuda_amd.build [--out PATH] [-D NAME=VALUE ...] [--flags "FLAGS"]

-D sets the kernels' compile-time tuning macros (e.g. CFD_FLUX_WAVES_PER_EU, CFD_FLUX_PREFETCH_CELL);
--flags appends compiler options to the HIP sources' command lines (they override the defaults);
--out writes a separately tuned library elsewhere (load it with CFD_HIP_LIB=PATH). HIP sources are compiled with floating-point contraction off, so that every FP64 kernel
reproduces the CPU restatement's rounding, and with packed FP32 ops disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcfdhip.so")
ARCH = os.environ.get("CFD_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["fv_kernels.hip", "fv_visc.hip", "fv_weno.hip", "fv_weno_visc.hip", "ref_kernels.hip",
               "mlp_kernels.hip", "capi_noalloc.hip", "fv_pos.hip", "fv_freeze.hip", "fv_persist.hip"]
CPP_SOURCES = ["mesh_host.cpp", "capi.cpp", "capi_batch.cpp", "capi_mlp.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          f"-I{os.path.join(ROOT, 'include')}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc() -> str:
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.sep not in p or os.path.exists(p)):
            return p
    return "hipcc"


def _compile(src: str, build_dir: str = BUILD, defines: tuple = (), flags: tuple = ()) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc()] + COMMON + [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd += HIP_FLAGS + list(flags)
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def sources() -> list[str]:
    return [s for s in HIP_SOURCES + CPP_SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = True, out: str = LIB, defines: tuple = (), flags: tuple = ()) -> str:
    build_dir = BUILD if out == LIB else os.path.join(os.path.dirname(os.path.abspath(out)), "_build")
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, tuple(defines), tuple(flags)), srcs))
    LIB_OUT = os.path.abspath(out)
    tmp = LIB_OUT + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_OUT)
    if verbose:
        print(f"built {LIB_OUT} from {', '.join(srcs)}" + (f" with {' '.join(defines)}" if defines else "")
              + (f" and {' '.join(flags)}" if flags else ""))
    return LIB_OUT


HOST_LIB = os.path.join(PKG, "libcfdhost.so")
HOST_TESTS = os.path.join(ROOT, "tests", "host")


def build_host(verbose: bool = True) -> str:
    """libcfdhost.so: the reference-named C++ host driver (include/cfd_host.hpp) over
    libcfdhip.so, and the C++ test drivers under tests/host (built into tests/host/_build)."""
    cmd = [_hipcc()] + COMMON + ["-shared", "-o", HOST_LIB, os.path.join(CSRC, "host_driver.cpp"),
                                 f"-L{PKG}", "-lcfdhip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    out_dir = os.path.join(HOST_TESTS, "_build")
    os.makedirs(out_dir, exist_ok=True)
    built = [HOST_LIB]
    for src in sorted(os.listdir(HOST_TESTS)):
        if not src.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, os.path.splitext(src)[0])
        cmd = [_hipcc()] + COMMON + ["-o", exe, os.path.join(HOST_TESTS, src), f"-L{PKG}", "-lcfdhost",
                                     "-lcfdhip", "-Wl,-rpath,$ORIGIN/../../../cfd_solver_withcuda_amd"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        built.append(exe)
    if verbose:
        print("built " + ", ".join(built))
    return HOST_LIB


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=LIB)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--flags", default="")
    a = ap.parse_args()
    build(out=a.out, defines=tuple(a.defines), flags=tuple(a.flags.split()))
    if a.out == LIB:
        build_host()
