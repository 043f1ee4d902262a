"""CPU tests of libcfdhip.so: it loads, exports every declared symbol, and its host-side mesh
processing (connectivity, renumbering, tiles) is bit-exact with the oracle. No GPU needed."""
import ctypes
import os
import re

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd._lib import LIB_PATH, lib
from cfd_solver_withcuda_amd.solver import Mesh
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b((?:cfd|launch|get)_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_declared_symbols():
    assert os.path.exists(LIB_PATH)
    L = lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert L.cfd_version() >= 1


MESHES = {
    "bump": lambda: meshgen.bump_mesh(40, 20, 0.1, seed=7),
    "ffs": lambda: meshgen.ffs_mesh(1.0 / 20, seed=3),
    "ramp": lambda: meshgen.ramp_mesh(30, 10, 15.0, seed=4),
    "ref_ramp": lambda: meshgen.read_vtk(os.path.join(ROOT, "tests", "golden", "Ramp_15o_52_18.vtk")),
    "ref_tri": lambda: meshgen.read_vtk(os.path.join(ROOT, "tests", "golden", "Test_Domain_Triangle.vtk")),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_connectivity_matches_oracle_bit_exact(name):
    m = MESHES[name]()
    a = Mesh(m.points, m.cells, tile=64, renumber=1, device=False).arrays()
    b = Oracle().mesh(m.points, m.cells).arrays()
    msk = np.arange(4)[None, :] < a.nf[:, None]
    for f in ("nodes", "nbr", "kind", "length", "dist"):
        assert np.array_equal(getattr(a, f)[msk], getattr(b, f)[msk]), f
    for f in ("nf", "ctr", "area", "inv_area", "inlet", "exit", "wall"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(a.nrm[np.repeat(msk, 2, axis=1)], b.nrm[np.repeat(msk, 2, axis=1)])


@pytest.mark.parametrize("tile,renumber,budget", [(32, 0, 0), (64, 1, 0), (256, 1, 0), (256, 1, 256), (128, 1, 192)])
def test_tiled_layout_is_consistent(tile, renumber, budget):
    """Every cell slot points at a face of its own tile whose L/R is that cell and whose other
    side is the reference neighbour; interior faces shared by two tiles appear in both."""
    m = meshgen.bump_mesh(40, 20, 0.1, seed=7)
    pm = Mesh(m.points, m.cells, tile=tile, renumber=renumber, device=False, face_budget=budget)
    a = pm.arrays()
    perm = pm.permutation()
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(len(perm))
    assert np.array_equal(np.sort(perm), np.arange(pm.nphys))
    tc, tf, fL, fR, slots = pm.layout()
    assert tf[0] == 0 and tf[-1] == pm.nface_entries and np.all(np.diff(tf) > 0)
    assert tc[0] == 0 and tc[-1] == pm.nphys and np.all(np.diff(tc) > 0) and np.diff(tc).max() <= tile
    assert np.diff(tf).max() == pm.max_tile_faces
    if budget:
        assert pm.max_tile_faces <= budget
    tile_of = np.repeat(np.arange(pm.ntiles), np.diff(tc))
    for k in range(pm.nphys):
        c = perm[k]
        t = tile_of[k]
        for f in range(a.nf[c]):
            s = int(slots[k, f])
            assert s != 0xFFFF
            e = tf[t] + (s >> 1)
            assert e < tf[t + 1]
            is_r = s & 1
            me, other = (fR[e], fL[e]) if is_r else (fL[e], fR[e])
            assert me == k
            nb = a.nbr[c, f]
            if nb < pm.nphys:
                assert other == iperm[nb]
            else:
                assert other == -a.kind[c, f]
        assert a.nf[c] == 4 or slots[k, 3] == 0xFFFF
    # faces: nfaces unique faces = interior + boundary
    nb_int = int((a.nbr[np.arange(4)[None, :] < a.nf[:, None]] < pm.nphys).sum()) // 2
    assert pm.nfaces == nb_int + pm.nghost


def test_mesh_rejects_bad_input():
    from cfd_solver_withcuda_amd._lib import CfdError
    pts = np.zeros((3, 3))
    with pytest.raises(CfdError):
        Mesh(pts, np.array([[0, 1, 2, -1]], np.int32), device=False)  # zero-length faces
    with pytest.raises(CfdError):
        Mesh(np.eye(3), np.array([[0, 1, 7, -1]], np.int32), device=False)  # node out of range


def test_batch_rejects_bad_tile_order():
    """xcd_remap outside 0..4096 is refused before any device work (the tile order's run length
    enters the workgroup-to-tile arithmetic)."""
    from cfd_solver_withcuda_amd._lib import CfdError
    from cfd_solver_withcuda_amd.batch import Batch
    from cfd_solver_withcuda_amd.solver import make_config

    m = meshgen.ramp_mesh(8, 4, 10.0, seed=1)
    for bad in (-1, 4097):
        cfg = make_config(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", rk=3, lts=1,
                          block=128, xcd_remap=bad)
        with pytest.raises(CfdError, match="xcd_remap"):
            Batch([m], np.array([[1.4, 2.0, 0.0, 1.0]]), cfg)


def test_mlp_rejects_bad_shapes():
    """cfd_mlp_create refuses shapes its kernels do not take (checked before any device work):
    more than 15 hidden layers (the per-step update treats at most 16 weight layers in one launch),
    more than 32 inputs or outputs."""
    from cfd_solver_withcuda_amd._lib import CfdError
    from cfd_solver_withcuda_amd.surrogate import MLP

    for args in ((4, 32, 16, 4), (33, 32, 2, 4), (4, 32, 2, 33), (4, 32, 0, 4)):
        with pytest.raises(CfdError, match="cfd_mlp_create"):
            MLP(*args, max_rows=512)


def test_host_driver_exports_declared_names():
    """libcfdhost.so defines every function include/cfd_host.hpp declares (the reference's driver
    names, including the dissipation plugin point)."""
    import shutil
    import subprocess

    host = os.path.join(os.path.dirname(LIB_PATH), "libcfdhost.so")
    assert os.path.exists(host)
    text = open(os.path.join(ROOT, "include", "cfd_host.hpp")).read()
    text = re.sub(r"//.*", "", text)
    names = set(re.findall(r"^(?:void|bool)\s+(\w+)\s*\(", text, flags=re.M))
    assert {"Calculate_Flux_For_All_Faces", "Evaluate_Time_Step", "ROE_2O", "Runge_Kutta_Method"} <= names
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", host], capture_output=True, text=True, check=True).stdout
    defined = {m.group(1) for m in re.finditer(r"\s[TW]\s+(\w+)\(", out)}
    assert not names - defined, sorted(names - defined)
