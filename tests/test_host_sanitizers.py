"""The host mesh processing (connectivity, tiled layouts, packed face records, malformed input)
built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer and run on generated mixed
meshes: no out-of-bounds access, leak or undefined behaviour, and the layout invariants hold.
CPU only; skipped where g++ lacks the sanitizers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "mesh_checks.cpp")
MESH_HOST = os.path.join(ROOT, "cfd_solver_withcuda_amd", "csrc", "mesh_host.cpp")


def write_mesh(path, m):
    pts = np.ascontiguousarray(m.points, np.float64)
    cells = np.ascontiguousarray(m.cells, np.int32)
    with open(path, "wb") as f:
        f.write(np.array([len(pts), len(cells)], np.int32).tobytes())
        f.write(pts.tobytes())
        f.write(cells.tobytes())


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("asan") / "mesh_checks")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", SRC, MESH_HOST, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "sanitize" in r.stderr or "asan" in r.stderr:
            pytest.skip("g++ without sanitizer runtimes")
        raise AssertionError(r.stderr)
    return out


def test_mesh_processing_under_sanitizers(checker, tmp_path):
    meshes = [meshgen.bump_mesh(60, 30, 0.1, seed=3), meshgen.ffs_mesh(1.0 / 40, seed=4),
              meshgen.ramp_mesh(50, 20, 15.0, seed=5), meshgen.shock_tube_mesh(40, 4, seed=6, jitter=0.1,
                                                                             tri_frac=0.5)]
    paths = []
    for i, m in enumerate(meshes):
        p = str(tmp_path / f"m{i}.bin")
        write_mesh(p, m)
        paths.append(p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker] + paths, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert r.stdout.count("checked") == len(meshes)
