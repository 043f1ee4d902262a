"""bench.py's command-line contract on a machine without a GPU: more than one GPU is refused
with a message, and the byte counts behind the roofline entry follow their stated formulas."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_algorithmic_byte_counts():
    C, F = 1000, 1800
    # flux pass: 152 B/cell (+8 with local time stepping, +32 with RK4) and 48 B/face
    assert bench.flux_pass_bytes(C, F, dict(lts=1, rk=3, order=2)) == 160 * C + 48 * F
    assert bench.flux_pass_bytes(C, F, dict(lts=0, rk=4, order=2)) == 184 * C + 48 * F
    # WENO5 adds the 16-byte line stencil per face and reads no gradients on its faces
    assert bench.flux_pass_bytes(C, F, dict(lts=1, rk=3, order=3)) == 96 * C + 64 * F
    # gradient pass: 104 B/cell (+8 for Green-Gauss's 1/area), 40 B/face WLSQ, 32 B/face GG
    assert bench.grad_pass_bytes(C, F, "wlsq") == 104 * C + 40 * F
    assert bench.grad_pass_bytes(C, F, "gg") == 112 * C + 32 * F


def test_committed_profiles_carry_the_traffic_of_the_default_launch():
    launch = "workload={w} tile=128 face_budget=256 block=128 grad_lds=0 xcd_remap=16"
    # the step-mesh workloads time their stages with the positivity fallback by default
    for w, extra in (("bump", ""), ("ffs", " positivity=1"), ("ffs_hllc", " positivity=1")):
        t = bench.measured_traffic((w, "muscl"), launch.format(w=w) + extra)
        assert t is not None and t > 0, w
