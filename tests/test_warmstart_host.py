"""CPU tests of the warm-start pipeline's host logic: the held-out family of config 4 (sizes,
slicing, determinism) and the freestream snapping of predicted states (with a stand-in for the
device network, whose predictions the test chooses)."""
import numpy as np

from cfd_solver_withcuda_amd import config4
from cfd_solver_withcuda_amd.solver import freestream
from cfd_solver_withcuda_amd.surrogate import GAMMA, predicted_state, prim_from_cons


class _Fixed:
    """Stands in for the MLP: predict() returns preset primitives."""

    def __init__(self, W):
        self.W = np.asarray(W, np.float32)

    def predict(self, feat):
        assert feat.shape == (len(self.W), 4)
        return self.W


def test_held_out_family_slices_and_sizes():
    a = config4.held_out_cases(n=12, first=0, count=12, min_cells=2000, max_cells=20000)
    b = config4.held_out_cases(n=12, first=4, count=5, min_cells=2000, max_cells=20000)
    assert len(a) == 12 and len(b) == 5
    for x, y in zip(a[4:9], b):
        assert (x.mach, x.param) == (y.mach, y.param)
        assert np.array_equal(x.mesh.cells, y.mesh.cells)
    sizes = np.array([len(c.mesh.cells) for c in a])
    assert sizes.min() >= 0.8 * 2000 and sizes.max() <= 1.25 * 20000
    machs = np.array([c.mach for c in a])
    thick = np.array([c.param for c in a])
    assert machs.min() >= 1.2 and machs.max() <= 3.0 and thick.min() >= 0.04 and thick.max() <= 0.12
    # the same seed gives the same family
    c = config4.held_out_cases(n=12, first=0, count=3, min_cells=2000, max_cells=20000)
    assert [(x.mach, x.param) for x in c] == [(x.mach, x.param) for x in a[:3]]


def test_snapping_to_the_freestream_state():
    fs = np.array(freestream(2.0))
    r, u, v, p = fs
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [2, 1, 0]], float)
    cells = np.array([[0, 1, 2, 3], [1, 4, 5, 2], [0, 1, 2, -1]])
    W = np.array([[r * 1.01, u * 0.995, v + 0.01 * u, p * 0.99],   # within 2 %: snaps
                  [r * 1.05, u, v, p],                              # density 5 % off: kept
                  [r, u, v, -1.0]], np.float64)                     # floored pressure, kept
    U = predicted_state(_Fixed(W), pts, cells, 2.0, 0.1, fs=fs, snap=0.02)
    Wp = prim_from_cons(U)
    assert np.allclose(Wp[0], fs, rtol=1e-15, atol=1e-15)
    assert np.allclose(Wp[1], np.float32(W[1]), rtol=1e-6)
    assert Wp[2, 3] > 0 and np.isclose(Wp[2, 3], 1e-3, rtol=1e-9)
    # without snapping the prediction is used as it is (after the floors)
    U0 = predicted_state(_Fixed(W), pts, cells, 2.0, 0.1)
    assert np.allclose(prim_from_cons(U0)[0], np.float32(W[0]), rtol=1e-6)
    # conserved form of the snapped cell
    assert np.isclose(U[0, 3], p / (GAMMA - 1) + 0.5 * r * (u * u + v * v), rtol=1e-15)
