"""What the trained surrogate gains over the method without it, on a small bump family: the
network is trained on the GPU by the project's own MFMA kernels from batched FP64 solutions, then
held-out cases start from its prediction instead of from freestream. The prediction must be much
closer to the converged field than freestream is (measured: 10-20x), and the warm march must start
from a smaller residual (3-7x after 50 steps). Steps to the 1e-8 residual fall much less (about
2 % in all here): the explicit march removes an error at the rate the flow carries it out of the
domain, whatever its size, so the test only requires that the warm start costs no more steps
(within 10 %)."""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import cases
from cfd_solver_withcuda_amd.batch import Batch
from cfd_solver_withcuda_amd.solver import make_config
from cfd_solver_withcuda_amd.surrogate import MLP, predicted_state, prim_from_cons
from cfd_solver_withcuda_amd.warmstart import SCHEME, freestream_states, training_set

pytestmark = pytest.mark.gpu

RANGES = dict(mach_range=(1.8, 3.0), thick_range=(0.04, 0.08))


def batch_of(cs):
    return Batch([c.mesh for c in cs], np.array([c.fs for c in cs]), make_config(block=128, **SCHEME))


def test_surrogate_warm_start_beats_freestream():
    print("case warm start: 12 training bumps of ~3k cells, 4 held-out of ~4k cells", flush=True)
    tr = cases.bump_family(12, seed=21, cells=3000, stream=400, **RANGES)
    bt = batch_of(tr)
    rt = bt.solve(tol=1e-8, max_steps=20000, check_every=100)
    assert rt["converged"], rt["status"]
    feat, targ = training_set(tr, bt.get_state())
    mlp = MLP(4, 256, 6, 4, max_rows=4096, seed=3)
    mlp.fit_normalization(feat, targ)
    hist = mlp.train(feat, targ, steps=1500, batch=4096, lr=2e-3, lr_final=2e-5, seed=4, log_every=100)
    assert hist[-1] < 0.05 * hist[0]

    ev = cases.bump_family(4, seed=22, cells=4000, stream=500, **RANGES)
    b = batch_of(ev)
    fs = freestream_states(ev)
    b.set_state(fs)
    rp = b.solve(tol=1e-8, max_steps=20000, check_every=50)
    assert rp["converged"], rp["status"]
    plain_first = b.history()[0]
    converged = b.get_state()
    warm = [predicted_state(mlp, c.mesh.points, c.mesh.cells, c.mach, c.param) for c in ev]
    b.set_state(warm)
    rw = b.solve(tol=1e-8, max_steps=20000, check_every=50)
    assert rw["converged"], rw["status"]
    warm_first = b.history()[0]
    for i, (U0, Uw, Uc) in enumerate(zip(fs, warm, converged)):
        Wc = prim_from_cons(Uc)
        e_fs = np.linalg.norm(prim_from_cons(U0) - Wc) / np.linalg.norm(Wc)
        e_nn = np.linalg.norm(prim_from_cons(Uw) - Wc) / np.linalg.norm(Wc)
        print(f"instance {i}: rel L2 to the converged field, freestream {e_fs:.3e}, surrogate {e_nn:.3e}; "
              f"Error[0] after 50 steps {plain_first[i]:.3e} vs {warm_first[i]:.3e}; "
              f"steps {rp['steps'][i]} vs {rw['steps'][i]}", flush=True)
        assert e_nn < 0.3 * e_fs, i
        assert warm_first[i] < plain_first[i], i
    assert rw["steps"].sum() <= 1.1 * rp["steps"].sum()


def test_warm_and_plain_fields_agree_without_freezing():
    """With a scheme whose steady march converges without limiter freezing (first-order HLLC on
    this family), the march has one fixed point: the batch started from the surrogate and the batch
    started from freestream end at the same field. Both run to Error[0] <= 1e-10 and their fields
    must agree to 1e-8 relative (the oracle's march gives 5e-12 to 7e-12 at 1e-11). With limiter
    freezing this does not hold: each march converges to the scheme frozen at its own state."""
    print("case warm start, first order: 8 training bumps of ~3k cells, 4 held-out of ~4k cells", flush=True)
    sch = dict(SCHEME, order=1, freeze_stall=0)
    mk = lambda cs: Batch([c.mesh for c in cs], np.array([c.fs for c in cs]), make_config(block=128, **sch))
    tr = cases.bump_family(8, seed=31, cells=3000, stream=400, **RANGES)
    bt = mk(tr)
    rt = bt.solve(tol=1e-8, max_steps=40000, check_every=100)
    assert rt["converged"], rt["status"]
    feat, targ = training_set(tr, bt.get_state())
    mlp = MLP(4, 256, 6, 4, max_rows=4096, seed=5)
    mlp.fit_normalization(feat, targ)
    mlp.train(feat, targ, steps=1000, batch=4096, lr=2e-3, lr_final=2e-5, seed=6, log_every=100)

    ev = cases.bump_family(4, seed=32, cells=4000, stream=500, **RANGES)
    b = mk(ev)
    b.set_state(freestream_states(ev))
    rp = b.solve(tol=1e-10, max_steps=100000, check_every=100)
    assert rp["converged"], rp["status"]
    plain = b.get_state()
    b.set_state([predicted_state(mlp, c.mesh.points, c.mesh.cells, c.mach, c.param) for c in ev])
    rw = b.solve(tol=1e-10, max_steps=100000, check_every=100)
    assert rw["converged"], rw["status"]
    warm = b.get_state()
    for i, (Up, Uw) in enumerate(zip(plain, warm)):
        d_l2 = np.linalg.norm(Uw - Up) / np.linalg.norm(Up)
        d_max = np.abs(Uw - Up).max() / np.abs(Up).max()
        print(f"instance {i}: steps {rp['steps'][i]} plain, {rw['steps'][i]} warm; fields rel L2 {d_l2:.3e}, "
              f"max {d_max:.3e}", flush=True)
        assert d_l2 <= 1e-8 and d_max <= 1e-8, i
