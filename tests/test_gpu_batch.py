"""GPU tests of the batched steady solver (cfd_batch_*): every instance of a batch evolves bit for
bit as the same mesh does alone (and as the oracle does), converged instances stop at the step
where their own Error[0] reaches the tolerance, and the autotuned launch shape changes nothing
in the results."""
import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.batch import Batch
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config
from oracle.oracle import Oracle, make_cfg

pytestmark = pytest.mark.gpu

RAMP_BCS = ("freestream", "exit", "exit", "wall", "wall")


def instances():
    specs = [(24, 12, 8.0, 2.0, 21), (40, 18, 15.0, 3.0, 22), (16, 10, 5.0, 1.6, 23), (33, 21, 19.0, 3.8, 24)]
    meshes = [meshgen.ramp_mesh(nx, ny, th, seed=sd) for nx, ny, th, _, sd in specs]
    fs = np.array([freestream(M) for *_, M, _ in specs])
    return meshes, fs


def kw(order=2):
    return dict(flux="ausm", order=order, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
                bc=RAMP_BCS)


def freestream_state(n, fs):
    return np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (n, 1))


def test_batch_steps_bit_exact_vs_single_solvers():
    meshes, fs = instances()
    print("case batch vs single, 4 ramps, 6 steps", flush=True)
    b = Batch(meshes, fs, make_config(block=128, **kw()))
    info = b.info()
    assert (info["tile"], info["face_budget"], info["block"]) == (128, 256, 128)
    r = b.solve(tol=-1.0, max_steps=6, check_every=6)
    assert not r["converged"] and np.all(r["steps"] == 6) and np.all(r["status"] == 1)
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(block=128, fs=fs[i], **kw()))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        err, _ = s.step(6, history=True)
        assert np.array_equal(Ub[i], s.get_state()), i
        assert np.allclose(r["err"][i], err[-1], rtol=1e-12, atol=0), (i, r["err"][i], err[-1])


def test_batch_steps_bit_exact_vs_oracle():
    meshes, fs = instances()
    meshes, fs = meshes[:2], fs[:2]
    b = Batch(meshes, fs, make_config(block=64, **kw()))
    b.solve(tol=-1.0, max_steps=2, check_every=2)
    Ub = b.get_state()
    orc = Oracle()
    for i, m in enumerate(meshes):
        om = orc.mesh(m.points, m.cells)
        U = np.ascontiguousarray(freestream_state(len(m.cells), fs[i]))
        cfg = make_cfg(flux="ausm", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, bc=RAMP_BCS, fs=fs[i])
        for _ in range(2):
            om.step(cfg, U, rk=3, lts=1)
        assert np.array_equal(Ub[i], U), i


def test_batch_convergence_masks():
    """Each instance stops at the first step whose own Error[0] <= tol, found by running it alone;
    its state is then frozen while the others march on."""
    meshes, fs = instances()
    tol = 1e-5
    b = Batch(meshes, fs, make_config(block=128, **kw(order=1)))
    r = b.solve(tol=tol, max_steps=20000, check_every=25)
    assert r["converged"], r
    assert len(set(r["steps"].tolist())) > 1  # instances converge at different steps
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(block=128, fs=fs[i], **kw(order=1)))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        k = int(r["steps"][i])
        err, _ = s.step(k, history=True)
        assert err[-1, 0] <= tol and np.all(err[:-1, 0] > tol), (i, k)
        assert np.array_equal(Ub[i], s.get_state()), i


def test_batch_autotune_keeps_results():
    meshes, fs = instances()
    b = Batch(meshes, fs, make_config(block=0, **kw()))
    info = b.info()
    print(f"autotune chose tile={info['tile']} budget={info['face_budget']} block={info['block']}", flush=True)
    assert info["block"] in (64, 128, 256) and info["tile"] <= info["block"]
    assert info["autotune_s"] > 0.0
    b.solve(tol=-1.0, max_steps=3, check_every=3)
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=info["tile"], face_budget=info["face_budget"]),
                   make_config(block=info["block"], fs=fs[i], **kw()))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        s.step(3)
        assert np.array_equal(Ub[i], s.get_state()), i


def test_batch_freezing_reaches_tolerance():
    """With limiter freezing on a stall (each instance on its own), MUSCL / Van Leer ramp solves,
    which stall under limiter switching, converge to 1e-8."""
    meshes, fs = instances()
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1, bc=RAMP_BCS)
    b0 = Batch(meshes, fs, make_config(block=128, **cfg))
    r0 = b0.solve(tol=1e-8, max_steps=3000, check_every=100)
    b1 = Batch(meshes, fs, make_config(block=128, freeze_stall=200, **cfg))
    r1 = b1.solve(tol=1e-8, max_steps=12000, check_every=100)
    print("plain", r0["status"].tolist(), r0["err"][:, 0].tolist(), "frozen", r1["status"].tolist(),
          r1["steps"].tolist(), flush=True)
    assert r1["converged"], r1
    assert np.all(r1["err"][:, 0] <= 1e-8)


def test_batch_history_cfl_and_first_step_freezing():
    """cfd_batch_history returns Error[0] per check (its last row is the solve's Error[0]) and the
    progress callback runs once per check;
    cfd_batch_set_cfl gives the same march as a batch created with that CFL; freezing mode 1
    (record on the first step, then frozen) matches a single solver put in the same mode, and
    leaves the other instances' marches unchanged."""
    meshes, fs = instances()
    meshes, fs = meshes[:3], fs[:3]
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", rk=3, lts=1, bc=RAMP_BCS,
               freeze_stall=50)
    print("case batch history / cfl / freeze mode 1, 3 ramps", flush=True)
    b = Batch(meshes, fs, make_config(block=128, cfl=0.8, **cfg))
    calls = []
    r = b.solve(tol=-1.0, max_steps=30, check_every=10, progress=lambda s, n: calls.append((s, n)),
                progress_every=1)
    assert calls == [(10, 3), (20, 3), (30, 3)]
    h = b.history()
    assert h.shape == (3, 3)
    assert np.array_equal(h[-1], r["err"][:, 0])
    assert np.all(np.isfinite(h)) and np.all(h > 0)
    # CFL override vs a batch built with it
    b.set_cfl(0.5)
    b.set_state([freestream_state(len(m.cells), f) for m, f in zip(meshes, fs)])
    b.solve(tol=-1.0, max_steps=20, check_every=20)
    b5 = Batch(meshes, fs, make_config(block=128, cfl=0.5, **cfg))
    b5.solve(tol=-1.0, max_steps=20, check_every=20)
    for u, v in zip(b.get_state(), b5.get_state()):
        assert np.array_equal(u, v)
    # first-step freezing on instance 0 only
    b.set_cfl(0.8)
    b.set_state([freestream_state(len(m.cells), f) for m, f in zip(meshes, fs)])
    b.set_freeze([1, 0, 0])
    b.solve(tol=-1.0, max_steps=12, check_every=12)
    Ub = b.get_state()
    b.set_freeze(None)
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256),
                   make_config(block=128, cfl=0.8, fs=fs[i], **cfg))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        if i == 0:
            s.set_freeze(1)
        s.step(12)
        assert np.array_equal(Ub[i], s.get_state()), i
    with pytest.raises(Exception):
        b.set_freeze([2, 0, 0])


def test_uneven_batch_of_32_bit_exact_vs_single_solvers_and_oracle(tmp_path, monkeypatch):
    """32 compression ramps of uneven sizes (about 1k to 100k cells, log-uniform) in one batch with
    an autotuned launch shape: after 4 steps every instance equals the same mesh solved alone, bit
    for bit, and the two smallest equal the oracle. A second batch of the same shape takes its launch
    shape from the autotune cache."""
    from cfd_solver_withcuda_amd import cases

    monkeypatch.setenv("CFD_AUTOTUNE_CACHE", str(tmp_path / "autotune.json"))
    cs = cases.ramp_batch(32, seed=41, min_cells=1000, max_cells=100000)
    sizes = np.array([len(c.mesh.cells) for c in cs])
    assert sizes.max() > 60000 and sizes.min() < 3000
    print(f"case uneven batch of 32, {sizes.sum()} cells", flush=True)
    fs = np.array([c.fs for c in cs])
    b = Batch([c.mesh for c in cs], fs, make_config(block=0, **kw()))
    info = b.info()
    assert info["shape_source"] == "autotuned"
    r = b.solve(tol=-1.0, max_steps=4, check_every=4)
    assert np.all(r["steps"] == 4)
    Ub = b.get_state()
    tile, budget, block = info["tile"], info["face_budget"], info["block"]
    for i, c in enumerate(cs):
        s = Solver(Mesh(c.mesh.points, c.mesh.cells, tile=tile, face_budget=budget),
                   make_config(block=block, fs=c.fs, **kw()))
        s.set_state(freestream_state(len(c.mesh.cells), c.fs))
        s.step(4)
        assert np.array_equal(Ub[i], s.get_state()), i
    orc = Oracle()
    for i in np.argsort(sizes)[:2]:
        c = cs[i]
        om = orc.mesh(c.mesh.points, c.mesh.cells)
        U = np.ascontiguousarray(freestream_state(len(c.mesh.cells), c.fs))
        for _ in range(4):
            om.step(make_cfg(flux="ausm", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, bc=RAMP_BCS,
                             fs=c.fs), U, rk=3, lts=1)
        assert np.array_equal(Ub[i], U), i
    b2 = Batch([c.mesh for c in cs], fs, make_config(block=0, **kw()))
    i2 = b2.info()
    assert i2["shape_source"] == "autotune cache"
    assert (i2["tile"], i2["face_budget"], i2["block"]) == (tile, budget, block)
    assert '"' in (tmp_path / "autotune.json").read_text()


def _solve_both(meshes, fs, cfg, tol, max_steps, check_every, freeze=None):
    out = []
    for mode in ("tiles", "persistent"):
        b = Batch(meshes, fs, make_config(**cfg))
        b.set_launch(mode)
        if freeze is not None:
            b.set_freeze(freeze)
        r = b.solve(tol=tol, max_steps=max_steps, check_every=check_every)
        out.append((r, b.get_state(), b.history()))
    return out


def test_persistent_launch_equals_tile_launch_as_instances_converge():
    """Persistent workgroups (cfd_batch_set_launch 1) give the same march as one workgroup per
    tile, bit for bit: 32 uneven ramps (more tiles than the persistent grid holds, so workgroups
    walk several tiles), first order to Error[0] <= 1e-4 so that instances drop out at different
    steps, and second order HLLC / Van Leer with first-step freezing on half the instances."""
    from cfd_solver_withcuda_amd import cases

    cs = cases.ramp_batch(32, seed=41, min_cells=1000, max_cells=100000)
    meshes, fs = [c.mesh for c in cs], np.array([c.fs for c in cs])
    print(f"case persistent vs tiles, 32 ramps, {sum(len(m.cells) for m in meshes)} cells, order 1", flush=True)
    (r0, U0, h0), (r1, U1, h1) = _solve_both(meshes, fs, dict(block=128, **kw(order=1)), 1e-4, 3000, 50)
    assert len(set(r0["steps"].tolist())) > 4  # instances converge at different steps
    assert np.array_equal(r0["steps"], r1["steps"]) and np.array_equal(r0["status"], r1["status"])
    assert np.array_equal(r0["err"], r1["err"]) and np.array_equal(h0, h1)
    for i in range(len(cs)):
        assert np.array_equal(U0[i], U1[i]), i
    # the instances that march longest, while most tiles have dropped out, equal the same meshes
    # solved alone (workgroups beyond the marching tiles must not take a tile a second time)
    for i in np.argsort(r0["steps"])[-3:]:
        c = cs[i]
        s = Solver(Mesh(c.mesh.points, c.mesh.cells, tile=128, face_budget=256),
                   make_config(block=128, fs=c.fs, **kw(order=1)))
        s.set_state(freestream_state(len(c.mesh.cells), c.fs))
        s.step(int(r0["steps"][i]))
        assert np.array_equal(U0[i], s.get_state()), i
    print("case persistent vs tiles, order 2 HLLC / Van Leer, freezing", flush=True)
    cfg = dict(block=64, flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
               bc=RAMP_BCS, freeze_stall=50)
    (r0, U0, _), (r1, U1, _) = _solve_both(meshes, fs, cfg, -1.0, 60, 20, freeze=[k % 2 for k in range(32)])
    assert np.array_equal(r0["err"], r1["err"])
    for i in range(len(cs)):
        assert np.array_equal(U0[i], U1[i]), i


def test_tile_order_over_xcds_changes_nothing_as_instances_converge():
    """The order in which tiles are dealt to the XCDs (xcd_remap: one contiguous run per XCD, runs
    of 3 or of 16 tiles in turn) changes neither which tiles a step computes nor its results, also
    once converged instances have left the compacted tile list and the launch grids have shrunk:
    32 uneven ramps, first order to Error[0] <= 1e-4, tile and persistent launches."""
    from cfd_solver_withcuda_amd import cases

    cs = cases.ramp_batch(32, seed=43, min_cells=1000, max_cells=60000)
    meshes, fs = [c.mesh for c in cs], np.array([c.fs for c in cs])
    print(f"case tile orders, 32 ramps, {sum(len(m.cells) for m in meshes)} cells", flush=True)
    runs = []
    for remap, mode in ((1, "tiles"), (3, "tiles"), (16, "tiles"), (3, "persistent")):
        b = Batch(meshes, fs, make_config(block=128, xcd_remap=remap, **kw(order=1)))
        b.set_launch(mode)
        r = b.solve(tol=1e-4, max_steps=3000, check_every=50)
        runs.append((r, b.get_state()))
    r0, U0 = runs[0]
    assert len(set(r0["steps"].tolist())) > 4  # instances converge at different steps
    for r, U in runs[1:]:
        assert np.array_equal(r0["steps"], r["steps"]) and np.array_equal(r0["err"], r["err"])
        for i in range(len(cs)):
            assert np.array_equal(U0[i], U[i]), i


def test_persistent_launch_rk4_and_limits():
    """RK4 stages under persistent workgroups equal the tile launch; order 3 and viscous batches
    refuse persistent workgroups, and unknown modes are refused."""
    meshes, fs = instances()
    cfg = dict(block=128, flux="roe", order=2, limiter="minmod", gradient="gg", cfl=0.6, rk=4, lts=1, bc=RAMP_BCS)
    (r0, U0, _), (r1, U1, _) = _solve_both(meshes, fs, cfg, -1.0, 8, 4)
    for i in range(len(meshes)):
        assert np.array_equal(U0[i], U1[i]), i
    b = Batch(meshes, fs, make_config(block=128, **kw()))
    with pytest.raises(Exception):
        b.set_launch(2)
    b3 = Batch(meshes[:1], fs[:1], make_config(block=128, **dict(kw(), order=3, flux="llf")))
    with pytest.raises(Exception):
        b3.set_launch("persistent")
    b3.set_launch("tiles")


STEP_BCS = ("freestream", "exit", "wall", "wall", "wall")


@pytest.mark.parametrize("rk,nsteps", [(3, 25), (4, 15)])
def test_batched_positivity_fallback_equals_single_solvers(rk, nsteps):
    """The positivity fallback in batched solves (the last stage writing a buffer of its own that is
    copied into the state over the marching tiles, the stages' marked-tile lists) gives each
    instance the march of the same mesh solved alone with the fallback, bit for bit: three forward
    steps of different sizes and Mach numbers from the impulsive start, where second-order stages
    turn non-physical at the step corner without it."""
    specs = [(0.025, 2, 3.0), (0.03, 5, 2.5), (0.02, 7, 3.2)]
    meshes = [meshgen.ffs_mesh(h, seed=sd) for h, sd, _ in specs]
    fs = np.array([freestream(M) for *_, M in specs])
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.4, rk=rk, lts=1, roe_bdiss=1,
               bc=STEP_BCS, positivity=1)
    print(f"case batched positivity rk{rk}, {sum(len(m.cells) for m in meshes)} cells", flush=True)
    b = Batch(meshes, fs, make_config(block=128, **cfg))
    r = b.solve(tol=-1.0, max_steps=nsteps, check_every=nsteps)
    assert np.all(r["steps"] == nsteps)
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(block=128, fs=fs[i], **cfg))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        err, _ = s.step(nsteps, history=True)
        assert np.array_equal(Ub[i], s.get_state()), i
        assert np.all(np.isfinite(Ub[i]))
    with pytest.raises(Exception):
        b.set_launch("persistent")


def test_batched_positivity_fallback_with_freezing_equals_single_solvers():
    """The fallback's kernel variants that carry the frozen limiter factors, batched: instances 0
    and 2 record their factors on the first step and march frozen, instance 1 evaluates its
    limiter; each equals the same mesh solved alone in the same mode, bit for bit."""
    specs = [(0.025, 2, 3.0), (0.03, 5, 2.5), (0.02, 7, 3.2)]
    meshes = [meshgen.ffs_mesh(h, seed=sd) for h, sd, _ in specs]
    fs = np.array([freestream(M) for *_, M in specs])
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.4, rk=3, lts=1, roe_bdiss=1,
               bc=STEP_BCS, positivity=1, freeze_stall=50)
    print("case batched positivity with freezing", flush=True)
    b = Batch(meshes, fs, make_config(block=128, **cfg))
    b.set_freeze([1, 0, 1])
    nsteps = 20
    b.solve(tol=-1.0, max_steps=nsteps, check_every=nsteps)
    Ub = b.get_state()
    for i, m in enumerate(meshes):
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(block=128, fs=fs[i], **cfg))
        s.set_state(freestream_state(len(m.cells), fs[i]))
        if i != 1:
            s.set_freeze(1)
        s.step(nsteps)
        assert np.array_equal(Ub[i], s.get_state()), i
