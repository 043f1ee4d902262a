"""Independent solvers on their own HIP streams: steps of two solvers issued alternately on two
streams, and a whole threaded march per solver, end bit-identical to the same solvers run one after
the other on the null stream (the library keeps no shared mutable state between handles)."""
import ctypes
import threading

import numpy as np
import pytest

from cfd_solver_withcuda_amd import meshgen
from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config

pytestmark = pytest.mark.gpu

BCS = ("freestream", "exit", "wall", "wall", "wall")


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return h


def solvers():
    out = []
    for nx, ny, t, M, flux in ((80, 40, 0.06, 2.0, "hllc"), (60, 30, 0.04, 1.7, "roe")):
        m = meshgen.bump_mesh(nx, ny, t, seed=nx)
        fs = freestream(M)
        s = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256),
                   make_config(flux=flux, order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
                               roe_bdiss=1, bc=BCS, fs=fs))
        s.set_state(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (len(m.cells), 1)))
        out.append(s)
    return out


def test_two_solvers_on_two_streams_match_sequential_runs():
    print("case two solvers, two streams, interleaved steps", flush=True)
    ref = solvers()
    for s in ref:
        s.step(40)
    want = [s.get_state() for s in ref]

    h = hip()
    st = [ctypes.c_void_p(), ctypes.c_void_p()]
    for x in st:
        assert h.hipStreamCreate(ctypes.byref(x)) == 0
    try:
        ss = solvers()
        for _ in range(40):
            for s, x in zip(ss, st):
                s.step(1, stream=x.value)
        for x in st:
            assert h.hipStreamSynchronize(x) == 0
        for s, w in zip(ss, want):
            assert np.array_equal(s.get_state(), w)

        print("case two solvers, two streams, one host thread each", flush=True)
        ss = solvers()
        errs = []

        def march(s, x):
            try:
                s.step(40, stream=x.value)
                h.hipStreamSynchronize(x)
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=march, args=(s, x)) for s, x in zip(ss, st)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
        for s, w in zip(ss, want):
            assert np.array_equal(s.get_state(), w)
    finally:
        for x in st:
            h.hipStreamDestroy(x)


def test_solvers_sharing_a_mesh_build_its_lazy_arrays_once():
    """Two solvers on one mesh, created from two host threads with WENO5 (whose line stencils the
    mesh builds on first use): both march like a solver made alone."""
    print("case shared mesh, two threads, order 3", flush=True)
    m = meshgen.bump_mesh(60, 30, 0.06, seed=5)
    fs = freestream(1.8)
    kw = dict(flux="hllc", order=3, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1, roe_bdiss=1,
              bc=BCS, fs=fs, block=128)
    U0 = np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (len(m.cells), 1))
    ref = Solver(Mesh(m.points, m.cells, tile=128, face_budget=256), make_config(**kw))
    ref.set_state(U0)
    ref.step(20)
    want = ref.get_state()
    mesh = Mesh(m.points, m.cells, tile=128, face_budget=256)
    out, errs = [None, None], []

    def run(i):
        try:
            s = Solver(mesh, make_config(**kw))
            s.set_state(U0)
            s.step(20)
            out[i] = s.get_state()
        except Exception as e:  # reported below
            errs.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for u in out:
        assert np.array_equal(u, want)


def test_set_state_after_async_steps_on_a_non_blocking_stream():
    """cfd_set_state after steps issued (without waiting) on a hipStreamNonBlocking stream: the
    new state replaces the old one only after those steps, so a march from it equals a march from
    the same state on a fresh solver."""
    print("case set_state after async steps on a non-blocking stream", flush=True)
    h = hip()
    st = ctypes.c_void_p()
    assert h.hipStreamCreateWithFlags(ctypes.byref(st), 1) == 0  # hipStreamNonBlocking
    try:
        a, b = solvers()[0], solvers()[0]
        U0 = a.get_state()
        Unew = U0.copy()
        Unew[:, 0] *= 1.01
        Unew[:, 3] *= 1.02
        for _ in range(5):
            a.step(20, stream=st.value)
            a.set_state(Unew)
        a.step(10, stream=st.value)
        assert h.hipStreamSynchronize(st) == 0
        b.set_state(Unew)
        b.step(10)
        assert np.array_equal(a.get_state(), b.get_state())
    finally:
        h.hipStreamDestroy(st)
