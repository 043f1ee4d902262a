"""Benchmark of the explicit finite-volume residual-and-update path on one MI355X.

Workload (default): steady supersonic flow over a 10%-thick circular-arc bump, channel
[-1,2]x[0,1], Mach 1.65, 1400x700 quads jittered by +-10% with 30% split into triangles
(~1.27M cells); HLLC flux, MUSCL with the Van Leer limiter, weighted-least-squares gradients,
local time stepping, TVD-RK3, FP64. One step = one RK3 step (3 gradient + 3 flux passes).

    python bench.py [--steps K] [--warmup W] [--workload bump|ffs|batch|sod|steady] [--no-cpu] [--warmstart 0|1]

Prints one JSON line. `value` is FP64 face-flux evaluations per second (unique faces times
residual evaluations per second) with the state resident in HBM. The roofline entry is the
fused flux/stage pass timed alone with HIP events on the stream it runs on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# rocprofv3 summaries of the flux kernel per workload, with calibrated FETCH_SIZE + WRITE_SIZE
PROFILES = {("ffs_hllc", "muscl"): "profiles/flux_tile_kernel_ffs4680x1560_hllc.md",
            ("bump", "muscl"): "profiles/flux_tile_kernel_bump1400x700.md",
            ("ffs", "muscl"): "profiles/flux_tile_kernel_ffs4680x1560.md",
            ("bump", "weno"): "profiles/flux_tile_kernel_bump1400x700_weno5.md",
            ("bump", "viscous"): "profiles/flux_tile_kernel_bump1400x700_viscous.md"}


def measured_traffic(key, config):
    """HBM bytes per launch of the flux kernel recorded in the committed profile summary for
    this workload and scheme, when that summary was taken with the same launch configuration."""
    if key not in PROFILES:
        return None
    path = os.path.join(ROOT, PROFILES[key])
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if line.startswith("traffic_"):
            k, _, v = line.partition(":")
            vals[k.strip()] = v.strip()
    if vals.get("traffic_config") != config or "traffic_bytes_per_launch" not in vals:
        return None
    return float(vals["traffic_bytes_per_launch"])


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None,
                   help="timed steps (default 40; 5 for the ffs workload, whose second-order march is "
                        "timed on a first-order-developed flow before it turns unstable)")
    p.add_argument("--warmup", type=int, default=None, help="untimed warm-up steps (default 5; 1 for ffs)")
    p.add_argument("--workload", default="ffs_hllc",
                   choices=["bump", "ffs", "ffs_hllc", "ffs_march", "batch", "sod", "steady"],
                   help="ffs_hllc (default): the 7.97M-cell forward step with HLLC + MUSCL/Van Leer + WLSQ, RK4 at "
                        "CFL 0.4 (the roofline run); ffs: the same mesh with the step configuration's Roe + "
                        "minmod + Green-Gauss; ffs_march: that configuration's full 2000-step march from the "
                        "impulsive start (positivity fallback on); bump: the 1.27M-cell mixed bump (HLLC + "
                        "MUSCL/Van Leer + WLSQ, LTS TVD-RK3); steady: one bump solve to 1e-8")
    p.add_argument("--positivity", type=int, default=None,
                   help="step-mesh workloads: 1 (default) times the stages with the positivity fallback on "
                        "(detecting and recomputing passes; without it the second-order march on the developed "
                        "step flow turns non-physical at the corner within a few steps), 0 the plain stages")
    p.add_argument("--march-steps", type=int, default=2000, help="ffs_march: steps from the impulsive start")
    p.add_argument("--with-bump", type=int, default=1,
                   help="step-mesh workloads: also measure the 1.27M-cell bump (sub-object bump)")
    p.add_argument("--thickness", type=float, default=0.04, help="steady workload: bump thickness")
    p.add_argument("--cfl", type=float, default=None,
                   help="steady and batch workloads: Courant number (default 1.2)")
    p.add_argument("--steady-max-steps", type=int, default=200000, help="steady workload: step limit")
    p.add_argument("--warmstart", type=int, default=1,
                   help="1 (default): also measure time to Error[0] <= 1e-8 of batched bump-family solves, "
                        "surrogate-warm vs freestream starts (sub-object time_to_residual); 0: skip")
    p.add_argument("--ws-train", type=int, default=32, help="training cases (bump family)")
    p.add_argument("--ws-eval", type=int, default=16, help="held-out cases, solved in one batch")
    p.add_argument("--ws-train-cells", type=int, default=20000)
    p.add_argument("--ws-eval-cells", type=int, default=40000)
    p.add_argument("--ws-mlp-steps", type=int, default=3000)
    p.add_argument("--ws-max-steps", type=int, default=30000)
    p.add_argument("--instances", type=int, default=32, help="batch workload: ramp instances")
    p.add_argument("--batch-max-steps", type=int, default=60000, help="batch workload: step limit")
    p.add_argument("--min-cells", type=int, default=20000)
    p.add_argument("--max-cells", type=int, default=100000)
    p.add_argument("--tile", type=int, default=128, help="max cells per tile")
    p.add_argument("--face-budget", type=int, default=256, help="max faces per tile")
    p.add_argument("--block", type=int, default=128)
    p.add_argument("--develop", type=int, default=-1,
                   help="steps marched from freestream before the warmup, so the timed steps see a "
                        "developed flow (-1: 2000 for bump, 1000 for the step mesh)")
    p.add_argument("--grad-lds", type=int, default=0)
    p.add_argument("--xcd-remap", type=int, default=16,
                   help="tile order over the XCDs: 0 block order, 1 one run per XCD, k runs of k tiles dealt in turn")
    p.add_argument("--gradient", default="", choices=["", "gg", "wlsq"],
                   help="override the workload's gradient (bump: wlsq, ffs: gg)")
    p.add_argument("--scheme", default="muscl", choices=["muscl", "weno", "viscous"],
                   help="bump/ffs workloads: MUSCL (default); WENO5 (order 3, MUSCL on faces without a "
                        "five-cell line of quadrilaterals; the bump then has no triangles); MUSCL with "
                        "laminar viscous terms (Re 1e4 per unit length) and no-slip walls")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--kernel-reps", type=int, default=20)
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 512 MiB device copy (read 512 MiB + write 512 MiB) five times, "
                        "to scale the FETCH_SIZE / WRITE_SIZE counters")
    a = p.parse_args()
    if a.steps is None:
        a.steps = 5 if a.workload.startswith("ffs") else 40
    if a.warmup is None:
        a.warmup = 1 if a.workload.startswith("ffs") else 5
    return a


def workload(name, scheme="muscl"):
    from cfd_solver_withcuda_amd import meshgen
    from cfd_solver_withcuda_amd.solver import freestream

    if name == "bump" and scheme == "weno":
        # all (jittered) quadrilaterals, so that WENO5 serves every face away from the boundary
        m = meshgen.bump_mesh(1400, 700, 0.10, seed=1, tri_frac=0.0)
        fs = freestream(1.65)
        cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
                   bc=("freestream", "exit", "wall", "wall", "wall"), fs=fs)
        desc = "bump_1400x700_quads_M1.65_hllc_muscl_vanleer_wlsq_lts_rk3"
    elif name == "bump":
        m = meshgen.bump_mesh(1400, 700, 0.10, seed=1)
        fs = freestream(1.65)
        cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.8, rk=3, lts=1,
                   bc=("freestream", "exit", "wall", "wall", "wall"), fs=fs)
        desc = "bump_1400x700_mixed_M1.65_hllc_muscl_vanleer_wlsq_lts_rk3"
    elif name == "ffs_hllc":
        # the roofline run on the step mesh: HLLC + MUSCL/Van Leer (WLSQ; --gradient gg for
        # Green-Gauss), with the step workload's RK4 at CFL 0.4
        m = meshgen.ffs_mesh(1.0 / 1560, seed=2)
        fs = freestream(3.0)
        cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=0.4, rk=4, lts=0,
                   bc=("freestream", "exit", "wall", "wall", "wall"), fs=fs)
        desc = "ffs_4680x1560_mixed_M3_hllc_muscl_vanleer_wlsq_rk4"
    else:
        # second-order Roe, minmod, Green-Gauss, RK4 at CFL 0.4
        m = meshgen.ffs_mesh(1.0 / 1560, seed=2)
        fs = freestream(3.0)
        cfg = dict(flux="roe", order=2, limiter="minmod", gradient="gg", cfl=0.4, rk=4, lts=0,
                   bc=("freestream", "exit", "wall", "wall", "wall"), fs=fs)
        desc = "ffs_4680x1560_mixed_M3_roe_muscl_minmod_gg_rk4"
    return m, cfg, desc


def flux_pass_bytes(C, F, cfg):
    """Algorithmic bytes of the first-stage launch of the fused flux/stage pass. Per cell: U (32)
    + gradients (64) + area, 1/area (16) + face slots (8) read; new U (32) written, plus dt/area
    (8) with local time stepping and the accumulator (32) with RK4. Per face: L,R (8) + normal,
    length (24) + centre offset (16), plus the line stencil (16) for WENO5, whose faces read no
    gradients (only the MUSCL faces without a five-cell line do; not counted). The viscous terms
    read nothing more."""
    weno = cfg["order"] == 3
    per_cell = 32 + (0 if weno else 64) + 16 + 8 + 32 + (8 if cfg["lts"] else 0) + (32 if cfg["rk"] == 4 else 0)
    return per_cell * C + (64 if weno else 48) * F


def resid_pass_bytes(C, F, cfg):
    """Algorithmic bytes of the flux pass in residual mode: per cell U (32) + gradients (64, MUSCL)
    + area (8) + face slots (8) read, R (32) + local dt (8) written; per face L,R (8) + normal,
    length (24) + centre offset (16)."""
    return (32 + (64 if cfg["order"] == 2 else 0) + 8 + 8 + 32 + 8) * C + 48 * F


def apply_scheme(cfg, desc, scheme):
    """The workload's configuration for --scheme (MUSCL as defined, WENO5, or viscous)."""
    if scheme == "weno":
        cfg = dict(cfg, order=3)
        desc = desc.replace("_muscl_", "_weno5_")
    elif scheme == "viscous":
        rho, u = cfg["fs"][0], float(np.hypot(cfg["fs"][1], cfg["fs"][2]))
        cfg = dict(cfg, viscous="constant", mu=rho * u / 1e4, prandtl=0.72, r_gas=1.0,
                   bc=("freestream", "exit", "symmetry", "noslip_wall", "noslip_wall"))
        desc = desc + "_laminar_Re1e4_noslip"
    return cfg, desc


def residual_contract(C, F, gradient, ms):
    """One MUSCL residual evaluation (gradient pass + flux pass, each timed alone) against the
    two-pass contract bytes B = 224 C + 80 F (Green-Gauss) or 224 C + 96 F (WLSQ), with C the
    physical plus ghost cells and F all faces."""
    b = 224 * C + (96 if gradient == "wlsq" else 80) * F
    gbs = b / (ms / 1e3) / 1e9
    return {"contract_bytes": b, "ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}


def grad_pass_bytes(C, F, gradient):
    # per cell U (32) + slots (8) + gradients out (64) [+ 1/area (8) for Green-Gauss];
    # per face L,R (8) + least-squares vectors (32) or normal, length (24)
    return (104 + (8 if gradient == "gg" else 0)) * C + (40 if gradient == "wlsq" else 32) * F


def cpu_baseline(m, cfg, U0, budget_s=15.0):
    """The oracle (the C restatement of the same scheme, OpenMP over faces and cells on all host
    threads, or OMP_NUM_THREADS) on the same mesh and developed state: repeated full-mesh residual
    evaluations for about `budget_s` seconds."""
    from oracle.oracle import Oracle, make_cfg

    orc = Oracle()
    om = orc.mesh(m.points, m.cells)
    visc = {k: cfg[k] for k in ("viscous", "mu", "prandtl", "r_gas") if k in cfg}
    ocfg = make_cfg(flux=cfg["flux"], order=cfg["order"], limiter=cfg["limiter"],
                    gradient=cfg["gradient"], roe_bdiss=1, cfl=cfg["cfl"], bc=cfg["bc"], fs=cfg["fs"], **visc)
    U = np.ascontiguousarray(U0.copy())
    n = 0
    t0 = time.perf_counter()
    while True:
        om.residual(ocfg, U)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    faces = om.nfaces
    threads = orc.threads()
    return dict(value=faces * n / dt, unit="face-flux evals/s", cores=threads, kind="port",
                sample=f"{n} residual evaluation(s) of the oracle on the full {om.nphys}-cell mesh "
                       f"({dt:.1f} s, {threads} OpenMP threads, os.cpu_count() = {os.cpu_count()})")


def warmstart_bench(args):
    """Time to Error[0] <= 1e-8 of held-out bump-family solves, batched, from the surrogate's
    field vs from freestream (the surrogate is trained first on the solver's own solutions)."""
    from cfd_solver_withcuda_amd import warmstart

    r = warmstart.run(n_train=args.ws_train, n_eval=args.ws_eval, train_cells=args.ws_train_cells,
                      eval_cells=args.ws_eval_cells, mlp_steps=args.ws_mlp_steps, max_steps=args.ws_max_steps,
                      log=lambda msg: print(msg, file=sys.stderr, flush=True))
    common = r.get("common") or {}
    return {"plain_s": r["plain"]["wall_s"], "warm_s": r["warm"]["wall_s"], "speedup": r["speedup"],
            "both_converged": {"instances": common.get("instances"), "plain_s": common.get("plain_s"),
                               "warm_s": common.get("warm_s"), "speedup": common.get("speedup"),
                               "plain_steps": common.get("plain_steps"), "warm_steps": common.get("warm_steps")},
            "plain_steps_max": r["plain"]["steps_max"], "warm_steps_max": r["warm"]["steps_max"],
            "plain_converged": r["plain"]["converged"], "warm_converged": r["warm"]["converged"],
            "instances": args.ws_eval, "cells_per_instance": args.ws_eval_cells,
            "fallbacks": r["warm"]["fallbacks"], "surrogate_rel_l2": r["surrogate_rel_l2"],
            "converged_fields_max_rel_diff": r["converged_fields_max_rel_diff"],
            "converged_fields_rel_l2_diff": r["converged_fields_rel_l2_diff"],
            "training": {"cases": r["train"]["cases"], "solve_s": r["train"]["wall_s"],
                         "mlp_steps": r["mlp"]["steps"], "mlp_s": r["mlp"]["train_s"],
                         "loss_first": r["mlp"]["loss_first"], "loss_last": r["mlp"]["loss_last"],
                         # forward + backward of the 331,268-parameter MLP: 6 flops per parameter
                         # and row, against the dense BF16 MFMA peak of 2.5 PFLOP/s
                         "mlp_tflops": 6 * 331268 * r["mlp"]["batch"] * r["mlp"]["steps"] / r["mlp"]["train_s"] / 1e12,
                         "mlp_frac_of_bf16_peak": 6 * 331268 * r["mlp"]["batch"] * r["mlp"]["steps"]
                                                  / r["mlp"]["train_s"] / 2.5e15},
            "scheme": "HLLC + MUSCL/Van Leer + WLSQ, LTS TVD-RK3 at CFL 1.2, limiter freezing on stall",
            "note": "plain_s / warm_s: wall time until every instance converged or reached the step limit; "
                    "both_converged: the instances both marches brought to 1e-8, time until the last of them; "
                    "a scaled-down run of the pipeline that fits the bench's minutes",
            "fields_note": "converged_fields_*: the two marches freeze the limiter at different states, so each "
                           "ends at the fixed point of its own frozen scheme; with a scheme that needs no "
                           "freezing (first order) the two fields agree to 5e-12 relative at Error[0] 1e-10 "
                           "(tests/test_gpu_warmstart.py::test_warm_and_plain_fields_agree_without_freezing)",
            # the pipeline at its stated size (2,000 training solves, 256 held-out solves of 50k-1M
            # cells) runs in parts of one GPU session each (python -m cfd_solver_withcuda_amd.config4)
            "full_size_run": "profiles/warmstart_bump_config4.md"}


def sod_bench(args):
    """Sod shock tube, 400x40 uniform quads on [0,1]x[0,0.1]: the reference's iteration (Roe
    first order with the Harten fix, TVD-RK3 in delta form, global dt, CFL 0.5) from the
    discontinuous start to t >= 0.2 (the last step is not clipped, as in the reference). The same
    steps run on the GPU, in the oracle and, where it is built, in the reference itself."""
    from cfd_solver_withcuda_amd import meshgen
    from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, make_config
    from cfd_solver_withcuda_amd.surrogate import cell_centres

    m = meshgen.shock_tube_mesh(400, 40)
    xc = cell_centres(m.points, m.cells)[:, 0]
    left = xc < 0.5
    U0 = np.ascontiguousarray(cons_from_prim(np.where(left, 1.0, 0.125), 0 * xc, 0 * xc, np.where(left, 1.0, 0.1)))
    mesh = Mesh(m.points, m.cells, tile=128)
    s = Solver(mesh, make_config(flux="roe", order=1, cfl=0.5, fs=(1.0, 0.0, 0.0, 1.0), rk=3, ref_mode=1))
    # the number of steps to t = 0.2 (the dt sequence is the reference's, so it is fixed)
    s.set_state(U0)
    t, n = 0.0, 0
    while t < 0.2:
        _, dts = s.step(50, history=True)
        for d in dts:
            if t < 0.2:
                t += d
                n += 1
    s.set_state(U0)
    s.get_state()
    t0 = time.perf_counter()
    s.step(n)
    U = s.get_state()
    gpu_s = time.perf_counter() - t0
    out = {
        "metric": "time to t=0.2, Sod shock tube (reference iteration)",
        "value": gpu_s, "unit": "s", "n_gpus": 1, "steps": n, "warmup": 0,
        "ms_per_step": 1e3 * gpu_s / n, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (Sod initial state)",
        "config": {"workload": "sod_400x40_quads_roe1o_harten_tvdrk3_cfl0.5_global_dt", "cells": int(len(m.cells)),
                   "end_time": t},
        "cell_updates_per_s": len(m.cells) * n / gpu_s,
    }
    if not args.no_cpu:
        from oracle.oracle import Oracle, Reference, make_cfg

        om = Oracle().mesh(m.points, m.cells)
        Uo = np.ascontiguousarray(U0.copy())
        ocfg = make_cfg(flux="roe", order=1, cfl=0.5, fs=(1.0, 0, 0, 1.0))
        t0 = time.perf_counter()
        for _ in range(n):
            om.step(ocfg, Uo, rk=3, lts=0, ref=True)
        port_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": port_s, "unit": "s", "cores": 1, "kind": "port",
                               "sample": f"the same {n} steps in the oracle (C restatement), 1 thread"}
        out["gpu_vs_oracle_max_abs_diff"] = float(np.abs(U - Uo).max())
        if Reference.available():
            import tempfile

            ref = Reference()
            with tempfile.TemporaryDirectory() as d:
                path = os.path.join(d, "sod.vtk")
                meshgen.write_vtk(path, m)
                ref.load_vtk(path)
            ref.configure(dissipation=3, second_order=0, time_accurate=1, lts=0, cfl=0.5, rho_in=1.0, p_in=1.0)
            ref.init_state(U0)
            t0 = time.perf_counter()
            for _ in range(n):
                ref.step()
            ref_s = time.perf_counter() - t0
            Ur = ref.state()[0][:len(m.cells)]
            out["cpu_baseline"] = {"value": ref_s, "unit": "s", "cores": 1, "kind": "reference",
                                   "sample": f"the same {n} steps in the reference's own CPU solver (built from its "
                                             f"sources), 1 thread; the oracle took {port_s:.2f} s"}
            out["gpu_vs_reference_max_rel_diff"] = float(np.abs(U - Ur).max() / np.abs(Ur).max())
    print(json.dumps(out), flush=True)
    return 0


def steady_bench(args):
    """One steady solve of the bump configuration at full size (1400x700 mixed, Mach 1.65, HLLC +
    MUSCL/Van Leer + WLSQ, LTS TVD-RK3, limiter freezing on a 300-step stall) from freestream to
    Error[0] <= 1e-8; --thickness picks the bump (0.10 does not converge with this scheme)."""
    from cfd_solver_withcuda_amd import meshgen
    from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, freestream, make_config

    m = meshgen.bump_mesh(1400, 700, args.thickness, seed=1)
    fs = freestream(1.65)
    cfl = args.cfl if args.cfl is not None else 1.2
    cfg = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=cfl, rk=3, lts=1,
               bc=("freestream", "exit", "wall", "wall", "wall"), fs=fs, freeze_stall=300,
               positivity=args.positivity or 0)
    mesh = Mesh(m.points, m.cells, tile=128, renumber=1, face_budget=256)
    s = Solver(mesh, make_config(roe_bdiss=1, block=128, xcd_remap=args.xcd_remap, **cfg))
    s.set_state(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (mesh.nphys, 1)))
    s.get_state()
    done, ok, err = 0, False, np.full(4, np.nan)
    t0 = time.perf_counter()
    while done < args.steady_max_steps:
        ok, n, err = s.run_steady(1e-8, min(2000, args.steady_max_steps - done), 100)
        done += n
        print(f"steady: {done} steps, Error[0] {err[0]:.3e}, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if ok or not np.isfinite(err[0]):
            break
    wall = time.perf_counter() - t0
    out = {
        "metric": "time to residual 1e-8, one steady solve", "value": wall, "unit": "s", "n_gpus": 1,
        "steps": done, "warmup": 0, "ms_per_step": 1e3 * wall / max(done, 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (freestream initial state)",
        "config": {"workload": f"bump_1400x700_mixed_t{args.thickness:g}_M1.65_hllc_muscl_vanleer_wlsq_lts_rk3"
                               f"_cfl{cfl:g}" + ("_positivity" if args.positivity else ""), "cells": mesh.nphys},
        "converged": bool(ok), "error": [float(x) for x in err],
    }
    print(json.dumps(out), flush=True)
    return 0


def ffs_march_bench(args):
    """The step configuration as specified: Mach 3 forward-facing step, 7.97M-cell mixed mesh, Roe +
    MUSCL/minmod + Green-Gauss, classical RK4 at CFL 0.4 with the global time step, 2000 steps from
    the impulsive (uniform freestream) start, with the positivity fallback (without it the
    second-order march turns non-physical at the step corner within about 25 steps)."""
    import torch

    from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, make_config

    m, cfg, desc = workload("ffs")
    nsteps = args.march_steps
    mesh = Mesh(m.points, m.cells, tile=args.tile, renumber=1, face_budget=args.face_budget)
    s = Solver(mesh, make_config(roe_bdiss=1, block=args.block, xcd_remap=args.xcd_remap, grad_lds=0, positivity=1, **cfg))
    fs = cfg["fs"]
    s.set_state(np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (mesh.nphys, 1)))
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, t_sim = 0, 0.0
    while done < nsteps:
        k = min(250, nsteps - done)
        err, dts = s.step(k, history=True, stream=stream)
        done += k
        t_sim += float(dts.sum())
        print(f"ffs_march: {done} steps, t = {t_sim:.4f}, Error[0] {err[-1][0]:.3e}, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    U = s.get_state()
    rho = U[:, 0]
    p = 0.4 * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / rho)
    C, F = mesh.nphys, mesh.nfaces
    step_bytes = 4 * (grad_pass_bytes(C, F, "gg") + flux_pass_bytes(C, F, cfg))
    out = {
        "metric": "step configuration march: cell-updates/s (2000 RK4 steps from the impulsive start)",
        "value": C * nsteps / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": nsteps, "warmup": 0,
        "ms_per_step": 1e3 * wall / nsteps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (uniform Mach 3 freestream at t = 0)",
        "config": {"workload": desc + "_positivity_fallback", "cells": C, "faces": F, "cfl": cfg["cfl"]},
        "simulated_time": t_sim, "finite": bool(np.all(np.isfinite(U))),
        "min_density": float(rho.min()), "min_pressure": float(p.min()),
        "achieved_GBs": step_bytes / (wall / nsteps) / 1e9,
        "achieved_note": "algorithmic bytes of 4 gradient + 4 fused flux/stage passes per step over the "
                         "wall time per step (host-timed, error norm read back every 250 steps)",
    }
    print(json.dumps(out), flush=True)
    return 0


def batch_bench(args):
    """Batched steady compression-ramp solves (AUSM + MUSCL/Van Leer + WLSQ, LTS TVD-RK3) to
    Error[0] <= 1e-8 per instance, in one batch with an autotuned launch shape."""
    from cfd_solver_withcuda_amd import cases
    from cfd_solver_withcuda_amd.batch import Batch
    from cfd_solver_withcuda_amd.solver import cons_from_prim, make_config

    t = time.perf_counter()
    cs = cases.ramp_batch(args.instances, min_cells=args.min_cells, max_cells=args.max_cells)
    cfl = args.cfl if args.cfl is not None else 1.2
    cfg = make_config(flux="ausm", order=2, limiter="vanleer", gradient="wlsq", cfl=cfl, rk=3, lts=1,
                      roe_bdiss=1, bc=cases.RAMP_BCS, block=0, freeze_stall=300)
    b = Batch([c.mesh for c in cs], np.array([c.fs for c in cs]), cfg)
    info = b.info()
    setup = time.perf_counter() - t
    ms = b.time_steps(args.steps)
    b.set_state([np.tile(cons_from_prim(np.array([c.fs[0]]), c.fs[1], c.fs[2], c.fs[3]), (len(c.mesh.cells), 1))
                 for c in cs])
    r = b.solve(tol=1e-8, max_steps=args.batch_max_steps, check_every=100,
                progress=lambda done, marching: print(f"batch: {done} steps, {marching} instances marching",
                                                      file=sys.stderr, flush=True),
                progress_every=20)
    out = {
        "metric": "time to residual 1e-8, batched plain FP64 solves",
        "value": r["wall_s"], "unit": "s", "n_gpus": 1, "steps": int(r["steps"].max()), "warmup": 0,
        "ms_per_step": ms / args.steps, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded ramp family, freestream initial states)",
        "config": {"workload": f"ramp_batch_{args.instances}x[{args.min_cells},{args.max_cells}]_ausm_muscl_vanleer_wlsq_lts_rk3_cfl{cfl:g}",
                   "cells": info["cells"], "tiles": info["tiles"], "autotuned": {k: info[k] for k in ("tile", "face_budget", "block")}},
        "cell_updates_per_s": info["cells"] / (ms / args.steps / 1e3),
        "converged": int((r["status"] == 0).sum()), "instances": args.instances,
        "last_converged_s": float(r["conv_s"][r["status"] == 0].max()) if (r["status"] == 0).any() else None,
        "not_converged": [{"mach": round(c.mach, 4), "angle_deg": round(c.param, 3), "status": int(st),
                           "error0": float(e[0])}
                          for c, st, e in zip(cs, r["status"], r["err"]) if st != 0],
        "steps_per_instance": {"min": int(r["steps"].min()), "max": int(r["steps"].max())},
        "setup_s": setup, "autotune_s": info["autotune_s"],
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU partitioning is a design note, not built).", file=sys.stderr)
        return 2
    import torch

    if args.copy_calibration:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b.copy_(a)
        e0.record()
        for _ in range(5):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        # the practical HBM ceiling of a streaming kernel: 512 MiB read + 512 MiB written per copy
        print(json.dumps({"copy_calibration_bytes_read": a.numel() * 8, "copies": 6,
                          "copy_ms": ms, "copy_GBs": 2 * a.numel() * 8 / (ms / 1e3) / 1e9}))
        return 0

    from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, make_config

    if args.workload == "batch":
        return batch_bench(args)
    if args.workload == "sod":
        return sod_bench(args)
    if args.workload == "steady":
        return steady_bench(args)
    if args.workload == "ffs_march":
        return ffs_march_bench(args)
    out = fv_bench(args, args.workload, args.scheme, args.gradient)
    if args.workload.startswith("ffs") and args.with_bump:
        # the 1.27M-cell mixed bump (the bump configuration's mesh and scheme), same measurements
        b = fv_bench(args, "bump", "muscl", "", sub=True)
        out["bump"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "config", "cell_updates_per_s",
                                         "roofline", "grad_pass", "resid_pass", "residual_contract")}
    if args.warmstart:
        out["time_to_residual"] = warmstart_bench(args)
    print(json.dumps(out), flush=True)
    return 0


def fv_bench(args, wl, scheme, gradient, sub=False):
    """Throughput of one workload: whole steps, the fused flux/stage pass, the gradient pass and the
    flux pass in residual mode, each timed alone with HIP events on the stream they run on."""
    import torch

    from cfd_solver_withcuda_amd.solver import Mesh, Solver, cons_from_prim, make_config

    m, cfg, desc = workload(wl, scheme)
    cfg, desc = apply_scheme(cfg, desc, scheme)
    if gradient and gradient != cfg["gradient"]:
        desc = desc.replace("_" + cfg["gradient"] + "_", "_" + gradient + "_")
        cfg["gradient"] = gradient
    t_setup = time.perf_counter()
    mesh = Mesh(m.points, m.cells, tile=args.tile, renumber=1, face_budget=args.face_budget)
    mk = lambda **kw: make_config(roe_bdiss=1, block=args.block, xcd_remap=args.xcd_remap, grad_lds=args.grad_lds, **kw)
    s = Solver(mesh, mk(**cfg))
    fs = cfg["fs"]
    U0 = np.tile(cons_from_prim(np.array([fs[0]]), fs[1], fs[2], fs[3]), (mesh.nphys, 1))
    s.set_state(U0)
    t_setup = time.perf_counter() - t_setup
    C, F = mesh.nphys, mesh.nfaces
    nres = 4 if cfg["rk"] == 4 else 3

    stream = torch.cuda.current_stream().cuda_stream
    # The timed steps run on a developed flow, not on the uniform start: on a uniform state every
    # limiter sees zero differences and skips its division, and every face takes the same flux
    # branch, which would flatter the kernels. The step mesh's impulsive start turns the
    # second-order march non-physical at the step corner within 25 steps, so it is developed with
    # the positivity fallback on, and the timed steps keep it on (the plain stages turn the developed
    # flow non-physical at the corner within a few steps; --positivity 0 times them anyway).
    develop = args.develop if args.develop >= 0 else {"bump": 2000, "ffs": 1000, "ffs_hllc": 1000}[wl]
    ffs = wl.startswith("ffs") and scheme == "muscl"
    pos = ffs and (args.positivity is None or args.positivity == 1)
    t_dev = time.perf_counter()
    if develop:
        if ffs:
            s.configure(mk(positivity=1, **cfg))
        s.step(develop, stream=stream)
        if ffs and not pos:
            s.configure(mk(**cfg))
    elif pos:
        s.configure(mk(positivity=1, **cfg))
    U_dev = s.get_state()
    t_dev = time.perf_counter() - t_dev
    s.step(args.warmup, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_steps = s.time_pass("step", args.steps, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_flux = s.time_pass("flux", args.kernel_reps, stream=stream) / args.kernel_reps
    ms_grad = s.time_pass("grad", args.kernel_reps, stream=stream) / args.kernel_reps
    ms_resid = s.time_pass("resid", args.kernel_reps, stream=stream) / args.kernel_reps
    torch.cuda.synchronize()
    err = s.step(1, history=True, stream=stream)[0][0]
    ok = bool(np.isfinite(err[0]))
    contract = residual_contract(mesh.nphys + mesh.nghost, F, cfg["gradient"], ms_grad + ms_resid)
    contract_other = None
    if ffs and not sub:
        # the same residual with the other gradient (the step-mesh roofline run is defined for both)
        other = "gg" if cfg["gradient"] == "wlsq" else "wlsq"
        s.configure(mk(**dict(cfg, gradient=other)))
        g2 = s.time_pass("grad", args.kernel_reps, stream=stream) / args.kernel_reps
        r2 = s.time_pass("resid", args.kernel_reps, stream=stream) / args.kernel_reps
        contract_other = dict(residual_contract(mesh.nphys + mesh.nghost, F, other, g2 + r2), gradient=other,
                              grad_ms=g2, resid_ms=r2)

    t_step = ms_steps / 1e3 / args.steps
    launch_cfg = (f"workload={wl} tile={args.tile} face_budget={args.face_budget} "
                  f"block={args.block} grad_lds={args.grad_lds} xcd_remap={args.xcd_remap}"
                  + (f" gradient={gradient}" if gradient else "")
                  + (f" scheme={scheme}" if scheme != "muscl" else "") + (" positivity=1" if pos else ""))
    value = F * nres / t_step
    # with the fallback the stage pass is its detecting pass, which also writes one flag byte per cell
    bflux = flux_pass_bytes(C, F, cfg) + (C if pos else 0)
    achieved = bflux / (ms_flux / 1e3) / 1e9
    bres = resid_pass_bytes(C, F, cfg)
    out = {
        "metric": "time to residual 1e-8: NN-warm vs plain (batched); FP64 face-flux evals/s",
        "value": value,
        "unit": "face-flux evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded generated mesh; the state marched {develop} steps from freestream, untimed, "
                f"before the warmup{', with the positivity fallback' if ffs and develop else ''}"
                f"{'; the timed steps with the positivity fallback' if pos else ''})",
        "value_is": "FP64 face-flux evals/s (unique faces x residual evaluations per s) of whole "
                    f"{'RK4' if nres == 4 else 'TVD-RK3'} steps; the time-to-residual part is the "
                    "time_to_residual object",
        "config": {"workload": desc, "cells": C, "faces": F, "tile": args.tile,
                   "face_budget": args.face_budget, "block": args.block, "grad_lds": args.grad_lds,
                   "residual_evals_per_step": nres, "positivity_fallback": pos,
                   "launch": launch_cfg},
        "cell_updates_per_s": C / t_step,
        "wall_s_timed": wall,
        "finite": ok,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic((wl, scheme), launch_cfg),
                     "kernel": ("flux_tile_kernel, the positivity fallback's detecting pass (fused face flux + "
                                "gather + RK stage + admissibility flags), first stage" if pos else
                                "flux_tile_kernel (fused face flux + gather + RK stage), first stage"),
                     "algorithmic_bytes_per_launch": bflux, "avg_launch_ms": ms_flux,
                     "traffic_source": PROFILES.get((wl, scheme))},
        "grad_pass": {"avg_launch_ms": ms_grad,
                      "algorithmic_bytes_per_launch": grad_pass_bytes(C, F, cfg["gradient"]),
                      "achieved_GBs": grad_pass_bytes(C, F, cfg["gradient"]) / (ms_grad / 1e3) / 1e9},
        "resid_pass": {"avg_launch_ms": ms_resid, "algorithmic_bytes_per_launch": bres,
                       "achieved_GBs": bres / (ms_resid / 1e3) / 1e9},
        "residual_contract": contract,
        "setup_s": t_setup, "develop_steps": develop, "develop_s": t_dev,
    }
    if contract_other:
        out["residual_contract_" + contract_other["gradient"]] = contract_other
    if not args.no_cpu and not sub:
        out["cpu_baseline"] = cpu_baseline(m, cfg, U_dev)
    return out

if __name__ == "__main__":
    sys.exit(main())
