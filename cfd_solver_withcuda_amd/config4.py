"""The learned warm start at its stated size, in resumable parts (each part fits one GPU session):

    python -m cfd_solver_withcuda_amd.config4 train --out DIR [--cases 2000 --cells 50000]
        2000 Latin-hypercube bump cases (Mach in [1.2, 3], thickness in [4 %, 12 %]) solved to
        Error[0] <= 1e-8 in batches on meshes of `cells` cells, a sample of each solution's cells
        kept, the coordinate MLP trained on them; writes DIR/surrogate.npz (weights and
        normalisation) and DIR/train.json.
    python -m cfd_solver_withcuda_amd.config4 eval --model DIR/surrogate.npz --out FILE
                                                  [--first 0 --count 256]
        256 held-out cases (another Latin hypercube) on meshes of 50k to 1M cells (log-uniform),
        instances [first, first + count) solved in one batch from freestream (plain) and from the
        surrogate's field (warm, with fallback to freestream where its first-step residual is
        larger); writes times, steps, surrogate error and the field agreement to FILE (JSON).

Both parts use the warm-start pipeline's scheme (warmstart.SCHEME).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import cases, meshgen
from .batch import Batch
from .solver import freestream, make_config
from .surrogate import MLP, features, predicted_state, prim_from_cons
from .warmstart import freestream_states, scheme


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def held_out_cases(n=256, seed=105, min_cells=50000, max_cells=1000000, stream=300, first=0, count=None):
    """Cases [first, first + count) of n bump cases, (Mach, thickness) by Latin hypercube, cell
    counts log-uniform in [min_cells, max_cells] (3:1 grids, 30 % of quads split)."""
    lh = cases.latin_hypercube(n, seed, stream)
    mach = 1.2 + 1.8 * lh[:, 0]
    thick = 0.04 + 0.08 * lh[:, 1]
    u = meshgen.uniform01(seed, stream + 7, np.arange(n, dtype=np.uint64))
    cells = np.exp(np.log(min_cells) + (np.log(max_cells) - np.log(min_cells)) * u)
    out = []
    for i in range(first, n if count is None else min(n, first + count)):
        quads = cells[i] / (1.0 + cases.TRI_FRACTION)
        ny = max(4, int(round(math.sqrt(quads / 3.0))))
        m = meshgen.bump_mesh(3 * ny, ny, float(thick[i]), seed=seed * 1000003 + i)
        out.append(cases.Case(m, tuple(freestream(float(mach[i]))), float(mach[i]), float(thick[i])))
    return out


def train(out_dir, n_cases=2000, cells=50000, chunk=250, sample=1024, tol=1e-8, max_steps=20000,
          mlp_steps=8000, mlp_batch=16384, seed=5):
    os.makedirs(out_dir, exist_ok=True)
    sch = scheme()
    t0 = time.perf_counter()
    tr = cases.bump_family(n_cases, seed=seed, cells=cells, stream=200)
    rng = np.random.default_rng(seed)
    feat, targ, summary = [], [], dict(cases=n_cases, cells_per_case=len(tr[0].mesh.cells), converged=0,
                                       used=0, solve_s=0.0, steps_max=0)
    for k0 in range(0, n_cases, chunk):
        part = tr[k0:k0 + chunk]
        b = Batch([c.mesh for c in part], np.array([c.fs for c in part]), make_config(block=128, **sch))
        b.set_state(freestream_states(part))
        r = b.solve(tol=tol, max_steps=max_steps, check_every=100,
                    progress=lambda done, marching: log(f"  chunk {k0 // chunk}: {done} steps, {marching} marching"),
                    progress_every=40)
        states = b.get_state()
        del b
        summary["converged"] += int((r["status"] == 0).sum())
        summary["solve_s"] += r["wall_s"]
        summary["steps_max"] = max(summary["steps_max"], int(r["steps"].max()))
        for c, U, st in zip(part, states, r["status"]):
            if st < 0:  # non-finite: not used
                continue
            summary["used"] += 1
            idx = rng.choice(len(U), size=min(sample, len(U)), replace=False)
            f = features(c.mesh.points, c.mesh.cells, c.mach, c.param)[idx]
            feat.append(f)
            targ.append(prim_from_cons(U[idx]).astype(np.float32))
        log(f"training solves {k0 + len(part)}/{n_cases}: {summary['converged']} converged, "
            f"{time.perf_counter() - t0:.0f} s")
    feat = np.concatenate(feat)
    targ = np.concatenate(targ)
    mlp = MLP(4, 256, 6, 4, max_rows=mlp_batch, seed=seed)
    mlp.fit_normalization(feat, targ)
    t1 = time.perf_counter()
    hist = mlp.train(feat, targ, steps=mlp_steps, batch=mlp_batch, lr=2e-3, lr_final=2e-5, seed=seed + 1,
                     log_every=max(1, mlp_steps // 20))
    summary["mlp"] = dict(samples=int(len(feat)), steps=mlp_steps, batch=mlp_batch,
                          train_s=time.perf_counter() - t1, loss_first=float(hist[0]), loss_last=float(hist[-1]))
    W, bvec = mlp.get_params()
    in_mean, in_scale, out_mean, out_scale = mlp.norm
    np.savez(os.path.join(out_dir, "surrogate.npz"), W=W, b=bvec, in_mean=in_mean, in_scale=in_scale,
             out_mean=out_mean, out_scale=out_scale)
    summary["total_s"] = time.perf_counter() - t0
    summary["scheme"] = {k: sch[k] for k in ("flux", "limiter", "gradient", "cfl", "freeze_stall")}
    with open(os.path.join(out_dir, "train.json"), "w") as f:
        json.dump(summary, f)
    log(json.dumps(summary))
    return summary


def load_model(path):
    d = np.load(path, allow_pickle=False)
    mlp = MLP(4, 256, 6, 4, max_rows=16384)
    mlp.set_params(d["W"], d["b"])
    mlp.set_normalization(d["in_mean"], d["in_scale"], d["out_mean"], d["out_scale"])
    return mlp


def evaluate(model, out_file, first=0, count=256, tol=1e-8, max_steps=60000, freeze_warm=True, snap=0.02,
             positivity=0):
    """freeze_warm: warm-started instances record their limiter factors on the first step and march
    frozen (else they freeze on a stall, as the plain march does). snap: cells predicted within
    that relative distance of the freestream state start at it exactly (predicted_state).
    positivity: the fallback of both marches (0 off, 1 sign and range check, 2 sign only)."""
    sch = scheme(positivity=positivity or None)
    mlp = load_model(model)
    t0 = time.perf_counter()
    ev = held_out_cases(first=first, count=count)
    log(f"held-out instances {first}..{first + len(ev) - 1}: {sum(len(c.mesh.cells) for c in ev)} cells, "
        f"meshes in {time.perf_counter() - t0:.0f} s")
    fs_states = freestream_states(ev)
    b = Batch([c.mesh for c in ev], np.array([c.fs for c in ev]), make_config(block=128, **sch))
    b.set_state(fs_states)
    progress = lambda done, marching: log(f"  {done} steps, {marching} marching")
    rp = b.solve(tol=tol, max_steps=max_steps, check_every=100, progress=progress, progress_every=50)
    plain = b.get_state()
    log(f"plain: {rp['wall_s']:.1f} s, {(rp['status'] == 0).sum()} converged")
    tw = time.perf_counter()
    warm = [predicted_state(mlp, c.mesh.points, c.mesh.cells, c.mach, c.param, fs=c.fs, snap=snap) for c in ev]
    t_inf = time.perf_counter() - tw
    b.set_state(warm)
    e_w = b.solve(tol=-1.0, max_steps=1, check_every=1)["err"][:, 0]
    b.set_state(fs_states)
    e_f = b.solve(tol=-1.0, max_steps=1, check_every=1)["err"][:, 0]
    use = np.isfinite(e_w) & (e_w <= e_f)
    b.set_state([w if u else f for w, f, u in zip(warm, fs_states, use)])
    if sch["freeze_stall"] > 0:
        b.set_freeze(use.astype(np.int32) if freeze_warm else None)
    t_check = time.perf_counter() - tw - t_inf
    rw = b.solve(tol=tol, max_steps=max_steps, check_every=100, progress=progress, progress_every=50)
    warm_wall = time.perf_counter() - tw
    warm_states = b.get_state()
    log(f"warm: {warm_wall:.1f} s, {(rw['status'] == 0).sum()} converged")
    both = (rp["status"] == 0) & (rw["status"] == 0)
    s_err, d_max, d_l2 = [], [], []
    for k in range(len(ev)):
        if rp["status"][k] == 0:  # surrogate error against the converged plain field
            Wp = prim_from_cons(plain[k])
            s_err.append(float(np.linalg.norm(prim_from_cons(warm[k]) - Wp) / np.linalg.norm(Wp)))
        if both[k]:
            d_max.append(float(np.abs(warm_states[k] - plain[k]).max() / np.abs(plain[k]).max()))
            d_l2.append(float(np.linalg.norm(warm_states[k] - plain[k]) / np.linalg.norm(plain[k])))
    out = dict(first=first, count=len(ev), cells=int(sum(len(c.mesh.cells) for c in ev)),
               scheme={k: sch.get(k, 0) for k in ("flux", "limiter", "gradient", "cfl", "freeze_stall", "positivity")},
               plain=dict(wall_s=rp["wall_s"], converged=int((rp["status"] == 0).sum()),
                          steps_max=int(rp["steps"].max()), steps_sum=int(rp["steps"].sum())),
               warm=dict(wall_s=warm_wall, march_s=rw["wall_s"], inference_s=t_inf, fallback_check_s=t_check,
                         converged=int((rw["status"] == 0).sum()), steps_max=int(rw["steps"].max()),
                         steps_sum=int(rw["steps"].sum()), fallbacks=int((~use).sum())),
               speedup=rp["wall_s"] / warm_wall,
               freeze_warm=bool(freeze_warm), snap=snap,
               surrogate_rel_l2=dict(instances=len(s_err), mean=float(np.mean(s_err)) if s_err else None,
                                     max=float(np.max(s_err)) if s_err else None),
               converged_fields=dict(instances=int(both.sum()), max_rel_diff=max(d_max) if d_max else None,
                                     rel_l2_mean=float(np.mean(d_l2)) if d_l2 else None,
                                     rel_l2_max=max(d_l2) if d_l2 else None))
    if both.any():
        out["both_converged"] = dict(plain_s=float(rp["conv_s"][both].max()),
                                     warm_s=float(rw["conv_s"][both].max() + t_inf + t_check),
                                     plain_steps=int(rp["steps"][both].sum()), warm_steps=int(rw["steps"][both].sum()))
        out["both_converged"]["speedup"] = out["both_converged"]["plain_s"] / out["both_converged"]["warm_s"]
    out["instances"] = [dict(mach=round(c.mach, 4), thickness=round(c.param, 4), cells=len(c.mesh.cells),
                             plain=[int(a), int(s1), float(e1)], warm=[int(w), int(s2), float(e2)],
                             used_surrogate=bool(u))
                        for c, a, s1, e1, w, s2, e2, u in zip(ev, rp["status"], rp["steps"], rp["err"][:, 0],
                                                             rw["status"], rw["steps"], rw["err"][:, 0], use)]
    with open(out_file, "w") as f:
        json.dump(out, f)
    log(json.dumps({k: v for k, v in out.items() if k != "instances"}))
    return out


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--cases", type=int, default=2000)
    t.add_argument("--cells", type=int, default=50000)
    t.add_argument("--chunk", type=int, default=250)
    t.add_argument("--mlp-steps", type=int, default=8000)
    t.add_argument("--max-steps", type=int, default=20000)
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--out", required=True)
    e.add_argument("--first", type=int, default=0)
    e.add_argument("--count", type=int, default=256)
    e.add_argument("--max-steps", type=int, default=60000)
    e.add_argument("--freeze-warm", type=int, default=1)
    e.add_argument("--snap", type=float, default=0.02)
    e.add_argument("--positivity", type=int, default=0)
    a = p.parse_args(argv)
    if a.cmd == "train":
        train(a.out, n_cases=a.cases, cells=a.cells, chunk=a.chunk, mlp_steps=a.mlp_steps, max_steps=a.max_steps)
    else:
        evaluate(a.model, a.out, first=a.first, count=a.count, max_steps=a.max_steps, freeze_warm=bool(a.freeze_warm),
                 snap=a.snap, positivity=a.positivity)
    return 0


if __name__ == "__main__":
    sys.exit(main())
