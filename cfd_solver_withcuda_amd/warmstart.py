"""Learned warm start of batched steady solves on the bump family.

  1. steady solutions of `n_train` bump cases (Latin hypercube in Mach x thickness) from the
     batched FP64 solver;
  2. a coordinate MLP (x, y, Mach, thickness) -> (rho, u, v, p) trained on their cell values
     with the BF16-MFMA kernels;
  3. `n_eval` held-out cases solved twice in one batch each: from freestream (plain) and from the
     surrogate's field (warm, with per-instance fallback to freestream when the warm start's
     first-step residual exceeds the freestream start's). Both marches run to Error[0] <= tol.

The scheme is HLLC + MUSCL/Van Leer + WLSQ gradients with limiter freezing on a 300-step stall
(without freezing, Van Leer marches on this family stall above 1e-8). Warm-started instances record
their limiter factors on the first step (their shocks are already in place); a frozen march whose
error grows tenfold returns to the evaluated limiter. A march frozen at one state converges to the
fixed point of that frozen scheme, so two marches frozen at different states end at slightly
different fields. With `limiter="vanalbada", freeze_stall=0` (the differentiable limiter, no
freezing) or `order=1, freeze_stall=0` (first order) both marches share one discrete solution
wherever the plain march converges. The plain
FP64 march decides the answer; the surrogate only moves its starting point. `field_tol` continues
both converged batches, untimed, to a tighter tolerance and compares the fields there too.
"""
from __future__ import annotations

import time

import numpy as np

from . import cases
from .batch import Batch
from .solver import cons_from_prim, make_config
from .surrogate import MLP, features, predicted_state, prim_from_cons

SCHEME = dict(flux="hllc", order=2, limiter="vanleer", gradient="wlsq", cfl=1.2, rk=3, lts=1,
              roe_bdiss=1, bc=cases.BUMP_BCS, freeze_stall=300)


def freestream_states(cs):
    return [np.tile(cons_from_prim(np.array([c.fs[0]]), c.fs[1], c.fs[2], c.fs[3]), (len(c.mesh.cells), 1))
            for c in cs]


def scheme(cfl=None, **over):
    kw = dict(SCHEME)
    if cfl is not None:
        kw["cfl"] = cfl
    kw.update({k: v for k, v in over.items() if v is not None})
    return kw


def solve_batch(cs, tol, max_steps, check_every, block=128, start=None, cfl=None, **over):
    b = Batch([c.mesh for c in cs], np.array([c.fs for c in cs]), make_config(block=block, **scheme(cfl, **over)))
    if start is not None:
        b.set_state(start)
    r = b.solve(tol=tol, max_steps=max_steps, check_every=check_every)
    return b, r


def training_set(cs, states):
    feat, targ = [], []
    for c, U in zip(cs, states):
        feat.append(features(c.mesh.points, c.mesh.cells, c.mach, c.param))
        targ.append(prim_from_cons(U).astype(np.float32))
    return np.concatenate(feat), np.concatenate(targ)


def run(n_train=64, n_eval=32, train_cells=30000, eval_cells=60000, tol=1e-8, train_tol=1e-8,
        max_steps=60000, check_every=100, mlp_steps=3000, mlp_batch=16384, seed=5, cfl=None,
        freeze_warm=True, warm_cfl=None, history=False, log=print, limiter=None, freeze_stall=None,
        field_tol=None, snap=0.02, order=None):
    out = {}
    over = dict(limiter=limiter, freeze_stall=freeze_stall, order=order)
    sch = scheme(cfl, **over)
    freeze_warm = freeze_warm and sch["freeze_stall"] > 0
    out["scheme"] = {k: sch[k] for k in ("flux", "order", "limiter", "gradient", "cfl", "freeze_stall")}
    t0 = time.perf_counter()
    tr = cases.bump_family(n_train, seed=seed, cells=train_cells, stream=200)
    b_tr, rt = solve_batch(tr, train_tol, max_steps, check_every, cfl=cfl, **over)
    states = b_tr.get_state()
    # converged solutions, and the finite end states of the ones still marching at max_steps
    ok = rt["status"] >= 0
    out["train"] = dict(cases=n_train, cells=int(sum(len(c.mesh.cells) for c in tr)),
                        converged=int((rt["status"] == 0).sum()), used=int(ok.sum()),
                        steps_max=int(rt["steps"].max()), wall_s=rt["wall_s"])
    log(f"training solves: {out['train']}")
    del b_tr
    feat, targ = training_set([c for c, k in zip(tr, ok) if k], [U for U, k in zip(states, ok) if k])
    mlp = MLP(4, 256, 6, 4, max_rows=mlp_batch, seed=seed)
    mlp.fit_normalization(feat, targ)
    t1 = time.perf_counter()
    hist = mlp.train(feat, targ, steps=mlp_steps, batch=mlp_batch, lr=2e-3, lr_final=2e-5, seed=seed + 1,
                     log_every=max(1, mlp_steps // 20))
    t_train = time.perf_counter() - t1
    out["mlp"] = dict(samples=int(len(feat)), steps=mlp_steps, batch=mlp_batch, train_s=t_train,
                      loss_first=float(hist[0]), loss_last=float(hist[-1]))
    log(f"mlp: {out['mlp']}")

    ev = cases.bump_family(n_eval, seed=seed + 100, cells=eval_cells, stream=300)
    fs_states = freestream_states(ev)
    # plain: freestream start
    b = Batch([c.mesh for c in ev], np.array([c.fs for c in ev]), make_config(block=128, **sch))
    b.set_state(fs_states)
    rp = b.solve(tol=tol, max_steps=max_steps, check_every=check_every)
    plain_states = b.get_state()
    if history:
        out["plain_history"] = b.history().tolist()
    out["plain"] = dict(wall_s=rp["wall_s"], steps_max=int(rp["steps"].max()),
                        steps_mean=float(rp["steps"].mean()), converged=int((rp["status"] == 0).sum()))
    log(f"plain: {out['plain']}")
    # warm: inference, fallback test on the first step's residual, march
    tw = time.perf_counter()
    warm = [predicted_state(mlp, c.mesh.points, c.mesh.cells, c.mach, c.param, fs=c.fs, snap=snap) for c in ev]
    t_inf = time.perf_counter() - tw
    b.set_state(warm)
    e_w = b.solve(tol=-1.0, max_steps=1, check_every=1)["err"][:, 0]
    b.set_state(fs_states)
    e_f = b.solve(tol=-1.0, max_steps=1, check_every=1)["err"][:, 0]
    use = np.isfinite(e_w) & (e_w <= e_f)
    start = [w if u else f for w, f, u in zip(warm, fs_states, use)]
    b.set_state(start)
    if warm_cfl is not None:
        b.set_cfl(warm_cfl)
    if freeze_warm:
        # a start close to the solution already has its shocks in place: record the limiter
        # factors on the first step and march frozen (a frozen march whose error grows tenfold
        # returns to the evaluated limiter)
        b.set_freeze(use.astype(np.int32))
    t_check = time.perf_counter() - tw - t_inf
    rw = b.solve(tol=tol, max_steps=max_steps, check_every=check_every)
    warm_wall = time.perf_counter() - tw
    warm_states = b.get_state()
    if freeze_warm:
        b.set_freeze(None)
    if history:
        out["warm_history"] = b.history().tolist()
    out["warm"] = dict(wall_s=warm_wall, march_s=rw["wall_s"], inference_s=t_inf, fallback_check_s=t_check,
                       steps_max=int(rw["steps"].max()), steps_mean=float(rw["steps"].mean()),
                       converged=int((rw["status"] == 0).sum()), fallbacks=int((~use).sum()))
    log(f"warm: {out['warm']}")
    # surrogate error against the converged field, and agreement of the two converged fields,
    # over the instances both marches converged
    both = (rp["status"] == 0) & (rw["status"] == 0)
    s_err, d_conv, d_l2 = [], [], []
    for k, (Uw0, Up, Uw) in enumerate(zip(warm, plain_states, warm_states)):
        if not both[k]:
            continue
        Wp = prim_from_cons(Up)
        s_err.append(np.linalg.norm(prim_from_cons(Uw0) - Wp) / np.linalg.norm(Wp))
        d_conv.append(np.abs(Uw - Up).max() / np.abs(Up).max())
        d_l2.append(np.linalg.norm(Uw - Up) / np.linalg.norm(Up))
    out["surrogate_rel_l2"] = dict(mean=float(np.mean(s_err)) if s_err else None,
                                   max=float(np.max(s_err)) if s_err else None)
    out["converged_fields_max_rel_diff"] = float(np.max(d_conv)) if d_conv else None
    out["converged_fields_rel_l2_diff"] = dict(mean=float(np.mean(d_l2)), max=float(np.max(d_l2))) if d_l2 else None
    if field_tol is not None and both.any():
        # both converged batches marched on (untimed) to field_tol: the distance between the two
        # fixed points, with less of the remaining march error in it
        fin = {}
        for name, states in (("plain", plain_states), ("warm", warm_states)):
            b.set_state(states)
            r2 = b.solve(tol=field_tol, max_steps=max_steps, check_every=check_every)
            fin[name] = (b.get_state(), r2["status"])
        d_max, d_l2t = [], []
        for k in range(len(ev)):
            if both[k] and fin["plain"][1][k] == 0 and fin["warm"][1][k] == 0:
                Up, Uw = fin["plain"][0][k], fin["warm"][0][k]
                d_max.append(np.abs(Uw - Up).max() / np.abs(Up).max())
                d_l2t.append(np.linalg.norm(Uw - Up) / np.linalg.norm(Up))
        out["fields_at_field_tol"] = dict(tol=field_tol, instances=len(d_max),
                                          max_rel_diff=float(np.max(d_max)) if d_max else None,
                                          rel_l2_diff_max=float(np.max(d_l2t)) if d_l2t else None)
    out["speedup"] = out["plain"]["wall_s"] / out["warm"]["wall_s"]
    # the instances both marches converged: time until the last of them converged, and steps
    if both.any():
        out["common"] = dict(instances=int(both.sum()),
                             plain_s=float(rp["conv_s"][both].max()),
                             warm_s=float(rw["conv_s"][both].max() + t_inf + t_check),
                             plain_steps=int(rp["steps"][both].sum()), warm_steps=int(rw["steps"][both].sum()))
        out["common"]["speedup"] = out["common"]["plain_s"] / out["common"]["warm_s"]
        out["common"]["steps_ratio"] = out["common"]["plain_steps"] / max(1, out["common"]["warm_steps"])
    out["instances"] = [dict(mach=c.mach, thickness=c.param, plain_status=int(a), plain_steps=int(b2),
                             warm_status=int(w), warm_steps=int(ws))
                        for c, a, b2, w, ws in zip(ev, rp["status"], rp["steps"], rw["status"], rw["steps"])]
    out["total_s"] = time.perf_counter() - t0
    log(f"surrogate rel L2 {out['surrogate_rel_l2']}, converged diff {out['converged_fields_max_rel_diff']}, "
        f"speedup {out['speedup']:.2f}, common {out.get('common')}")
    return out


__all__ = ["run", "solve_batch", "training_set", "freestream_states", "SCHEME"]
