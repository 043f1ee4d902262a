"""Learned warm start for the steady FP64 march: a coordinate MLP (x, y, Mach, thickness) ->
(rho, u, v, p) trained on the solver's own steady solutions, on libcfdhip.so's BF16-MFMA kernels
(cfd_mlp_*). The reference has no counterpart. The plain march stays the arbiter: the predicted
field only initialises it, and a start whose first-step residual exceeds the freestream start's
falls back to freestream (the fallback test is in warmstart.run).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, fptr, lib

GAMMA = 1.4


class MLP:
    """Handle of a device MLP: n_in -> depth x width (SiLU) -> n_out."""

    def __init__(self, n_in=4, width=256, depth=6, n_out=4, max_rows=65536, seed=0):
        h = ctypes.c_void_p()
        check(lib().cfd_mlp_create(n_in, width, depth, n_out, max_rows, seed, ctypes.byref(h)), "cfd_mlp_create")
        self.h = h
        self.n_in, self.width, self.depth, self.n_out = n_in, width, depth, n_out
        nw, nb = ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().cfd_mlp_num_params(h, ctypes.byref(nw), ctypes.byref(nb)), "cfd_mlp_num_params")
        self.nw, self.nb = nw.value, nb.value
        self.norm = (np.zeros(n_in, np.float32), np.ones(n_in, np.float32),
                     np.zeros(n_out, np.float32), np.ones(n_out, np.float32))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().cfd_mlp_destroy(h)
            except Exception:
                pass
            self.h = None

    def set_normalization(self, in_mean, in_scale, out_mean, out_scale):
        self.norm = tuple(np.ascontiguousarray(a, np.float32) for a in (in_mean, in_scale, out_mean, out_scale))
        check(lib().cfd_mlp_set_normalization(self.h, *[fptr(a) for a in self.norm]), "cfd_mlp_set_normalization")

    def fit_normalization(self, feat, targ):
        feat = np.asarray(feat, np.float64)
        targ = np.asarray(targ, np.float64)
        lo, hi = feat.min(axis=0), feat.max(axis=0)
        self.set_normalization(0.5 * (lo + hi), 2.0 / np.maximum(hi - lo, 1e-12),
                               targ.mean(axis=0), 1.0 / np.maximum(targ.std(axis=0), 1e-12))

    def get_params(self):
        W = np.zeros(self.nw, np.float32)
        b = np.zeros(self.nb, np.float32)
        check(lib().cfd_mlp_get_params(self.h, fptr(W), fptr(b)), "cfd_mlp_get_params")
        return W, b

    def set_params(self, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        check(lib().cfd_mlp_set_params(self.h, fptr(W), fptr(b)), "cfd_mlp_set_params")

    def gradients(self, feat, targ):
        feat = np.ascontiguousarray(feat, np.float32)
        targ = np.ascontiguousarray(targ, np.float32)
        gW = np.zeros(self.nw, np.float32)
        gb = np.zeros(self.nb, np.float32)
        loss = ctypes.c_double()
        check(lib().cfd_mlp_gradients(self.h, fptr(feat), fptr(targ), len(feat), fptr(gW), fptr(gb),
                                      ctypes.byref(loss)), "cfd_mlp_gradients")
        return loss.value, gW, gb

    def train(self, feat, targ, steps, batch=16384, lr=2e-3, lr_final=5e-5, seed=1, log_every=50):
        feat = np.ascontiguousarray(feat, np.float32)
        targ = np.ascontiguousarray(targ, np.float32)
        hist = np.zeros((steps + log_every - 1) // log_every, np.float32)
        check(lib().cfd_mlp_train(self.h, fptr(feat), fptr(targ), len(feat), steps, batch, lr, lr_final, seed,
                                  fptr(hist), log_every), "cfd_mlp_train")
        return hist

    def set_launch(self, mode):
        """Training steps: "kernels" (launched one by one; default) or "graph" (one captured step,
        replayed). Results are identical."""
        m = {"kernels": 0, "graph": 1}[mode] if isinstance(mode, str) else int(mode)
        check(lib().cfd_mlp_set_launch(self.h, m), "cfd_mlp_set_launch")

    def predict(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((len(feat), self.n_out), np.float32)
        check(lib().cfd_mlp_predict(self.h, fptr(feat), len(feat), fptr(out)), "cfd_mlp_predict")
        return out

    def time_step(self, batch, reps):
        ms = ctypes.c_float()
        check(lib().cfd_mlp_time_step(self.h, batch, reps, ctypes.byref(ms)), "cfd_mlp_time_step")
        return float(ms.value)


def cell_centres(points, cells):
    """Vertex-average centres of triangles/quads (cells padded with -1), as the reference's
    Compute_Centroid (src/Grid_Computations.cpp:957-980)."""
    pts = np.asarray(points, np.float64)[:, :2]
    c = np.asarray(cells)
    valid = c >= 0
    n = valid.sum(axis=1)
    xy = np.where(valid[..., None], pts[np.where(valid, c, 0)], 0.0).sum(axis=1)
    return xy / n[:, None]


def features(points, cells, mach, thickness):
    ctr = cell_centres(points, cells)
    n = len(ctr)
    return np.column_stack([ctr, np.full(n, mach), np.full(n, thickness)]).astype(np.float32)


def prim_from_cons(U):
    rho = U[:, 0]
    u = U[:, 1] / rho
    v = U[:, 2] / rho
    p = (GAMMA - 1.0) * (U[:, 3] - 0.5 * rho * (u * u + v * v))
    return np.column_stack([rho, u, v, p])


def cons_from_prim_arr(W):
    rho, u, v, p = W[:, 0], W[:, 1], W[:, 2], W[:, 3]
    return np.column_stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)])


def predicted_state(mlp: MLP, points, cells, mach, thickness, floor=1e-3, fs=None, snap=0.0):
    """Conserved FP64 state from the surrogate's primitive prediction (density and pressure
    floored so the state is admissible). With `fs` (rho, u, v, p) and snap > 0, cells whose
    prediction lies within `snap` of that state (relative to rho, |u| and p of it) take it exactly:
    ahead of the leading shock a supersonic flow is that state, and the network's small errors
    there would otherwise have to be convected out of the domain."""
    W = mlp.predict(features(points, cells, mach, thickness)).astype(np.float64)
    W[:, 0] = np.maximum(W[:, 0], floor)
    W[:, 3] = np.maximum(W[:, 3], floor)
    if fs is not None and snap > 0.0:
        r, u, v, p = (float(x) for x in fs)
        q = max(np.hypot(u, v), 1e-300)
        near = ((np.abs(W[:, 0] - r) <= snap * r) & (np.abs(W[:, 1] - u) <= snap * q) &
                (np.abs(W[:, 2] - v) <= snap * q) & (np.abs(W[:, 3] - p) <= snap * p))
        W[near] = (r, u, v, p)
    return cons_from_prim_arr(W)


__all__ = ["MLP", "features", "cell_centres", "prim_from_cons", "cons_from_prim_arr", "predicted_state"]


def _time_training(steps, batch=16384, samples=1 << 20, launch="kernels"):
    """Wall time per training step of the 4 -> 6 x 256 -> 4 network on random samples of a smooth
    field (5 untimed steps first)."""
    import time

    rng = np.random.default_rng(0)
    feat = rng.uniform(0, 1, (samples, 4)).astype(np.float32)
    targ = np.column_stack([np.sin(3 * feat[:, 0]), feat[:, 1] * feat[:, 2], np.cos(feat[:, 3]),
                            feat[:, 0] ** 2]).astype(np.float32)
    m = MLP(4, 256, 6, 4, max_rows=batch, seed=1)
    m.fit_normalization(feat, targ)
    m.set_launch(launch)
    m.train(feat, targ, steps=5, batch=batch, log_every=1000)
    t = time.perf_counter()
    m.train(feat, targ, steps=steps, batch=batch, log_every=1 << 30)
    dt = (time.perf_counter() - t) / steps
    flops = 6 * (m.nw + m.nb) * batch
    return dict(ms_per_step=1e3 * dt, tflops=flops / dt / 1e12, batch=batch, launch=launch)


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="time the surrogate's training step on the GPU")
    ap.add_argument("--time", type=int, default=200, help="timed training steps")
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--launch", default="kernels", choices=["graph", "kernels"])
    a = ap.parse_args()
    print(json.dumps(_time_training(a.time, a.batch, launch=a.launch)))
