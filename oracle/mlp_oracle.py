"""FP64 NumPy restatement of the coordinate-MLP surrogate (the checker of the BF16-MFMA kernels).

TEST INFRASTRUCTURE ONLY: imported by tests/. The reference has no MLP; this module is pinned
by finite differences and by PyTorch autograd (tests/test_oracle_mlp.py).

Network: x -> depth hidden layers z = W x + b, y = silu(z) -> output z = W x + b. Parameters are
concatenated per layer as in the C-ABI: W_l row-major [out x in], then b_l. The loss is the mean
squared error of the normalised outputs, (v - mean) * scale. With bf16=True the values the
device stores in BF16 (inputs, weights, hidden activations, hidden pre-activations as the
backward pass reads them, output gradients and hidden pre-activation gradients) are rounded to
BF16 (round to nearest even) at the same points; all arithmetic stays FP64.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


class MLP:
    def __init__(self, n_in, width, depth, n_out):
        self.dims = [(n_in if l == 0 else width, n_out if l == depth else width) for l in range(depth + 1)]
        self.nw = sum(k * m for k, m in self.dims)
        self.nb = sum(m for _, m in self.dims)

    def split(self, W, b):
        Ws, bs, ow, ob = [], [], 0, 0
        for k, m in self.dims:
            Ws.append(np.asarray(W[ow:ow + k * m], np.float64).reshape(m, k))
            bs.append(np.asarray(b[ob:ob + m], np.float64))
            ow += k * m
            ob += m
        return Ws, bs

    def forward(self, W, b, feat, norm, bf16=False):
        """Returns (outputs in normalised units [n, n_out], cache)."""
        in_mean, in_scale = norm[0], norm[1]
        q = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
        Ws, bs = self.split(W, b)
        X = q((np.asarray(feat, np.float64) - in_mean) * in_scale)
        xs, zs = [X], []
        for l, (Wl, bl) in enumerate(zip(Ws, bs)):
            Z = xs[-1] @ q(Wl).T + bl
            zs.append(Z)
            if l < len(Ws) - 1:
                xs.append(q(silu(Z)))
        return zs[-1], (xs, zs, Ws)

    def predict(self, W, b, feat, norm, bf16=False):
        out, _ = self.forward(W, b, feat, norm, bf16)
        return out / norm[3] + norm[2]

    def loss_and_grads(self, W, b, feat, targ, norm, bf16=False):
        q = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
        out, (xs, zs, Ws) = self.forward(W, b, feat, norm, bf16)
        T = (np.asarray(targ, np.float64) - norm[2]) * norm[3]
        n, k = T.shape
        E = out - T
        loss = float(np.mean(E * E))
        dZ = q(2.0 * E / (n * k))
        gW, gb = [None] * len(Ws), [None] * len(Ws)
        for l in range(len(Ws) - 1, -1, -1):
            gW[l] = dZ.T @ xs[l]
            gb[l] = dZ.sum(axis=0)
            if l > 0:
                dZ = q((dZ @ q(Ws[l])) * dsilu(q(zs[l - 1])))
        return loss, np.concatenate([g.ravel() for g in gW]), np.concatenate(gb)

    def init(self, seed=0):
        rng = np.random.default_rng(seed)
        W = np.concatenate([rng.uniform(-1, 1, m * k) * np.sqrt(6.0 / k) for k, m in self.dims])
        b = rng.uniform(-0.1, 0.1, self.nb)
        return W.astype(np.float32), b.astype(np.float32)
