"""GPU tests of the BF16-MFMA surrogate kernels (cfd_mlp_*) against the FP64 MLP oracle with
BF16 rounding emulated at the points the device stores BF16: forward predictions, loss and every
layer's weight and bias gradients of the full 4 -> 6 x 256 -> 4 network; and Adam training that
fits a smooth field."""
import numpy as np
import pytest

from cfd_solver_withcuda_amd.surrogate import MLP
from oracle.mlp_oracle import MLP as OMLP

pytestmark = pytest.mark.gpu


def data(n, seed=0):
    rng = np.random.default_rng(seed)
    feat = np.column_stack([rng.uniform(-1, 2, n), rng.uniform(0, 1, n), rng.uniform(1.2, 3.0, n),
                            rng.uniform(0.04, 0.12, n)]).astype(np.float32)
    x, y, M, t = feat.T.astype(np.float64)
    targ = np.column_stack([1.4 * (1 + 0.3 * np.sin(3 * x) * y * t * 8), M * (1 - 0.2 * np.exp(-((x - 0.5) ** 2) * 8)),
                            0.3 * np.sin(4 * x) * (1 - y), 1.0 + 0.5 * np.cos(2 * x + y) * M / 3]).astype(np.float32)
    return feat, targ


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def net():
    m = MLP(4, 256, 6, 4, max_rows=4096, seed=11)
    feat, targ = data(4096)
    m.fit_normalization(feat, targ)
    return m


def test_parameter_count(net):
    assert net.nw + net.nb == 331268


def test_gradients_vs_oracle(net):
    print("case mlp gradients 1024 rows", flush=True)
    feat, targ = data(1024, seed=1)
    W, b = net.get_params()
    loss, gW, gb = net.gradients(feat, targ)
    o = OMLP(4, 256, 6, 4)
    norm = [a.astype(np.float64) for a in net.norm]
    oloss, ogW, ogb = o.loss_and_grads(W.astype(np.float64), b.astype(np.float64), feat.astype(np.float64),
                                       targ.astype(np.float64), norm, bf16=True)
    assert loss == pytest.approx(oloss, rel=2e-3)
    ow = ob = 0
    for l, (k, m) in enumerate(o.dims):
        e_w = rel(gW[ow:ow + k * m], ogW[ow:ow + k * m])
        e_b = rel(gb[ob:ob + m], ogb[ob:ob + m])
        print(f"layer {l}: rel err dW {e_w:.2e} db {e_b:.2e}", flush=True)
        assert e_w < 3e-2 and e_b < 3e-2, (l, e_w, e_b)
        ow += k * m
        ob += m


def test_predict_vs_oracle(net):
    feat, _ = data(1500, seed=2)  # not a multiple of 512: the tail chunk is padded
    W, b = net.get_params()
    out = net.predict(feat)
    o = OMLP(4, 256, 6, 4)
    norm = [a.astype(np.float64) for a in net.norm]
    want = o.predict(W.astype(np.float64), b.astype(np.float64), feat.astype(np.float64), norm, bf16=True)
    scale = np.abs(want).max(axis=0)
    assert np.all(np.abs(out - want).max(axis=0) <= 2e-2 * scale), (np.abs(out - want).max(axis=0), scale)


def test_training_fits_a_smooth_field():
    m = MLP(4, 256, 6, 4, max_rows=8192, seed=5)
    feat, targ = data(20000, seed=3)
    m.fit_normalization(feat, targ)
    hist = m.train(feat, targ, steps=400, batch=4096, lr=2e-3, lr_final=1e-4, seed=9, log_every=50)
    print("loss history", hist, flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < hist[0] / 50
    test_feat, test_targ = data(3000, seed=4)
    pred = m.predict(test_feat)
    err = np.abs(pred - test_targ).mean(axis=0) / np.abs(test_targ).mean(axis=0)
    print("relative mean abs error per output", err, flush=True)
    assert np.all(err < 0.05)


def test_captured_training_step_equals_kernel_launches():
    """Training from one captured step replayed (the default) gives the same weights and loss
    history, bit for bit, as launching each kernel, over two train calls in a row (the second
    continues Adam's step count, with a smaller batch and a different data set); the host does its
    own synchronised work between the calls."""
    feat, targ = data(12000, seed=7)
    feat2, targ2 = data(5000, seed=8)
    out = []
    for mode in ("kernels", "graph"):
        m = MLP(4, 256, 6, 4, max_rows=4096, seed=21)
        m.fit_normalization(feat, targ)
        m.set_launch(mode)
        h1 = m.train(feat, targ, steps=30, batch=2048, lr=2e-3, lr_final=1e-4, seed=3, log_every=5)
        W1, b1 = m.get_params()
        p1 = m.predict(feat[:600])
        h2 = m.train(feat2, targ2, steps=17, batch=1024, lr=1e-3, lr_final=1e-4, seed=4, log_every=4)
        W2, b2 = m.get_params()
        out.append((h1, W1, b1, p1, h2, W2, b2))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert not np.array_equal(out[0][1], out[0][5])
    with pytest.raises(Exception):
        MLP(4, 32, 1, 4, max_rows=512).set_launch(2)
