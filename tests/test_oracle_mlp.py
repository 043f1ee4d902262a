"""Pins of the FP64 MLP oracle (oracle/mlp_oracle.py): gradients against central finite
differences and against PyTorch autograd (CPU, float64); BF16 rounding against known values."""
import numpy as np
import pytest

from oracle.mlp_oracle import MLP, bf16_round


def small():
    net = MLP(4, 16, 3, 4)
    W, b = net.init(seed=3)
    rng = np.random.default_rng(5)
    feat = rng.uniform(-1, 1, (64, 4))
    targ = rng.normal(size=(64, 4))
    norm = (np.array([0.1, -0.2, 0.0, 0.3]), np.array([1.5, 0.7, 1.0, 2.0]),
            np.array([0.2, 0.0, -0.1, 0.5]), np.array([1.0, 2.0, 0.5, 1.5]))
    return net, W.astype(np.float64), b.astype(np.float64), feat, targ, norm


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-3, 1.0 + 2.0 ** -7 + 2.0 ** -8])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -7, -2.5, np.float32(3.0e-3), 1.0 + 2.0 ** -6])
    got = bf16_round(vals)
    assert got[0] == 1.0 and got[1] == 1.0          # tie rounds to even
    assert got[2] == 1.0 + 2.0 ** -7                # above the tie rounds up
    assert got[3] == -2.5 and got[5] == want[5]     # tie with odd mantissa rounds up
    assert abs(got[4] - 3.0e-3) / 3.0e-3 < 2.0 ** -8


def test_mlp_gradients_vs_finite_differences():
    net, W, b, feat, targ, norm = small()
    loss, gW, gb = net.loss_and_grads(W, b, feat, targ, norm)
    rng = np.random.default_rng(0)
    h = 1e-6
    for i in rng.choice(len(W), 25, replace=False):
        Wp, Wm = W.copy(), W.copy()
        Wp[i] += h
        Wm[i] -= h
        fd = (net.loss_and_grads(Wp, b, feat, targ, norm)[0] - net.loss_and_grads(Wm, b, feat, targ, norm)[0]) / (2 * h)
        assert fd == pytest.approx(gW[i], rel=1e-5, abs=1e-9)
    for i in range(len(b)):
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        fd = (net.loss_and_grads(W, bp, feat, targ, norm)[0] - net.loss_and_grads(W, bm, feat, targ, norm)[0]) / (2 * h)
        assert fd == pytest.approx(gb[i], rel=1e-5, abs=1e-9)


def test_mlp_gradients_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    net, W, b, feat, targ, norm = small()
    loss, gW, gb = net.loss_and_grads(W, b, feat, targ, norm)
    Wt = torch.tensor(W, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    x = (torch.tensor(feat) - torch.tensor(norm[0])) * torch.tensor(norm[1])
    ow = ob = 0
    for l, (k, m) in enumerate(net.dims):
        z = x @ Wt[ow:ow + k * m].reshape(m, k).T + bt[ob:ob + m]
        ow += k * m
        ob += m
        x = torch.nn.functional.silu(z) if l < len(net.dims) - 1 else z
    T = (torch.tensor(targ) - torch.tensor(norm[2])) * torch.tensor(norm[3])
    lt = torch.mean((x - T) ** 2)
    lt.backward()
    assert float(lt) == pytest.approx(loss, rel=1e-12)
    assert np.allclose(Wt.grad.numpy(), gW, rtol=1e-10, atol=1e-14)
    assert np.allclose(bt.grad.numpy(), gb, rtol=1e-10, atol=1e-14)
