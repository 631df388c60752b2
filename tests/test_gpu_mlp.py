"""GPU: the hand-written BF16 MFMA MLP kernels against the FP64 oracle.

forward: relative error of the outputs <= 2e-2 (BF16 activations);
gradients (hadros_mlp_grad, every parameter): per-layer relative error of the
weight and bias gradients <= 3e-2 and cosine similarity >= 0.999 against the
FP64 oracle run on the same BF16-rounded weights and inputs with BF16-rounded
activations and back-propagated gradients."""
import numpy as np
import pytest

from hadros_amd import ffi
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def setup_net(n_in=14, width=256, depth=4, n_out=16, n=3000, seed=1):
    import torch

    net = ffi.MLP(n_in, width, depth, n_out, train_batch=4096, infer_batch=2048, seed=seed)
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, n_in)).astype(np.float32)
    t = rng.normal(size=(n, n_out)).astype(np.float32)
    xm, xs = x.mean(0), x.std(0) + 0.1
    tm, ts = t.mean(0), t.std(0) + 0.1
    net.set_norm(xm, xs, tm, ts)
    X = torch.from_numpy(x).cuda()
    T = torch.from_numpy(t).cuda()
    layers = mo.unflatten(mo.bf16(net.params()), n_in, width, depth, n_out)
    xn = mo.bf16((x - xm) / xs)
    tn = (t - tm) / ts
    return net, X, T, layers, xn, tn, (xm, xs, tm, ts)


def test_forward_matches_oracle():
    net, X, T, layers, xn, tn, (xm, xs, tm, ts) = setup_net()
    y = net.forward(X).cpu().numpy().astype(np.float64)
    yr, _ = mo.forward(layers, xn, round_act=True)
    yr = yr * ts + tm
    err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    assert err < 2e-2, err


def test_gradients_match_oracle():
    n_in, width, depth, n_out = 14, 256, 4, 16
    net, X, T, layers, xn, tn, _ = setup_net(n=4096)
    g, loss = net.grad(X, T)
    lr, gr = mo.loss_grad(layers, xn, tn, round_act=True)
    assert abs(loss - lr) <= 2e-2 * abs(lr)
    sizes = [n_in] + [width] * depth + [n_out]
    q = 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        for cnt in (a * b, b):
            ga, gb = g[q:q + cnt], gr[q:q + cnt]
            rel = np.linalg.norm(ga - gb) / max(np.linalg.norm(gb), 1e-30)
            cos = float(ga @ gb / max(np.linalg.norm(ga) * np.linalg.norm(gb), 1e-30))
            assert rel < 3e-2 and cos > 0.999, (a, b, cnt, rel, cos)
            q += cnt
    assert q == g.size


def test_training_fits_a_smooth_function():
    import torch

    net = ffi.MLP(14, 256, 4, 16, train_batch=8192, infer_batch=65536, seed=3)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, size=(65536, 14)).astype(np.float32)
    W = rng.normal(size=(14, 16)).astype(np.float32)
    t = np.sin(x @ W * 0.7).astype(np.float32)
    net.set_norm(np.zeros(14), np.ones(14), np.zeros(16), np.ones(16))
    X, T = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    losses = []
    for step in range(400):
        i = (step * 8192) % 65536
        losses.append(net.train_step(X[i:i + 8192], T[i:i + 8192], 1e-3))
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:5]), (losses[:5], losses[-5:])
    # padded first-layer rows stay exactly zero and inference agrees with training forward
    y = net.forward(X[:4096]).cpu().numpy()
    assert np.mean((y - t[:4096]) ** 2) < 0.2 * np.mean(t ** 2)
