"""TEST INFRASTRUCTURE ONLY: FP64 NumPy restatement of the surrogate MLP
(n_in -> width x depth SiLU -> n_out linear, MSE on standardised targets),
used to check the GPU forward and backward kernels.

Pinned by central finite differences of its own loss (tests/test_mlp_oracle.py).
The GPU keeps weights, inputs and activations in BF16; `bf16()` reproduces
that rounding so the comparison isolates the kernels' arithmetic.
"""
import numpy as np


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p, n_in, width, depth, n_out):
    """Flat parameters W1 (n_in x width), b1, ..., W_L, b_L -> list of (W, b)."""
    sizes = [n_in] + [width] * depth + [n_out]
    out, q = [], 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        W = np.asarray(p[q:q + a * b], np.float64).reshape(a, b)
        q += a * b
        bb = np.asarray(p[q:q + b], np.float64)
        q += b
        out.append((W, bb))
    assert q == len(p)
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(layers, xn, round_act=False):
    """xn: standardised inputs [n x n_in].  Returns (outputs, cache)."""
    h = xn
    cache = []
    for li, (W, b) in enumerate(layers):
        z = h @ W + b
        cache.append((h, z))
        if li < len(layers) - 1:
            h = silu(z)
            if round_act:
                h = bf16(h)
        else:
            h = z
    return h, cache


def loss_grad(layers, xn, tn, round_act=False):
    """MSE (mean over all outputs) and its gradient, flat in the same order."""
    y, cache = forward(layers, xn, round_act)
    n, k = y.shape
    d = y - tn
    loss = float(np.mean(d * d))
    g = 2.0 * d / (n * k)
    if round_act:
        g = bf16(g)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        h, z = cache[li]
        W, b = layers[li]
        grads.append((h.T @ g, g.sum(0)))
        if li > 0:
            gp = (g @ W.T) * dsilu(cache[li - 1][1])
            g = bf16(gp) if round_act else gp
    grads.reverse()
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat
