"""CPU: pin the FP64 MLP oracle's gradient with central finite differences."""
import numpy as np

from oracle import mlp_oracle as mo


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, width, depth, n_out = 5, 16, 2, 3
    sizes = [n_in] + [width] * depth + [n_out]
    n_p = sum(a * b + b for a, b in zip(sizes[:-1], sizes[1:]))
    p = rng.normal(size=n_p) * 0.4
    x = rng.normal(size=(11, n_in))
    t = rng.normal(size=(11, n_out))
    loss, g = mo.loss_grad(mo.unflatten(p, n_in, width, depth, n_out), x, t)
    idx = rng.choice(n_p, 60, replace=False)
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        lp, _ = mo.loss_grad(mo.unflatten(pp, n_in, width, depth, n_out), x, t)
        lm, _ = mo.loss_grad(mo.unflatten(pm, n_in, width, depth, n_out), x, t)
        num = (lp - lm) / (2 * e)
        assert abs(num - g[i]) <= 1e-7 + 1e-6 * abs(num), (i, num, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 0.0], np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[3] == np.float32(-3.140625)
    # 8 significant bits: relative error <= 2^-9
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-45)
    assert r[2] == 1.0 + 2 ** -7   # ties go to even: 1 + 3*2^-9 -> 1 + 4*2^-9
