"""Pins the CPU oracle of the learned-start network (oracle/mlp.py): its float64 gradients
against central differences of its own loss, its BF16 rounding against the IEEE bit patterns,
and its parameter layout against the library's parameter count."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp  # noqa: E402


def _problem(seed=0, m=24):
    rng = np.random.default_rng(seed)
    n_params = sum(a * b + b for a, b in zip(mlp.DIMS[:-1], mlp.DIMS[1:]))
    p = rng.normal(0, 0.15, n_params)
    X = rng.normal(0, 1, (m, 16))
    T = rng.normal(0, 1, (m, 32))
    w = rng.uniform(0.1, 1.0, 32)
    return p, X, T, w


def test_gradient_matches_central_differences():
    p, X, T, w = _problem()
    _, g = mlp.loss_grad(p, X, T, w, bf16_ops=False)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 60, replace=False)
    # include biases of every layer and the first weights of each layer
    offs, o = [], 0
    for a, b in zip(mlp.DIMS[:-1], mlp.DIMS[1:]):
        offs += [o, o + a * b]
        o += a * b + b
    for i in list(idx) + offs:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, X, T, w, bf16_ops=False)[0] - mlp.loss_grad(pm, X, T, w, bf16_ops=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -9, -3.14159, 65504.0, 1e-30], np.float32)
    got = mlp.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins; 1 + 2^-7 + 2^-8 ties to 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2 ** -6) and got[3] == 1.0
    assert abs(got[4] - np.float32(-3.140625)) == 0.0
    assert (got.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_forward_is_close_to_float64():
    p, X, T, w = _problem(2)
    y64, _ = mlp.forward(p, X, bf16_ops=False)
    y16, _ = mlp.forward(p, X, bf16_ops=True)
    assert np.abs(y16 - y64).max() <= 0.05 * np.abs(y64).max()


def test_parameter_count_matches_library():
    import assist_amd.learned as L
    assert L.num_params() == sum(a * b + b for a, b in zip(mlp.DIMS[:-1], mlp.DIMS[1:]))
