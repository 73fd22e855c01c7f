"""CPU oracle of the learned-start network (test infrastructure: only tests/, smoke() and the
benchmark's checks use it; the product runs assist_amd/csrc/mlp.hip).

The network is 16 -> 128 -> 128 -> 128 -> 32 with tanh on the hidden layers, weights W[in][out]
and biases, loss = sum_j w_j (Y - T)^2 / M.  ``forward``/``loss_grad`` restate it in numpy with
the GPU's numerics when ``bf16=True``: every operand of a product (inputs, activations, the
weights, and the gradients fed back through the products) is rounded to BF16 (round to nearest
even) and accumulated in FP32 or better; biases stay FP32.  With ``bf16=False`` the same
formulas run in float64, which the tests pin against central differences.
"""
import numpy as np

DIMS = [16, 128, 128, 128, 32]


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even); returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32).reshape(x.shape)
    nan = np.isnan(x)
    if nan.any():
        out = np.where(nan, x, out)
    return out


def unpack(params):
    """Flat parameter vector -> list of (W, b)."""
    out, o = [], 0
    for i in range(4):
        k, n = DIMS[i], DIMS[i + 1]
        W = params[o:o + k * n].reshape(k, n)
        o += k * n
        b = params[o:o + n]
        o += n
        out.append((W, b))
    return out


def _mm(a, b, low):
    if low:
        return (bf16(a).astype(np.float64) @ bf16(b).astype(np.float64)).astype(np.float32)
    return a @ b


def forward(params, X, bf16_ops=True):
    """Returns (Y, hidden activations H1..H3)."""
    layers = unpack(np.asarray(params, np.float32 if bf16_ops else np.float64))
    h = np.asarray(X, np.float32 if bf16_ops else np.float64)
    hs = []
    for i, (W, b) in enumerate(layers):
        z = _mm(h, W, bf16_ops) + b
        if i < 3:
            h = np.tanh(z)
            hs.append(h)
        else:
            h = z
    return h, hs


def loss_grad(params, X, T, w, bf16_ops=True):
    """(loss, flat gradient) with the GPU's ordering of operations."""
    dt = np.float32 if bf16_ops else np.float64
    X = np.asarray(X, dt)
    T = np.asarray(T, dt)
    w = np.asarray(w, dt)
    M = X.shape[0]
    Y, hs = forward(params, X, bf16_ops)
    d = Y - T
    loss = float(np.sum(w * d * d) / M)
    dcur = (2.0 * w * d / M).astype(dt)
    layers = unpack(np.asarray(params, dt))
    grads = [None] * 4
    for l in range(3, -1, -1):
        hin = X if l == 0 else hs[l - 1]
        W, _ = layers[l]
        gW = _mm(hin.T, dcur, bf16_ops)
        gb = dcur.sum(axis=0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dh = _mm(dcur, W.T, bf16_ops)
        dcur = (dh * (1.0 - hs[l - 1] * hs[l - 1])).astype(dt)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return loss, flat
