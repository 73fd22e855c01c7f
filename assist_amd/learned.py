"""Learned IAS15 start: step traces, training targets, GPU training and prediction.

IAS15 starts every system from a small trial step (0.001 d by default) and lets its step
control grow it at most fourfold per step; the first step also starts its predictor-corrector
from b = 0.  ``StepModel`` -- a 16 -> 128 -> 128 -> 128 -> 32 perceptron with BF16 weights on
the matrix cores (libassist Part 3) -- predicts from a system's state

* the step IAS15 settles on (output 0: log10 of IAS15's own next-step proposal, in days), and
* the first step's b coefficients (outputs 1..21), as b_k,c (k+1)! / (|a0| (c s)^(k+1)) with
  s the sign of the step and c = (7! eps)^(1/7) = 0.175 for eps = 1e-9: IAS15 sizes its steps
  so that |b6| / |a| is about eps, and b_k ~ a^(k+1) dt^(k+1) / (k+1)! for a smooth
  acceleration, so this scaling makes every b_k O(1) at the settled step.

IAS15's own error test still accepts or rejects every step: a step that is too long is
rejected and retried shorter exactly as in the plain integrator.

Inputs (features, computed on the GPU): heliocentric position / 3 au, heliocentric velocity
x 40 d/au, log10 of the distances to the three nearest of the 11 planetary bodies, log10 of
the last step.  At the start of an integration the last step is unknown; prediction iterates
the network from a last step of 1 day, feeding its own step back (steady-state traces, which
the network is trained on, have the last step close to the next one).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from ._libassist import clibassist
from .batch import BatchConfig, BatchExtra, BatchStats, DeviceArray, _check, _opts_ptr

N_IN, N_HIDDEN, N_OUT = 16, 128, 32
TRACE_WORDS = 34
N_USED_OUT = 22
B_CLIP = 20.0
WARM_SCALE = 0.175  # (7! eps)^(1/7) for eps = 1e-9; the same constant is in kernels.hip
# loss weights of the outputs: the step dominates, the 21 warm-start values follow
OUTPUT_WEIGHTS = np.array([1.0] + [0.05] * 21 + [0.0] * 10, dtype=np.float32)

clibassist.assist_mlp_create.restype = C.c_void_p
clibassist.assist_mlp_create.argtypes = [C.c_int, C.c_ulonglong]
clibassist.assist_mlp_free.argtypes = [C.c_void_p]
clibassist.assist_mlp_max_batch.argtypes = [C.c_void_p]
clibassist.assist_mlp_set_params.argtypes = [C.c_void_p, C.c_void_p]
clibassist.assist_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p]
clibassist.assist_mlp_set_output_weights.argtypes = [C.c_void_p, C.c_void_p]
clibassist.assist_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
clibassist.assist_mlp_predict.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
clibassist.assist_mlp_loss_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float),
                                            C.c_void_p, C.c_void_p]
clibassist.assist_mlp_adam_step.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                            C.c_void_p]
clibassist.assist_step_features.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
clibassist.assist_batch_predict_start.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                  C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]


def _p(d):
    return None if d is None else C.c_void_p(d.ptr)


def num_params() -> int:
    return clibassist.assist_mlp_num_params()


def param_shapes():
    """(W shape, b shape) per layer, in the order of the flat parameter vector."""
    dims = [N_IN, N_HIDDEN, N_HIDDEN, N_HIDDEN, N_OUT]
    return [((dims[i], dims[i + 1]), (dims[i + 1],)) for i in range(4)]


class StepModel:
    """The network on the GPU (libassist's assist_mlp)."""

    def __init__(self, max_batch: int = 65536, seed: int = 0):
        self.ptr = clibassist.assist_mlp_create(int(max_batch), int(seed))
        if not self.ptr:
            raise RuntimeError("assist_mlp_create failed (no GPU?)")
        self.max_batch = int(max_batch)
        self.set_output_weights(OUTPUT_WEIGHTS)

    def free(self):
        if getattr(self, "ptr", None):
            clibassist.assist_mlp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()

    def params(self) -> np.ndarray:
        out = np.empty(num_params(), dtype=np.float32)
        _check(clibassist.assist_mlp_get_params(self.ptr, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (num_params(),)
        _check(clibassist.assist_mlp_set_params(self.ptr, p.ctypes.data))

    def set_output_weights(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.shape == (N_OUT,)
        _check(clibassist.assist_mlp_set_output_weights(self.ptr, w.ctypes.data))

    def forward(self, d_X: DeviceArray, m: int, d_Y: DeviceArray, row0: int = 0):
        _check(clibassist.assist_mlp_forward(self.ptr, C.c_void_p(d_X.ptr + row0 * N_IN * 4), int(m),
                                             C.c_void_p(d_Y.ptr + row0 * N_OUT * 4), None))

    def predict(self, d_X: DeviceArray, m: int, d_Y: DeviceArray):
        """The forward pass of the prediction path (any m, no activations kept)."""
        _check(clibassist.assist_mlp_predict(self.ptr, C.c_void_p(d_X.ptr), int(m), C.c_void_p(d_Y.ptr), None))

    def loss_grad(self, d_X: DeviceArray, d_T: DeviceArray, m: int, row0: int = 0, want_grad: bool = False):
        loss = C.c_float(0.0)
        g = np.empty(num_params(), dtype=np.float32) if want_grad else None
        _check(clibassist.assist_mlp_loss_grad(self.ptr, C.c_void_p(d_X.ptr + row0 * N_IN * 4),
                                               C.c_void_p(d_T.ptr + row0 * N_OUT * 4), int(m), C.byref(loss),
                                               None if g is None else g.ctypes.data, None))
        return (loss.value, g) if want_grad else loss.value

    def adam(self, lr: float, step: int, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(clibassist.assist_mlp_adam_step(self.ptr, lr, beta1, beta2, eps, int(step), None))

    def save(self, path: str, **meta):
        np.savez(path, params=self.params(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str, max_batch: int = 65536) -> "StepModel":
        with np.load(path, allow_pickle=False) as z:
            m = cls(max_batch=max_batch)
            m.set_params(z["params"])
        return m


def default_model_path() -> str:
    """ASSIST_STEP_MODEL if set, else the model shipped with the package."""
    return os.environ.get("ASSIST_STEP_MODEL") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                                                               "step_mlp.npz")


# ---- traces --------------------------------------------------------------------------------------

def integrate_traced(ephem, t0, t1, states, params=None, config: BatchConfig | None = None, capacity=1_000_000,
                     every=1, skip=8, extras=None):
    """Plain IAS15 over a batch with step traces; returns (records [count, 34], final states,
    status, steps, evals)."""
    n, nv, _ = states.shape
    d_t0, d_t1 = DeviceArray(np.asarray(t0, float)), DeviceArray(np.asarray(t1, float))
    d_st = DeviceArray(np.asarray(states, float))
    d_pp = DeviceArray(np.asarray(params, float)) if params is not None else None
    d_status = DeviceArray(np.zeros(n, np.int32))
    d_steps = DeviceArray(np.zeros(n, np.int64))
    d_evals = DeviceArray(np.zeros(n, np.int64))
    d_trace = DeviceArray(np.zeros((max(capacity, 1), TRACE_WORDS)))
    d_count = DeviceArray(np.zeros(1, np.int32))
    ex = BatchExtra.new(trace=d_trace.ptr, trace_count=d_count.ptr, trace_capacity=int(capacity),
                        trace_every=int(every), trace_skip=int(skip))
    stats = BatchStats()
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate_device_ex(
        C.byref(ephem), _opts_ptr(extras), C.byref(cfg), n, nv - 1, _p(d_t0), _p(d_t1), _p(d_st), _p(d_pp),
        _p(d_status), _p(d_steps), _p(d_evals), None, C.byref(stats), C.byref(ex)))
    count = min(int(d_count.download()[0]), capacity)
    rec = d_trace.download()[:count]
    return rec, d_st.download(), d_status.download(), d_steps.download(), d_evals.download()


def features(ephem, rows: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """GPU features of trace records (or any [t, x, v, dt_last, ...] rows): (X [n, 16], omega [n])."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    n, stride = rows.shape
    d_rows = DeviceArray(rows)
    d_X = DeviceArray(np.zeros((n, N_IN), np.float32))
    d_w = DeviceArray(np.zeros(n))
    _check(clibassist.assist_step_features(C.byref(ephem), n, _p(d_rows), stride, _p(d_X), _p(d_w)))
    return d_X.download(), d_w.download()


def targets(records: np.ndarray, omega: np.ndarray | None = None) -> np.ndarray:
    """Training targets [n, 32] of trace records: log10 of IAS15's next-step proposal and the
    scaled b of the accepted step."""
    n = records.shape[0]
    T = np.zeros((n, N_OUT), np.float32)
    T[:, 0] = np.clip(np.log10(np.abs(records[:, 9])), -3.0, 3.0)
    sign = np.sign(records[:, 8])
    amag = np.linalg.norm(records[:, 10:13], axis=1)
    b = records[:, 13:34].reshape(n, 7, 3)
    for k in range(7):
        scale = amag * (WARM_SCALE * sign) ** (k + 1) / math.factorial(k + 1)
        with np.errstate(divide="ignore", invalid="ignore"):
            v = b[:, k, :] / scale[:, None]
        T[:, 1 + 3 * k:4 + 3 * k] = np.clip(np.nan_to_num(v, nan=0.0, posinf=B_CLIP, neginf=-B_CLIP), -B_CLIP, B_CLIP)
    return T


def train(model: StepModel, X: np.ndarray, T: np.ndarray, epochs=4, batch=8192, lr=1e-3, lr_end=1e-4, seed=0,
          log=None):
    """Adam on the GPU over (X, T), minibatches in a random order each epoch; returns the mean
    loss of each epoch."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(X.shape[0])
    X, T = X[perm], T[perm]
    d_X, d_T = DeviceArray(np.ascontiguousarray(X, np.float32)), DeviceArray(np.ascontiguousarray(T, np.float32))
    n = X.shape[0]
    nb = max(1, n // batch)
    total = epochs * nb
    hist = []
    step = 0
    for ep in range(epochs):
        losses = []
        for bi in rng.permutation(nb):
            step += 1
            frac = (step - 1) / max(total - 1, 1)
            cur_lr = lr * (lr_end / lr) ** frac
            m = min(batch, n - bi * batch)
            losses.append(model.loss_grad(d_X, d_T, m, row0=int(bi) * batch))
            model.adam(cur_lr, step)
        hist.append(float(np.mean(losses)))
        if log:
            log("epoch %d: loss %.5f" % (ep, hist[-1]))
    return hist


def predict_start(ephem, model: StepModel, n, n_var, d_t0, d_t1, d_states, min_dt=0.0, iterations=3, warm=True,
                  out=None):
    """Device arrays (dt0 [n], bw [n, 21] or None) for assist_batch_integrate_device_ex.  `out`
    (dt0, bw) reuses buffers of a previous call (every element is written), so that repeated
    launches keep the same device pointers, as a captured graph replay needs."""
    if out is not None:
        d_dt0, d_bw = out
        if d_dt0.nbytes < 8 * n or (warm and (d_bw is None or d_bw.nbytes < 8 * 21 * n)):
            raise ValueError("predict_start: the output buffers are smaller than the batch")
        if not warm:
            d_bw = None
    else:
        d_dt0 = DeviceArray.empty(n)
        d_bw = DeviceArray.empty((n, 21)) if warm else None
    _check(clibassist.assist_batch_predict_start(C.byref(ephem), C.c_void_p(model.ptr), int(n), int(n_var), _p(d_t0),
                                                 _p(d_t1), _p(d_states), float(min_dt), int(iterations), _p(d_dt0),
                                                 _p(d_bw)))
    return d_dt0, d_bw


def integrate_device_learned(ephem, n, n_var, d_t0, d_t1, d_states, d_params, d_status, d_steps, d_evals, d_dt0,
                             d_bw=None, extras=None, config: BatchConfig | None = None) -> BatchStats:
    """IAS15 on device buffers from the learned start (d_dt0, d_bw from predict_start)."""
    ex = BatchExtra.new(dt0=d_dt0.ptr, bw=d_bw.ptr if d_bw is not None else None)
    stats = BatchStats()
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate_device_ex(
        C.byref(ephem), _opts_ptr(extras), C.byref(cfg), int(n), int(n_var), _p(d_t0), _p(d_t1), _p(d_states),
        _p(d_params), _p(d_status), _p(d_steps), _p(d_evals), None, C.byref(stats), C.byref(ex)))
    return stats
