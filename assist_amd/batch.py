"""numpy interface to the batched GPU entry points of libassist.

A *system* is one real test particle followed by its first-order variational
particles; ``states`` has shape (n, 1 + n_var, 6) with x, y, z, vx, vy, vz in
au and au/day (barycentric, equatorial J2000, like the reference), ``params``
shape (n, 1 + n_var, 3) with the Marsden A1, A2, A3 of the real particle and
their variations for the variational ones (reference particle_params layout,
src/forces.c:839-841, 1030-1032).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._libassist import clibassist

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)

PRECISION_FP64 = 0
PRECISION_MIXED = 1


class BatchConfig(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("min_dt", C.c_double), ("dt_initial", C.c_double),
                ("max_steps", C.c_longlong), ("precision", C.c_int), ("lanes_per_system", C.c_int),
                ("block_size", C.c_int), ("blocks_per_cu", C.c_int), ("waves_per_simd", C.c_int),
                ("use_graph", C.c_int), ("substep_cache_off", C.c_int)]

    @classmethod
    def default(cls, **kw):
        c = cls()
        clibassist.assist_batch_config_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class BatchStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("rejected", C.c_longlong), ("evals", C.c_longlong),
                ("kernel_ms", C.c_double), ("lanes_per_system", C.c_int), ("block_size", C.c_int),
                ("grid_blocks", C.c_int), ("waves_per_simd", C.c_int)]


clibassist.assist_batch_error_string.restype = C.c_char_p
clibassist.assist_device_malloc.restype = C.c_void_p
clibassist.assist_device_malloc.argtypes = [C.c_size_t]
clibassist.assist_device_free.argtypes = [C.c_void_p]
clibassist.assist_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
clibassist.assist_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
clibassist.assist_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]


class BatchError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise BatchError(clibassist.assist_batch_error_string(rc).decode())


def gpu_count() -> int:
    return int(clibassist.assist_gpu_device_count())


def _d(a):
    return a.ctypes.data_as(_dp)


def _opts_ptr(extras):
    """Force options: an Extras instance, a dict of fields, or None (defaults)."""
    if extras is None:
        return None
    from .extras import Extras
    if isinstance(extras, Extras):
        return C.byref(extras)
    ex = Extras.__new__(Extras)
    C.Structure.__init__(ex)
    ex._forces = 0x7F
    ex.gr_eih_sources = 1
    ex.alpha, ex.nk, ex.nm, ex.nn, ex.r0 = 1.0, 0.0, 2.0, 5.093, 1.0
    for k, v in extras.items():
        setattr(ex, "_forces" if k == "forces" else k, v)
    _opts_ptr._keep = ex
    return C.byref(ex)


def forces(ephem, t, states, params=None, extras=None, precision=PRECISION_FP64):
    """Accelerations (n, 1+n_var, 3) and per-system status of the force model at times t."""
    states = np.ascontiguousarray(states, dtype=np.float64)
    n, nv, six = states.shape
    assert six == 6
    t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (n,)))
    acc = np.zeros((n, nv, 3))
    st = np.zeros(n, dtype=np.int32)
    pp = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.shape == (n, nv, 3)
        pp = _d(params)
    _check(clibassist.assist_batch_forces(C.byref(ephem), _opts_ptr(extras), n, nv - 1, _d(t), _d(states), pp,
                                          _d(acc), st.ctypes.data_as(_ip), precision))
    return acc, st


@dataclass
class IntegrateResult:
    states: np.ndarray
    status: np.ndarray
    steps: np.ndarray
    stats: BatchStats


def get_particles(ephem, body, t):
    """Batched ephemeris queries on the GPU: (out (n, 10) = GM, x, y, z, vx, vy, vz, ax, ay, az,
    status (n,)) for ASSIST bodies ``body[q]`` at times ``t[q]`` (broadcast against each other)."""
    body, t = np.broadcast_arrays(np.asarray(body, dtype=np.int32), np.asarray(t, dtype=np.float64))
    body = np.ascontiguousarray(body.ravel())
    t = np.ascontiguousarray(t.ravel())
    n = body.size
    out = np.zeros((n, 10))
    st = np.zeros(n, dtype=np.int32)
    _check(clibassist.assist_batch_get_particles(C.byref(ephem), n, body.ctypes.data_as(_ip), _d(t), _d(out),
                                                 st.ctypes.data_as(_ip)))
    return out, st


def integrate(ephem, t0, t1, states, params=None, extras=None, config: BatchConfig | None = None) -> IntegrateResult:
    """IAS15 for every system from t0[i] to t1[i] on the GPU (host arrays; copies included)."""
    states = np.array(states, dtype=np.float64, copy=True, order="C")
    n, nv, six = states.shape
    assert six == 6
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
    t1 = np.ascontiguousarray(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
    st = np.zeros(n, dtype=np.int32)
    steps = np.zeros(n, dtype=np.int64)
    stats = BatchStats()
    pp = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.shape == (n, nv, 3)
        pp = _d(params)
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate(C.byref(ephem), _opts_ptr(extras), C.byref(cfg), n, nv - 1, _d(t0), _d(t1),
                                             _d(states), pp, st.ctypes.data_as(_ip), steps.ctypes.data_as(_llp),
                                             C.byref(stats)))
    return IntegrateResult(states, st, steps, stats)


@dataclass
class DenseResult:
    outputs: list          # per system: (m_i, 1 + n_var, 6) states at its output times
    states: np.ndarray     # end of each system's last step
    status: np.ndarray
    steps: np.ndarray
    stats: BatchStats


def integrate_dense(ephem, t0, out_times, states, params=None, extras=None,
                    config: BatchConfig | None = None) -> DenseResult:
    """Dense output (the batched counterpart of the reference's integrate_or_interpolate,
    reference src/assist.c:642-680): system i is integrated from t0[i] through its output times
    out_times[i] (sorted in one direction, not before t0[i]) and its state at each of them is
    interpolated inside the IAS15 step that covers it."""
    states = np.array(states, dtype=np.float64, copy=True, order="C")
    n, nv, six = states.shape
    assert six == 6 and len(out_times) == n
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
    counts = np.array([len(np.atleast_1d(o)) for o in out_times], dtype=np.int64)
    off = np.zeros(n + 1, dtype=np.int32)
    off[1:] = np.cumsum(counts)
    flat = np.ascontiguousarray(np.concatenate([np.atleast_1d(np.asarray(o, dtype=np.float64)) for o in out_times])
                                if n else np.zeros(0))
    out = np.zeros((max(int(off[-1]), 1), nv, 6))
    st = np.zeros(n, dtype=np.int32)
    steps = np.zeros(n, dtype=np.int64)
    stats = BatchStats()
    pp = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.shape == (n, nv, 3)
        pp = _d(params)
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate_dense(
        C.byref(ephem), _opts_ptr(extras), C.byref(cfg), n, nv - 1, _d(t0), _d(states), pp,
        off.ctypes.data_as(_ip), _d(flat), _d(out), st.ctypes.data_as(_ip), steps.ctypes.data_as(_llp), C.byref(stats)))
    outputs = [out[off[i]:off[i + 1]] for i in range(n)]
    return DenseResult(outputs, states, st, steps, stats)


class DeviceArray:
    """A device buffer owned by libassist (for inputs kept resident in HBM)."""

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = clibassist.assist_device_malloc(max(self.nbytes, 8))
        if not self.ptr:
            raise BatchError("device allocation failed")
        self.upload(host)

    @classmethod
    def empty(cls, shape, dtype=np.float64) -> "DeviceArray":
        """An uninitialised device buffer (for outputs the library writes in full)."""
        self = cls.__new__(cls)
        self.shape, self.dtype = tuple(np.atleast_1d(shape)), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = clibassist.assist_device_malloc(max(self.nbytes, 8))
        if not self.ptr:
            raise BatchError("device allocation failed")
        return self

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        _check(clibassist.assist_memcpy_h2d(self.ptr, host.ctypes.data, self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(clibassist.assist_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        _check(clibassist.assist_memcpy_d2d(self.ptr, other.ptr, self.nbytes))

    def free(self):
        if self.ptr:
            clibassist.assist_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def synchronize():
    _check(clibassist.assist_device_synchronize())


def integrate_device(ephem, n, n_var, d_t0, d_t1, d_states, d_params, d_status, d_steps, d_evals, extras=None,
                     config: BatchConfig | None = None, stream=None) -> BatchStats:
    """IAS15 on device-resident buffers; returns the stats (kernel_ms from HIP events)."""
    stats = BatchStats()
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate_device(
        C.byref(ephem), _opts_ptr(extras), C.byref(cfg), n, n_var, C.c_void_p(d_t0.ptr), C.c_void_p(d_t1.ptr),
        C.c_void_p(d_states.ptr), C.c_void_p(d_params.ptr) if d_params is not None else None,
        C.c_void_p(d_status.ptr), C.c_void_p(d_steps.ptr) if d_steps is not None else None,
        C.c_void_p(d_evals.ptr) if d_evals is not None else None, C.c_void_p(stream) if stream else None,
        C.byref(stats)))
    return stats
