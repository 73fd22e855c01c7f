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

# per-system status codes (enum ASSIST_SYSTEM_STATUS)
SYSTEM_OK = 0
SYSTEM_COVERAGE = 5
SYSTEM_NONFINITE = -1
SYSTEM_MAX_STEPS = -2
SYSTEM_DT_UNDERFLOW = -3
SYSTEM_PAUSED = -4

# snapshot layout (ASSIST_SNAPSHOT_HEADER, ASSIST_SNAPSHOT_VECTOR)
SNAPSHOT_HEADER = 8
SNAPSHOT_VECTOR = 57


class BatchConfig(C.Structure):
    _fields_ = [("size", C.c_size_t), ("epsilon", C.c_double), ("min_dt", C.c_double), ("dt_initial", C.c_double),
                ("max_steps", C.c_longlong), ("precision", C.c_int), ("lanes_per_system", C.c_int),
                ("block_size", C.c_int), ("blocks_per_cu", C.c_int), ("waves_per_simd", C.c_int),
                ("use_graph", C.c_int), ("substep_cache_off", C.c_int), ("reserved_", C.c_int),
                ("corrector_tolerance", C.c_double), ("first_pass_tolerance", C.c_double),
                ("first_pass_accel", C.c_double)]

    @classmethod
    def default(cls, **kw):
        c = cls()
        clibassist.assist_batch_config_default(C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class BatchExtra(C.Structure):
    """ctypes mirror of ``struct assist_batch_extra`` (every field, so that the library never
    reads past the end of a Python-built struct); build it with :meth:`new`."""
    _fields_ = [("size", C.c_size_t), ("dt0", C.c_void_p), ("bw", C.c_void_p), ("trace", C.c_void_p),
                ("trace_count", C.c_void_p), ("trace_capacity", C.c_int), ("trace_every", C.c_int),
                ("trace_skip", C.c_int), ("reserved_", C.c_int), ("out_offset", C.c_void_p), ("out_t", C.c_void_p),
                ("out_states", C.c_void_p), ("snapshot_in", C.c_void_p), ("snapshot_out", C.c_void_p),
                ("t_pause", C.c_void_p), ("archive", C.c_void_p), ("archive_count", C.c_void_p),
                ("archive_capacity", C.c_int), ("reserved2_", C.c_int)]

    @classmethod
    def new(cls, **kw):
        """assist_batch_extra_init, then the given fields (DeviceArray values give their pointer)."""
        e = cls()
        clibassist.assist_batch_extra_init(C.byref(e))
        for k, v in kw.items():
            setattr(e, k, v.ptr if isinstance(v, DeviceArray) else v)
        return e


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


def _ptr(d):
    return None if d is None else C.c_void_p(d.ptr)


def integrate_device(ephem, n, n_var, d_t0, d_t1, d_states, d_params, d_status, d_steps, d_evals, extras=None,
                     config: BatchConfig | None = None, stream=None, extra: BatchExtra | None = None) -> BatchStats:
    """IAS15 on device-resident buffers; returns the stats (kernel_ms from HIP events).  ``extra``
    (a :class:`BatchExtra`) adds the optional inputs and outputs of assist_batch_integrate_device_ex."""
    stats = BatchStats()
    cfg = config or BatchConfig.default()
    _check(clibassist.assist_batch_integrate_device_ex(
        C.byref(ephem), _opts_ptr(extras), C.byref(cfg), n, n_var, _ptr(d_t0), _ptr(d_t1), _ptr(d_states),
        _ptr(d_params), _ptr(d_status), _ptr(d_steps), _ptr(d_evals), C.c_void_p(stream) if stream else None,
        C.byref(stats), C.byref(extra) if extra is not None else None))
    return stats


def snapshot_words(nvec: int) -> int:
    """Doubles per system of a snapshot (ASSIST_SNAPSHOT_WORDS)."""
    return SNAPSHOT_HEADER + SNAPSHOT_VECTOR * nvec


@dataclass
class CheckpointResult:
    states: np.ndarray     # each system's state where it stopped
    status: np.ndarray     # 0 finished, ASSIST_SYSTEM_PAUSED, or an error
    steps: np.ndarray      # accepted steps since the system's start (a resumed system counts on)
    evals: np.ndarray
    snapshots: np.ndarray  # (n, snapshot_words(1 + n_var)): each system's IAS15 state where it stopped
    stats: BatchStats


def integrate_checkpoint(ephem, t0, t1, states, params=None, extras=None, config: BatchConfig | None = None,
                         t_pause=None, resume=None) -> CheckpointResult:
    """IAS15 from t0[i] (or from the snapshot ``resume[i]``) towards t1[i], stopping each system
    at t1[i] or, with ``t_pause``, at its first step boundary at or past t_pause[i] (status
    SYSTEM_PAUSED).  The returned snapshots resume the integration with the same steps as an
    uninterrupted one."""
    states = np.array(states, dtype=np.float64, copy=True, order="C")
    n, nv, six = states.shape
    assert six == 6
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
    t1 = np.ascontiguousarray(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
    w = snapshot_words(nv)
    bufs = dict(t0=DeviceArray(t0), t1=DeviceArray(t1), st=DeviceArray(states),
                status=DeviceArray(np.zeros(n, dtype=np.int32)), steps=DeviceArray(np.zeros(n, dtype=np.int64)),
                evals=DeviceArray(np.zeros(n, dtype=np.int64)), snap=DeviceArray.empty((n, w)))
    kw = dict(snapshot_out=bufs["snap"])
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.shape == (n, nv, 3)
        bufs["params"] = DeviceArray(params)
    if t_pause is not None:
        bufs["pause"] = DeviceArray(np.ascontiguousarray(np.broadcast_to(np.asarray(t_pause, dtype=np.float64), (n,))))
        kw["t_pause"] = bufs["pause"]
    if resume is not None:
        resume = np.ascontiguousarray(resume, dtype=np.float64)
        assert resume.shape == (n, w)
        bufs["resume"] = DeviceArray(resume)
        kw["snapshot_in"] = bufs["resume"]
    stats = integrate_device(ephem, n, nv - 1, bufs["t0"], bufs["t1"], bufs["st"], bufs.get("params"), bufs["status"],
                             bufs["steps"], bufs["evals"], extras=extras, config=config, extra=BatchExtra.new(**kw))
    return CheckpointResult(bufs["st"].download(), bufs["status"].download(), bufs["steps"].download(),
                            bufs["evals"].download(), bufs["snap"].download(), stats)


class StepArchive:
    """Snapshots of every system at its start and after each accepted IAS15 step, kept on the
    device (the counterpart of a REBOUND SimulationArchive saved every step, which reference
    src/assist.c:599-633 interpolates); :meth:`interpolate` gives states at arbitrary times."""

    def __init__(self, n, nvec, capacity, d_archive, d_count, result):
        self.n, self.nvec, self.capacity = n, nvec, capacity
        self.d_archive, self.d_count = d_archive, d_count
        self.result = result
        self.counts = d_count.download()
        if (self.counts > capacity).any():
            raise BatchError("step archive capacity %d exceeded (a system took %d snapshots)"
                             % (capacity, int(self.counts.max())))

    def snapshots(self, i: int) -> np.ndarray:
        """The snapshots of system i, (count_i, snapshot_words(nvec))."""
        a = self.d_archive.download().reshape(self.n, self.capacity, -1)
        return a[i, :self.counts[i]]

    def interpolate(self, systems, t):
        """States (nq, nvec, 6) and per-query status of systems[q] at t[q] (broadcast)."""
        systems, t = np.broadcast_arrays(np.asarray(systems, dtype=np.int32), np.asarray(t, dtype=np.float64))
        systems = np.ascontiguousarray(systems.ravel())
        t = np.ascontiguousarray(t.ravel())
        nq = systems.size
        ds, dt = DeviceArray(systems), DeviceArray(t)
        out = DeviceArray.empty((max(nq, 1), self.nvec, 6))
        st = DeviceArray.empty(max(nq, 1), dtype=np.int32)
        _check(clibassist.assist_batch_archive_interpolate_device(
            self.n, self.nvec - 1, C.c_void_p(self.d_archive.ptr), self.capacity, C.c_void_p(self.d_count.ptr), nq,
            C.c_void_p(ds.ptr), C.c_void_p(dt.ptr), C.c_void_p(out.ptr), C.c_void_p(st.ptr), None))
        return out.download()[:nq], st.download()[:nq]


def integrate_archive(ephem, t0, t1, states, params=None, extras=None, config: BatchConfig | None = None,
                      capacity: int = 4096) -> StepArchive:
    """IAS15 from t0[i] to t1[i] saving every system's snapshot at every step boundary."""
    states = np.array(states, dtype=np.float64, copy=True, order="C")
    n, nv, six = states.shape
    assert six == 6
    t0 = np.ascontiguousarray(np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)))
    t1 = np.ascontiguousarray(np.broadcast_to(np.asarray(t1, dtype=np.float64), (n,)))
    d_arch = DeviceArray.empty((n, capacity, snapshot_words(nv)))
    d_count = DeviceArray(np.zeros(n, dtype=np.int32))
    d_st = DeviceArray(states)
    d_status = DeviceArray(np.zeros(n, dtype=np.int32))
    d_steps = DeviceArray(np.zeros(n, dtype=np.int64))
    d_pp = None
    if params is not None:
        params = np.ascontiguousarray(params, dtype=np.float64)
        assert params.shape == (n, nv, 3)
        d_pp = DeviceArray(params)
    stats = integrate_device(ephem, n, nv - 1, DeviceArray(t0), DeviceArray(t1), d_st, d_pp, d_status, d_steps, None,
                             extras=extras, config=config,
                             extra=BatchExtra.new(archive=d_arch, archive_count=d_count, archive_capacity=capacity))
    res = IntegrateResult(d_st.download(), d_status.download(), d_steps.download(), stats)
    return StepArchive(n, nv, capacity, d_arch, d_count, res)
