"""The allocation-free integration entry point (assist_batch_integrate_device_noalloc) against
assist_batch_integrate_device, through the C-ABI: two calls in a row, the second on a smaller
batch, each bit for bit equal to the allocating entry point's results (the static scratch keeps
the last call's contents; the kernel takes its records only after writing them)."""
import ctypes as C

import numpy as np
import pytest

from conftest import population

pytestmark = pytest.mark.gpu


def _device_run(ephem, entry, t0, t1, states, params, cfg):
    import assist_amd
    from assist_amd import batch
    n, nv = states.shape[:2]
    d = dict(t0=batch.DeviceArray(t0), t1=batch.DeviceArray(t1), st=batch.DeviceArray(states),
             pp=batch.DeviceArray(params) if params is not None else None,
             status=batch.DeviceArray(np.full(n, 7, np.int32)), steps=batch.DeviceArray(np.zeros(n, np.int64)),
             evals=batch.DeviceArray(np.zeros(n, np.int64)))
    p = lambda x: None if x is None else C.c_void_p(x.ptr)  # noqa: E731
    args = [C.byref(ephem), None, C.byref(cfg), n, nv - 1, p(d["t0"]), p(d["t1"]), p(d["st"]), p(d["pp"]),
            p(d["status"]), p(d["steps"]), p(d["evals"]), None]
    lib = assist_amd.clibassist
    if entry == "noalloc":
        rc = lib.assist_batch_integrate_device_noalloc(*args)
    else:
        stats = batch.BatchStats()
        rc = lib.assist_batch_integrate_device(*args, C.byref(stats))
    assert rc == 0, batch.clibassist.assist_batch_error_string(rc)
    return {k: d[k].download() for k in ("st", "status", "steps", "evals")}


@pytest.mark.parametrize("n_var,precision,lanes", [(6, 0, 0), (0, 1, 1)])
def test_noalloc_matches_the_allocating_entry_point(gpu, ephem, model, n_var, precision, lanes):
    from assist_amd import batch
    # the ephemeris must be resident on the GPU before the allocation-free call
    _, st = batch.get_particles(ephem, [0], [0.0])
    assert st[0] == 0
    cfg = batch.BatchConfig.default(precision=precision, lanes_per_system=lanes, min_dt=0.001)
    for n, seed, span in ((3000, 701, 150.0), (700, 702, 60.0)):
        states, params, _ = population(model, n, seed=seed, n_var=n_var, ng_fraction=0.2)
        rng = np.random.default_rng(seed)
        t0 = rng.uniform(-50.0, 50.0, n)
        t1 = t0 + rng.uniform(-span, span, n)
        a = _device_run(ephem, "noalloc", t0, t1, states, params, cfg)
        b = _device_run(ephem, "alloc", t0, t1, states, params, cfg)
        assert (a["status"] == 0).all()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_noalloc_rejects_what_would_allocate(gpu, ephem, model):
    import assist_amd
    from assist_amd import batch
    lib = assist_amd.clibassist
    n = 4
    states, _, _ = population(model, n, seed=703, n_var=0)
    d_t0, d_t1, d_st = batch.DeviceArray(np.zeros(n)), batch.DeviceArray(np.ones(n)), batch.DeviceArray(states)
    d_status = batch.DeviceArray(np.zeros(n, np.int32))
    cfg = batch.BatchConfig.default(use_graph=1)
    rc = lib.assist_batch_integrate_device_noalloc(C.byref(ephem), None, C.byref(cfg), n, 0, C.c_void_p(d_t0.ptr),
                                                   C.c_void_p(d_t1.ptr), C.c_void_p(d_st.ptr), None,
                                                   C.c_void_p(d_status.ptr), None, None, None)
    assert rc == 102
    # an ephemeris that was never made resident on the GPU
    other = assist_amd.Ephem(*batch_paths())
    rc = lib.assist_batch_integrate_device_noalloc(C.byref(other), None, None, n, 0, C.c_void_p(d_t0.ptr),
                                                   C.c_void_p(d_t1.ptr), C.c_void_p(d_st.ptr), None,
                                                   C.c_void_p(d_status.ptr), None, None, None)
    assert rc == 102


def batch_paths():
    from assist_amd import synthetic
    return synthetic.ensure_synthetic_ephemeris(0)
