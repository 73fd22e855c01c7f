"""The integration launched as a captured HIP graph (assist_batch_config.use_graph): every replay
gives the eager launch's results bit for bit, with eager work and synchronisations between the
replays (the first call captures, later calls with the same device buffers replay)."""
import numpy as np
import pytest

from conftest import population

pytestmark = pytest.mark.gpu


def _run(ephem, n, bufs, states, cfg, learned=None):
    from assist_amd import batch
    import assist_amd.learned as L
    d_t0, d_t1, d_st, d_status, d_steps, d_evals = bufs
    d_st.upload(states)
    if learned is None:
        st = batch.integrate_device(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, config=cfg)
    else:
        d_dt0, d_bw = learned
        st = L.integrate_device_learned(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, d_dt0, d_bw,
                                        config=cfg)
    return d_st.download(), d_status.download(), d_steps.download(), d_evals.download(), st


@pytest.mark.parametrize("form", ["fp64", "mixed+learned"])
def test_graph_replays_equal_eager(gpu, ephem, model, form):
    from assist_amd import batch
    import assist_amd.learned as L
    n = 3000
    states, _, _ = population(model, n, seed=111, n_var=6)
    t0, t1 = np.zeros(n), np.full(n, 200.0)
    bufs = (batch.DeviceArray(t0), batch.DeviceArray(t1), batch.DeviceArray(states),
            batch.DeviceArray(np.zeros(n, np.int32)), batch.DeviceArray(np.zeros(n, np.int64)),
            batch.DeviceArray(np.zeros(n, np.int64)))
    prec = batch.PRECISION_MIXED if form == "mixed+learned" else batch.PRECISION_FP64
    learned = None
    if form == "mixed+learned":
        net = L.StepModel.load(L.default_model_path(), max_batch=n)
        learned = L.predict_start(ephem, net, n, 6, bufs[0], bufs[1], batch.DeviceArray(states), min_dt=0.001)
    eager = _run(ephem, n, bufs, states, batch.BatchConfig.default(min_dt=0.001, precision=prec), learned)
    other, _, _ = population(model, 64, seed=112, n_var=6)
    for k in range(4):
        got = _run(ephem, n, bufs, states, batch.BatchConfig.default(min_dt=0.001, precision=prec, use_graph=1), learned)
        for a, b in zip(eager[:4], got[:4]):
            assert np.array_equal(a, b), (form, k)
        assert got[4].kernel_ms > 0
        # eager work of the program between the replays
        batch.forces(ephem, 10.0 * k, other)
        batch.integrate(ephem, 0.0, 5.0 + k, other)
        batch.synchronize()
    print("%s: 4 graph replays equal the eager launch; last replay %.2f ms" % (form, got[4].kernel_ms))
