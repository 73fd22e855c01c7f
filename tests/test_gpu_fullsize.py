"""Parity at the benchmark configurations' full sizes.

The 1 m bar is stated for orbits without close approaches.  Every system is held to it unless it
is outside the bar *and* its own measured trajectory comes within 3 Hill radii, or 0.1 au
(planets) / 0.05 au (asteroids), of a perturber (tests/encounters.py); the exemptions are
printed, and they must stay under 5% of the configuration.

* Configuration 2 in full (4 096 main-belt systems x (1 + 6 variational), 3 650 days) against
  the oracle, system by system: <= 1 m in position, variational states to 1e-6; and the mixed form
  with the learned start against the same oracle run, <= 1 m in position.
* Configuration 4 in full (262 144 systems x (1 + 6 variational), 365 days, 70/30 main-belt/NEO,
  seed 4), through size-independent properties:
  - two FP64 launches are bit-identical (the work queue hands systems out in a different order);
  - the mixed+learned form agrees with the FP64 form within 1 m;
  - forward then backward returns the systems to their start within 1 m.
The oracle runs on the host cores; the whole module takes about a minute on one MI355X box.
"""
import os

import numpy as np
import pytest

from conftest import AU_M, rel_err
from encounters import bar_with_measured_exemptions

pytestmark = pytest.mark.gpu

NTHREADS = min(len(os.sched_getaffinity(0)), 16)


def test_configuration_2_full_size_against_oracle(gpu, ephem, oracle, model):
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(2)
    n = 4096
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r = batch.integrate(ephem, 0.0, 3650.0, states)
    o = oracle.integrate(0.0, 3650.0, states, nthreads=NTHREADS)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    dvar = rel_err(r.states[:, 1:], o["states"][:, 1:])
    print("configuration 2: variational max rel %.3e; steps differ by at most %d"
          % (dvar.max(), np.abs(r.steps - o["steps"]).max()))
    bar_with_measured_exemptions(ephem, "configuration 2 in full vs oracle", dpos, np.zeros(n), np.full(n, 3650.0),
                                  states, None)
    assert np.percentile(dpos, 99) <= 0.1
    assert dvar.max() <= 1e-6
    # the mixed form with the learned start against the same FP64 oracle: the benchmark's own
    # accuracy statement (<= 1 m after 10 years), with measured exemptions
    import assist_amd.learned as L
    net = L.StepModel.load(L.default_model_path(), max_batch=65536)
    m, mst, _, mev = _run_device(ephem, states, np.zeros(n), np.full(n, 3650.0),
                                 batch.BatchConfig.default(precision=batch.PRECISION_MIXED), learned=net)
    assert (mst == 0).all()
    mpos = np.linalg.norm(m[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    mvar = rel_err(m[:, 1:], o["states"][:, 1:])
    print("configuration 2, mixed+learned against the oracle: p99 |dr| %.3e m, variational max rel %.3e, p99 %.3e; "
          "%d evaluations against the oracle's %d" % (np.percentile(mpos, 99), mvar.max(), np.percentile(mvar, 99),
                                                      mev.sum(), o["evals"].sum()))
    bar_with_measured_exemptions(ephem, "configuration 2 in full, mixed+learned vs oracle", mpos, np.zeros(n),
                                 np.full(n, 3650.0), states, None)
    assert np.percentile(mpos, 99) <= 0.1
    # (the mixed form evaluates the variational partial derivatives in FP32: measured 8e-5)
    assert mvar.max() <= 5e-4


@pytest.fixture(scope="module")
def c4(model):
    from assist_amd import synthetic
    rng = np.random.default_rng(4)
    n = 262144
    st, is_neo = synthetic.mixed_population(model, n, rng, 0.3, 0.0)
    states = np.zeros((n, 7, 6))
    states[:, 0] = st
    states[:, 1:] = synthetic.variational_identity(n, 6)
    return states, is_neo


def _run_device(ephem, states, t0, t1, cfg, learned=None):
    from assist_amd import batch
    import assist_amd.learned as L
    n = states.shape[0]
    d_t0, d_t1, d_st = batch.DeviceArray(t0), batch.DeviceArray(t1), batch.DeviceArray(states)
    d_status = batch.DeviceArray(np.zeros(n, np.int32))
    d_steps = batch.DeviceArray(np.zeros(n, np.int64))
    d_evals = batch.DeviceArray(np.zeros(n, np.int64))
    if learned is None:
        batch.integrate_device(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, config=cfg)
    else:
        d_dt0, d_bw = L.predict_start(ephem, learned, n, 6, d_t0, d_t1, batch.DeviceArray(states), min_dt=0.001)
        L.integrate_device_learned(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, d_dt0, d_bw,
                                   config=cfg)
    return d_st.download(), d_status.download(), d_steps.download(), d_evals.download()


def test_configuration_4_full_size_properties(gpu, ephem, oracle, model, c4):
    from assist_amd import batch
    import assist_amd.learned as L
    states, is_neo = c4
    n = states.shape[0]
    t0, t1 = np.zeros(n), np.full(n, 365.0)
    f64 = batch.BatchConfig.default(min_dt=0.001, max_steps=200000)
    a = _run_device(ephem, states, t0, t1, f64)
    b = _run_device(ephem, states, t0, t1, f64)
    assert (a[1] == 0).all()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    net = L.StepModel.load(L.default_model_path(), max_batch=65536)
    m = _run_device(ephem, states, t0, t1, batch.BatchConfig.default(min_dt=0.001, max_steps=200000,
                                                                    precision=batch.PRECISION_MIXED), net)
    assert (m[1] == 0).all()
    d = np.linalg.norm(m[0][:, 0, :3] - a[0][:, 0, :3], axis=-1) * AU_M
    print("configuration 4 in full: evaluations mixed+learned %d vs FP64 %d; main belt p99 %.3e m, NEO p99 %.3e m"
          % (m[3].sum(), a[3].sum(), np.percentile(d[~is_neo], 99), np.percentile(d[is_neo], 99)))
    bar_with_measured_exemptions(ephem, "configuration 4 in full, mixed+learned vs FP64", d, t0, t1, states, None)
    # time reversal: back from the end states to t0
    back = _run_device(ephem, a[0], t1, t0, f64)
    assert (back[1] == 0).all()
    dr = np.linalg.norm(back[0][:, 0, :3] - states[:, 0, :3], axis=-1) * AU_M
    bar_with_measured_exemptions(ephem, "configuration 4 in full, 365 d out and back", dr, t0, t1, states, None)
