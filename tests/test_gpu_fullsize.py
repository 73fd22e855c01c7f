"""Parity at the benchmark configurations' full sizes.

The 1 m bar is stated for orbits without close approaches: it is applied to the orbits whose
initial heliocentric perihelion stays outside Mars's aphelion (+ 0.08 au) and aphelion inside
3.95 au; the planet-crossing ones are reported (and held to p99 <= 0.1 m where compared with the
oracle).

* Configuration 2 in full (4 096 main-belt systems x (1 + 6 variational), 3 650 days) against
  the oracle, system by system: <= 1 m in position, variational states to 1e-6.
* Configuration 4 in full (262 144 systems x (1 + 6 variational), 365 days, 70/30 main-belt/NEO,
  seed 4), through size-independent properties:
  - two FP64 launches are bit-identical (the work queue hands systems out in a different order);
  - the mixed+learned form agrees with the FP64 form within 1 m;
  - forward then backward returns the systems to their start within 1 m.
The oracle runs on all host cores; the whole module takes under a minute on one MI355X box.
"""
import os

import numpy as np
import pytest

from conftest import AU_M, rel_err

pytestmark = pytest.mark.gpu


def _planet_crossers(oracle, states):
    """Systems whose initial heliocentric orbit reaches inside Mars's aphelion (plus 0.08 au) or
    beyond 3.95 au: those can approach Mars or feel Jupiter strongly, and the 1 m bar is stated
    for orbits without close approaches."""
    from assist_amd import synthetic
    sun = oracle.all_ephem(0, 0.0)[1]
    r = states[:, 0, :3] - sun[1:4]
    v = states[:, 0, 3:6] - sun[4:7]
    a, e = synthetic.state_to_elements(sun[0], r, v)[:2]
    return (a * (1 - e) < 1.666 + 0.08) | (a * (1 + e) > 3.95), a, e


def test_configuration_2_full_size_against_oracle(gpu, ephem, oracle, model):
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(2)
    n = 4096
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r = batch.integrate(ephem, 0.0, 3650.0, states)
    o = oracle.integrate(0.0, 3650.0, states, nthreads=os.cpu_count() or 8)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    dvar = rel_err(r.states[:, 1:], o["states"][:, 1:])
    cross, a, e = _planet_crossers(oracle, states)
    print("configuration 2, all 4096 systems: max |dr| %.3e m, p99 %.3e m; without planet-crossing orbits "
          "(%d of %d) max %.3e m; variational max rel %.3e; steps differ by at most %d"
          % (dpos.max(), np.percentile(dpos, 99), (~cross).sum(), n, dpos[~cross].max(), dvar.max(),
             np.abs(r.steps - o["steps"]).max()))
    for i in np.argsort(-dpos)[:5]:
        print("  system %d: |dr| %.3e m, a %.3f au, e %.3f, q %.3f au, Q %.3f au, crosser %s"
              % (i, dpos[i], a[i], e[i], a[i] * (1 - e[i]), a[i] * (1 + e[i]), bool(cross[i])))
    assert dpos[~cross].max() <= 1.0
    assert np.percentile(dpos, 99) <= 0.1
    assert dvar.max() <= 1e-6


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
    quiet = ~is_neo & ~_planet_crossers(oracle, states)[0]
    print("configuration 4 in full: mixed+learned vs FP64: main belt without planet crossers (%d) max %.3e m, "
          "all main belt p99 %.3e m, NEO p99 %.3e m; evaluations %d vs %d"
          % (quiet.sum(), d[quiet].max(), np.percentile(d[~is_neo], 99), np.percentile(d[is_neo], 99), m[3].sum(),
             a[3].sum()))
    assert d[quiet].max() <= 1.0
    # time reversal: back from the end states to t0
    back = _run_device(ephem, a[0], t1, t0, f64)
    assert (back[1] == 0).all()
    dr = np.linalg.norm(back[0][:, 0, :3] - states[:, 0, :3], axis=-1) * AU_M
    print("configuration 4 in full, 365 d out and back: main belt without planet crossers max %.3e m, "
          "main belt p99 %.3e m" % (dr[quiet].max(), np.percentile(dr[~is_neo], 99)))
    assert dr[quiet].max() <= 1.0
