"""Per-system failure codes of the batched integrator, forced from the host so that every kernel
exits normally (enum ASSIST_SYSTEM_STATUS in include/assist.h; the oracle uses the same codes):

* ASSIST_SYSTEM_NONFINITE (-1): a NaN or infinite component in a system's state;
* ASSIST_SYSTEM_MAX_STEPS (-2): max_steps accepted steps without reaching t1;
* ASSIST_SYSTEM_DT_UNDERFLOW (-3): a first trial step of 0 or NaN (the batch's dt_initial, or one
  system's learned first step);
* ASSIST_SYSTEM_COVERAGE (5): a time outside the ephemeris.

In each case the GPU reports what the oracle reports for the same input, the state it leaves is
the oracle's (the last accepted step), and the other systems of the batch end bit for bit where
they end in a batch without the failing one.
"""
import numpy as np
import pytest

from conftest import population

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, t0, t1, states, **cfg):
    return oracle.integrate(t0, t1, states, cfg=oracle.cfg(**cfg))


def test_nonfinite_state(gpu, ephem, oracle, model):
    from assist_amd import batch
    n = 6
    states, _, _ = population(model, n, seed=401, n_var=2, neo_fraction=0.0)
    bad = states.copy()
    bad[1, 0, 0] = np.nan         # real particle position
    bad[3, 2, 4] = np.inf         # a variational particle's velocity
    t0, t1 = np.zeros(n), np.full(n, 50.0)
    r = batch.integrate(ephem, t0, t1, bad)
    o = _oracle_run(oracle, t0, t1, bad)
    np.testing.assert_array_equal(r.status, o["status"])
    np.testing.assert_array_equal(r.status, [0, -1, 0, -1, 0, 0])
    good = batch.integrate(ephem, t0, t1, states)
    ok = r.status == 0
    np.testing.assert_array_equal(r.states[ok], good.states[ok])
    np.testing.assert_array_equal(r.steps[ok], good.steps[ok])


@pytest.mark.parametrize("max_steps", [1, 5])
def test_max_steps(gpu, ephem, oracle, model, max_steps):
    """max_steps accepted steps, then ASSIST_SYSTEM_MAX_STEPS with the state of the last accepted
    step.  The first steps of IAS15 from a 0.001-day trial step are sized by round-off in b6, so
    the GPU and the oracle stop at slightly different times: the GPU's state is checked against
    the oracle integrated to the time the GPU stopped (from its snapshot)."""
    from assist_amd import batch
    n = 8
    states, _, _ = population(model, n, seed=402, n_var=1)
    t0 = np.zeros(n)
    t1 = np.full(n, 200.0)
    t1[2] = 1e-4  # reached in one step: finishes before the limit
    cfg = batch.BatchConfig.default(max_steps=max_steps)
    r = batch.integrate_checkpoint(ephem, t0, t1, states, config=cfg)
    o = _oracle_run(oracle, t0, t1, states, max_steps=max_steps)
    np.testing.assert_array_equal(r.status, o["status"])
    expect = np.full(n, -2)
    expect[2] = 0
    np.testing.assert_array_equal(r.status, expect)
    np.testing.assert_array_equal(r.steps, o["steps"])
    assert (r.steps[r.status == -2] == max_steps).all()
    t_stop = r.snapshots[:, 0]
    assert (t_stop > 0).all() and (t_stop < t1).sum() == n - 1
    ref = oracle.integrate(t0, t_stop, states)
    np.testing.assert_allclose(r.states, ref["states"], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dt", [0.0, np.nan])
def test_dt_underflow_from_the_first_step(gpu, ephem, oracle, model, dt):
    from assist_amd import batch
    n = 4
    states, _, _ = population(model, n, seed=403, n_var=0)
    t0, t1 = np.zeros(n), np.full(n, 30.0)
    r = batch.integrate(ephem, t0, t1, states, config=batch.BatchConfig.default(dt_initial=dt))
    o = _oracle_run(oracle, t0, t1, states, dt=dt)
    np.testing.assert_array_equal(o["status"], -3)
    np.testing.assert_array_equal(r.status, o["status"])
    np.testing.assert_array_equal(r.steps, 0)
    np.testing.assert_array_equal(r.states, states)


def test_dt_underflow_of_one_system_learned_first_step(gpu, ephem, model):
    """A zero learned first step fails that system only."""
    from assist_amd import batch
    import assist_amd.learned as L
    n = 5
    states, _, _ = population(model, n, seed=404, n_var=0)
    t0, t1 = np.zeros(n), np.full(n, 40.0)
    dt0 = np.array([0.01, 0.0, 0.02, 0.05, 0.01])

    def run(dt0):
        d_st = batch.DeviceArray(states)
        d_status = batch.DeviceArray(np.zeros(n, np.int32))
        d_steps = batch.DeviceArray(np.zeros(n, np.int64))
        L.integrate_device_learned(ephem, n, 0, batch.DeviceArray(t0), batch.DeviceArray(t1), d_st, None, d_status,
                                   d_steps, None, batch.DeviceArray(dt0), None)
        return d_st.download(), d_status.download(), d_steps.download()

    s, st, steps = run(dt0)
    np.testing.assert_array_equal(st, [0, -3, 0, 0, 0])
    assert steps[1] == 0
    np.testing.assert_array_equal(s[1], states[1])
    fixed = dt0.copy()
    fixed[1] = 0.01
    s2, st2, steps2 = run(fixed)
    ok = st == 0
    np.testing.assert_array_equal(s[ok], s2[ok])
    np.testing.assert_array_equal(steps[ok], steps2[ok])


def test_coverage_in_a_batch(gpu, ephem, oracle, model):
    """Systems whose integration leaves the ephemeris stop with ASSIST_ERROR_COVERAGE where the
    oracle stops; the others are unaffected."""
    from assist_amd import batch
    n = 4
    states, _, _ = population(model, n, seed=405, n_var=1)
    t0 = np.array([9990.0, 0.0, -9995.0, 100.0])
    t1 = np.array([10030.0, 60.0, -10020.0, 120.0])
    r = batch.integrate(ephem, t0, t1, states)
    o = _oracle_run(oracle, t0, t1, states)
    np.testing.assert_array_equal(r.status, [5, 0, 5, 0])
    np.testing.assert_array_equal(r.status, o["status"])
    # (the stopped systems end at their last accepted step, whose time depends on the step
    # sequence; the others are held to the oracle)
    ok = r.status == 0
    np.testing.assert_allclose(r.states[ok], o["states"][ok], rtol=1e-12, atol=1e-13)
