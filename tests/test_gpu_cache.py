"""The substep ephemeris cache of 8-lane groups leaves every result unchanged, bit for bit.

The first predictor-corrector iteration of a step attempt stores the raw Chebyshev positions of
its 7 substeps, the later iterations load them (the reference's per-body cache keyed by the exact
time, reference src/forces.c:175-263, serves the same repeated substep times), and the groups
that need fresh positions share the Chebyshev rounds of their wave.  Each case runs the same
batch with the cache on and off (assist_batch_config.substep_cache_off = -1 and 1: every trip
then evaluates every group's positions in the wave-wide rounds) and compares states, status, step
counts and the launch's evaluation and rejection totals exactly.  (Agreement with the oracle,
with the cache on, is the subject of the parity tests.)
"""
import numpy as np
import pytest

from conftest import population

pytestmark = pytest.mark.gpu


# lanes per system and variational particles per system of each case: 8-lane groups with 6
# variational particles, and with the real particle only (7 lanes of each group own no vector)
SHAPES = [(8, 6), (8, 0)]


def _pair(ephem, t0, t1, states, params, lanes, **cfg):
    from assist_amd import batch
    on = batch.integrate(ephem, t0, t1, states, params,
                         config=batch.BatchConfig.default(lanes_per_system=lanes, substep_cache_off=-1, **cfg))
    off = batch.integrate(ephem, t0, t1, states, params,
                          config=batch.BatchConfig.default(lanes_per_system=lanes, substep_cache_off=1, **cfg))
    return on, off


def _same(on, off, lanes):
    assert (on.status == 0).all() and (off.status == 0).all()
    assert np.array_equal(on.states, off.states)
    assert np.array_equal(on.steps, off.steps)
    assert (on.stats.evals, on.stats.rejected) == (off.stats.evals, off.stats.rejected)
    assert on.stats.lanes_per_system == lanes


@pytest.mark.parametrize("lanes,n_var", SHAPES)
@pytest.mark.parametrize("precision", [0, 1])
def test_cache_equal_spans(gpu, ephem, model, precision, lanes, n_var):
    states, _, _ = population(model, 3000, 41, n_var=n_var)
    n = states.shape[0]
    on, off = _pair(ephem, np.zeros(n), np.full(n, 120.0), states, None, lanes, precision=precision, min_dt=0.001)
    print("equal spans, %d lanes, precision %d: %d steps, %d evaluations, %d rejected"
          % (lanes, precision, on.stats.steps, on.stats.evals, on.stats.rejected))
    _same(on, off, lanes)


@pytest.mark.parametrize("lanes,n_var", SHAPES)
def test_cache_uneven_spans_marsden(gpu, ephem, model, lanes, n_var):
    # ragged tail: groups leave their wave at different times (partly active waves), half of
    # the systems integrate backward, 30% carry a Marsden model
    rng = np.random.default_rng(42)
    states, params, _ = population(model, 2500, 42, n_var=n_var, ng_fraction=0.3)
    n = states.shape[0]
    span = np.exp(rng.uniform(np.log(5.0), np.log(400.0), n)) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    t0 = rng.uniform(-200.0, 200.0, n)
    on, off = _pair(ephem, t0, t0 + span, states, params, lanes, min_dt=0.001)
    print("uneven spans, %d lanes: %d steps, %d evaluations, %d rejected"
          % (lanes, on.stats.steps, on.stats.evals, on.stats.rejected))
    _same(on, off, lanes)


@pytest.mark.parametrize("lanes,n_var", SHAPES)
def test_cache_few_systems(gpu, ephem, model, lanes, n_var):
    # fewer systems than groups in one wave, so most lanes of the grid never start
    states, _, _ = population(model, 3, 43, n_var=n_var)
    on, off = _pair(ephem, np.zeros(3), np.array([30.0, -45.0, 10.0]), states, None, lanes)
    _same(on, off, lanes)


@pytest.mark.parametrize("lanes,n_var", SHAPES)
def test_cache_dense_output(gpu, ephem, model, lanes, n_var):
    from assist_amd import batch
    states, _, _ = population(model, 600, 44, n_var=n_var)
    n = states.shape[0]
    rng = np.random.default_rng(44)
    outs = [np.sort(rng.uniform(0.0, 90.0, 5)) for _ in range(n)]
    on = batch.integrate_dense(ephem, np.zeros(n), outs, states,
                               config=batch.BatchConfig.default(lanes_per_system=lanes, substep_cache_off=-1))
    off = batch.integrate_dense(ephem, np.zeros(n), outs, states,
                                config=batch.BatchConfig.default(lanes_per_system=lanes, substep_cache_off=1))
    assert (on.status == 0).all()
    assert np.array_equal(on.states, off.states)
    for a, b in zip(on.outputs, off.outputs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("lanes,n_var", SHAPES)
def test_cache_rejected_steps(gpu, ephem, model, lanes, n_var):
    # a 500-day first trial step (a third of a main-belt orbit) is rejected and retried with a
    # smaller one, so every system starts a second attempt of its first step, which must refill
    # the cache of its substeps
    states, _, _ = population(model, 1500, 45, n_var=n_var, neo_fraction=0.0)
    n = states.shape[0]
    on, off = _pair(ephem, np.zeros(n), np.full(n, 600.0), states, None, lanes, dt_initial=500.0)
    print("rejections, %d lanes: %d steps, %d evaluations, %d rejected"
          % (lanes, on.stats.steps, on.stats.evals, on.stats.rejected))
    assert on.stats.rejected >= n
    _same(on, off, lanes)
