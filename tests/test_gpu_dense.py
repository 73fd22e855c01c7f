"""Dense output (the batched counterpart of reference assist_integrate_or_interpolate,
src/assist.c:556-597, 642-680), checked as the reference's own test does
(unit_tests/onthefly_interpolation/problem.c: 20 output times 40 days apart, interpolated states
against integrating directly to each time, |dx| < 2e-13 au), plus backward and ragged output
lists, outputs at t0, and the drop-in Extras.integrate_or_interpolate."""
import numpy as np
import pytest

from conftest import AU_M, population

pytestmark = pytest.mark.gpu


def _main_belt(model, n, seed, n_var=0):
    from assist_amd import synthetic
    rng = np.random.default_rng(seed)
    states = np.zeros((n, 1 + n_var, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    if n_var:
        states[:, 1:] = synthetic.variational_identity(n, n_var)
    return states


def test_dense_output_matches_direct_integration(gpu, ephem, model):
    from assist_amd import batch
    n = 16
    states = _main_belt(model, n, seed=51, n_var=2)
    times = 40.0 * (np.arange(20) + 1)
    r = batch.integrate_dense(ephem, np.zeros(n), [times] * n, states)
    assert (r.status == 0).all()
    worst = 0.0
    for j, tj in enumerate(times):
        d = batch.integrate(ephem, np.zeros(n), np.full(n, tj), states)
        assert (d.status == 0).all()
        dx = np.abs(np.stack([o[j] for o in r.outputs])[:, 0, 0] - d.states[:, 0, 0])
        worst = max(worst, dx.max())
        # variational vectors interpolate with the same polynomial
        dv = np.abs(np.stack([o[j] for o in r.outputs])[:, 1:, :3] - d.states[:, 1:, :3]).max()
        assert dv <= 1e-9 * np.abs(d.states[:, 1:, :3]).max()
    print("dense output vs direct integration: max |dx| = %.3e au (%.3e m)" % (worst, worst * AU_M))
    assert worst < 2e-13


def test_dense_output_backward_ragged_and_at_start(gpu, ephem, model):
    from assist_amd import batch
    states, _, _ = population(model, 6, seed=52, n_var=0)
    t0 = np.array([0.0, 100.0, -50.0, 10.0, 0.0, 300.0])
    outs = [np.array([0.0, 5.0, 17.5, 400.0]),          # first output at t0
            np.array([90.0, 20.0, -300.0]),             # backward
            np.array([]),                               # none
            np.array([10.0 + 1e-6]),                    # a single, tiny step away
            np.array([123.456]),
            np.array([300.0, 299.0, 298.0, 250.0, -700.0])]
    r = batch.integrate_dense(ephem, t0, outs, states)
    assert (r.status == 0).all()
    assert r.outputs[2].shape[0] == 0
    np.testing.assert_array_equal(r.outputs[0][0, 0], states[0, 0])
    for i in range(6):
        for j, tj in enumerate(outs[i]):
            d = batch.integrate(ephem, t0[i:i + 1], np.array([tj]), states[i:i + 1])
            dx = np.linalg.norm(r.outputs[i][j, 0, :3] - d.states[0, 0, :3]) * AU_M
            assert dx < 0.01, (i, j, dx)


def test_dense_output_rejects_unsorted_times(gpu, ephem, model):
    from assist_amd import batch
    states, _, _ = population(model, 2, seed=53, n_var=0)
    with pytest.raises(batch.BatchError):
        batch.integrate_dense(ephem, np.zeros(2), [np.array([10.0, 5.0]), np.array([1.0])], states)
    with pytest.raises(batch.BatchError):
        batch.integrate_dense(ephem, np.zeros(2), [np.array([-10.0, 5.0]), np.array([1.0])], states)


def test_extras_integrate_or_interpolate(gpu, ephem, model):
    import assist_amd
    from assist_amd import batch, synthetic
    x = synthetic.config1_particle(model)
    sim = assist_amd.Simulation()
    sim.add(x=x[0], y=x[1], z=x[2], vx=x[3], vy=x[4], vz=x[5])
    ex = assist_amd.Extras(sim, ephem)
    ex.integrate_or_interpolate(123.25)
    assert sim.t == 123.25
    d = batch.integrate(ephem, np.zeros(1), np.array([123.25]), x.reshape(1, 1, 6))
    p = sim.particles[0]
    assert np.linalg.norm(np.array([p.x, p.y, p.z]) - d.states[0, 0, :3]) * AU_M < 0.01
    ex.detach(sim)
