"""Checkpoint/resume and interpolation in a step archive (the batched counterparts of REBOUND's
SimulationArchive as ASSIST uses it, reference src/assist.c:599-633 assist_create_interpolated_
simulation and :682-752 assist_interpolate_simulation, tested by the reference at
unit_tests/interpolation_spk/problem.c with an archive saved every step).

* A system paused at a step boundary and resumed from its snapshot ends where an uninterrupted
  integration ends, bit for bit, with the same step count: the snapshot holds everything IAS15
  carries between steps (state, compensations, last step's b and e, step sizes).
* States interpolated from the archive agree with dense output and, to the reference's bar for
  on-the-fly interpolation (unit_tests/onthefly_interpolation/problem.c:54, 2e-13 au), with
  integrating directly to each time.
"""
import numpy as np
import pytest

from conftest import AU_M, population

pytestmark = pytest.mark.gpu


def _spans(n, rng, lo=150.0, hi=400.0):
    t1 = rng.uniform(lo, hi, n)
    t1[1::2] *= -1.0
    return t1


@pytest.mark.parametrize("n_var,lanes,precision", [(6, 0, 0), (6, 0, 1), (0, 1, 0), (0, 1, 1), (2, 4, 0)])
def test_resume_equals_uninterrupted_run(gpu, ephem, model, n_var, lanes, precision):
    from assist_amd import batch
    n = 64
    states, params, _ = population(model, n, seed=301 + n_var + lanes, n_var=n_var, ng_fraction=0.25)
    rng = np.random.default_rng(7)
    t0 = rng.uniform(-100.0, 100.0, n)
    t1 = t0 + _spans(n, rng)
    cfg = batch.BatchConfig.default(lanes_per_system=lanes, precision=precision, min_dt=0.001)
    full = batch.integrate_checkpoint(ephem, t0, t1, states, params, config=cfg)
    assert (full.status == 0).all()
    # pause twice, resuming from the snapshot each time (the snapshots pick the lanes per system)
    cfg_resume = batch.BatchConfig.default(precision=precision, min_dt=0.001)
    p1 = batch.integrate_checkpoint(ephem, t0, t1, states, params, config=cfg, t_pause=t0 + 0.3 * (t1 - t0))
    assert (p1.status == batch.SYSTEM_PAUSED).all()
    assert (p1.snapshots[:, 7] == float(p1.stats.lanes_per_system)).all()
    # the pause is the first boundary at or past the pause time
    tp = p1.snapshots[:, 0]
    assert ((tp - (t0 + 0.3 * (t1 - t0))) * np.sign(t1 - t0) >= 0).all()
    assert (np.abs(tp - t0) < np.abs(t1 - t0)).all()
    p2 = batch.integrate_checkpoint(ephem, t0, t1, p1.states, params, config=cfg_resume,
                                    t_pause=t0 + 0.7 * (t1 - t0), resume=p1.snapshots)
    assert (p2.status == batch.SYSTEM_PAUSED).all()
    assert p2.stats.lanes_per_system == full.stats.lanes_per_system
    p3 = batch.integrate_checkpoint(ephem, t0, t1, p2.states, params, config=cfg_resume, resume=p2.snapshots)
    assert (p3.status == 0).all()
    np.testing.assert_array_equal(p3.states, full.states)
    np.testing.assert_array_equal(p3.steps, full.steps)
    np.testing.assert_array_equal(p3.evals, full.evals)
    np.testing.assert_array_equal(p3.snapshots, full.snapshots)
    # the state in a snapshot is the state the entry point returns
    w = batch.snapshot_words(1 + n_var)
    assert w == p1.snapshots.shape[1]
    sv = p1.snapshots[:, 8:].reshape(n, 1 + n_var, 57)[:, :, :6]
    np.testing.assert_array_equal(sv, p1.states)


def test_resume_from_snapshot_at_start_and_at_end(gpu, ephem, model):
    """A snapshot taken before the first step (a zero pause) resumes like a fresh start; one of a
    finished system resumes to a zero-length integration."""
    from assist_amd import batch
    n = 8
    states, _, _ = population(model, n, seed=311, n_var=3)
    t0 = np.zeros(n)
    t1 = np.full(n, 120.0)
    full = batch.integrate_checkpoint(ephem, t0, t1, states)
    start = batch.integrate_checkpoint(ephem, t0, t1, states, t_pause=t0)
    # the first boundary at or past t0 is the end of the first step
    assert (start.status == batch.SYSTEM_PAUSED).all() and (start.steps == 1).all()
    again = batch.integrate_checkpoint(ephem, t0, t1, start.states, resume=start.snapshots)
    np.testing.assert_array_equal(again.states, full.states)
    done = batch.integrate_checkpoint(ephem, t0, t1, full.states, resume=full.snapshots)
    assert (done.status == 0).all()
    np.testing.assert_array_equal(done.states, full.states)
    np.testing.assert_array_equal(done.steps, full.steps)


def test_archive_records_every_step(gpu, ephem, model):
    from assist_amd import batch
    n = 12
    states, _, _ = population(model, n, seed=321, n_var=2)
    rng = np.random.default_rng(3)
    t0 = np.zeros(n)
    t1 = _spans(n, rng, 60.0, 200.0)
    arc = batch.integrate_archive(ephem, t0, t1, states, capacity=512)
    assert (arc.result.status == 0).all()
    np.testing.assert_array_equal(arc.counts, arc.result.steps + 1)
    plain = batch.integrate(ephem, t0, t1, states)
    np.testing.assert_array_equal(arc.result.states, plain.states)
    for i in range(n):
        s = arc.snapshots(i)
        # start state, then one snapshot per accepted step, ending at t1 with the final state
        np.testing.assert_array_equal(s[0, 8:].reshape(3, 57)[:, :6], states[i])
        assert s[0, 0] == t0[i] and s[0, 2] == 0.0
        np.testing.assert_array_equal(s[-1, 8:].reshape(3, 57)[:, :6], plain.states[i])
        assert s[-1, 0] == t1[i]
        dts = s[1:, 2]
        assert (np.sign(dts) == np.sign(t1[i])).all()
        np.testing.assert_allclose(np.diff(s[:, 0]), dts, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(s[:, 3], np.arange(len(s)))


def test_archive_interpolation_matches_dense_output_and_direct_integration(gpu, ephem, model):
    """20 times 40 days apart (the reference's on-the-fly interpolation test shape), forward and
    backward: the archive gives the dense-output states, and agrees with integrating directly to
    each time within 2e-13 au."""
    from assist_amd import batch
    n = 16
    states, _, _ = population(model, n, seed=331, n_var=2, neo_fraction=0.0)
    times = 40.0 * (np.arange(20) + 1)
    for sign in (1.0, -1.0):
        t0 = np.zeros(n)
        t1 = np.full(n, sign * 900.0)
        arc = batch.integrate_archive(ephem, t0, t1, states, capacity=1024)
        assert (arc.result.status == 0).all()
        dense = batch.integrate_dense(ephem, t0, [sign * times] * n, states)
        sys = np.repeat(np.arange(n), len(times))
        tq = np.tile(sign * times, n)
        out, st = arc.interpolate(sys, tq)
        assert (st == 0).all()
        ref = np.stack(dense.outputs).reshape(-1, 3, 6)
        np.testing.assert_allclose(out, ref, rtol=1e-15, atol=1e-15)
        worst = 0.0
        for j, tj in enumerate(sign * times):
            d = batch.integrate(ephem, t0, np.full(n, tj), states)
            dx = np.abs(out.reshape(n, len(times), 3, 6)[:, j, 0, :3] - d.states[:, 0, :3]).max()
            worst = max(worst, dx)
        print("archive interpolation vs direct integration (%+d): max |dx| = %.3e au (%.3e m)"
              % (sign, worst, worst * AU_M))
        assert worst < 2e-13


def test_archive_interpolation_edges(gpu, ephem, model):
    from assist_amd import batch
    n = 4
    states, _, _ = population(model, n, seed=341, n_var=0)
    t0 = np.array([0.0, 10.0, 0.0, 5.0])
    t1 = np.array([50.0, -40.0, 0.0, 5.5])
    arc = batch.integrate_archive(ephem, t0, t1, states, capacity=256)
    assert (arc.result.status == 0).all()
    assert arc.counts[2] == 1  # zero-length span: the start snapshot only
    s0 = arc.snapshots(0)
    # at every archived time the archive returns that snapshot's state exactly
    out, st = arc.interpolate(np.zeros(len(s0), dtype=np.int32), s0[:, 0])
    assert (st == 0).all()
    np.testing.assert_array_equal(out[:, 0, :6], s0[:, 8:14])
    # the first step interpolates too; ends and outside times
    qs = np.array([0, 0, 0, 1, 1, 2, 2, 3, 0, 9, -1])
    qt = np.array([1e-3, 50.0, 50.5, -40.0, 10.5, 0.0, 1.0, 5.25, -1e-9, 0.0, 0.0])
    out, st = arc.interpolate(qs, qt)
    np.testing.assert_array_equal(st, [0, 0, 5, 0, 5, 0, 5, 0, 5, 100, 100])
    np.testing.assert_array_equal(out[st != 0], 0.0)
    np.testing.assert_array_equal(out[1, 0], arc.result.states[0, 0])
    np.testing.assert_array_equal(out[3, 0], arc.result.states[1, 0])
    np.testing.assert_array_equal(out[5, 0], states[2, 0])


def test_archive_capacity_overflow_is_reported(gpu, ephem, model):
    from assist_amd import batch
    states, _, _ = population(model, 2, seed=351, n_var=0)
    with pytest.raises(batch.BatchError, match="capacity"):
        batch.integrate_archive(ephem, np.zeros(2), np.full(2, 300.0), states, capacity=3)
