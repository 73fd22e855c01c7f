"""GPU tests of the learned IAS15 start (assist_amd/csrc/mlp.hip, assist_amd/learned.py):

* the MFMA forward pass and the hand-written backward pass against the CPU oracle
  (oracle/mlp.py, which emulates the BF16 operand rounding), on shapes that are not
  multiples of the 16 x 16 tiles;
* Adam training on the GPU fits a teacher network;
* the feature kernel against features computed from the oracle's ephemeris;
* step traces of the plain integrator are consistent with its step counts;
* the trained network (assist_amd/models/step_mlp.npz) against the plain start: IAS15's own
  error test still accepts or rejects every step, so final states agree with the plain run to
  the integrator's accuracy,
  and the learned start saves force evaluations.
"""
import os
import sys

import numpy as np
import pytest

from conftest import AU_M, population

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp as mlp_oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def _random_params(seed, scale=0.2):
    import assist_amd.learned as L
    rng = np.random.default_rng(seed)
    return (rng.normal(0, scale, L.num_params())).astype(np.float32)


def test_mlp_forward_matches_oracle(gpu):
    import assist_amd.learned as L
    from assist_amd.batch import DeviceArray
    net = L.StepModel(max_batch=512, seed=1)
    p = _random_params(3)
    net.set_params(p)
    rng = np.random.default_rng(4)
    m = 301
    X = rng.normal(0, 1, (m, 16)).astype(np.float32)
    d_X, d_Y = DeviceArray(X), DeviceArray(np.zeros((m, 32), np.float32))
    net.forward(d_X, m, d_Y)
    y = d_Y.download()
    ref, _ = mlp_oracle.forward(p, X, bf16_ops=True)
    scale = np.abs(ref).max()
    err = np.abs(y - ref)
    print("forward: max |err| / max |y| = %.3g, mean %.3g" % (err.max() / scale, err.mean() / scale))
    assert err.max() <= 1e-2 * scale
    assert err.mean() <= 1e-3 * scale


@pytest.mark.parametrize("m", [1, 301, 4096, 70001])
def test_mlp_predict_equals_forward(gpu, m):
    """The fused prediction kernel (all four layers in one launch, weights staged in LDS) gives
    the forward pass's outputs bit for bit, on ragged row counts and with the shipped network;
    it also takes batches beyond the network's workspace (70 001 rows against 65 536)."""
    import assist_amd.learned as L
    from assist_amd.batch import DeviceArray
    rng = np.random.default_rng(40 + m)
    X = rng.normal(0, 1, (m, 16)).astype(np.float32)
    for net in (L.StepModel(max_batch=min(m, 65536), seed=5), L.StepModel.load(L.default_model_path(), max_batch=min(m, 65536))):
        d_X, d_P = DeviceArray(X), DeviceArray(np.zeros((m, 32), np.float32))
        net.predict(d_X, m, d_P)
        yp = d_P.download()
        yf = np.zeros_like(yp)
        for lo in range(0, m, 65536):
            hi = min(m, lo + 65536)
            d_x, d_y = DeviceArray(X[lo:hi]), DeviceArray(np.zeros((hi - lo, 32), np.float32))
            net.forward(d_x, hi - lo, d_y)
            yf[lo:hi] = d_y.download()
        np.testing.assert_array_equal(yp, yf)


def test_mlp_gradients_match_oracle(gpu):
    import assist_amd.learned as L
    from assist_amd.batch import DeviceArray
    net = L.StepModel(max_batch=2048, seed=2)
    p = _random_params(5)
    net.set_params(p)
    w = np.linspace(0.2, 1.0, 32).astype(np.float32)
    net.set_output_weights(w)
    rng = np.random.default_rng(6)
    m = 1500  # several reduction chunks of the weight gradients, ragged last tile
    X = rng.normal(0, 1, (m, 16)).astype(np.float32)
    T = rng.normal(0, 1, (m, 32)).astype(np.float32)
    loss, g = net.loss_grad(DeviceArray(X), DeviceArray(T), m, want_grad=True)
    ref_loss, ref_g = mlp_oracle.loss_grad(p, X, T, w, bf16_ops=True)
    rel = np.linalg.norm(g - ref_g) / np.linalg.norm(ref_g)
    print("gradients: loss %.6g vs %.6g, relative L2 error %.3g" % (loss, ref_loss, rel))
    assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss)
    assert rel <= 1e-2
    # every layer's weight and bias block agrees on its own
    o = 0
    for (ws, bs) in L.param_shapes():
        for shape in (ws, bs):
            k = int(np.prod(shape))
            blk, rblk = g[o:o + k], ref_g[o:o + k]
            assert np.linalg.norm(blk - rblk) <= 2e-2 * np.linalg.norm(rblk) + 1e-12, (shape, o)
            o += k


def test_mlp_training_fits_a_teacher(gpu):
    import assist_amd.learned as L
    rng = np.random.default_rng(7)
    teacher = _random_params(8, scale=0.15)
    m = 16384
    X = rng.normal(0, 1, (m, 16)).astype(np.float32)
    X[:, 10:] = 0.0
    T, _ = mlp_oracle.forward(teacher, X, bf16_ops=False)
    T = T.astype(np.float32)
    net = L.StepModel(max_batch=4096, seed=9)
    net.set_output_weights(np.ones(32, np.float32))
    hist = L.train(net, X, T, epochs=12, batch=1024, lr=3e-3, lr_end=3e-4)
    print("teacher fit: loss per epoch", ["%.4g" % h for h in hist])
    assert hist[-1] < 0.3 * hist[0]
    assert all(b <= a * 1.02 for a, b in zip(hist[1:], hist[2:]))


def test_features_match_oracle_ephemeris(gpu, ephem, oracle, model):
    import assist_amd.learned as L
    states, _, _ = population(model, 64, seed=21, n_var=0)
    rng = np.random.default_rng(22)
    rows = np.zeros((64, 8))
    rows[:, 0] = rng.uniform(-5000, 5000, 64)
    rows[:, 1:7] = states[:, 0]
    rows[:, 7] = rng.uniform(0.01, 30.0, 64)
    X, omega = L.features(ephem, rows)
    for i in range(64):
        t = rows[i, 0]
        sun = oracle.all_ephem(0, t)[1]
        xh = rows[i, 1:4] - sun[1:4]
        vh = rows[i, 4:7] - sun[4:7]
        d = sorted(np.linalg.norm(rows[i, 1:4] - oracle.all_ephem(b, t)[1][1:4]) for b in range(11))
        ref = np.concatenate([xh / 3.0, vh * 40.0, np.log10(d[:3]), [np.log10(rows[i, 7])]])
        np.testing.assert_allclose(X[i, :10], ref, rtol=2e-6, atol=2e-6)
        assert (X[i, 10:] == 0).all()
        np.testing.assert_allclose(omega[i], np.sqrt(sun[0] / np.linalg.norm(xh) ** 3), rtol=1e-12)


def test_traces_follow_the_integration(gpu, ephem, model):
    import assist_amd.learned as L
    from assist_amd.batch import BatchConfig
    states, _, _ = population(model, 32, seed=23, n_var=0, neo_fraction=0.0)
    t1 = np.full(32, 400.0)
    rec, _, status, steps, _ = L.integrate_traced(ephem, np.zeros(32), t1, states, capacity=100000, every=1, skip=3,
                                                  config=BatchConfig.default())
    assert (status == 0).all()
    assert rec.shape[0] == int(np.sum(np.maximum(steps - 3, 0)))
    # records of one system chain: t_next = t + dts
    order = np.lexsort((rec[:, 0], rec[:, 1]))
    assert np.isfinite(rec).all()
    assert (rec[:, 8] > 0).all() and (rec[:, 7] > 0).all()
    # IAS15 never grows the step more than fourfold, and the proposal is the uncapped one
    assert (rec[:, 8] <= 4.0 * rec[:, 7] * (1 + 1e-12)).all()
    # during the start-up ramp the error estimate asks for far more than the fourfold cap
    assert (rec[:, 9] > 4.0 * rec[:, 8]).any()
    assert order.size == rec.shape[0]


def _learned_model_or_skip():
    import assist_amd.learned as L
    path = L.default_model_path()
    if not os.path.exists(path):
        pytest.skip("no trained step model at %s (python -m assist_amd.train_step_model)" % path)
    return L.StepModel.load(path, max_batch=65536)


def test_learned_start_ablation(gpu, ephem, model):
    """The trained network against the plain start on 2048 systems x (1 + 6 variational),
    365 d, the 70/30 main-belt/NEO mix."""
    import assist_amd.learned as L
    from assist_amd import batch
    net = _learned_model_or_skip()
    n = 2048
    states, _, is_neo = population(model, n, seed=31, n_var=6)
    t0, t1 = np.zeros(n), np.full(n, 365.0)
    cfg = batch.BatchConfig.default(min_dt=0.001, max_steps=200000)
    d_t0, d_t1 = batch.DeviceArray(t0), batch.DeviceArray(t1)
    d_init = batch.DeviceArray(states)

    def run(learned, warm):
        d_st = batch.DeviceArray(states)
        d_status = batch.DeviceArray(np.zeros(n, np.int32))
        d_steps = batch.DeviceArray(np.zeros(n, np.int64))
        d_evals = batch.DeviceArray(np.zeros(n, np.int64))
        if learned:
            d_dt0, d_bw = L.predict_start(ephem, net, n, 6, d_t0, d_t1, d_init, min_dt=0.001, warm=warm)
            L.integrate_device_learned(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, d_dt0, d_bw,
                                       config=cfg)
        else:
            batch.integrate_device(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, config=cfg)
        return d_st.download(), d_status.download(), d_evals.download()

    s0, st0, e0 = run(False, False)
    s1, st1, e1 = run(True, False)
    s2, st2, e2 = run(True, True)
    assert (st0 == 0).all() and (st1 == 0).all() and (st2 == 0).all()
    for s in (s1, s2):
        d = np.linalg.norm(s[:, 0, :3] - s0[:, 0, :3], axis=1) * AU_M
        print("position difference to the plain start: main belt max %.3g m, NEO max %.3g m"
              % (d[~is_neo].max(), d[is_neo].max()))
        assert d[~is_neo].max() < 1.0
    print("force evaluations: plain %d, learned step %d (%.3fx), learned step + warm start %d (%.3fx)"
          % (e0.sum(), e1.sum(), e0.sum() / e1.sum(), e2.sum(), e0.sum() / e2.sum()))
    assert e1.sum() < e0.sum()


def test_mixed_learned_10yr_position_error(gpu, ephem, oracle, model):
    """The mixed-precision form started by the trained network, against the FP64 oracle: 64
    main-belt systems x (1 + 6 variational), 3650 days, <= 1 m."""
    import assist_amd.learned as L
    from assist_amd import batch, synthetic
    net = _learned_model_or_skip()
    rng = np.random.default_rng(2)
    n = 64
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    t0, t1 = np.zeros(n), np.full(n, 3650.0)
    d_t0, d_t1, d_init = batch.DeviceArray(t0), batch.DeviceArray(t1), batch.DeviceArray(states)
    d_st = batch.DeviceArray(states)
    d_status = batch.DeviceArray(np.zeros(n, np.int32))
    d_steps = batch.DeviceArray(np.zeros(n, np.int64))
    d_evals = batch.DeviceArray(np.zeros(n, np.int64))
    d_dt0, d_bw = L.predict_start(ephem, net, n, 6, d_t0, d_t1, d_init)
    L.integrate_device_learned(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, d_dt0, d_bw,
                               config=batch.BatchConfig.default(precision=batch.PRECISION_MIXED))
    o = oracle.integrate(t0, t1, states, None)
    assert (d_status.download() == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(d_st.download()[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("mixed + learned, 10 yr: max |dr| = %.3e m, evaluations %d (oracle %d)"
          % (dpos.max(), d_evals.download().sum(), o["evals"].sum()))
    assert dpos.max() <= 1.0
