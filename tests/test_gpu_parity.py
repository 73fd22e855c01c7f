"""GPU parity against the CPU oracle (HIP kernels through the C-ABI of libassist.so).

Bars (from the hot-path contract):
  * FP64 force evaluation: relative error <= 1e-12 per evaluated vector, for the
    full model and for every term on its own;
  * FP64 integration: real-particle position within 1 m of the oracle after the
    span for orbits without close approaches (10 years for the main belt).
"""
import numpy as np
import pytest

from conftest import AU_M, population, rel_err

pytestmark = pytest.mark.gpu

FORCE_TERMS = {
    "SUN": 0x01, "PLANETS": 0x02, "ASTEROIDS": 0x04, "NON_GRAVITATIONAL": 0x08, "EARTH_HARMONICS": 0x10,
    "SUN_HARMONICS": 0x20, "GR_EIH": 0x40, "GR_SIMPLE": 0x80, "GR_POTENTIAL": 0x100,
}


def _forces_both(ephem, oracle, t, states, params, forces, precision=0, **shape):
    from assist_amd import batch
    opts = dict(forces=forces, **shape)
    g, gs = batch.forces(ephem, t, states, params, extras=opts, precision=precision)
    o, os_ = oracle.forces(t, states, params, oracle.opts(forces=forces, **shape))
    assert (gs == os_).all()
    return g, o


def _sample(model, n, seed, n_var=6, ng=0.5):
    rng = np.random.default_rng(seed)
    states, params, _ = population(model, n, seed, n_var=n_var, ng_fraction=ng)
    t = rng.uniform(-9500.0, 9500.0, n)
    if n_var:
        # generic variational vectors, and Marsden-parameter variations on the last one
        states[:, 1:] = rng.normal(size=states[:, 1:].shape)
        params[:, n_var] = rng.normal(size=(n, 3)) * 1e-12
    return t, states, params


def test_forces_full_model_fp64(gpu, ephem, oracle, model):
    t, states, params = _sample(model, 512, 11)
    g, o = _forces_both(ephem, oracle, t, states, params, 0x7F)
    e_real = rel_err(g[:, 0], o[:, 0])
    e_var = rel_err(g[:, 1:], o[:, 1:])
    print("full model: max rel err real %.3e var %.3e" % (e_real.max(), e_var.max()))
    assert e_real.max() <= 1e-12
    assert e_var.max() <= 1e-12


@pytest.mark.parametrize("term", list(FORCE_TERMS))
def test_forces_each_term_fp64(gpu, ephem, oracle, model, term):
    t, states, params = _sample(model, 256, 12)
    g, o = _forces_both(ephem, oracle, t, states, params, FORCE_TERMS[term])
    e_real = rel_err(g[:, 0], o[:, 0])
    e_var = rel_err(g[:, 1:], o[:, 1:])
    nz = np.linalg.norm(o[:, 0], axis=-1) > 0
    print("%s: max rel err real %.3e var %.3e" % (term, e_real[nz].max() if nz.any() else 0, e_var.max()))
    assert np.abs(g[~nz, 0]).max(initial=0.0) == 0.0
    assert e_real.max() <= 1e-12
    assert e_var.max() <= 1e-12


@pytest.mark.parametrize("n_var", [0, 1, 3, 7])
def test_forces_variational_counts(gpu, ephem, oracle, model, n_var):
    t, states, params = _sample(model, 128, 13, n_var=n_var)
    g, o = _forces_both(ephem, oracle, t, states, params, 0x7F)
    assert rel_err(g, o).max() <= 1e-12


def test_forces_coverage_status(gpu, ephem, oracle, model):
    from assist_amd import batch
    states, _, _ = population(model, 4, 3, n_var=0)
    t = np.array([0.0, 10001.0, -10001.0, 9999.0])
    _, st = batch.forces(ephem, t, states)
    _, ost = oracle.forces(t, states)
    assert list(st) == [0, 5, 5, 0]
    assert list(ost) == [0, 5, 5, 0]


def test_forces_mixed_precision_error(gpu, ephem, oracle, model):
    """Mixed form: real acceleration stays at FP64 level; variational partials are FP32."""
    t, states, params = _sample(model, 512, 14)
    g, o = _forces_both(ephem, oracle, t, states, params, 0x7F, precision=1)
    e_real = rel_err(g[:, 0], o[:, 0]).max()
    e_var = rel_err(g[:, 1:], o[:, 1:]).max()
    print("mixed: max rel err real %.3e var %.3e" % (e_real, e_var))
    assert e_real <= 1e-11
    assert e_var <= 1e-5


# Marsden g(r) = alpha (r/r0)^-nm (1 + (r/r0)^nn)^-nk: the default 1/r^2, the water-ice law of
# Marsden, Sekanina & Yeomans (1973), and another nk = 0 shape (the kernels skip the powers that
# nk = 0 makes exactly 1 or -0)
MARSDEN_SHAPES = {
    "default": {},
    "comet": dict(alpha=0.1112620426, nk=4.6142, nm=2.15, nn=5.093, r0=2.808),
    "nk0": dict(alpha=0.5, nk=0.0, nm=2.6, nn=3.0, r0=1.5),
}


@pytest.mark.parametrize("shape", list(MARSDEN_SHAPES))
def test_forces_marsden_shapes(gpu, ephem, oracle, model, shape):
    t, states, params = _sample(model, 256, 15, ng=1.0)
    kw = MARSDEN_SHAPES[shape]
    g, o = _forces_both(ephem, oracle, t, states, params, 0x08, **kw)
    e_term = rel_err(g, o).max()
    g, o = _forces_both(ephem, oracle, t, states, params, 0x7F, **kw)
    e_full = rel_err(g, o).max()
    gm, om = _forces_both(ephem, oracle, t, states, params, 0x7F, precision=1, **kw)
    e_mixed = rel_err(gm[:, 0], om[:, 0]).max()
    print("Marsden %s: term %.3e, full model %.3e, mixed real %.3e" % (shape, e_term, e_full, e_mixed))
    assert e_term <= 1e-12 and e_full <= 1e-12
    assert e_mixed <= 1e-11


def _integrate_both(ephem, oracle, t0, t1, states, params=None, **cfg):
    from assist_amd import batch
    r = batch.integrate(ephem, t0, t1, states, params, config=batch.BatchConfig.default(**cfg))
    ocfg = oracle.cfg(**{k: v for k, v in cfg.items() if k in ("epsilon", "min_dt")})
    if "dt_initial" in cfg:
        ocfg.dt = cfg["dt_initial"]
    o = oracle.integrate(t0, t1, states, params, cfg=ocfg)
    return r, o


def test_integrate_main_belt_10yr_with_variations(gpu, ephem, oracle, model):
    """Configuration-2 shape at reduced count: 64 main-belt systems x (1+6), 3650 days."""
    from assist_amd import synthetic
    rng = np.random.default_rng(2)
    n = 64
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r, o = _integrate_both(ephem, oracle, 0.0, 3650.0, states)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    dvar = rel_err(r.states[:, 1:], o["states"][:, 1:])
    print("10 yr: max |dr| = %.3e m, steps gpu %d oracle %d, var rel %.3e" % (
        dpos.max(), r.steps.sum(), o["steps"].sum(), dvar.max()))
    assert dpos.max() <= 1.0
    assert dvar.max() <= 1e-6
    assert np.abs(r.steps - o["steps"]).max() <= 2


def test_integrate_uneven_spans_both_directions(gpu, ephem, oracle, model):
    """Configuration-3 shape at reduced count: mixed population, log-uniform spans, half backward, 10% NG."""
    from assist_amd import synthetic
    rng = np.random.default_rng(3)
    n = 96
    states, params, is_neo = population(model, n, 3, n_var=0, ng_fraction=0.1)
    span = synthetic.log_uniform_spans(n, rng, 30.0, 3000.0)
    r, o = _integrate_both(ephem, oracle, 0.0, span, states, params)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    # orbits that pass close to a planet amplify round-off; compare those on a relative scale
    print("uneven: max |dr| main belt %.3e m, NEO %.3e m" % (dpos[~is_neo].max(), dpos[is_neo].max()))
    assert dpos[~is_neo].max() <= 1.0
    assert np.median(dpos[is_neo]) <= 1.0


@pytest.mark.parametrize("shape", ["comet", "nk0"])
def test_integrate_marsden_shapes(gpu, ephem, oracle, model, shape):
    """Main-belt orbits with strong non-gravitational forces of a non-default shape, 1000 days."""
    from assist_amd import batch
    n = 32
    states, params, _ = population(model, n, 16, n_var=6, ng_fraction=1.0, neo_fraction=0.0)
    params[:, 0] = np.random.default_rng(16).normal(size=(n, 3)) * 1e-10
    kw = MARSDEN_SHAPES[shape]
    r = batch.integrate(ephem, 0.0, 1000.0, states, params, extras=dict(forces=0x7F, **kw))
    o = oracle.integrate(0.0, 1000.0, states, params, opts=oracle.opts(forces=0x7F, **kw))
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    dvar = rel_err(r.states[:, 1:], o["states"][:, 1:])
    print("Marsden %s, 1000 d: max |dr| = %.3e m, var rel %.3e" % (shape, dpos.max(), dvar.max()))
    assert dpos.max() <= 1.0
    assert dvar.max() <= 1e-6


@pytest.mark.parametrize("mode,span,tol_m", [
    ("fixed", 100.0, 1e-4), ("fixed", 1000.0, 1e-3), ("fixed", 10000.0, 5e-3),
    ("adaptive", 100.0, 1e-2), ("adaptive", 1000.0, 5e-2), ("adaptive", 10000.0, 0.1)])
def test_integrate_round_trip(gpu, ephem, oracle, model, mode, span, tol_m):
    """Out and back (reference unit_tests/roundtrip_spk: fixed 10-day steps with epsilon = 0,
    and roundtrip_adaptive_spk), for a main-belt orbit like the reference's, with the reference's
    tolerances; the GPU round trip also matches the oracle's."""
    from assist_amd import batch, synthetic
    x = synthetic.config1_particle(model).reshape(1, 1, 6)
    kw = dict(epsilon=0.0, dt_initial=10.0) if mode == "fixed" else {}
    cfg = batch.BatchConfig.default(**kw)
    out = batch.integrate(ephem, 0.0, span, x, config=cfg)
    back = batch.integrate(ephem, span, 0.0, out.states, config=cfg)
    d = np.linalg.norm(back.states[0, 0, :3] - x[0, 0, :3]) * AU_M
    ocfg = dict(epsilon=0.0, dt=10.0) if mode == "fixed" else {}
    oo = oracle.integrate(0.0, span, x, cfg=oracle.cfg(**ocfg))
    ob = oracle.integrate(span, 0.0, oo["states"], cfg=oracle.cfg(**ocfg))
    do = np.linalg.norm(ob["states"][0, 0, :3] - x[0, 0, :3]) * AU_M
    dgo = np.linalg.norm(back.states[0, 0, :3] - ob["states"][0, 0, :3]) * AU_M
    print("%s round trip %g d: gpu %.3e m, oracle %.3e m, gpu-oracle %.3e m" % (mode, span, d, do, dgo))
    assert d < tol_m and do < tol_m
    assert dgo < tol_m


def test_integrate_edge_cases(gpu, ephem, model):
    from assist_amd import batch
    states, _, _ = population(model, 4, 5, n_var=0)
    r = batch.integrate(ephem, np.array([0.0, 0.0, 0.0, 9990.0]), np.array([0.0, 10.0, -10.0, 10020.0]), states)
    assert r.status[0] == 0 and r.steps[0] == 0
    assert np.array_equal(r.states[0], states[0])
    assert r.status[1] == 0 and r.status[2] == 0
    assert r.status[3] == 5  # leaves the ephemeris coverage
    empty = batch.integrate(ephem, np.zeros(0), np.zeros(0), np.zeros((0, 1, 6)))
    assert empty.states.shape == (0, 1, 6)


def test_variational_matches_shadow_particle(gpu, ephem, model):
    """Variational particle vs a shadow particle displaced by 10 m (reference unit_tests/variational_spk)."""
    from assist_amd import batch, synthetic
    x = synthetic.config1_particle(model)
    dx = 10.0 / AU_M
    states = np.zeros((2, 2, 6))
    states[:, 0] = x
    states[1, 0, 0] += dx
    states[:, 1, 0] = 1.0
    for span in (1.0, 101.0):
        r = batch.integrate(ephem, 0.0, span, states, config=batch.BatchConfig.default(lanes_per_system=2))
        diff1 = r.states[1, 0, 0] - r.states[0, 0, 0]
        diff2 = r.states[0, 1, 0] * dx
        rel = abs((diff1 - diff2) / diff1)
        print("span %g: relative difference %.3e" % (span, rel))
        assert rel < 1e-5


def test_simulation_dropin_matches_oracle(gpu, ephem, oracle, model):
    """Reference-style use: Simulation + Extras, integrate, compare with the oracle."""
    import assist_amd
    from assist_amd import synthetic
    x = synthetic.config1_particle(model)
    sim = assist_amd.Simulation()
    ex = assist_amd.Extras(sim, ephem)
    sim.t = 0.0
    sim.add(x=x[0], y=x[1], z=x[2], vx=x[3], vy=x[4], vz=x[5])
    sim.integrate(1000.0)
    o = oracle.integrate(0.0, 1000.0, x.reshape(1, 1, 6))
    p = sim.particles[0]
    d = np.linalg.norm(np.array([p.x, p.y, p.z]) - o["states"][0, 0, :3]) * AU_M
    print("config 1 (1000 d): |dr| = %.3e m, t = %g" % (d, sim.t))
    assert sim.t == 1000.0
    assert d <= 1.0
    del ex


@pytest.mark.parametrize("lanes,block,waves", [(8, 256, 2), (8, 128, 1), (8, 128, 2), (8, 64, 1)])
def test_integrate_launch_shapes_agree(gpu, ephem, model, lanes, block, waves):
    """Every launch shape the autotuner may pick (block size, 1- or 2-wave kernel variant) runs
    the same arithmetic per system: results equal the default shape's bit for bit."""
    from assist_amd import batch
    states, _, _ = population(model, 300, seed=41, n_var=6)
    t0, t1 = np.zeros(300), np.linspace(-400.0, 400.0, 300)
    ref = batch.integrate(ephem, t0, t1, states, config=batch.BatchConfig.default(min_dt=0.001))
    cfg = batch.BatchConfig.default(min_dt=0.001, lanes_per_system=lanes, block_size=block, waves_per_simd=waves)
    out = batch.integrate(ephem, t0, t1, states, config=cfg)
    assert out.stats.waves_per_simd == waves and out.stats.block_size == block
    assert (out.status == ref.status).all() and (out.steps == ref.steps).all()
    assert np.array_equal(out.states, ref.states)


def test_integrate_is_deterministic(gpu, ephem, model):
    """Two launches of the same batch (the work queue hands systems to groups in a different
    order each time) give bit-identical states, step and evaluation counts."""
    from assist_amd import batch
    n = 6000
    states, _, _ = population(model, n, seed=43, n_var=6)
    t0, t1 = np.zeros(n), np.full(n, 200.0)
    cfg = batch.BatchConfig.default(min_dt=0.001)
    d_t0, d_t1 = batch.DeviceArray(t0), batch.DeviceArray(t1)
    outs = []
    for _ in range(2):
        d_st = batch.DeviceArray(states)
        d_status = batch.DeviceArray(np.zeros(n, np.int32))
        d_steps = batch.DeviceArray(np.zeros(n, np.int64))
        d_evals = batch.DeviceArray(np.zeros(n, np.int64))
        batch.integrate_device(ephem, n, 6, d_t0, d_t1, d_st, None, d_status, d_steps, d_evals, config=cfg)
        outs.append((d_st.download(), d_steps.download(), d_evals.download()))
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])
    assert np.array_equal(outs[0][0], outs[1][0])


@pytest.mark.parametrize("n_var,ng,precision", [(6, 0.0, 0), (0, 0.3, 0), (6, 0.3, 1)])
def test_results_do_not_depend_on_wave_mates(gpu, ephem, model, n_var, ng, precision):
    """8-lane groups whose queue ran dry stay in their wave as helpers and take over Chebyshev
    items of the groups still integrating.  A system integrated alone (seven helper groups in
    its wave from the start) or in a small group of systems ends bit for bit where it ends
    inside a full uneven batch (few helpers until the tail)."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(47)
    n = 4000
    states, params, _ = population(model, n, 47, n_var=n_var, ng_fraction=ng)
    span = synthetic.log_uniform_spans(n, rng, 5.0, 600.0)
    t0 = np.zeros(n)
    cfg = batch.BatchConfig.default(min_dt=0.001, lanes_per_system=8, precision=precision)
    full = batch.integrate(ephem, t0, span, states, params, config=cfg)
    assert (full.status == 0).all()
    for idx in ([int(np.argmax(span))], [0], [3, 7, 11]):
        sub = batch.integrate(ephem, t0[idx], span[idx], states[idx], None if params is None else params[idx], config=cfg)
        assert np.array_equal(sub.states, full.states[idx])
        assert np.array_equal(sub.steps, full.steps[idx])


@pytest.mark.parametrize("lanes,block", [(1, 256), (1, 128)])
def test_integrate_one_lane_per_system(gpu, ephem, oracle, model, lanes, block):
    """Real-only systems on one lane each (the 1-lane kernel evaluates the ephemeris slots one at
    a time): 64 main-belt systems, 3650 days, within 1 m of the oracle; and the same steps as the
    8-lane shape up to round-off in the group sums."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(51)
    n = 64
    states = np.zeros((n, 1, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    cfg = batch.BatchConfig.default(lanes_per_system=lanes, block_size=block)
    r = batch.integrate(ephem, 0.0, 3650.0, states, config=cfg)
    assert r.stats.lanes_per_system == lanes
    o = oracle.integrate(0.0, 3650.0, states)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("1 lane per system, 10 yr: max |dr| = %.3e m" % dpos.max())
    assert dpos.max() <= 1.0
    assert np.abs(r.steps - o["steps"]).max() <= 2


def test_uneven_batch_queue_order_keeps_results(gpu, ephem, model):
    """Batches of 16 384 systems or more hand out the longest spans first; every system's result
    is the one it gets in a small (unordered) batch, bit for bit."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(121)
    n = 20000
    states, _, _ = population(model, n, seed=121, n_var=6)
    span = synthetic.log_uniform_spans(n, rng, 1.0, 100.0)
    cfg = batch.BatchConfig.default(min_dt=0.001)
    big = batch.integrate(ephem, 0.0, span, states, config=cfg)
    for lo in (0, 10000):
        sl = slice(lo, lo + 10000)
        small = batch.integrate(ephem, 0.0, span[sl], states[sl], config=cfg)
        assert np.array_equal(small.states, big.states[sl])
        assert (small.steps == big.steps[sl]).all() and (small.status == big.status[sl]).all()


def test_corrector_tolerance_against_oracle(gpu, ephem, oracle, model):
    """The relaxed predictor-corrector stop (the mixed form's default: the real particle's
    max|db6| / max|a| below epsilon / 100) run in FP64 against the oracle with the same option:
    64 main-belt systems x (1 + 6 variational), 10 years, within 1 m and the same step counts to
    within 2; and its own effect, against IAS15's standard stop, well inside the bar."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(71)
    n = 64
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r = batch.integrate(ephem, 0.0, 3650.0, states, config=batch.BatchConfig.default(corrector_tolerance=1e-11))
    o = oracle.integrate(0.0, 3650.0, states, cfg=oracle.cfg(pc_tol=1e-11))
    std = batch.integrate(ephem, 0.0, 3650.0, states)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    d = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    e = np.linalg.norm(r.states[:, 0, :3] - std.states[:, 0, :3], axis=-1) * AU_M
    print("corrector tolerance 1e-11, 10 yr: vs oracle max %.3e m; vs the standard stop max %.3e m; "
          "evaluations %d vs %d (%.3fx)" % (d.max(), e.max(), r.stats.evals, std.stats.evals,
                                          std.stats.evals / r.stats.evals))
    assert d.max() <= 1.0
    assert np.abs(r.steps - o["steps"]).max() <= 2
    assert e.max() <= 0.1
    assert std.stats.evals / r.stats.evals > 1.3



@pytest.mark.parametrize("accel", [0.0, -1.0])
def test_first_pass_tolerance_against_oracle(gpu, ephem, oracle, model, accel):
    """The mixed form's first-pass stop (an attempt ends after its first predictor-corrector pass
    when that pass changed the real particle's b6 by less than 3 epsilon of its acceleration, times
    max(0.1, 6e-5 au/day^2 / max|a|) where max|a| exceeds 6e-5; accel = -1: no scale) run in FP64
    against the oracle with the same
    options, on the configuration-4 mix of main-belt and NEO-like orbits x (1 + 6 variational),
    10 years: within 1 m and the same step counts to within 2; and its own effect against IAS15's
    standard stop."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(72)
    n = 64
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.mixed_population(model, n, rng, 0.3)[0]
    states[:, 1:] = synthetic.variational_identity(n, 6)
    span = 3652.5
    cfg = batch.BatchConfig.default(corrector_tolerance=1e-11, first_pass_tolerance=3e-9, first_pass_accel=accel,
                                    min_dt=0.001)
    r = batch.integrate(ephem, 0.0, span, states, config=cfg)
    o = oracle.integrate(0.0, span, states, cfg=oracle.cfg(pc_tol=1e-11, pc_tol1=3e-9, pc_aref=6e-5 if accel == 0 else 0.0,
                                                           min_dt=0.001))
    std = batch.integrate(ephem, 0.0, span, states, config=batch.BatchConfig.default(min_dt=0.001))
    tol = batch.integrate(ephem, 0.0, span, states, config=batch.BatchConfig.default(corrector_tolerance=1e-11,
                                                                                       min_dt=0.001))
    assert (r.status == 0).all() and (o["status"] == 0).all()
    d = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    e = np.linalg.norm(r.states[:, 0, :3] - std.states[:, 0, :3], axis=-1) * AU_M
    dv = np.abs(r.states[:, 1:] - o["states"][:, 1:]).max() / np.abs(o["states"][:, 1:]).max()
    print("first-pass tolerance 3e-9, accel %g, 10 yr: vs oracle max %.3e m (variational %.1e); vs the standard "
          "stop max %.3e m; evaluations %d vs %d standard (%.3fx; oracle %d), %d with the corrector tolerance alone"
          % (accel, d.max(), dv, e.max(), r.stats.evals, std.stats.evals, std.stats.evals / r.stats.evals,
             o["evals"].sum(), tol.stats.evals))
    assert d.max() <= 1.0
    assert dv <= 1e-6
    assert np.abs(r.steps - o["steps"]).max() <= 2
    assert abs(r.stats.evals / o["evals"].sum() - 1) < 0.01
    assert e.max() <= 1.0
    assert r.stats.evals < tol.stats.evals
