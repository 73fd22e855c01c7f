"""GPU tests of the Chebyshev series length: slots whose records all have at most 11 coefficients
(the synthetic asteroid file's) run a short form of the series, the others the full 14-term form
of the padded record (assist_amd/csrc/force_model.h, cheb_series and wave_positions).

An asteroid file with 13-coefficient records takes the full form for the asteroid slots on one
lane per system and in the wave-wide rounds of 8-lane groups; both shapes are held to the oracle
(which evaluates each record's own coefficient count, reference src/spk.c:514-530) at the 1 m bar
of the other 10-year parity tests, in FP64 and in the mixed form.
"""
import numpy as np
import pytest

from conftest import AU_M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def long_asteroids(tmp_path_factory):
    """(Ephem, Oracle, model) of the synthetic files with 16-day, 13-coefficient asteroid records."""
    import assist_amd
    from assist_amd import synthetic
    from oracle_lib import Oracle
    d = tmp_path_factory.mktemp("ast13")
    pp, ap = str(d / "p.bsp"), str(d / "a13.bsp")
    model = synthetic.write_synthetic_ephemeris(pp, ap, 0, asteroid_record=(16.0, 13))
    return assist_amd.Ephem(pp, ap), Oracle(pp, ap), model


@pytest.mark.parametrize("lanes", [1, 8])
@pytest.mark.parametrize("precision", [0, 1])
def test_full_series_for_long_asteroid_records(gpu, long_asteroids, lanes, precision):
    from assist_amd import batch, synthetic
    ephem, oracle, model = long_asteroids
    rng = np.random.default_rng(71)
    n = 64
    states = np.zeros((n, 1, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    cfg = batch.BatchConfig.default(lanes_per_system=lanes, precision=precision)
    r = batch.integrate(ephem, 0.0, 3650.0, states, config=cfg)
    assert r.stats.lanes_per_system == lanes
    o = oracle.integrate(0.0, 3650.0, states)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("13-coefficient asteroid records, %d lane(s), precision %d, 10 yr: max |dr| = %.3e m"
          % (lanes, precision, dpos.max()))
    assert dpos.max() <= 1.0


@pytest.fixture(scope="module")
def planets_only():
    """(Ephem, Oracle) of the synthetic planets file alone (no asteroid file)."""
    import assist_amd
    from assist_amd import synthetic
    from oracle_lib import Oracle
    pp, _ = synthetic.ensure_synthetic_ephemeris(0)
    return assist_amd.Ephem(pp), Oracle(pp, None)


@pytest.mark.parametrize("lanes", [1, 8])
def test_missing_asteroid_slots(gpu, planets_only, model, lanes):
    """Without an asteroid file the 16 asteroid slots point at the Sun's 14-coefficient records and
    no term uses their positions, so they run the short series; the results are the oracle's
    (every force but the asteroids', 10 years, 1 m)."""
    from assist_amd import batch, synthetic
    ephem, oracle = planets_only
    rng = np.random.default_rng(72)
    n = 64
    states = np.zeros((n, 1, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    forces = 0x7F & ~0x04
    cfg = batch.BatchConfig.default(lanes_per_system=lanes)
    r = batch.integrate(ephem, 0.0, 3650.0, states, extras={"forces": forces}, config=cfg)
    o = oracle.integrate(0.0, 3650.0, states, opts=oracle.opts(forces=forces))
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("no asteroid file, %d lane(s), 10 yr: max |dr| = %.3e m" % (lanes, dpos.max()))
    assert dpos.max() <= 1.0
