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
