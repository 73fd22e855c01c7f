import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

AU_M = 149597870700.0


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ephem_paths():
    from assist_amd import synthetic
    return synthetic.ensure_synthetic_ephemeris(0)


@pytest.fixture(scope="session")
def short_asteroid_paths(tmp_path_factory):
    """The seed-0 planets with an asteroid file covering only JD 2451545 +/- 3000 days."""
    from assist_amd import synthetic
    d = tmp_path_factory.mktemp("short_asteroids")
    pp, ap = str(d / "planets.bsp"), str(d / "asteroids.bsp")
    synthetic.write_synthetic_ephemeris(pp, ap, 0, asteroid_coverage_days=3000.0)
    return pp, ap


@pytest.fixture(scope="session")
def model():
    from assist_amd import synthetic
    return synthetic.SyntheticModel(0)


@pytest.fixture(scope="session")
def oracle(ephem_paths):
    from oracle_lib import Oracle
    return Oracle(*ephem_paths)


@pytest.fixture(scope="session")
def ephem(ephem_paths):
    import assist_amd
    return assist_amd.Ephem(*ephem_paths)


@pytest.fixture(scope="session")
def gpu():
    from assist_amd import batch
    if batch.gpu_count() < 1:
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")
    return True


def rel_err(a, b):
    """|a - b| / |b| per vector (last axis)."""
    num = np.linalg.norm(a - b, axis=-1)
    den = np.linalg.norm(b, axis=-1)
    return num / np.where(den > 0, den, 1.0)


def population(model, n, seed, n_var=6, t0=0.0, neo_fraction=0.3, ng_fraction=0.0):
    """Mixed main-belt / NEO states (n, 1+n_var, 6) with identity variations, and params."""
    from assist_amd import synthetic
    rng = np.random.default_rng(seed)
    st, is_neo = synthetic.mixed_population(model, n, rng, neo_fraction, t0)
    states = np.zeros((n, 1 + n_var, 6))
    states[:, 0] = st
    if n_var:
        states[:, 1:] = synthetic.variational_identity(n, n_var)
    params = None
    if ng_fraction > 0:
        params = np.zeros((n, 1 + n_var, 3))
        params[:, 0] = synthetic.marsden_params(n, rng, ng_fraction)
    return states, params, is_neo
