"""GPU parity with JPL binary (.440/.441) planets (reference src/ascii_ephem.c): the planet
columns are re-laid into the device's record format at first use, so the force and
integration kernels are the SPK ones; these tests hold them to the same bars against the
oracle reading the same file (oracle/ascii.c):

  * FP64 force evaluation: relative error <= 1e-12 per vector (full model, 6 variational);
  * FP64 integration: position within 1 m of the oracle after 10 years (main belt);
  * coverage status at and beyond the file's ends.
"""
import numpy as np
import pytest

from conftest import AU_M, population, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jpl_paths():
    from assist_amd import synthetic
    _, ap = synthetic.ensure_synthetic_ephemeris(0)
    return synthetic.ensure_synthetic_ascii(0), ap


@pytest.fixture(scope="module")
def jpl_oracle(jpl_paths):
    from oracle_lib import Oracle
    return Oracle(*jpl_paths)


@pytest.fixture(scope="module")
def jpl_ephem(jpl_paths):
    import assist_amd
    return assist_amd.Ephem(*jpl_paths)


def test_forces_full_model_jpl_binary(gpu, jpl_ephem, jpl_oracle, model):
    from assist_amd import batch
    rng = np.random.default_rng(61)
    n = 512
    states, params, _ = population(model, n, 61, n_var=6, ng_fraction=0.5)
    states[:, 1:] = rng.normal(size=states[:, 1:].shape)
    params[:, 6] = rng.normal(size=(n, 3)) * 1e-12
    # random times, plus block and sub-interval boundaries (multiples of 4 days from the start)
    t = rng.uniform(-9990.0, 9990.0, n)
    t[:64] = -10000.0 + 4.0 * rng.integers(1, 4999, 64)
    g, gs = batch.forces(jpl_ephem, t, states, params, extras=dict(forces=0x7F))
    o, os_ = jpl_oracle.forces(t, states, params, jpl_oracle.opts(forces=0x7F))
    assert (gs == 0).all() and (os_ == 0).all()
    e_real, e_var = rel_err(g[:, 0], o[:, 0]), rel_err(g[:, 1:], o[:, 1:])
    print("JPL binary planets: max rel err real %.3e var %.3e" % (e_real.max(), e_var.max()))
    assert e_real.max() <= 1e-12
    assert e_var.max() <= 1e-12


def test_forces_coverage_jpl_binary(gpu, jpl_ephem, jpl_oracle, model):
    from assist_amd import batch
    states, _, _ = population(model, 5, 62, n_var=0)
    t = np.array([0.0, 10000.5, -10000.5, 10000.0, -10000.0])
    g, st = batch.forces(jpl_ephem, t, states)
    o, ost = jpl_oracle.forces(t, states)
    assert list(st) == [0, 5, 5, 0, 0] and list(ost) == [0, 5, 5, 0, 0]
    ok = st == 0
    assert rel_err(g[ok, 0], o[ok, 0]).max() <= 1e-12


def test_integrate_main_belt_10yr_jpl_binary(gpu, jpl_ephem, jpl_oracle, model):
    """64 main-belt systems x (1 + 6 variational), 3650 days, against the oracle."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(63)
    n = 64
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r = batch.integrate(jpl_ephem, 0.0, 3650.0, states)
    o = jpl_oracle.integrate(0.0, 3650.0, states)
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    dvar = rel_err(r.states[:, 1:], o["states"][:, 1:])
    print("JPL binary 10 yr: max |dr| = %.3e m, var rel %.3e" % (dpos.max(), dvar.max()))
    assert dpos.max() <= 1.0
    assert dvar.max() <= 1e-6


def test_simulation_dropin_jpl_binary(gpu, jpl_ephem, jpl_oracle, model):
    """Reference-style use with a .440 planets file: Simulation + Extras, integrate."""
    import assist_amd
    from assist_amd import synthetic
    x = synthetic.config1_particle(model)
    sim = assist_amd.Simulation()
    ex = assist_amd.Extras(sim, jpl_ephem)
    sim.t = 0.0
    sim.add(x=x[0], y=x[1], z=x[2], vx=x[3], vy=x[4], vz=x[5])
    sim.integrate(1000.0)
    o = jpl_oracle.integrate(0.0, 1000.0, x.reshape(1, 1, 6))
    p = sim.particles[0]
    d = np.linalg.norm(np.array([p.x, p.y, p.z]) - o["states"][0, 0, :3]) * AU_M
    print("config 1 with JPL binary planets (1000 d): |dr| = %.3e m" % d)
    assert sim.t == 1000.0 and d <= 1.0
    del ex
