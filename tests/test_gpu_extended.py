"""GPU parity of the extended force model against the oracle: several EIH sources
(gr_eih_sources up to 11, reference src/forces.c:1288-1983 with the source loop) and the
geocentric frame (reference src/forces.c:72-87 offsets, 157-172 indirect term).  These run in
FP64 in kernels of their own (the default kernels are unchanged); the bars are the FP64 ones:

  * force evaluation: relative error <= 1e-12 per vector, with 6 variational vectors;
  * integration: position within 1 m of the oracle.
"""
import numpy as np
import pytest

from conftest import AU_M, population, rel_err

pytestmark = pytest.mark.gpu


def _near_earth(oracle, n, seed, n_var=6):
    """Geocentric states 0.002-0.05 au from the Earth with heliocentric-orbit speeds."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(-3000, 3000, n)
    st = np.zeros((n, 1 + n_var, 6))
    for i in range(n):
        d = rng.normal(size=3)
        st[i, 0, :3] = d / np.linalg.norm(d) * rng.uniform(0.002, 0.05)
        st[i, 0, 3:] = rng.normal(size=3) * 0.003
    if n_var:
        st[:, 1:] = rng.normal(size=st[:, 1:].shape)
    return t, st


def _both(ephem, oracle, t, states, params, **opts):
    from assist_amd import batch
    g, gs = batch.forces(ephem, t, states, params, extras=opts)
    o, os_ = oracle.forces(t, states, params, oracle.opts(**opts))
    assert (gs == 0).all() and (os_ == 0).all()
    return g, o


@pytest.mark.parametrize("nsrc", [2, 5, 11])
@pytest.mark.parametrize("forces", [0x40, 0x7F])
def test_forces_several_eih_sources(gpu, ephem, oracle, model, nsrc, forces):
    rng = np.random.default_rng(70 + nsrc)
    n = 256
    states, params, _ = population(model, n, 70 + nsrc, n_var=6, ng_fraction=0.5)
    states[:, 1:] = rng.normal(size=states[:, 1:].shape)
    params[:, 6] = rng.normal(size=(n, 3)) * 1e-12
    t = rng.uniform(-9000.0, 9000.0, n)
    g, o = _both(ephem, oracle, t, states, params, forces=forces, gr_eih_sources=nsrc)
    e_real, e_var = rel_err(g[:, 0], o[:, 0]), rel_err(g[:, 1:], o[:, 1:])
    print("EIH %d sources, forces 0x%X: max rel err real %.3e var %.3e" % (nsrc, forces, e_real.max(), e_var.max()))
    assert e_real.max() <= 1e-12 and e_var.max() <= 1e-12


@pytest.mark.parametrize("nsrc", [1, 11])
@pytest.mark.parametrize("n_var", [0, 6])
def test_forces_geocentric(gpu, ephem, oracle, nsrc, n_var):
    t, states = _near_earth(oracle, 256, 80 + nsrc + n_var, n_var=n_var)
    g, o = _both(ephem, oracle, t, states, None, forces=0x7F, gr_eih_sources=nsrc, geocentric=1)
    e_real = rel_err(g[:, 0], o[:, 0])
    e_var = rel_err(g[:, 1:], o[:, 1:]) if n_var else np.zeros(1)
    print("geocentric, %d EIH sources, %d var: max rel err real %.3e var %.3e" % (nsrc, n_var, e_real.max(), e_var.max()))
    assert e_real.max() <= 1e-12 and e_var.max() <= 1e-12


def test_integrate_main_belt_eih11(gpu, ephem, oracle, model):
    """32 main-belt systems x (1 + 6 variational), 3650 days, 11 EIH sources."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(90)
    n = 32
    states = np.zeros((n, 7, 6))
    states[:, 0] = synthetic.main_belt(model, n, rng)
    states[:, 1:] = synthetic.variational_identity(n, 6)
    r = batch.integrate(ephem, 0.0, 3650.0, states, extras=dict(gr_eih_sources=11))
    o = oracle.integrate(0.0, 3650.0, states, opts=oracle.opts(gr_eih_sources=11))
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("11 EIH sources, 10 yr: max |dr| = %.3e m" % dpos.max())
    assert dpos.max() <= 1.0
    assert rel_err(r.states[:, 1:], o["states"][:, 1:]).max() <= 1e-6


def test_integrate_geocentric(gpu, ephem, oracle, model):
    """Geocentric integration of 16 near-Earth objects for 60 days (1 + 6 variational), against
    the oracle, and against a barycentric integration of the same orbits."""
    from assist_amd import batch, synthetic
    rng = np.random.default_rng(91)
    n = 16
    bary, _ = synthetic.mixed_population(model, n, rng, 1.0, 0.0)
    e0 = oracle.all_ephem(3, 0.0)[1]
    states = np.zeros((n, 7, 6))
    states[:, 0] = bary - e0[1:7]
    states[:, 1:] = synthetic.variational_identity(n, 6)
    cfg = batch.BatchConfig.default(min_dt=0.001)
    r = batch.integrate(ephem, 0.0, 60.0, states, extras=dict(geocentric=1), config=cfg)
    o = oracle.integrate(0.0, 60.0, states, opts=oracle.opts(geocentric=1), cfg=oracle.cfg(min_dt=0.001))
    assert (r.status == 0).all() and (o["status"] == 0).all()
    dpos = np.linalg.norm(r.states[:, 0, :3] - o["states"][:, 0, :3], axis=-1) * AU_M
    print("geocentric 60 d: max |dr| to the oracle = %.3e m" % dpos.max())
    assert dpos.max() <= 1.0
    # the same orbits integrated barycentrically without the frame-dependent EIH term
    nog = dict(forces=0x3F)
    rg = batch.integrate(ephem, 0.0, 60.0, states, extras=dict(nog, geocentric=1), config=cfg)
    sb = states.copy()
    sb[:, 0] = bary
    rb = batch.integrate(ephem, 0.0, 60.0, sb, extras=nog, config=cfg)
    e1 = oracle.all_ephem(3, 60.0)[1]
    d = np.linalg.norm(rg.states[:, 0, :3] + e1[1:4] - rb.states[:, 0, :3], axis=-1) * AU_M
    print("geocentric vs barycentric frame, 60 d: max |dr| = %.3e m" % d.max())
    assert d.max() <= 10.0


def test_extended_model_rejects_mixed_precision(gpu, ephem, model):
    from assist_amd import batch
    states, _, _ = population(model, 4, 92, n_var=0)
    with pytest.raises(batch.BatchError):
        batch.forces(ephem, 0.0, states, extras=dict(geocentric=1), precision=1)
    with pytest.raises(batch.BatchError):
        batch.integrate(ephem, 0.0, 10.0, states, extras=dict(gr_eih_sources=3),
                        config=batch.BatchConfig.default(precision=batch.PRECISION_MIXED))
