"""Pins the oracle's force model (reference src/forces.c) term by term against
independent statements of the physics:

* point masses: Newton's law summed in numpy;
* Earth J2, J3, J4 and solar J2: minus the gradient of the zonal potential
  GM sum_n J_n R^n P_n(sin(lat)) / r^(n+1) in the body's pole frame, by
  central differences;
* EIH: the Einstein-Infeld-Hoffmann 1PN acceleration (Newhall, Standish &
  Williams 1983, beta = gamma = 1) with the Sun as the only source;
* simple GR: the 1PN Schwarzschild test-particle acceleration; potential GR:
  -6 GM^2 r / (c^2 r^4);
* Marsden forces with the default g(r) = 1/r^2 and with a comet-like shape
  g(r) = alpha (r/r0)^-nm (1 + (r/r0)^nn)^-nk (Marsden, Sekanina & Yeomans 1973);
* every term's variational partials: central finite differences of the real
  acceleration (positions, velocities and the Marsden A's);
* reference quirks: the variational point-mass term ignores the force flags
  (reference src/forces.c:359) and all-zero A's skip the Marsden term
  (src/forces.c:861-862).
"""
import os

import numpy as np
import pytest

from assist_amd import synthetic

TERMS = {"SUN": 0x01, "PLANETS": 0x02, "ASTEROIDS": 0x04, "NON_GRAVITATIONAL": 0x08, "EARTH_HARMONICS": 0x10,
         "SUN_HARMONICS": 0x20, "GR_EIH": 0x40, "GR_SIMPLE": 0x80, "GR_POTENTIAL": 0x100}


def _states(model, n, seed, close_to_earth=False):
    rng = np.random.default_rng(seed)
    st, _ = synthetic.mixed_population(model, n, rng, 0.5, 0.0)
    t = rng.uniform(-5000, 5000, n)
    return t, st


def _acc(oracle, t, x, forces, params=None, **opts):
    acc, st = oracle.forces(t, x[:, None, :], None if params is None else params[:, None, :], oracle.opts(forces=forces, **opts))
    assert (st == 0).all()
    return acc[:, 0]


def _body(oracle, b, t):
    f, o = oracle.all_ephem(b, t)
    assert f == 0
    return o


def test_point_masses_are_newton(oracle, model):
    t, x = _states(model, 64, 1)
    a = _acc(oracle, t, x, 0x07)
    for i in range(len(t)):
        ref = np.zeros(3)
        for b in range(27):
            o = _body(oracle, b, t[i])
            d = x[i, :3] - o[1:4]
            ref -= o[0] * d / np.linalg.norm(d) ** 3
        assert np.linalg.norm(a[i] - ref) <= 1e-14 * np.linalg.norm(ref)


@pytest.mark.parametrize("flag,body", [(0x01, [0]), (0x02, list(range(1, 11))), (0x04, list(range(11, 27)))])
def test_force_flags_select_bodies(oracle, model, flag, body):
    t, x = _states(model, 16, 2)
    a = _acc(oracle, t, x, flag)
    for i in range(len(t)):
        ref = np.zeros(3)
        for b in body:
            o = _body(oracle, b, t[i])
            d = x[i, :3] - o[1:4]
            ref -= o[0] * d / np.linalg.norm(d) ** 3
        assert np.linalg.norm(a[i] - ref) <= 1e-14 * np.linalg.norm(ref)


def _legendre(n, s):
    return {2: (3 * s**2 - 1) / 2, 3: (5 * s**3 - 3 * s) / 2, 4: (35 * s**4 - 30 * s**2 + 3) / 8}[n]


def _zonal_acc(gm, R, J, rot, d, h=1e-7):
    """-grad of GM sum_n J_n R^n P_n(z'/r) / r^(n+1), z' along the pole (rows of rot)."""
    def f(p):
        q = rot @ p
        r = np.linalg.norm(q)
        return gm * sum(Jn * R**n * _legendre(n, q[2] / r) / r ** (n + 1) for n, Jn in J.items())
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h * np.linalg.norm(d)
        g[k] = (f(d + e) - f(d - e)) / (2 * e[k])
    return -g


def _rot(ra_deg, dec_deg):
    a, dlt = np.radians(ra_deg), np.radians(dec_deg)
    ca, sa, cd, sd = np.cos(a), np.sin(a), np.cos(dlt), np.sin(dlt)
    return np.array([[-sa, ca, 0.0], [-ca * sd, -sa * sd, cd], [ca * cd, sa * cd, sd]])


@pytest.fixture(scope="module")
def zonal_ephems(tmp_path_factory):
    """Ephemerides with only one of J2E, J3E, J4E non-zero."""
    from oracle_lib import Oracle
    d = tmp_path_factory.mktemp("zonal")
    out = {}
    for n in (2, 3, 4):
        c = {"J2E": 0.0, "J3E": 0.0, "J4E": 0.0}
        c["J%dE" % n] = synthetic.CONSTANTS["J%dE" % n]
        pp, ap = str(d / ("p%d.bsp" % n)), str(d / ("a%d.bsp" % n))
        synthetic.write_synthetic_ephemeris(pp, ap, 0, constants=c)
        out[n] = Oracle(pp, ap)
    return out


@pytest.mark.parametrize("n", [2, 3, 4])
def test_earth_zonal_harmonics_are_potential_gradients(zonal_ephems, model, n):
    """Each of J2, J3, J4 (reference src/forces.c:516-539) against the potential gradient; the
    particles sit 0.001-0.05 au from the Earth so the terms are well resolved."""
    o = zonal_ephems[n]
    rng = np.random.default_rng(n)
    for k in range(12):
        t = rng.uniform(-3000, 3000)
        e = _body(o, 3, t)
        d = rng.normal(size=3)
        d *= rng.uniform(1e-3, 5e-2) / np.linalg.norm(d)
        x = np.concatenate([e[1:4] + d, e[4:7]])[None]
        a = _acc(o, np.array([t]), x, 0x10)[0]
        ref = _zonal_acc(e[0], o.ephem.Re_eq, {n: getattr(o.ephem, "J%dE" % n)}, _rot(0.0, 90.0), d)
        assert np.linalg.norm(a - ref) <= 1e-6 * np.linalg.norm(ref), (n, a, ref)


def test_solar_j2_is_potential_gradient(oracle, model):
    rng = np.random.default_rng(5)
    for k in range(12):
        t = rng.uniform(-3000, 3000)
        s = _body(oracle, 0, t)
        d = rng.normal(size=3)
        d *= rng.uniform(0.2, 3.0) / np.linalg.norm(d)
        x = np.concatenate([s[1:4] + d, [0.01, 0.0, 0.0]])[None]
        a = _acc(oracle, np.array([t]), x, 0x20)[0]
        ref = _zonal_acc(s[0], oracle.ephem.Rs_eq, {2: oracle.ephem.J2SUN}, _rot(286.13, 63.87), d)
        assert np.linalg.norm(a - ref) <= 1e-6 * np.linalg.norm(ref)


def test_eih_matches_textbook_1pn(oracle, model):
    """EIH with the Sun as source (gr_eih_sources = 1): Newhall, Standish & Williams (1983)."""
    t, x = _states(model, 24, 6)
    a = _acc(oracle, t, x, 0x40)
    c2 = oracle.ephem.c_squared
    for i in range(len(t)):
        bodies = [_body(oracle, b, t[i]) for b in range(11)]
        sun = bodies[0]
        ri, vi = x[i, :3], x[i, 3:]
        rj, vj, gmj = sun[1:4], sun[4:7], sun[0]
        aj = np.zeros(3)
        for k in range(1, 11):
            dk = rj - bodies[k][1:4]
            aj -= bodies[k][0] * dk / np.linalg.norm(dk) ** 3
        rij = ri - rj
        r = np.linalg.norm(rij)
        pot_i = sum(b[0] / np.linalg.norm(ri - b[1:4]) for b in bodies)
        pot_j = sum(b[0] / np.linalg.norm(rj - b[1:4]) for b in bodies[1:])
        bracket = (-4 * pot_i - pot_j + vi @ vi + 2 * vj @ vj - 4 * vi @ vj - 1.5 * (rij @ vj / r) ** 2 + 0.5 * (-rij) @ aj) / c2
        ref = -gmj * rij / r**3 * bracket + gmj / r**3 * (rij @ (4 * vi - 3 * vj)) * (vi - vj) / c2 + 3.5 * gmj * aj / r / c2
        assert np.linalg.norm(a[i] - ref) <= 1e-12 * np.linalg.norm(ref)


def test_gr_simple_and_potential_closed_forms(oracle, model):
    t, x = _states(model, 16, 7)
    a_s = _acc(oracle, t, x, 0x80)
    a_p = _acc(oracle, t, x, 0x100)
    c2 = oracle.ephem.c_squared
    for i in range(len(t)):
        s = _body(oracle, 0, t[i])
        r, v, gm = x[i, :3] - s[1:4], x[i, 3:] - s[4:7], s[0]
        rn = np.linalg.norm(r)
        ref_s = gm / (c2 * rn**3) * ((4 * gm / rn - v @ v) * r + 4 * (r @ v) * v)
        ref_p = -6 * gm**2 * r / (c2 * rn**4)
        assert np.linalg.norm(a_s[i] - ref_s) <= 1e-13 * np.linalg.norm(ref_s)
        assert np.linalg.norm(a_p[i] - ref_p) <= 1e-13 * np.linalg.norm(ref_p)


def test_marsden_default_shape(oracle, model):
    t, x = _states(model, 16, 8)
    rng = np.random.default_rng(8)
    A = rng.normal(size=(16, 3)) * 1e-12
    a = _acc(oracle, t, x, 0x08, A)
    for i in range(len(t)):
        s = _body(oracle, 0, t[i])
        r, v = x[i, :3] - s[1:4], x[i, 3:] - s[4:7]
        rn = np.linalg.norm(r)
        h = np.cross(r, v)
        tt = np.cross(h, r)
        ref = (A[i, 0] * r / rn + A[i, 1] * tt / np.linalg.norm(tt) + A[i, 2] * h / np.linalg.norm(h)) / rn**2
        assert np.linalg.norm(a[i] - ref) <= 1e-13 * np.linalg.norm(ref)
    a0 = _acc(oracle, t, x, 0x08, np.zeros((16, 3)))
    assert (a0 == 0).all()


def _jacobian_cols(oracle, t, x, forces, params=None, **opts):
    """Variational accelerations for unit variations of x, y, z, vx, vy, vz (and A1..A3)."""
    n = len(t)
    nv = 6 if params is None else 9
    st = np.zeros((n, 1 + nv, 6))
    st[:, 0] = x
    for k in range(6):
        st[:, 1 + k, k] = 1.0
    pp = None
    if params is not None:
        pp = np.zeros((n, 1 + nv, 3))
        pp[:, 0] = params
        for k in range(3):
            pp[:, 7 + k, k] = 1.0
    acc, s = oracle.forces(t, st, pp, oracle.opts(forces=forces, **opts))
    assert (s == 0).all()
    return acc[:, 1:]


# water-ice sublimation law of Marsden, Sekanina & Yeomans (1973)
COMET_SHAPE = dict(alpha=0.1112620426, nk=4.6142, nm=2.15, nn=5.093, r0=2.808)


def test_marsden_comet_shape(oracle, model):
    t, x = _states(model, 12, 12)
    rng = np.random.default_rng(12)
    A = rng.normal(size=(12, 3)) * 1e-9
    a = _acc(oracle, t, x, 0x08, A, **COMET_SHAPE)
    c = COMET_SHAPE
    for i in range(len(t)):
        s = _body(oracle, 0, t[i])
        r, v = x[i, :3] - s[1:4], x[i, 3:] - s[4:7]
        rn = np.linalg.norm(r)
        q = rn / c["r0"]
        g = c["alpha"] * q ** -c["nm"] * (1 + q ** c["nn"]) ** -c["nk"]
        h = np.cross(r, v)
        tt = np.cross(h, r)
        ref = g * (A[i, 0] * r / rn + A[i, 1] * tt / np.linalg.norm(tt) + A[i, 2] * h / np.linalg.norm(h))
        assert np.linalg.norm(a[i] - ref) <= 1e-13 * np.linalg.norm(ref)
    # partials of the shaped term, including dg/dr's (1 + q^nn) factor, by central differences
    params = rng.normal(size=(4, 3)) * 1e-9
    t, x = t[:4], x[:4]
    cols = _jacobian_cols(oracle, t, x, 0x08, params, **COMET_SHAPE)
    for i in range(4):
        for k in range(6):
            h = (1e-6 if k < 3 else 1e-8) * (np.linalg.norm(x[i, :3]) if k < 3 else np.linalg.norm(x[i, 3:]))
            xp, xm = x[i].copy(), x[i].copy()
            xp[k] += h
            xm[k] -= h
            ap = _acc(oracle, np.array([t[i]] * 2), np.stack([xp, xm]), 0x08, np.stack([params[i]] * 2), **COMET_SHAPE)
            fd = (ap[0] - ap[1]) / (2 * h)
            assert np.abs(cols[i, k] - fd).max() <= 1e-6 * np.abs(cols[i]).max(), (i, k)


@pytest.mark.parametrize("term", ["DIRECT", "NON_GRAVITATIONAL", "EARTH_HARMONICS", "SUN_HARMONICS", "GR_EIH",
                                  "GR_SIMPLE", "GR_POTENTIAL"])
def test_variational_partials_are_finite_differences(oracle, model, term):
    forces = 0x07 if term == "DIRECT" else TERMS[term]
    t, x = _states(model, 6, 9)
    if term == "EARTH_HARMONICS":  # move the particles near the Earth so the term is resolved
        for i in range(len(t)):
            e = _body(oracle, 3, t[i])
            x[i, :3] = e[1:4] + np.array([0.004, -0.003, 0.002]) * (1 + i)
    params = np.random.default_rng(9).normal(size=(len(t), 3)) * 1e-11 if term == "NON_GRAVITATIONAL" else None
    cols = _jacobian_cols(oracle, t, x, forces, params)
    for i in range(len(t)):
        for k in range(6 + (3 if params is not None else 0)):
            if k < 6:
                h = (1e-6 if k < 3 else 1e-8) * (np.linalg.norm(x[i, :3]) if k < 3 else np.linalg.norm(x[i, 3:]))
                xp, xm = x[i].copy(), x[i].copy()
                xp[k] += h
                xm[k] -= h
                pp = None if params is None else np.stack([params[i], params[i]])
                ap = _acc(oracle, np.array([t[i]] * 2), np.stack([xp, xm]), forces, pp)
            else:
                h = 1e-3 * abs(params[i, k - 6])
                pp = np.stack([params[i], params[i]])
                pp[0, k - 6] += h
                pp[1, k - 6] -= h
                ap = _acc(oracle, np.array([t[i]] * 2), np.stack([x[i], x[i]]), forces, pp)
            fd = (ap[0] - ap[1]) / (2 * h)
            scale = np.abs(cols[i]).max()
            assert np.abs(cols[i, k] - fd).max() <= 1e-6 * scale, (term, i, k, cols[i, k], fd)


def test_variational_point_masses_ignore_flags(oracle, model):
    """The reference's variational point-mass pass uses every body whatever the flags
    (reference src/forces.c:352-359)."""
    t, x = _states(model, 8, 10)
    full = _jacobian_cols(oracle, t, x, 0x07)
    for flag in (0x01, 0x02, 0x04):
        assert np.array_equal(_jacobian_cols(oracle, t, x, flag), full)


def test_zero_marsden_parameters_skip_variations(oracle, model):
    """A1 = A2 = A3 = 0 skips the Marsden term for the particle and its variations
    (reference src/forces.c:861-862), even when the A-variations are non-zero."""
    t, x = _states(model, 4, 11)
    cols = _jacobian_cols(oracle, t, x, 0x08, np.zeros((4, 3)))
    assert (cols == 0).all()


# ---- extended model: several EIH sources, geocentric frame (reference src/forces.c:72-87,
# 157-172, 1288-1983) ----

def _eih_textbook(bodies, ri, vi, nsrc, c2):
    """EIH acceleration of a massless particle with sources j < nsrc (Newhall, Standish &
    Williams 1983, beta = gamma = 1): every source's bracket uses the particle's potential from
    all 11 bodies and the source's from the others; the source's Newtonian acceleration comes
    from the other 10 bodies."""
    pot_i = sum(b[0] / np.linalg.norm(ri - b[1:4]) for b in bodies)
    a = np.zeros(3)
    t78 = np.zeros(3)
    for j in range(nsrc):
        rj, vj, gmj = bodies[j][1:4], bodies[j][4:7], bodies[j][0]
        aj = np.zeros(3)
        pot_j = 0.0
        for k in range(11):
            if k == j:
                continue
            dk = rj - bodies[k][1:4]
            aj -= bodies[k][0] * dk / np.linalg.norm(dk) ** 3
            pot_j += bodies[k][0] / np.linalg.norm(dk)
        rij = ri - rj
        r = np.linalg.norm(rij)
        bracket = (-4 * pot_i - pot_j + vi @ vi + 2 * vj @ vj - 4 * vi @ vj - 1.5 * (rij @ vj / r) ** 2
                   + 0.5 * (-rij) @ aj) / c2
        a += -gmj * rij / r**3 * bracket
        t78 += gmj / r**3 * (rij @ (4 * vi - 3 * vj)) * (vi - vj) / c2 + 3.5 * gmj * aj / r / c2
    return a + t78


@pytest.mark.parametrize("nsrc", [2, 5, 11])
def test_eih_several_sources_match_textbook(oracle, model, nsrc):
    t, x = _states(model, 12, 20 + nsrc)
    a = _acc(oracle, t, x, 0x40, gr_eih_sources=nsrc)
    c2 = oracle.ephem.c_squared
    for i in range(len(t)):
        bodies = [_body(oracle, b, t[i]) for b in range(11)]
        ref = _eih_textbook(bodies, x[i, :3], x[i, 3:], nsrc, c2)
        assert np.linalg.norm(a[i] - ref) <= 1e-12 * np.linalg.norm(ref), (nsrc, i)
    # the sources add: the Sun alone is the first term of the sum
    a1 = _acc(oracle, t, x, 0x40, gr_eih_sources=1)
    assert np.abs(a - a1).max() > 0


def _near_earth(oracle, model, n, seed):
    """Geocentric states 0.002-0.05 au from the Earth moving at heliocentric-orbit speeds."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(-3000, 3000, n)
    x = np.zeros((n, 6))
    for i in range(n):
        d = rng.normal(size=3)
        x[i, :3] = d / np.linalg.norm(d) * rng.uniform(0.002, 0.05)
        x[i, 3:] = rng.normal(size=3) * 0.003
    return t, x


def test_geocentric_frame_is_the_shifted_barycentric_one(oracle, model):
    """For the position-dependent terms and the Marsden and simple-GR terms (which use velocities
    relative to the Sun) the geocentric acceleration plus the Earth's equals the barycentric
    acceleration at the shifted state (the reference's frame offset and indirect term)."""
    t, x = _near_earth(oracle, model, 16, 31)
    A = np.random.default_rng(32).normal(size=(16, 3)) * 1e-12
    for forces in (0x07, 0x10, 0x20, 0x08, 0x80, 0x100, 0x3F):
        ag = _acc(oracle, t, x, forces, A if forces & 0x08 else None, geocentric=1)
        xb = x.copy()
        ae = np.zeros((len(t), 3))
        for i in range(len(t)):
            e = _body(oracle, 3, t[i])
            xb[i, :3] += e[1:4]
            xb[i, 3:] += e[4:7]
            ae[i] = e[7:10]
        ab = _acc(oracle, t, xb, forces, A if forces & 0x08 else None)
        # the indirect term (the Sun's pull on the Earth, ~3e-4 au/d^2) is subtracted and added
        # back, so its rounding sets the scale for the small terms
        err = np.linalg.norm(ag + ae - ab, axis=-1) / (np.linalg.norm(ab, axis=-1) + np.linalg.norm(ae, axis=-1))
        assert err.max() <= 1e-13, (forces, err.max())


@pytest.mark.parametrize("case", ["eih11", "geocentric", "geocentric_eih11"])
def test_extended_partials_are_finite_differences(oracle, model, case):
    opts = dict(gr_eih_sources=11) if "eih11" in case else {}
    forces = 0x40 if case == "eih11" else 0x7F & ~0x08
    if "geocentric" in case:
        opts["geocentric"] = 1
        t, x = _near_earth(oracle, model, 4, 33)
    else:
        t, x = _states(model, 4, 34)
    n = len(t)
    st = np.zeros((n, 7, 6))
    st[:, 0] = x
    for k in range(6):
        st[:, 1 + k, k] = 1.0
    acc, s = oracle.forces(t, st, None, oracle.opts(forces=forces, **opts))
    assert (s == 0).all()
    cols = acc[:, 1:]
    for i in range(n):
        for k in range(6):
            h = (1e-6 if k < 3 else 1e-8) * (np.linalg.norm(x[i, :3]) if k < 3 else np.linalg.norm(x[i, 3:]))
            xp, xm = x[i].copy(), x[i].copy()
            xp[k] += h
            xm[k] -= h
            ap = _acc(oracle, np.array([t[i]] * 2), np.stack([xp, xm]), forces, **opts)
            fd = (ap[0] - ap[1]) / (2 * h)
            scale = np.abs(cols[i]).max()
            assert np.abs(cols[i, k] - fd).max() <= 1e-6 * scale, (case, i, k, cols[i, k], fd)
