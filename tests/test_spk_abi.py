"""CPU tests of the SPK reader's C interface (include/spk.h, reference src/spk.h:15-122), through
ctypes, on the synthetic planet and asteroid files.  They restate the reference's own unit tests
of this interface (which need DE440 and sb441-n16):

* unit_tests/spk_init/problem.c:10-62 -- target counts, and code 0 in targets[] past `num`;
* unit_tests/spk_join_masses/problem.c -- every target but 199/299 gets a mass;
* unit_tests/spk_load_constants/problem.c -- every constant is found;
* unit_tests/spk_planets_calc/problem.c -- the Sun, Earth and Moon evaluate (finite, success);

and check the values against the oracle (bit for bit) and against the constants and masses the
synthetic files were written with.
"""
import ctypes as C
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def spk():
    from assist_amd import spk
    return spk


def _calc_planets(fn, ephem, jd_ref, jd_rel, code):
    v = [C.c_double() for _ in range(10)]
    f = fn(C.byref(ephem), jd_ref, jd_rel, code, *[C.byref(x) for x in v])
    return f, np.array([x.value for x in v])


def test_spk_init_targets(spk, ephem_paths):
    pp, ap = ephem_paths
    pl = spk.spk_init(pp)
    assert pl
    p = pl.contents
    # Sun, Mercury..Pluto barycentres, Earth and Moon (DE440 adds 199 and 299: 14 there)
    assert p.num == 12
    assert sorted(p.targets[i].code for i in range(p.num)) == [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 301, 399]
    assert p.allocated_num >= 16
    assert sum(1 for i in range(16) if p.targets[i].code != 0) == 12
    for i in range(p.num):
        t = p.targets[i]
        assert t.mass == 0.0  # until the masses are joined
        assert t.ind >= 0 and t.allocated_ind > t.ind
        assert t.end > t.beg and t.res > 0
        assert all(1 <= t.one[k] <= t.two[k] for k in range(t.ind + 1))
    assert p.len > 0 and p.map
    assert spk.spk_free(pl) == 0
    al = spk.spk_init(ap)
    assert al and al.contents.num == 16
    assert sum(1 for i in range(20) if al.contents.targets[i].code != 0) == 16
    assert spk.spk_free(al) == 0
    assert not spk.spk_init(pp + ".missing")
    assert spk.spk_free(None) == -1


def test_spk_load_constants_and_join_masses(spk, ephem_paths):
    import assist_amd
    from assist_amd import synthetic
    pp, ap = ephem_paths
    lib = assist_amd.clibassist
    data = spk.load_constants_and_masses(pp)
    for k in ("AU", "EMRAT", "J2E", "J3E", "J4E", "J2SUN", "RE", "CLIGHT", "ASUN"):
        assert getattr(data, k) != 0.0, k
        assert getattr(data, k) == pytest.approx(synthetic.CONSTANTS[k], rel=1e-15), k
    names = {data.masses.names[i].decode(): data.masses.values[i] for i in range(data.masses.count)}
    assert names["GMS"] == pytest.approx(synthetic.GM_PLANETS["GMS"], rel=1e-15)
    assert sum(1 for k in names if k.startswith("MA")) == 16
    ephem = assist_amd.Ephem.__new__(assist_amd.Ephem)
    C.memset(C.addressof(ephem), 0, C.sizeof(ephem))
    lib.assist_apply_spk_constants(C.byref(ephem), C.byref(data))
    assert ephem.AU == data.AU and ephem.CLIGHT == data.CLIGHT
    assert ephem.Re_eq == ephem.RE / ephem.AU and ephem.Rs_eq == ephem.ASUN / ephem.AU
    assert ephem.c_AU_per_day == (ephem.CLIGHT / ephem.AU) * 86400
    assert ephem.over_c_squared == 1.0 / (ephem.c_AU_per_day * ephem.c_AU_per_day)
    pl, al = spk.spk_init(pp), spk.spk_init(ap)
    lib.assist_spk_join_masses(pl, C.byref(data.masses), ephem.EMRAT)
    lib.assist_spk_join_masses(al, C.byref(data.masses), ephem.EMRAT)
    for f in (pl, al):
        for i in range(f.contents.num):
            t = f.contents.targets[i]
            assert t.mass != 0.0, t.code
    # GMB split by EMRAT between the Earth (399) and the Moon (301)
    gmb = names["GMB"]
    earth = lib.assist_spk_find_target(pl, 399).contents.mass
    moon = lib.assist_spk_find_target(pl, 301).contents.mass
    assert earth == gmb * (ephem.EMRAT / (1.0 + ephem.EMRAT)) and moon == gmb * (1.0 / (1.0 + ephem.EMRAT))
    # a mass already set is kept
    tgt = lib.assist_spk_find_target(al, 2000000 + 1)
    if not tgt:
        tgt = C.pointer(al.contents.targets[0])
    tgt.contents.mass = 42.0
    lib.assist_spk_join_masses(al, C.byref(data.masses), ephem.EMRAT)
    assert tgt.contents.mass == 42.0
    assert not lib.assist_spk_find_target(pl, 12345)
    lib.assist_free_spk_constants_and_masses(C.byref(data))
    assert data.masses.count == 0 and not data.masses.names
    spk.spk_free(pl)
    spk.spk_free(al)
    # an unreadable file gives all-zero constants
    bad = spk.load_constants_and_masses(pp + ".missing")
    assert bad.AU == 0.0 and bad.masses.count == 0


def test_spk_calc_planets_and_asteroids_match_the_oracle(spk, ephem, oracle, ephem_paths):
    """assist_spk_calc_planets(_by_assist), assist_spk_calc and assist_spk_target_pos against the
    oracle's evaluation of the same records, bit for bit (the Sun, Earth and Moon as
    unit_tests/spk_planets_calc/problem.c asks, and every other body)."""
    import assist_amd
    lib = assist_amd.clibassist
    jd_ref = ephem.jd_ref
    naif = [10, 1, 2, 399, 301, 4, 5, 6, 7, 8, 9]
    for t in (0.0, 123.25, -4321.5):
        for body, code in enumerate(naif):
            f, v = _calc_planets(lib.assist_spk_calc_planets, ephem, jd_ref, t, code)
            assert f == 0 and all(math.isfinite(x) for x in v)
            fb, vb = _calc_planets(lib.assist_spk_calc_planets_by_assist, ephem, jd_ref, t, body)
            assert fb == 0
            np.testing.assert_array_equal(v, vb)
            fo, o = oracle.all_ephem(body, t)
            assert fo == 0
            np.testing.assert_array_equal(v, o)
        for m in range(16):
            gm, x, y, z = C.c_double(), C.c_double(), C.c_double(), C.c_double()
            f = lib.assist_spk_calc(C.cast(ephem.spk_asteroids, C.POINTER(spk.SpkFile)),
                                    jd_ref, t, m, C.byref(gm), C.byref(x), C.byref(y), C.byref(z))
            assert f == 0
            # the oracle's asteroid query is barycentric: heliocentric + the Sun
            fo, o = oracle.all_ephem(11 + m, t)
            _, sun = oracle.all_ephem(0, t)
            assert gm.value == o[0]
            np.testing.assert_allclose([x.value + sun[1], y.value + sun[2], z.value + sun[3]], o[1:4], rtol=0,
                                       atol=1e-15)
    pl = C.cast(ephem.spk_planets, C.POINTER(spk.SpkFile))
    sun = lib.assist_spk_find_target(pl, 10)
    for t in (0.0, 777.0):
        pos = lib.assist_spk_target_pos(pl, sun, jd_ref, t)
        fo, o = oracle.spk_eval(0, int(C.addressof(sun.contents) - C.addressof(pl.contents.targets[0]))
                                // C.sizeof(spk.SpkTarget), t)
        assert fo == 0
        np.testing.assert_array_equal(list(pos.u) + list(pos.v) + list(pos.w), o)
    # errors: unknown code, outside coverage, asteroid index past the file
    f, _ = _calc_planets(lib.assist_spk_calc_planets, ephem, jd_ref, 0.0, 12345)
    assert f == 4  # ASSIST_ERROR_NEPHEM
    f, _ = _calc_planets(lib.assist_spk_calc_planets, ephem, jd_ref, 1e6, 10)
    assert f == 5  # ASSIST_ERROR_COVERAGE
    f, _ = _calc_planets(lib.assist_spk_calc_planets_by_assist, ephem, jd_ref, 0.0, 11)
    assert f == 4
    d = C.c_double()
    al = C.cast(ephem.spk_asteroids, C.POINTER(spk.SpkFile))
    assert lib.assist_spk_calc(al, jd_ref, 0.0, 16, *[C.byref(d)] * 4) == 3  # ASSIST_ERROR_NAST
    assert lib.assist_spk_calc(None, jd_ref, 0.0, 0, *[C.byref(d)] * 4) == 2  # ASSIST_ERROR_AST_FILE
