"""CPU tests of the JPL binary (.440/.441) planets backend (reference src/ascii_ephem.c,
src/assist.c:66-135 detection, :291-327 constants and asteroid masses).

No DE440 file is available, so the file is the synthetic solar system written in the JPL
binary layout with DE440's block shape (32-day blocks, DE440's coefficient and sub-interval
counts, 8 144-byte blocks, more than 400 constants).  What pins each part:

* the oracle's evaluation (oracle/ascii.c) against numpy's Chebyshev series
  (numpy.polynomial.chebyshev.chebval / chebder) on the raw coefficients of the block and
  sub-interval that cover each time: an independent evaluation of the same file;
* the file against the Kepler orbits it was fitted to (the synthetic model);
* the library (assist_amd/csrc/ascii.cpp) against the oracle, bit for bit, through the
  reference's own entry points (assist_get_particle, assist_ascii_find_constant);
* the reference's format-detection byte samples are covered in test_library.py.
"""
import ctypes as C
import os
import struct
import tempfile

import numpy as np
import pytest
from numpy.polynomial import chebyshev as cheb

AU_KM = 149597870.7


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


def _read_file(path):
    """Header, constants and blocks of a JPL binary file, parsed with struct/numpy."""
    raw = open(path, "rb").read()
    beg, end, inc, num, cau, cem = struct.unpack_from("<dddidd", raw, 0x0A5C)
    cols = [struct.unpack_from("<iii", raw, 0x0A5C + 44 + 12 * k) for k in range(12)]
    ver, = struct.unpack_from("<i", raw, 0x0A5C + 44 + 144)
    cols.append(struct.unpack_from("<iii", raw, 0x0A5C + 44 + 148))
    names = [raw[0x00FC + 6 * p:0x00FC + 6 * p + 6].decode() for p in range(400)]
    o = 0x0B28
    for p in range(400, num):
        names.append(raw[o:o + 6].decode())
        o += 6
    cols += [struct.unpack_from("<iii", raw, o), struct.unpack_from("<iii", raw, o + 12)]
    ncm = [3] * 15
    ncm[11], ncm[14] = 2, 1
    rec_words = 2 + sum(c[1] * c[2] * m for c, m in zip(cols, ncm))
    words = np.frombuffer(raw, dtype="<f8")
    values = words[rec_words:rec_words + num]
    return dict(beg=beg, end=end, inc=inc, cau=cau, cem=cem, ver=ver, cols=cols, ncm=ncm, names=names,
                constants=dict(zip([n.strip() for n in names], values)), rec_words=rec_words,
                blocks=words[2 * rec_words:].reshape(-1, rec_words))


def _column_state(F, col, t):
    """Position (km), velocity (km/d) and acceleration (km/d^2) of one column at JD 2451545 + t,
    from numpy's Chebyshev routines (the block fraction is formed from the day offsets, not from
    the full JD, which would round it to 5e-10 d)."""
    off, ncf, niv = F["cols"][col]
    t_beg = F["beg"] - 2451545.0
    blk = min(int((t - t_beg) // F["inc"]), F["blocks"].shape[0] - 1)
    f = (t - (t_beg + blk * F["inc"])) / F["inc"]
    k = min(int(f * niv), niv - 1)
    z = 2.0 * (f * niv - k) - 1.0
    dzdt = 2.0 * niv / F["inc"]
    cf = F["blocks"][blk, off - 1 + ncf * 3 * k: off - 1 + ncf * 3 * (k + 1)].reshape(3, ncf)
    pos = np.array([cheb.chebval(z, c) for c in cf])
    vel = np.array([cheb.chebval(z, cheb.chebder(c)) for c in cf]) * dzdt
    acc = np.array([cheb.chebval(z, cheb.chebder(c, 2)) for c in cf]) * dzdt ** 2
    return pos, vel, acc


# ASSIST body -> column (Earth and Moon: EMB and geocentric Moon)
BODY_COL = {0: 10, 1: 0, 2: 1, 5: 3, 6: 4, 7: 5, 8: 6, 9: 7, 10: 8}


def test_file_has_the_de440_block_layout(jpl_paths):
    import assist_amd
    F = _read_file(jpl_paths[0])
    assert 8 * F["rec_words"] == 8144  # DE440/441 block size
    assert F["ver"] == 440 and F["inc"] == 32.0
    assert [c[0] for c in F["cols"][:13]] == [3, 171, 231, 309, 342, 366, 387, 405, 423, 441, 753, 819, 899]
    assert len(F["names"]) > 400 and F["names"][:3] == ["DENUM ", "LENUM ", "TDATEF"]
    assert F["cem"] == F["constants"]["EMRAT"] and F["cau"] == F["constants"]["AU"]
    fd = os.open(jpl_paths[0], os.O_RDONLY)
    try:
        assert assist_amd.clibassist.assist_detect_ephemeris_file_format(fd) == 2  # FILE_FORMAT_ASCII_BIN
    finally:
        os.close(fd)


def test_oracle_matches_numpy_chebyshev(jpl_paths, jpl_oracle):
    """The oracle's T/S/U recurrences, block and sub-interval selection and unit conversion
    against numpy's Chebyshev series on the raw coefficients."""
    F = _read_file(jpl_paths[0])
    rng = np.random.default_rng(5)
    times = list(rng.uniform(-9990.0, 9990.0, 40)) + [0.0, 1.0, 16.0, 31.999, -10000.0 + 1e-9]
    cem = F["cem"]
    for body in range(11):
        for t in times:
            f, o = jpl_oracle.all_ephem(body, t)
            assert f == 0
            if body in BODY_COL:
                p, v, a = _column_state(F, BODY_COL[body], t)
            else:
                pe, ve, ae = _column_state(F, 2, t)
                pl, vl, al = _column_state(F, 9, t)
                k = -1.0 / (1.0 + cem) if body == 3 else cem / (1.0 + cem)
                p, v, a = pe + k * pl, ve + k * vl, ae + k * al
            scale = np.abs(p).max()
            assert np.abs(o[1:4] * F["cau"] - p).max() <= 1e-14 * scale + 1e-9, (body, t)
            assert np.abs(o[4:7] * F["cau"] - v).max() <= 1e-12 * np.abs(v).max() + 1e-12, (body, t)
            assert np.abs(o[7:10] * F["cau"] - a).max() <= 1e-10 * np.abs(a).max() + 1e-12, (body, t)


def test_file_follows_the_kepler_orbits(jpl_oracle, model):
    """The written file reproduces the orbits it was fitted to (fit error only)."""
    rng = np.random.default_rng(6)
    for t in rng.uniform(-9990.0, 9990.0, 25):
        for body, naif in ((0, None), (1, 1), (2, 2), (5, 4), (6, 5), (7, 6), (8, 7), (9, 8), (10, 9)):
            r, v = model.sun(t) if naif is None else model.planet(naif, t)
            f, o = jpl_oracle.all_ephem(body, t)
            assert f == 0
            assert np.abs(o[1:4] - r).max() < 1e-9, (body, t)
            assert np.abs(o[4:7] - v).max() < 1e-9, (body, t)
        emb = model.planet(3, t)[0]
        assert np.abs(jpl_oracle.all_ephem(3, t)[1][1:4] - (emb + model.earth_wrt_emb(t)[0])).max() < 1e-12
        assert np.abs(jpl_oracle.all_ephem(4, t)[1][1:4] - (emb + model.moon_wrt_emb(t)[0])).max() < 1e-12


def test_library_matches_oracle_bitwise(jpl_ephem, jpl_oracle):
    rng = np.random.default_rng(7)
    times = list(rng.uniform(-9999.0, 9999.0, 15)) + [-10000.0, 10000.0, 0.0, 48.0]
    for body in range(27):
        for t in times:
            p = jpl_ephem.get_particle(body, t)
            f, o = jpl_oracle.all_ephem(body, t)
            assert f == 0
            assert (p.m, p.x, p.y, p.z) == tuple(o[:4]), (body, t)
            if body < 11:
                assert (p.vx, p.vy, p.vz, p.ax, p.ay, p.az) == tuple(o[4:10]), (body, t)


def test_constants_masses_and_frame(jpl_paths, jpl_ephem, model):
    import assist_amd
    from assist_amd import synthetic
    F = _read_file(jpl_paths[0])
    assert jpl_ephem.planets_source == 2
    assert jpl_ephem.time_bounds() == (-10000.0, 10000.0)
    assert jpl_ephem.AU == F["constants"]["AU"] and jpl_ephem.EMRAT == F["cem"]
    assert jpl_ephem.J2E == synthetic.CONSTANTS["J2E"] and jpl_ephem.J4E == synthetic.CONSTANTS["J4E"]
    assert jpl_ephem.Re_eq == synthetic.CONSTANTS["RE"] / synthetic.CONSTANTS["AU"]
    gmb, emrat = synthetic.GM_PLANETS["GMB"], F["cem"]
    assert jpl_ephem.get_particle(3, 0.0).m == (emrat / (1. + emrat)) * gmb
    assert jpl_ephem.get_particle(4, 0.0).m == 1. / (1 + emrat) * gmb
    # asteroid masses come from the file's MAnnnn constants, which sit past the first 400 names
    for k, a in enumerate(model.asteroids):
        assert jpl_ephem.get_particle(11 + k, 0.0).m == a["gm"]
    ap = C.cast(jpl_ephem.ascii_planets, C.c_void_p)
    val = C.c_double(-1.0)
    lib = assist_amd.clibassist
    lib.assist_ascii_find_constant.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
    assert lib.assist_ascii_find_constant(ap, b"AU", C.byref(val)) == 1 and val.value == F["constants"]["AU"]
    assert lib.assist_ascii_find_constant(ap, b"MA0004", C.byref(val)) == 1 and val.value == model.asteroids[15]["gm"]
    assert lib.assist_ascii_find_constant(ap, b"MA000", C.byref(val)) == 0 and val.value == 0.0
    assert lib.assist_ascii_find_constant(ap, b"GMX", C.byref(val)) == 0


def test_coverage_and_body_errors(jpl_ephem):
    with pytest.raises(RuntimeError):
        jpl_ephem.get_particle(0, 10000.5)
    with pytest.raises(RuntimeError):
        jpl_ephem.get_particle(5, -10000.5)
    # at the last instant of coverage the last sub-interval's end is evaluated (the data are
    # continuous there, so the state is the limit from inside)
    a = jpl_ephem.get_particle(1, 10000.0)
    b = jpl_ephem.get_particle(1, 10000.0 - 1e-9)
    assert abs(a.x - b.x) < 1e-10 and abs(a.vx - b.vx) < 1e-10


def test_malformed_files_are_rejected(jpl_paths):
    import assist_amd
    raw = bytearray(open(jpl_paths[0], "rb").read())
    with tempfile.TemporaryDirectory() as d:
        cases = {
            "truncated.440": raw[:8144 * 10],                                   # fewer blocks than the coverage
            "wide.440": raw[:0x0A5C + 44] + struct.pack("<iii", 3, 40, 4) + raw[0x0A5C + 56:],  # 40 coefficients
            "offset.440": raw[:0x0A5C + 44] + struct.pack("<iii", 1020, 14, 4) + raw[0x0A5C + 56:],  # past the block
        }
        for name, data in cases.items():
            p = os.path.join(d, name)
            open(p, "wb").write(bytes(data))
            with pytest.raises(RuntimeError):
                assist_amd.Ephem(p, jpl_paths[1])
            init = assist_amd.clibassist.assist_ascii_init
            init.restype = C.c_void_p
            assert init(p.encode()) is None


def test_agrees_with_the_spk_file_of_the_same_orbits(jpl_ephem, ephem):
    """The JPL binary and SPK files of the synthetic solar system describe the same orbits; they
    differ by their fits only (Mercury's 32-day SPK records are the coarsest)."""
    rng = np.random.default_rng(8)
    for t in rng.uniform(-9000.0, 9000.0, 20):
        for body in range(27):
            p, q = jpl_ephem.get_particle(body, t), ephem.get_particle(body, t)
            tol = 1e-6 if body == 1 else 1e-9
            assert abs(p.x - q.x) + abs(p.y - q.y) + abs(p.z - q.z) < tol, (body, t)
