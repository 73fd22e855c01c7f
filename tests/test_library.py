"""CPU tests of the C-ABI boundary and the Python drop-in (no GPU calls):

* libassist.so loads and exports every function include/assist.h declares;
* format detection on the byte samples of the reference's own test
  (unit_tests/spk_detection/problem.c:14-46, fixtures in tests/golden);
* host-side ephemeris queries agree with the oracle bit for bit;
* the Python wrapper's names, layouts and errors mirror the reference wrapper
  (assist/ephem.py, assist/extras.py, assist/test/test_forces.py:10-22,
  assist/test/test_basic.py:34-41);
* GPU entry points fail loudly (no CPU fallback) when no GPU is visible.
"""
import ctypes as C
import os
import re
import struct
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    """Names of the functions include/*.h declare (function-pointer members excluded)."""
    inc = os.path.join(ROOT, "include")
    txt = "\n".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    txt = re.sub(r"^\s*#[^\n]*", "", txt, flags=re.M)
    names = set()
    for chunk in txt.split(";"):
        chunk = chunk.split("{")[-1].split("}")[-1]
        if "(" not in chunk or "(*" in chunk.replace(" ", "") or "typedef" in chunk:
            continue
        m = re.search(r"(\w+)\s*\(", chunk)
        if m:
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import assist_amd
    lib = assist_amd.clibassist
    names = _declared_functions()
    assert len(names) >= 38
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    for data in ("assist_version_str", "assist_build_str", "assist_githash_str", "assist_error_messages",
                 "assist_error_messages_N"):
        assert C.c_void_p.in_dll(lib, data) is not None


def test_version_and_error_messages():
    import assist_amd
    assert assist_amd.__version__ == "1.2.0"
    assert assist_amd.assist_error_messages(5).startswith("The requested time is outside the coverage")
    with pytest.raises(RuntimeError):
        assist_amd.assist_error_messages(6)


def _detect(path):
    import assist_amd
    fd = os.open(path, os.O_RDONLY)
    try:
        return assist_amd.clibassist.assist_detect_ephemeris_file_format(fd)
    finally:
        os.close(fd)


def test_format_detection_on_reference_samples(ephem_paths):
    assert _detect(os.path.join(GOLDEN, "detect_valid_bsp_sample.bin")) == 0      # FILE_FORMAT_VALID_BSP
    assert _detect(os.path.join(GOLDEN, "detect_ascii_bin_sample.bin")) == 2      # FILE_FORMAT_ASCII_BIN
    assert _detect(os.path.join(GOLDEN, "detect_ascii_source_sample.txt")) == 3   # FILE_FORMAT_UNKNOWN
    assert _detect(ephem_paths[0]) == 0


def test_discovery_and_missing_files():
    import assist_amd
    buf = C.create_string_buffer(1024)
    with tempfile.TemporaryDirectory() as d:
        assert assist_amd.clibassist.assist_discover_planets_path(buf, 1024, d.encode()) == 0
        os.makedirs(os.path.join(d, "data"))
        open(os.path.join(d, "data", "de440.bsp"), "wb").close()
        assert assist_amd.clibassist.assist_discover_planets_path(buf, 1024, d.encode()) == 1
        assert buf.value.decode().endswith("/data/de440.bsp")
    with pytest.raises(RuntimeError):
        assist_amd.Ephem("/nonexistent.bsp", None)


def test_ephem_dropin(ephem, oracle):
    import assist_amd
    assert ephem.jd_ref == 2451545.0
    assert ephem.time_bounds() == (-10000.0, 10000.0)
    for body, t in ((0, 0.0), (1, 100.0), (3, -2500.5), (4, 77.7), (20, 200.0)):
        p = ephem.get_particle(body, t)
        f, o = oracle.all_ephem(body, t)
        assert f == 0
        assert (p.m, p.x, p.y, p.z) == (o[0], o[1], o[2], o[3])
        if body < 11:
            assert (p.vx, p.vy, p.vz, p.ax, p.ay, p.az) == tuple(o[4:10])
    assert ephem.get_particle("Sun", 0).x == ephem.get_particle(0, 0).x
    assert ephem.get_particle("vesta", 10).x == ephem.get_particle(26, 10).x
    with pytest.raises(ValueError):
        ephem.get_particle("Planet 9", 0)
    with pytest.raises(RuntimeError):
        ephem.get_particle(0, 1e6)
    assert len(assist_amd.ASSIST_BODY_IDS) == 27


def test_extras_dropin(ephem):
    import assist_amd
    sim = assist_amd.Simulation()
    ex = assist_amd.Extras(sim, ephem)
    assert ex.forces == ["SUN", "PLANETS", "ASTEROIDS", "NON_GRAVITATIONAL", "EARTH_HARMONICS", "SUN_HARMONICS", "GR_EIH"]
    assert ex.gr_eih_sources == 1
    assert (ex.alpha, ex.nk, ex.nm, ex.nn, ex.r0) == (1.0, 0.0, 2.0, 5.093, 1.0)
    with pytest.raises(AttributeError):
        ex.forces = "no array"
    with pytest.raises(AttributeError):
        ex.forces = ["Magic"]
    with pytest.raises(AttributeError):
        ex.forces = [1, 2, 3]
    f = ex.forces
    f.remove("GR_EIH")
    ex.forces = f
    assert "GR_EIH" not in ex.forces and len(ex.forces) == 6
    ex.forces = ["sun"]  # accepted (upper-cased check) but, as in the reference, matches nothing
    assert ex.forces == []
    with pytest.raises(AttributeError):
        _ = ex.particle_params
    ex.particle_params = np.array([1e-12, 0.0, 0.0])
    assert sim.ri_ias15.adaptive_mode == 1
    ex.detach(sim)


def test_struct_layouts_match_reference_offsets():
    """Field offsets of the drop-in structs equal the reference layouts (LP64)."""
    import assist_amd
    from assist_amd.extras import Extras
    E = assist_amd.Ephem
    assert E.jd_ref.offset == 0 and E.spk_planets.offset == 8 and E.planets_source.offset == 32
    assert E.planets_calc.offset == 40 and E.spk_target_index.offset == 48 and E.spk_emb_index.offset == 92
    assert E.AU.offset == 96 and E.over_c_squared.offset == 96 + 13 * 8
    assert Extras.steps_done.offset == 56 and Extras._forces.offset == 60 and Extras.gr_eih_sources.offset == 64
    assert Extras.alpha.offset == 72 and Extras.r0.offset == 104


def _c_layout(struct_name, fields):
    """(sizeof, {field: offsetof}) of a struct of include/assist.h, from gcc."""
    import json
    import subprocess
    src = ("#include <stdio.h>\n#include <stddef.h>\n#include \"assist.h\"\n#include \"spk.h\"\n"
           "int main(void){printf(\"{\\\"__sizeof__\\\": %zu\", sizeof(struct " + struct_name + "));")
    for f in fields:
        src += "printf(\", \\\"%s\\\": %%zu\", offsetof(struct %s, %s));" % (f, struct_name, f)
    src += "printf(\"}\\n\"); return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "layout")
        subprocess.run(["gcc", "-x", "c", "-I", os.path.join(ROOT, "include"), "-o", exe, "-"], input=src.encode(),
                       check=True)
        out = json.loads(subprocess.run([exe], check=True, capture_output=True).stdout)
    return out.pop("__sizeof__"), out


def test_struct_mirrors_match_the_header():
    """Every ctypes mirror of a struct of include/assist.h has the C struct's fields at the C
    offsets and the C size (a short mirror makes the library read or write past the end of the
    caller's struct).  Mirror fields named with a leading underscore wrap the C field."""
    import assist_amd
    from assist_amd.batch import BatchConfig, BatchExtra, BatchStats
    from assist_amd.extras import Extras
    from assist_amd import spk
    from assist_amd.simulation import IntegratorIAS15, Particle, Simulation, VariationalConfiguration
    mirrors = ((BatchConfig, "assist_batch_config"), (BatchStats, "assist_batch_stats"),
               (BatchExtra, "assist_batch_extra"), (assist_amd.Ephem, "assist_ephem"), (Extras, "assist_extras"),
               (Particle, "reb_particle"), (Simulation, "reb_simulation"),
               (VariationalConfiguration, "reb_variational_configuration"),
               (IntegratorIAS15, "reb_simulation_integrator_ias15"),
               (spk.SpkFile, "spk_s"), (spk.SpkTarget, "spk_target"), (spk.MassData, "mass_data"),
               (spk.ConstantsAndMasses, "spk_constants_and_masses"), (spk.Mpos, "mpos_s"))
    for cls, name in mirrors:
        fields = [f[0] for f in cls._fields_]
        cnames = [f.lstrip("_") if f != "reserved_" else f for f in fields]
        size, offs = _c_layout(name, cnames)
        assert C.sizeof(cls) == size, name
        for f, c in zip(fields, cnames):
            assert getattr(cls, f).offset == offs[c], (name, f)


def test_batch_config_default_sets_every_field():
    """assist_batch_config_default defines the whole struct, whatever the memory held before (C
    callers declare it on the stack)."""
    import assist_amd
    from assist_amd.batch import BatchConfig
    c = BatchConfig()
    C.memset(C.addressof(c), 0x5A, C.sizeof(c))
    assist_amd.clibassist.assist_batch_config_default(C.byref(c))
    assert c.size == C.sizeof(BatchConfig)
    assert (c.epsilon, c.min_dt, c.dt_initial, c.max_steps, c.precision) == (1e-9, 0.0, 0.001, 1000000, 0)
    assert (c.lanes_per_system, c.block_size, c.blocks_per_cu, c.waves_per_simd, c.use_graph,
            c.substep_cache_off) == (0, 0, 0, 0, 0, 0)


def test_structs_of_another_size_are_rejected(ephem):
    """A config or extra struct whose size field is not the library's (a caller built against
    another layout) is rejected with ASSIST_BATCH_ERROR_ARG before anything is read past it or
    any device is touched."""
    import assist_amd
    from assist_amd.batch import BatchConfig, BatchExtra
    lib = assist_amd.clibassist
    e = BatchExtra.new()
    assert e.size == C.sizeof(BatchExtra) and e.trace_every == 1
    assert all(getattr(e, f) in (None, 0) for f, _ in BatchExtra._fields_ if f not in ("size", "trace_every"))
    t = np.zeros(1)
    st = np.zeros((1, 1, 6))
    status = np.zeros(1, dtype=np.int32)
    dp = C.POINTER(C.c_double)
    for bad in (0, C.sizeof(BatchConfig) - 8, C.sizeof(BatchConfig) + 8):
        c = BatchConfig.default()
        c.size = bad
        rc = lib.assist_batch_integrate(C.byref(ephem), None, C.byref(c), 1, 0, t.ctypes.data_as(dp),
                                        t.ctypes.data_as(dp), st.ctypes.data_as(dp), None,
                                        status.ctypes.data_as(C.POINTER(C.c_int)), None, None)
        assert rc == 100
        rc = lib.assist_batch_integrate_device_ex(C.byref(ephem), None, C.byref(c), 1, 0, C.c_void_p(8), C.c_void_p(8),
                                                  C.c_void_p(8), None, C.c_void_p(8), None, None, None, None, None)
        assert rc == 100
        x = BatchExtra.new()
        x.size = bad
        rc = lib.assist_batch_integrate_device_ex(C.byref(ephem), None, None, 1, 0, C.c_void_p(8), C.c_void_p(8),
                                                  C.c_void_p(8), None, C.c_void_p(8), None, None, None, None,
                                                  C.byref(x))
        assert rc == 100
    # an archive needs its counts and a capacity; a resume cannot take dense output
    x = BatchExtra.new(archive=8, archive_capacity=0, archive_count=8)
    assert lib.assist_batch_integrate_device_ex(C.byref(ephem), None, None, 1, 0, C.c_void_p(8), C.c_void_p(8),
                                                C.c_void_p(8), None, C.c_void_p(8), None, None, None, None,
                                                C.byref(x)) == 100
    x = BatchExtra.new(snapshot_in=8, out_t=8, out_offset=8, out_states=8)
    assert lib.assist_batch_integrate_device_ex(C.byref(ephem), None, None, 1, 0, C.c_void_p(8), None,
                                                C.c_void_p(8), None, C.c_void_p(8), None, None, None, None,
                                                C.byref(x)) == 102


def test_gpu_entry_points_fail_without_gpu(ephem):
    from assist_amd import batch
    if batch.gpu_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(batch.BatchError):
        batch.forces(ephem, 0.0, np.zeros((1, 1, 6)))
    with pytest.raises(batch.BatchError):
        batch.integrate(ephem, 0.0, 1.0, np.zeros((1, 1, 6)))


def test_synthetic_file_satisfies_reader_assumptions(ephem_paths):
    for path in ephem_paths:
        buf = open(path, "rb").read()
        assert buf[:8] == b"DAF/SPK "
        nd, ni = struct.unpack_from("<ii", buf, 8)
        assert (nd, ni) == (2, 6)
        fward = struct.unpack_from("<i", buf, 76)[0]
        off = (fward - 1) * 1024
        nxt, prev, nsum = struct.unpack_from("<ddd", buf, off)
        assert nxt == 0.0 and prev == 0.0 and buf[off + 8] == 0
        assert len(buf) % 1024 == 0


def test_elements_round_trip(model):
    from assist_amd import synthetic
    rng = np.random.default_rng(1)
    a = rng.uniform(0.8, 3.5, 100)
    e = rng.uniform(0, 0.6, 100)
    inc = rng.uniform(0.05, 0.6, 100)
    node, argp, M = (rng.uniform(0, 2 * np.pi, 100) for _ in range(3))
    r, v = synthetic.kepler_state(model.gms, a, e, inc, node, argp, M)
    a2, e2, i2, n2, w2, M2 = synthetic.state_to_elements(model.gms, r, v)
    np.testing.assert_allclose(a2, a, rtol=1e-12)
    np.testing.assert_allclose(e2, e, atol=1e-12)
    np.testing.assert_allclose(i2, inc, atol=1e-12)
    d = lambda x, y: np.abs((x - y + np.pi) % (2 * np.pi) - np.pi)
    assert d(n2, node).max() < 1e-10 and d(w2 + M2, argp + M).max() < 1e-9


def test_dense_output_checks_time_order_before_the_gpu(ephem):
    """assist_batch_integrate_dense validates each system's output times (sorted in one
    direction from t0) on the host, before any device work."""
    import ctypes as C
    import assist_amd
    lib = assist_amd.clibassist
    from assist_amd import batch
    states = np.zeros((1, 1, 6))
    states[0, 0, 0] = 2.0
    t0 = np.zeros(1)
    off = np.array([0, 2], dtype=np.int32)
    out = np.zeros((2, 1, 6))
    st = np.zeros(1, dtype=np.int32)
    steps = np.zeros(1, dtype=np.int64)
    dp = C.POINTER(C.c_double)
    for times in (np.array([10.0, 5.0]), np.array([-1.0, 5.0])):
        rc = lib.assist_batch_integrate_dense(C.byref(ephem), None, None, 1, 0, t0.ctypes.data_as(dp),
                                              states.ctypes.data_as(dp), None, off.ctypes.data_as(C.POINTER(C.c_int)),
                                              times.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                              st.ctypes.data_as(C.POINTER(C.c_int)),
                                              steps.ctypes.data_as(C.POINTER(C.c_longlong)), None)
        assert rc == 100


def test_batched_queries_fail_without_gpu(ephem):
    from assist_amd import batch
    if batch.gpu_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(batch.BatchError):
        batch.get_particles(ephem, [0], [0.0])


def test_operation_count_of_the_chebyshev_series():
    """The bench's algorithmic count leaves out the Chebyshev series of the evaluations that the
    substep cache serves; pin that share against a hand count of the evaluation formula.  A
    position-only evaluation of P coefficients costs 2 (record index) + 2 (z) + 1 (2z) + 3 x 2
    (first fused term) + (P - 2) x (1 fma for T + 3 fma) x 2 = 11 + 8 (P - 2) operations; the
    velocity adds (P - 2) x (2 fma for S + 3 fma) x 2 = 10 (P - 2) and 3 scalings.  The benchmark
    ephemeris has 10 slots of 14 coefficients, Earth and Moon of 13 and 16 asteroids of 11."""
    sys.path.insert(0, ROOT)
    import bench
    oc = bench.opcounts()
    shapes = [14] * 10 + [13] * 2 + [11] * 16
    no_vel = sum(11 + 8 * (p - 2) for p in shapes)
    assert oc["cheb"]["no_sun_vel"] == no_vel == 2596
    assert oc["cheb"]["sun_vel"] == no_vel + 10 * (14 - 2) + 3
    # the series are part of the full evaluation count they are subtracted from
    assert oc["cheb"]["sun_vel"] < oc["eval"]["nvec7_ng0_fp64"]["fp64"]


def test_cpu_baseline_protocol(ephem_paths):
    """bench.py's CPU baseline on a tiny sample: several timed runs on all threads (mean and
    standard deviation), one run on one thread, and the reference's benchmark shape (10
    particles sharing one IAS15 step for 10 years), with the oracle's results returned for the
    parity check beside the GPU."""
    sys.path.insert(0, ROOT)
    import bench
    from assist_amd import synthetic
    cfg = dict(bench.CONFIGS["c4"])
    cfg["n"] = 64
    t0, t1, states, params = bench.make_workload(cfg, synthetic.SyntheticModel(0))
    cb, r, m = bench.cpu_baseline(ephem_paths, cfg, t0, t1, states, params, sample_s=0.05, samples=3)
    assert cb["kind"] == "port" and cb["unit"] == "particle-days/s" and cb["samples"] == 3
    assert cb["value"] > 0 and cb["std"] >= 0 and cb["cores"] >= 1
    assert cb["single_thread"]["value"] > 0
    shape = cb["reference_benchmark_shape"]
    assert shape["samples"] == 5 and shape["steps"] > 0 and shape["evals"] >= 8 * shape["steps"]
    assert r["states"].shape == (m, 7, 6) and (r["status"] == 0).all()


# force flags (enum ASSIST_FORCES) of the coverage cases below, and the status the reference's
# rule gives at a time the planets cover and the asteroids do not: the direct term reports
# coverage errors (its variational loop fetches the first asteroid whatever the flags, reference
# src/forces.c:351-367), the other terms do not check them
COVERAGE_CASES = [(0x7F, 5), (0x03, 5), (0x04, 5), (0x10 | 0x40, 0), (0x20 | 0x08, 0), (0x00, 0)]


def test_coverage_windows_per_file(short_asteroid_paths, oracle):
    """With an asteroid file shorter than the planets file: time bounds are the intersection (as
    reference assist_ephem_time_bounds), planet queries succeed where only the planets cover,
    asteroid queries fail there, and the oracle's force model follows the reference's rule."""
    import assist_amd
    from oracle_lib import Oracle
    pp, ap = short_asteroid_paths
    e = assist_amd.Ephem(pp, ap)
    o = Oracle(pp, ap)
    assert e.time_bounds() == (-3000.0, 3000.0)
    for t in (5000.0, -9000.0):
        p = e.get_particle(5, t)
        f, ref = o.all_ephem(5, t)
        assert f == 0 and (p.x, p.y, p.z) == tuple(ref[1:4])
        with pytest.raises(RuntimeError):
            e.get_particle(12, t)
        assert o.all_ephem(12, t)[0] == 5
    st = np.zeros((1, 2, 6))
    st[0, 0] = [2.5, 0.1, 0.05, -0.001, 0.01, 0.002]
    st[0, 1, 0] = 1.0
    for forces, expect in COVERAGE_CASES:
        _, s_in = o.forces(2000.0, st, opts=o.opts(forces=forces))
        _, s_out = o.forces(5000.0, st, opts=o.opts(forces=forces))
        assert (s_in[0], s_out[0]) == (0, expect), forces
