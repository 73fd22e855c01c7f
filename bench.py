#!/usr/bin/env python3
"""Particle-days per second of the batched IAS15 + ASSIST force model on one MI355X.

Workload (default): 262,144 systems -- 70% main-belt, 30% NEO-like, seed 4 --
each a real test particle with 6 first-order variational particles, integrated
365 days forward from JD 2451545.0 with IAS15 (epsilon 1e-9), all ASSIST
forces on, on the seed-0 synthetic ephemeris.  One "step" is one integration
of the whole batch; inputs are resident in HBM before the timed region.

Prints one JSON line with the metric, the roofline of the integration kernel
(algorithmic FLOPs from the operation count of the GPU algorithm divided by the
kernel's HIP-event time) and the CPU baseline (the C oracle on the host cores,
a bounded sample of the same workload).
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

# Populations with NEO-like orbits use min_dt = 0.001 d, as the reference does for encounters
# (reference unit_tests/apophis_spk/problem.c:51): in a close planetary encounter the
# round-off of the barycentric difference dominates IAS15's error estimate and would
# otherwise drive the step to zero.
CONFIGS = {
    "c4": dict(n=262144, n_var=6, span=365.0, neo=0.3, seed=4, uneven=False, min_dt=0.001,
               workload="262144 systems (70% main-belt, 30% NEO) x (1 + 6 variational), 365 d, eps 1e-9, "
                        "min_dt 0.001 d, forces 0x7F"),
    "c2": dict(n=4096, n_var=6, span=3650.0, neo=0.0, seed=2, uneven=False, min_dt=0.0,
               workload="4096 main-belt systems x (1 + 6 variational), 3650 d, eps 1e-9, forces 0x7F"),
    "c5": dict(n=1048576, n_var=0, span=365.0, neo=0.3, seed=5, uneven=False, min_dt=0.001,
               workload="1048576 systems (70% main-belt, 30% NEO), 365 d, eps 1e-9, min_dt 0.001 d, forces 0x7F"),
    "c3": dict(n=65536, n_var=0, span=None, neo=0.3, seed=3, uneven=True, min_dt=0.001,
               workload="65536 systems (70% main-belt, 30% NEO), log-uniform 30-10000 d spans, half backward, "
                        "10% Marsden, eps 1e-9, min_dt 0.001 d"),
}


def opcounts():
    """Operation counts of the GPU algorithm (assist_amd/csrc/opcount.cpp)."""
    exe = os.path.join(ROOT, "assist_amd", "_build", "opcount")
    src = os.path.join(ROOT, "assist_amd", "csrc", "opcount.cpp")
    model = os.path.join(ROOT, "assist_amd", "csrc", "force_model.h")
    if not os.path.exists(exe) or os.path.getmtime(exe) < max(os.path.getmtime(src), os.path.getmtime(model)):
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe, src], check=True)
    return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


def algorithmic_flops(oc, nvec, ng, mixed, evals, steps, rejected, cached=False):
    """FP64 and FP32 operations of one launch from per-system eval/step counts; ``ng`` is the
    per-system flag of a non-zero Marsden model, ``rejected`` per-system or total rejections.
    With ``cached`` (the substep ephemeris cache of 8-lane groups), only the evaluations that
    compute fresh positions (each step's first one and the first iteration of every attempt)
    count the Chebyshev series."""
    form = "mixed" if mixed else "fp64"
    ev0 = oc["eval"]["nvec%d_ng0_%s" % (nvec, form)]
    ev1 = oc["eval"].get("nvec%d_ng1_%s" % (nvec, form), ev0)
    ias = oc["ias15"]
    comps = 3 * nvec
    ng = np.zeros(len(evals), dtype=bool) if ng is None else np.asarray(ng, dtype=bool)
    ev_tot = float(np.sum(evals))
    ev_ng = float(np.sum(evals[ng]))
    st_tot = float(np.sum(steps))
    rej_tot = float(np.sum(rejected)) if rejected is not None else 0.0
    substeps = ev_tot - st_tot
    f64 = (ev_tot - ev_ng) * ev0["fp64"] + ev_ng * ev1["fp64"]
    f32 = (ev_tot - ev_ng) * ev0["fp32"] + ev_ng * ev1["fp32"]
    f64 += substeps * (ias["substep_scalar"] + comps * (ias["predict_comp"] + ias["correct_comp_7substeps"] / 7.0))
    f64 += (st_tot + rej_tot) * (comps * ias["setup_comp"] + ias["control"])
    f64 += st_tot * (comps * ias["accept_comp"] + ias.get("accept_scalar", 0))
    if cached:
        fresh = st_tot + 7.0 * (st_tot + rej_tot)
        f64 -= max(ev_tot - fresh, 0.0) * oc["cheb"]["sun_vel"]
    return f64, f32


def make_workload(cfg, model):
    from assist_amd import synthetic
    rng = np.random.default_rng(cfg["seed"])
    n, nv = cfg["n"], 1 + cfg["n_var"]
    st, _ = synthetic.mixed_population(model, n, rng, cfg["neo"], 0.0)
    states = np.zeros((n, nv, 6))
    states[:, 0] = st
    if cfg["n_var"]:
        states[:, 1:] = synthetic.variational_identity(n, cfg["n_var"])
    t0 = np.zeros(n)
    params = None
    if cfg["uneven"]:
        t1 = synthetic.log_uniform_spans(n, rng)
        params = np.zeros((n, nv, 3))
        params[:, 0] = synthetic.marsden_params(n, rng, 0.1)
    else:
        t1 = np.full(n, cfg["span"])
    return t0, t1, states, params


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _cgroup_cpus():
    """CPUs granted by the cgroup v2 (or v1) CPU quota, or None when unlimited or unreadable."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            return float(quota) / float(period)
    except (OSError, ValueError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
            quota = float(f.read())
        with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
            period = float(f.read())
        if quota > 0:
            return quota / period
    except (OSError, ValueError):
        pass
    return None


def host_cpus():
    """The host cores this process may run on: its CPU affinity set, limited by the cgroup CPU
    quota when one is set (threads beyond the quota only time-share it).  Returns the thread count
    to use and the facts it came from."""
    affinity = len(os.sched_getaffinity(0))
    quota = _cgroup_cpus()
    threads = affinity
    if quota is not None:
        threads = max(1, min(threads, int(quota + 0.5)))
    return threads, {"nproc": os.cpu_count(), "affinity": affinity,
                     "cgroup_cpu_quota": quota, "cpu_model": _cpu_model()}


# Initial state of the reference's benchmark particle (reference unit_tests/benchmark/problem.c:72-76);
# particle i is offset by i / 1e-10 au in x, as there.
REFERENCE_BENCHMARK_STATE = [-2.724183384883979E+00, -3.523994546329214E-02, 9.036596202793466E-02,
                             -1.374545432301129E-04, -1.027075301472321E-02, -4.195690627695180E-03]


def cpu_reference_shape(o, samples=5):
    """The reference's own benchmark shape on the C oracle: 10 particles in one simulation sharing
    one IAS15 step, 10 years, default forces, one thread, mean and standard deviation of 5 samples
    (reference unit_tests/benchmark/problem.c:60-95, 130-149).  It starts at t = 0 (JD 2451545.0)
    rather than t = 8416.5, which with 3652.5 days lies beyond the synthetic ephemeris' +-10 000 d."""
    reals = np.tile(np.array(REFERENCE_BENCHMARK_STATE), (10, 1))
    reals[:, 0] += np.arange(10) / 1e-10
    secs = []
    for _ in range(samples):
        ts = time.perf_counter()
        _, st = o.integrate_simulation(0.0, 3652.5, reals)
        secs.append(time.perf_counter() - ts)
    mean, std = float(np.mean(secs)), float(np.std(secs, ddof=1))
    return dict(seconds_mean=mean, seconds_std=std, samples=samples, particle_days_per_s=10 * 3652.5 / mean,
                steps=int(st.steps), evals=int(st.evals),
                note="synthetic ephemeris, t0 = 0; the reference reports 0.182 s on DE440 on its authors' laptop")


def cpu_baseline(paths, cfg, t0, t1, states, params, sample_s=3.0, samples=5):
    """The C oracle on the host cores over a bounded sample of the same workload: `samples` timed
    runs of the first m systems on all threads (mean and standard deviation, the reference's
    protocol), one run on one thread, and the reference's benchmark shape."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(*paths)
    cores, host = host_cpus()
    prm = (lambda sl: None) if params is None else (lambda sl: params[sl])
    ocfg = o.cfg(min_dt=cfg["min_dt"])
    probe = 16
    tp = time.perf_counter()
    o.integrate(t0[:probe], t1[:probe], states[:probe], prm(slice(0, probe)), nthreads=1, cfg=ocfg)
    per_sys = (time.perf_counter() - tp) / probe
    m = int(min(len(t0), max(cores, cores * sample_s / max(per_sys, 1e-6))))
    sel = slice(0, m)
    pdays = float(np.sum(np.abs(t1[sel] - t0[sel])))
    rates = []
    for k in range(samples):
        ts = time.perf_counter()
        rk = o.integrate(t0[sel], t1[sel], states[sel], prm(sel), nthreads=cores, cfg=ocfg)
        rates.append(pdays / (time.perf_counter() - ts))
        if k == 0:
            r = rk
    m1 = int(min(len(t0), max(1, sample_s / max(per_sys, 1e-6))))
    s1 = slice(0, m1)
    ts = time.perf_counter()
    o.integrate(t0[s1], t1[s1], states[s1], prm(s1), nthreads=1, cfg=ocfg)
    rate1 = float(np.sum(np.abs(t1[s1] - t0[s1]))) / (time.perf_counter() - ts)
    return dict(value=float(np.mean(rates)), std=float(np.std(rates, ddof=1)), samples=samples,
                unit="particle-days/s", cores=cores, kind="port", host=host,
                sample="first %d systems of the workload (%.0f particle-days) on %d threads (the process's "
                       "affinity set of %d CPUs%s), %d runs, %s"
                       % (m, pdays, cores, host["affinity"],
                          "" if host["cgroup_cpu_quota"] is None else ", cgroup quota %.1f CPUs" % host["cgroup_cpu_quota"],
                          samples, host["cpu_model"]),
                single_thread={"value": rate1, "sample": "first %d systems on 1 thread" % m1},
                reference_benchmark_shape=cpu_reference_shape(o)), r, m


def copy_calibration(mib):
    """One plain device-to-device copy of ``mib`` MiB (the runtime's copy kernel reads and writes
    that many bytes), for scaling the HBM traffic counters of the integration kernel measured in
    the same profiled process."""
    from assist_amd import batch
    n = mib * (1 << 20) // 8
    src = batch.DeviceArray(np.ones(n))
    dst = batch.DeviceArray(np.zeros(n))
    batch.synchronize()
    dst.copy_from(src)
    batch.synchronize()
    print("copy calibration: %d MiB read, %d MiB written" % (mib, mib), file=sys.stderr, flush=True)
    src.free()
    dst.free()


def measured_traffic(config, form, n):
    """HBM bytes per launch of the integration kernel from the committed counter profile
    (profiles/ias15_kernel_traffic.md), or None when this workload and form were not profiled."""
    path = os.path.join(ROOT, "profiles", "ias15_kernel_traffic.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    for line in txt.splitlines():
        if line.startswith("traffic_json:"):
            for row in json.loads(line.split(":", 1)[1]):
                if row["config"] == config and row["form"] == form and row["systems"] == n:
                    return row["bytes_per_launch"]
    return None


FORMS = {
    "fp64": dict(precision=0, learned=False),
    "mixed": dict(precision=1, learned=False),
    "fp64+learned": dict(precision=0, learned=True),
    "mixed+learned": dict(precision=1, learned=True),
}


def run_form(form, args, cfg, ephem, bufs, n, nv, step_model):
    """Warmup, then `steps` timed passes of one form over the batch; returns its measurements."""
    from assist_amd import batch, learned
    spec = FORMS[form]
    bc = batch.BatchConfig.default(precision=spec["precision"], lanes_per_system=args.lanes, block_size=args.block,
                                   waves_per_simd=args.waves, min_dt=cfg["min_dt"], max_steps=200000,
                                   use_graph=1 if args.launch == "graph" else 0)
    d_t0, d_t1, d_init, d_states, d_params, d_status, d_steps, d_evals = bufs

    # the learned start's outputs are written in full by every prediction; one pair of buffers
    # keeps the device pointers of the captured launch the same from step to step
    start_bufs = (batch.DeviceArray.empty(n), batch.DeviceArray.empty((n, 21))) if spec["learned"] else None

    def one_step():
        d_states.copy_from(d_init)
        if spec["learned"]:
            d_dt0, d_bw = learned.predict_start(ephem, step_model, n, nv - 1, d_t0, d_t1, d_init, min_dt=cfg["min_dt"],
                                                out=start_bufs)
            return learned.integrate_device_learned(ephem, n, nv - 1, d_t0, d_t1, d_states, d_params, d_status, d_steps,
                                                    d_evals, d_dt0, d_bw, config=bc)
        return batch.integrate_device(ephem, n, nv - 1, d_t0, d_t1, d_states, d_params, d_status, d_steps, d_evals,
                                      config=bc)

    for _ in range(args.warmup):
        one_step()
        print("%s: warmup step done" % form, file=sys.stderr, flush=True)
    batch.synchronize()
    kernel_ms = []
    ts = time.perf_counter()
    for k in range(args.steps):
        st = one_step()
        kernel_ms.append(st.kernel_ms)
        print("%s: step %d: integration kernel %.1f ms" % (form, k, st.kernel_ms), file=sys.stderr, flush=True)
    batch.synchronize()
    elapsed = time.perf_counter() - ts
    return dict(form=form, elapsed=elapsed, kernel_ms=float(np.mean(kernel_ms)), stats=st,
                status=d_status.download(), steps=d_steps.download(), evals=d_evals.download(),
                states=d_states.download(), mixed=spec["precision"] == 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--form", default="mixed+learned", choices=sorted(FORMS),
                    help="the form reported as `value`; the plain FP64 form is always measured beside it")
    ap.add_argument("--n", type=int, default=0, help="override the number of systems")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0, help="kernel variant: 1 or 2 waves per SIMD (0 = choose)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--launch", default="graph", choices=["graph", "direct"],
                    help="integration launches: a HIP graph captured in the warmup and replayed, or direct")
    ap.add_argument("--copy-kernel-mib", type=int, default=0,
                    help="also copy this many MiB device-to-device before the timed region, to calibrate the "
                         "FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: every system is independent, so N GPUs would each take 1/N of the "
              "batch; this build measures the single-GPU shard (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import assist_amd
    from assist_amd import batch, learned, synthetic
    if batch.gpu_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_kernel_mib:
        copy_calibration(args.copy_kernel_mib)
    cfg = dict(CONFIGS[args.config])
    if args.n:
        cfg["n"] = args.n
    paths = synthetic.ensure_synthetic_ephemeris(0)
    ephem = assist_amd.Ephem(*paths)
    model = synthetic.SyntheticModel(0)
    t0, t1, states, params = make_workload(cfg, model)
    n, nv = states.shape[0], states.shape[1]
    step_model = None
    if FORMS[args.form]["learned"]:
        path = learned.default_model_path()
        if not os.path.exists(path):
            print("the learned forms need the trained step model %s (python -m assist_amd.train_step_model)" % path,
                  file=sys.stderr)
            sys.exit(1)
        step_model = learned.StepModel.load(path, max_batch=65536)
    bufs = (batch.DeviceArray(t0), batch.DeviceArray(t1), batch.DeviceArray(states), batch.DeviceArray(states),
            batch.DeviceArray(params) if params is not None else None, batch.DeviceArray(np.zeros(n, dtype=np.int32)),
            batch.DeviceArray(np.zeros(n, dtype=np.int64)), batch.DeviceArray(np.zeros(n, dtype=np.int64)))
    pdays = float(np.sum(np.abs(t1 - t0)))
    oc = opcounts()
    ng = None if params is None else np.any(params[:, 0] != 0, axis=-1)
    forms = [args.form] if args.form == "fp64" else ["fp64", args.form]
    res = {f: run_form(f, args, cfg, ephem, bufs, n, nv, step_model) for f in forms}
    main_res, plain = res[args.form], res["fp64"]
    value = pdays * args.steps / main_res["elapsed"]
    plain_value = pdays * args.steps / plain["elapsed"]
    st = main_res["stats"]
    # (the device entry points report no rejection total; a rejected attempt's 7 fresh evaluations
    # then count as cached ones, so the count is a lower bound; these workloads reject rarely)
    f64, f32 = algorithmic_flops(oc, nv, ng, main_res["mixed"], main_res["evals"], main_res["steps"], st.rejected,
                                 cached=st.lanes_per_system == 8)
    kms = main_res["kernel_ms"]
    ach64 = f64 / (kms * 1e-3) / 1e12
    ach32 = f32 / (kms * 1e-3) / 1e12
    frac = ach64 / FP64_PEAK_TFLOPS + ach32 / FP32_PEAK_TFLOPS
    traffic = measured_traffic(args.config, args.form, n)
    dpos = np.linalg.norm(main_res["states"][:, 0, :3] - plain["states"][:, 0, :3], axis=-1) * 149597870700.0
    result = {
        "metric": "particle-days/sec (plain FP64 vs mixed+learned) at <=1 m 10-yr position error",
        "value": value,
        "unit": "particle-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if main_res["mixed"] else "f64",
        "data": "synthetic ephemeris (seed 0) and synthetic particle states",
        "config": {"workload": cfg["workload"], "systems": n, "vectors_per_system": nv, "form": args.form,
                   "particle_days_per_step": pdays, "lanes_per_system": st.lanes_per_system,
                   "block_size": st.block_size, "grid_blocks": st.grid_blocks, "waves_per_simd": st.waves_per_simd,
                   "launch": args.launch},
        "plain_fp64": {"value": plain_value, "ms_per_step": plain["elapsed"] / args.steps * 1e3,
                       "kernel_ms": plain["kernel_ms"], "evals_total": int(plain["evals"].sum())},
        "speedup_vs_plain_fp64": value / plain_value,
        "roofline": {"bound": "fp64-valu", "achieved": ach64 + ach32, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": frac, "traffic": traffic,
                     "traffic_source": None if traffic is None else
                     "committed counter profile profiles/ias15_kernel_traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                     "WRITE_SIZE of this workload and form, scaled by a copy kernel), not measured in this run",
                     "kernel": "ias15_kernel", "kernel_ms": kms, "flops_fp64_per_launch": f64,
                     "flops_fp32_per_launch": f32, "achieved_fp64": ach64, "achieved_fp32": ach32,
                     "hbm_gbs": None if traffic is None else traffic / (kms * 1e-3) / 1e9},
        "integration": {"steps_total": int(main_res["steps"].sum()), "evals_total": int(main_res["evals"].sum()),
                        "status_nonzero": int((main_res["status"] != 0).sum()),
                        "max_position_diff_to_plain_fp64_m": float(dpos.max()),
                        "p99_position_diff_to_plain_fp64_m": float(np.percentile(dpos, 99))},
    }
    if not args.no_cpu_baseline:
        cb, r, m = cpu_baseline(paths, cfg, t0, t1, states, params)
        dcpu = np.linalg.norm(plain["states"][:m, 0, :3] - r["states"][:, 0, :3], axis=-1) * 149597870700.0
        result["cpu_baseline"] = cb
        result["parity_vs_oracle_sample"] = {"systems": m, "form": "fp64", "max_position_diff_m": float(dcpu.max())}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
