"""Benchmark of the force-and-step hot path on one MI355X.

Default workload (the direct-sum scale-up configuration): Plummer sphere,
N=65536 equal masses (M=1, G=1, scale radius 1, seeded), softening 0.005,
leapfrog dt=5e-4, REB_GRAVITY_BASIC in FP64.  A "step" is one leapfrog step
(drift, direct force on all N(N-1) ordered pairs, kick, drift).  The headline
`value` is the whole-run pair-interaction rate (Gpair/s) with the state
resident in HBM; the JSON line also carries the COMPENSATED and split-precision
modes, the accuracy of each mode against the bit-exact FP64 path, the roofline
of the dominant kernel (timed with HIP events on its own stream) and the
reference's OpenMP CPU path timed on this host (oracle/_ref/ref_bench).

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu] [--quick]
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

FP64_PEAK_TFLOPS = 78.6        # MI355X dense FP64 vector
# Algorithmic FLOP per ordered pair of the one-sided kernels (the reference's
# per-target loop, gravity.c:215-232 and :315-341; sqrt = div = 1).
FLOP_PER_PAIR = {"fp64": 20, "compensated": 28, "split": 28, "exact": 20}
# Algorithmic FLOP per unordered pair of the pair-symmetric kernel: the
# reference's serial loop, gravity.c:195-214 (3 sub, 3 mul + 3 add for r^2
# with softening, sqrt, 2 mul for r^3, div, 2 mul by the masses, 6 mul + 6 add
# into both bodies) = 27; split precision 27 + 8 (the FP32 part counted apart
# as in the one-sided kernel's 28 = 15 FP64 + 13 FP32 per ordered pair).
FLOP_PER_SYM_PAIR = {"fp64": 27, "split": 35}


def profiled(what, kernel, workload_tag):
    """A per-launch figure of `kernel` on this workload recorded in the
    committed rocprofv3 summary profiles/<workload_tag>.md on a line
    `<what> <kernel>: <value>`: `traffic_bytes_per_launch` (FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for a device-to-device
    copy of known size) or `hw_flops_per_launch` (FP64 VALU FLOPs, FMA = 2).
    (None, None) when not recorded."""
    path = os.path.join(ROOT, "profiles", workload_tag + ".md")
    if not os.path.exists(path):
        return None, None
    key = "%s %s:" % (what, kernel)
    with open(path) as f:
        for line in f:
            line = line.strip().strip("`")
            if line.startswith(key):
                return float(line[len(key):].split()[0]), os.path.relpath(path, ROOT)
    return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def plummer_sim(amd, n, seed, soft, dt, precision, gravity="basic"):
    sim = amd.Sim()
    sim.s.rand_seed = seed
    sim.s.softening = soft
    sim.s.G = 1.0
    sim.add_plummer(n, 1.0, 1.0)
    sim.move_to_com()
    sim.set_integrator("leapfrog")
    sim.s.dt = dt
    sim.s.gravity = amd.GRAVITY[gravity]
    sim.precision = precision
    return sim


def run_mode(amd, args, name, precision, gravity, steps, warmup):
    sim = plummer_sim(amd, args.n, args.seed, args.softening, args.dt, precision, gravity)
    L = amd.lib()
    e0 = sim.energy()
    L.reb_amd_set_resident(sim.ptr, 1)
    sim.steps(warmup)
    L.reb_amd_synchronize_device(sim.ptr)
    L.reb_amd_set_profiling(sim.ptr, 1)
    t0 = time.perf_counter()
    sim.steps(steps)
    L.reb_amd_synchronize_device(sim.ptr)
    t1 = time.perf_counter()
    import ctypes
    ktot = ctypes.c_double(0.0)
    kn = ctypes.c_int(0)
    L.reb_amd_kernel_time(sim.ptr, ctypes.byref(ktot), ctypes.byref(kn))
    kpairs = L.reb_amd_kernel_pairs(sim.ptr)
    L.reb_amd_set_profiling(sim.ptr, 0)
    L.reb_amd_set_resident(sim.ptr, 0)
    e1 = sim.energy()
    sim.check()
    n = args.n
    pairs = float(n) * float(n - 1)
    wall = t1 - t0
    kernel_ms = ktot.value / max(kn.value, 1)
    # ordered pairs the timed kernel evaluates per launch: N(N-1) for the
    # one-sided kernels; for the pair-symmetric one the pairs between
    # superblocks, each unordered pair once (counted here as two ordered)
    kp = kpairs / max(kn.value, 1)
    sym = kp < pairs
    if sym:
        units, flop_unit, unit = kp / 2.0, FLOP_PER_SYM_PAIR[name], "unordered pair"
        kernel = "k_direct_sym<%s>" % ("FM_SPLIT" if name == "split" else "FM_FP64")
    else:
        units, flop_unit, unit = kp, FLOP_PER_PAIR[name], "ordered pair"
        kernel = "k_exact" if name == "exact" else "k_direct_fast<%s>" % (
            {"fp64": "FM_FP64", "split": "FM_SPLIT", "compensated": "FM_COMP"}[name])
    res = {
        "precision": precision, "gravity": gravity,
        "steps": steps, "seconds": wall, "ms_per_step": 1e3 * wall / steps,
        "gpairs_per_s": pairs * steps / wall / 1e9,
        "kernel": kernel, "kernel_ms": kernel_ms, "kernel_launches": kn.value,
        "kernel_ordered_pairs_per_launch": kp,
        "kernel_gpairs_per_s": kp / (kernel_ms * 1e-3) / 1e9,
        "work_unit": unit, "units_per_launch": units, "flop_per_unit": flop_unit,
        "algorithmic_tflops": flop_unit * units / (kernel_ms * 1e-3) / 1e12,
        "rel_energy_drift": (e1 - e0) / abs(e0),
    }
    log("%-12s %8.2f ms/step  %8.1f Gpair/s (kernel %.3f ms, %.1f TFLOP/s algorithmic)" % (
        name, res["ms_per_step"], res["gpairs_per_s"], kernel_ms, res["algorithmic_tflops"]))
    return res, sim


def accuracy(amd, args):
    """Per-particle error of each mode's initial accelerations, on sampled
    rows, against (1) the bit-exact FP64 path (identical to the reference's
    BASIC sum) and (2) the exactly summed reference forces: the reference's
    per-pair terms (gravity.c:222-230) summed in extended precision.  The
    normalised error divides by sum_j |a_ij| (conditioning of the sum)."""
    def acc_of(prec, gravity="basic"):
        sim = plummer_sim(amd, args.n, args.seed, args.softening, args.dt, prec, gravity)
        a = sim.update_acceleration()
        st = sim.state()
        sim.free()
        return a, st

    ref, st = acc_of("exact")
    x = st[:, :3]
    m = st[:, 9]
    rows = np.sort(np.random.default_rng(0).choice(args.n, min(args.n, args.acc_rows), replace=False))
    truth = np.zeros((len(rows), 3))
    cond = np.zeros(len(rows))
    soft2 = args.softening * args.softening
    for b in range(0, len(rows), 64):
        r = rows[b:b + 64]
        d = x[r, None, :] - x[None, :, :]
        rr = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2] + soft2)
        with np.errstate(divide="ignore", invalid="ignore"):
            pre = -1.0 / (rr * rr * rr) * m[None, :]
        pre[np.arange(len(r)), r] = 0.0
        t = pre[..., None] * d
        truth[b:b + 64] = t.astype(np.longdouble).sum(1).astype(np.float64)
        cond[b:b + 64] = np.sqrt((t * t).sum(2)).sum(1)
    tn = np.sqrt((truth ** 2).sum(1))

    def stats(a):
        da = np.sqrt(((a[rows] - truth) ** 2).sum(1))
        return {"rel_max": float((da / tn).max()), "rel_p99": float(np.quantile(da / tn, 0.99)),
                "rel_median": float(np.median(da / tn)), "normalised_max": float((da / cond).max())}

    out = {"exact": {"vs_exact_sum": stats(ref)}}
    log("accuracy %-12s vs exactly summed: rel max %.2e p99 %.2e | normalised max %.2e" % (
        "exact(ref)", out["exact"]["vs_exact_sum"]["rel_max"], out["exact"]["vs_exact_sum"]["rel_p99"],
        out["exact"]["vs_exact_sum"]["normalised_max"]))
    for name, prec, grav in (("fp64", "fp64", "basic"), ("compensated", "fp64", "compensated"),
                             ("split", "split", "basic")):
        a, _ = acc_of(prec, grav)
        da = np.sqrt(((a[rows] - ref[rows]) ** 2).sum(1))
        rel = da / np.sqrt((ref[rows] ** 2).sum(1))
        out[name] = {"vs_reference": {"rel_max": float(rel.max()), "rel_p99": float(np.quantile(rel, 0.99)),
                                      "rel_median": float(np.median(rel)), "normalised_max": float((da / cond).max())},
                     "vs_exact_sum": stats(a), "sampled_rows": int(len(rows))}
        v = out[name]
        log("accuracy %-12s vs reference: rel max %.2e p99 %.2e | vs exactly summed: rel max %.2e p99 %.2e "
            "normalised max %.2e" % (name, v["vs_reference"]["rel_max"], v["vs_reference"]["rel_p99"],
                                     v["vs_exact_sum"]["rel_max"], v["vs_exact_sum"]["rel_p99"],
                                     v["vs_exact_sum"]["normalised_max"]))
    return out


def host_cpu():
    """Model name, physical cores and logical CPUs of this host (/proc/cpuinfo),
    and the CPUs this process may run on."""
    model, cores, logical = None, set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            phys = core = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "processor":
                    logical += 1
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "physical_cores": len(cores) or None, "logical_cpus": logical or None,
            "usable_cpus": usable}


def cpu_baseline(args):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(exe):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    steps = args.cpu_steps
    cmd = [exe, "plummer", str(args.n), str(steps), repr(args.softening), repr(args.dt), "basic", str(args.seed)]
    log("cpu baseline:", " ".join(cmd), "threads", threads)
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        log(out.stderr)
        return None
    res = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": res["pairs_per_s"] / 1e9, "unit": "Gpair/s", "cores": res["threads"], "kind": "reference",
            "host": host_cpu(),
            "sample": "reference OpenMP build (gravity.c:215-232 per-target loop), Plummer N=%d, %d leapfrog "
                      "steps after 1 warm-up step, %.1f s" % (args.n, steps, res["seconds"])}


def ensemble_leg(amd, args):
    """Batched WHFast ensemble: 16384 systems (star + K planets, K uniform in
    {2..8}), democratic heliocentric, safe_mode 0, dt = inner period / 20,
    `ens_steps` steps in one request (state resident in HBM)."""
    import ctypes
    nb, st, dt = amd.planetary_systems(args.ens_systems, args.ens_seed, 2, 8)
    ens = amd.Ensemble(nb, st, dt, safe_mode=0)
    e0 = ens.energy()
    tune_ms, shape = ens.tune(probe_steps=args.ens_probe)
    log("ensemble: %d systems, %d bodies, %d tasks; tuned launch %s (%.2f ms / %d steps)" % (
        len(nb), int(nb.sum()), shape["tasks"], shape, tune_ms, args.ens_probe))
    L = amd.lib()
    hip = ctypes.CDLL("libamdhip64.so")
    t0 = time.perf_counter()
    ens.steps(args.ens_steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    e1 = ens.energy()
    rel = np.abs((e1 - e0) / e0)
    flags = ens.flags()
    sys_steps = float(len(nb)) * args.ens_steps
    res = {
        "metric": "ensemble systems*steps/s (WHFast DH, 16384 systems, K in 2..8)",
        "value": sys_steps / wall, "unit": "system-steps/s", "seconds": wall, "steps": args.ens_steps,
        "systems": int(len(nb)), "bodies": int(nb.sum()), "launch": shape,
        "us_per_step": 1e6 * wall / args.ens_steps,
        "rel_energy_error": {"median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)),
                             "max": float(rel.max()), "mean": float(rel.mean())},
        "kepler_period_warnings": int(((flags >> 1) & 1).sum()),
        # systems whose Kepler solves took the quartic iteration / fell back
        # to bisection at least once during the request
        "kepler_quartic_systems": int(((flags >> 3) & 1).sum()),
        "kepler_bisection_systems": int(((flags >> 4) & 1).sum()),
    }
    log("ensemble: %.3e system-steps/s (%.2f us/step, %.2f s), rel energy error median %.2e mean %.2e; "
        "systems with quartic Kepler solves %d, with bisection %d" % (
            res["value"], res["us_per_step"], wall, np.median(rel), rel.mean(), res["kepler_quartic_systems"],
            res["kepler_bisection_systems"]))
    del hip, L
    if not args.no_cpu:
        cpu = ensemble_cpu(args)
        if cpu is not None:
            # the same systems on the GPU for the CPU's step count: energy errors must match
            ens2 = amd.Ensemble(nb, st, dt, safe_mode=0)
            ens2.steps(args.ens_cpu_steps)
            r2 = np.abs((ens2.energy() - e0) / e0)
            ens2.free()
            cpu["gpu_at_same_steps"] = {"max_rel_energy_error": float(r2.max()),
                                        "mean_rel_energy_error": float(r2.mean())}
            cpu["energy_error_matches"] = bool(
                "%.3e" % r2.max() == "%.3e" % cpu["max_rel_energy_error"] and
                "%.3e" % r2.mean() == "%.3e" % cpu["mean_rel_energy_error"])
            res["cpu_baseline"] = cpu
            res["speedup_vs_cpu"] = res["value"] / cpu["value"]
            # the full horizon on a sample: the first `ens_full_systems`
            # systems, integrated for all `ens_steps` steps by the reference
            full = ensemble_cpu(args, systems=args.ens_full_systems, steps=args.ens_steps)
            if full is not None:
                k = args.ens_full_systems
                gmax, gmean = float(rel[:k].max()), float(rel[:k].mean())
                full["gpu"] = {"max_rel_energy_error": gmax, "mean_rel_energy_error": gmean}
                full["energy_error_matches"] = bool("%.3e" % gmax == "%.3e" % full["max_rel_energy_error"] and
                                                    "%.3e" % gmean == "%.3e" % full["mean_rel_energy_error"])
                res["full_horizon_check"] = full
                log("ensemble full horizon (%d systems, %d steps): cpu max %.3e mean %.3e, gpu max %.3e mean "
                    "%.3e; match: %s" % (k, args.ens_steps, full["max_rel_energy_error"],
                                         full["mean_rel_energy_error"], gmax, gmean, full["energy_error_matches"]))
            # algorithmic work per system-step, counted by the CPU restatement
            # integrating the first 512 systems for 200 steps
            sys.path.insert(0, ROOT)
            import oracle
            k = min(512, len(nb))
            fl = oracle.whfast_flops_per_system_step(nb[:k], st[:int(nb[:k].sum())], dt[:k], 200)
            achieved = fl * res["value"] / 1e12
            res["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                               "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "k_whfast_dh",
                               "flop_per_system_step": fl,
                               "note": "latency-bound: 1.3 wavefronts per SIMD at 16384 systems"}
            log("ensemble cpu: %.3e system-steps/s on %d threads -> GPU %.0fx; energy errors match: %s" % (
                cpu["value"], cpu["cores"], res["speedup_vs_cpu"], cpu["energy_error_matches"]))
    ens.free()
    return res


def ensemble_cpu(args, systems=None, steps=None):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(exe):
        return None
    systems = args.ens_systems if systems is None else systems
    steps = args.ens_cpu_steps if steps is None else steps
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [exe, "ensemble", str(systems), str(steps), str(args.ens_seed), "0", "2", "8"]
    log("ensemble cpu baseline:", " ".join(cmd), "threads", threads)
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        log(out.stderr)
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["system_steps_per_s"], "unit": "system-steps/s", "cores": r["threads"], "kind": "reference",
            "sample": "reference OpenMP build, omp parallel for over the same %d systems (one reb_simulation "
                      "each, reb_simulation_steps), %d steps, %.1f s" % (systems, steps, r["seconds"]),
            "max_rel_energy_error": r["max_rel_energy_error"], "mean_rel_energy_error": r["mean_rel_energy_error"]}


def megno_leg(amd, args):
    """MEGNO map: the ensemble's systems (seed `ens_seed`) in Jacobi
    coordinates with safe_mode 1, each with reb_simulation_init_megno_seed(r, s)
    drawn for system s, `megno_steps` steps in one request; the first
    systems' <Y> compared with the reference's at the CPU sample's step count."""
    nb, st, dt = amd.planetary_systems(args.megno_systems, args.ens_seed, 2, 8)
    first = np.concatenate([[0], np.cumsum(nb)[:-1]])
    var0 = np.zeros((int(nb.sum()), 10))
    for s in range(len(nb)):
        sim = amd.Sim()
        sim.set_state(st[first[s]:first[s] + nb[s]])
        sim.init_megno(seed=s)
        v = sim.var_particles()
        var0[first[s]:first[s] + nb[s]] = np.stack([v[f] for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay",
                                                                      "az", "m")], axis=1)
        sim.free()

    def make():
        e = amd.Ensemble(nb, st, dt, safe_mode=1, coordinates="jacobi")
        e.init_megno(var0)
        return e

    ens = make()
    tune_ms, shape = ens.tune(probe_steps=args.ens_probe)
    t0 = time.perf_counter()
    ens.steps(args.megno_steps)
    wall = time.perf_counter() - t0
    meg, lya = ens.megno()
    ens.free()
    res = {"metric": "MEGNO map system*steps/s (WHFast Jacobi + variational set, %d systems, K in 2..8)" %
                     len(nb),
           "value": float(len(nb)) * args.megno_steps / wall, "unit": "system-steps/s", "seconds": wall,
           "steps": args.megno_steps, "systems": int(len(nb)), "launch": shape,
           "us_per_step": 1e6 * wall / args.megno_steps,
           # systems whose orbits went non-finite have no MEGNO (as in the reference)
           "megno": {"median": float(np.nanmedian(meg)), "p90": float(np.nanquantile(meg, 0.9)),
                     "chaotic_fraction_Y_gt_4": float((meg > 4.).mean()),
                     "non_finite_systems": int((~np.isfinite(meg)).sum())}}
    log("megno: %.3e system-steps/s (%.2f us/step), <Y> median %.3f, %.1f%% with <Y> > 4, %d non-finite" % (
        res["value"], res["us_per_step"], res["megno"]["median"], 100. * res["megno"]["chaotic_fraction_Y_gt_4"],
        res["megno"]["non_finite_systems"]))
    # the reference's serial build (its sums are the GPU's), OpenMP across systems
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench_serial")
    if not args.no_cpu and os.path.exists(exe):
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        cmd = [exe, "megno", str(len(nb)), str(args.megno_cpu_steps), str(args.ens_seed), "2", "8"]
        log("megno cpu baseline:", " ".join(cmd), "threads", threads)
        out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        if out.returncode == 0:
            r = json.loads(out.stdout.strip().splitlines()[-1])
            e2 = make()
            e2.steps(args.megno_cpu_steps)
            m2, _ = e2.megno()
            e2.free()
            same = ["%.17g" % a for a in m2[:len(r["megno_first"])]] == ["%.17g" % b for b in r["megno_first"]]
            res["cpu_baseline"] = {
                "value": r["system_steps_per_s"], "unit": "system-steps/s", "cores": r["threads"],
                "kind": "reference",
                "sample": "reference serial build, omp parallel for over the same %d systems with init_megno_seed, "
                          "%d steps, %.1f s" % (
                    len(nb), args.megno_cpu_steps, r["seconds"]),
                "megno_first_systems_bitwise_equal": same}
            res["speedup_vs_cpu"] = res["value"] / r["system_steps_per_s"]
            log("megno cpu: %.3e system-steps/s on %d threads -> GPU %.0fx; first systems' <Y> bitwise equal: %s" % (
                r["system_steps_per_s"], r["threads"], res["speedup_vs_cpu"], same))
        else:
            log(out.stderr)
    return res


def serving_jobs(args):
    """Poisson arrivals (seeded): 70% planetary systems (star + K planets, K
    uniform in {2..10}, generated as the ensemble's) for log-uniform 1e3..1e5
    inner orbits (20 steps per orbit), 30% Plummer clusters N uniform in
    {256..8192} for 100..2000 leapfrog steps in split precision."""
    rng = np.random.default_rng(args.serve_seed)
    t = 0.0
    jobs = []
    for k in range(args.serve_jobs):
        t += rng.exponential(1.0 / args.serve_rate)
        seed = 100000 + k
        if rng.random() < 0.7:
            orbits = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
            jobs.append({"kind": "whfast", "t": t, "seed": seed, "steps": int(round(20 * orbits))})
        else:
            jobs.append({"kind": "nbody", "t": t, "seed": seed, "N": int(rng.integers(256, 8193)),
                         "steps": int(rng.integers(100, 2001)), "softening": 0.01, "dt": 1e-3})
    return jobs


def serving_leg(amd, args):
    jobs = serving_jobs(args)
    # inputs are made before the clock starts
    for j in jobs:
        if j["kind"] == "whfast":
            nb, st, dt = amd.planetary_systems(1, j["seed"], 2, 10)
            j["state"], j["dt"] = st, float(dt[0])
        else:
            sim = amd.Sim()
            sim.s.rand_seed = j["seed"]
            sim.add_plummer(j["N"], 1.0, 1.0)
            sim.move_to_com()
            j["state"] = sim.state()
            sim.free()
    srv = amd.Server(memory_budget=args.serve_budget, max_nbody_jobs=args.serve_workers,
                     tick_steps=args.serve_tick)
    t0 = time.perf_counter()
    ids = []
    for j in jobs:
        delay = j["t"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        if j["kind"] == "whfast":
            ids.append(srv.submit_whfast(j["state"], j["dt"], steps=j["steps"]))
        else:
            ids.append(srv.submit_nbody(j["state"], softening=j["softening"], dt=j["dt"], steps=j["steps"],
                                        precision="split"))
    lat, derr = {"whfast": [], "nbody": []}, {"whfast": [], "nbody": []}
    t_first, t_last = None, 0.0
    for j, jid in zip(jobs, ids):
        if srv.wait(jid, 600) != "done":
            raise RuntimeError("serving job failed: " + srv.job_error(jid))
        _, (e0, e1), (ts, ta, td) = srv.result(jid)
        lat[j["kind"]].append(td - ts)
        derr[j["kind"]].append(abs((e1 - e0) / e0))
        t_first = ts if t_first is None else min(t_first, ts)
        t_last = max(t_last, td)
    stats = srv.stats()
    srv.close()
    all_lat = np.array(lat["whfast"] + lat["nbody"])
    span = t_last - t_first
    res = {
        "metric": "serving jobs/s and p50/p99 latency (70% WHFast K 2..10, 30% Plummer N 256..8192 split)",
        "jobs": len(jobs), "jobs_per_s": len(jobs) / span, "span_s": span,
        "arrival_rate": args.serve_rate,
        "latency_s": {"p50": float(np.quantile(all_lat, 0.5)), "p99": float(np.quantile(all_lat, 0.99)),
                      "max": float(all_lat.max())},
        "latency_by_kind_s": {k: {"p50": float(np.quantile(v, 0.5)), "p99": float(np.quantile(v, 0.99)), "n": len(v)}
                              for k, v in lat.items() if v},
        "rel_energy_error": {k: {"median": float(np.median(v)), "max": float(np.max(v))} for k, v in derr.items() if v},
        "device_pool_busy_s": stats["pool_busy_s"], "pool_ticks": int(stats["ticks"]),
        "peak_bytes": stats["peak_bytes"], "budget": stats["budget"],
    }
    log("serving: %d jobs in %.2f s -> %.2f jobs/s; latency p50 %.3f s p99 %.3f s" % (
        len(jobs), span, res["jobs_per_s"], res["latency_s"]["p50"], res["latency_s"]["p99"]))
    if not args.no_cpu:
        res["cpu_baseline"] = serving_cpu(args, jobs)
        if res["cpu_baseline"]:
            res["speedup_vs_cpu"] = res["jobs_per_s"] / res["cpu_baseline"]["value"]
            log("serving cpu: %.3f jobs/s projected on %d threads" % (res["cpu_baseline"]["value"],
                                                                      res["cpu_baseline"]["cores"]))
    return res


def serving_cpu(args, jobs):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(exe):
        return None
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".jobs", delete=False) as f:
        for j in jobs:
            if j["kind"] == "whfast":
                f.write("whfast %d %d\n" % (j["seed"], j["steps"]))
            else:
                f.write("nbody %d %d %d %r %r\n" % (j["seed"], j["N"], j["steps"], j["softening"], j["dt"]))
        path = f.name
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [exe, "serve", path, repr(float(args.serve_cpu_scale))]
    log("serving cpu baseline:", " ".join(cmd), "threads", threads)
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    os.unlink(path)
    if out.returncode != 0:
        log(out.stderr)
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["jobs_per_s"], "unit": "jobs/s", "cores": r["threads"], "kind": "reference",
            "sample": "reference OpenMP build, the same %d jobs (N-body in FP64 BASIC: the reference has no split "
                      "mode), one job per thread, every job for 1/%g of its steps (%.1f s), projected to the full "
                      "steps" % (r["jobs"], r["scale"], r["seconds"])}


def small_leg(amd, args):
    """The plain method at the reference's CPU-sized case: Plummer N=1024,
    softening 0.01, leapfrog dt=1e-3, 1000 steps, BASIC in EXACT precision (bit-
    identical to the reference, so the energy drift is the reference's)."""
    L = amd.lib()

    def run(precision):
        sim = plummer_sim(amd, 1024, 1, 0.01, 1e-3, precision)
        e0 = sim.energy()
        L.reb_amd_set_resident(sim.ptr, 1)
        sim.steps(10)
        L.reb_amd_synchronize_device(sim.ptr)
        t0 = time.perf_counter()
        sim.steps(1000)
        L.reb_amd_synchronize_device(sim.ptr)
        wall = time.perf_counter() - t0
        L.reb_amd_set_resident(sim.ptr, 0)
        e1 = sim.energy()
        sim.free()
        return wall, (e1 - e0) / abs(e0)

    wall, drift = run("exact")
    wall64, drift64 = run("fp64")
    pairs = 1024.0 * 1023.0
    res = {"metric": "Plummer N=1024 leapfrog steps/s (EXACT, bit-identical)", "value": 1000 / wall,
           "unit": "steps/s", "us_per_step": 1e6 * wall / 1000, "gpairs_per_s": pairs * 1000 / wall / 1e9,
           "rel_energy_drift_1010_steps": drift,
           "fp64": {"value": 1000 / wall64, "unit": "steps/s", "us_per_step": 1e6 * wall64 / 1000,
                    "rel_energy_drift_1010_steps": drift64}}
    log("small: N=1024 EXACT %.0f steps/s (%.1f us/step), FP64 %.0f steps/s" % (
        res["value"], res["us_per_step"], 1000 / wall64))
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not args.no_cpu and os.path.exists(exe):
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        out = subprocess.run([exe, "plummer", "1024", "1000", "0.01", "0.001", "basic", "1"], env=env,
                             capture_output=True, text=True, timeout=600)
        if out.returncode == 0:
            r = json.loads(out.stdout.strip().splitlines()[-1])
            res["cpu_baseline"] = {"value": r["steps"] / r["seconds"], "unit": "steps/s", "cores": r["threads"],
                                   "kind": "reference",
                                   "sample": "reference OpenMP build, the same case, 1000 steps after 1 warm-up step"}
            res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
            log("small cpu: %.0f steps/s on %d threads" % (res["cpu_baseline"]["value"], r["threads"]))
    return res


def sheet_leg(amd, args):
    """Tree gravity + tree collisions: shearing sheet of N=200000 hard spheres
    (10 x 10 root boxes of 136 radii: optical depth ~1), SEI, 2 ghost boxes,
    opening angle 0.5, restitution 0.5; `sheet_steps` timed steps."""
    import ctypes
    L = amd.lib()
    t_gen = time.perf_counter()
    sim = amd.shearing_sheet(args.sheet_n, args.sheet_root, args.sheet_nroot, args.sheet_seed)
    log("sheet: N=%d generated in %.1f s" % (args.sheet_n, time.perf_counter() - t_gen))
    L.reb_amd_set_resident(sim.ptr, 1)
    sim.steps(1)                                   # warm-up (allocations, first tree)
    L.reb_amd_synchronize_device(sim.ptr)
    # the CPU baseline's window: the steps right after the warm-up step (the
    # densest collision steps), timed alone, then the rest of the run
    n0 = sim.s.collisions_log_n
    window = max(1, min(args.sheet_cpu_steps, args.sheet_steps))
    t0 = time.perf_counter()
    sim.steps(window)
    L.reb_amd_synchronize_device(sim.ptr)
    wall_window = time.perf_counter() - t0
    ncoll_window = sim.s.collisions_log_n - n0
    if args.sheet_steps > window:
        sim.steps(args.sheet_steps - window)
    L.reb_amd_synchronize_device(sim.ptr)
    wall = time.perf_counter() - t0
    ncoll = sim.s.collisions_log_n - n0
    inter = L.reb_amd_tree_interactions(sim.ptr)
    L.reb_amd_set_resident(sim.ptr, 0)
    sim.check()
    res = {"metric": "shearing sheet steps/s (tree gravity + tree collisions, N=%d)" % args.sheet_n,
           "value": args.sheet_steps / wall, "unit": "steps/s", "steps": args.sheet_steps, "seconds": wall,
           "ms_per_step": 1e3 * wall / args.sheet_steps, "collisions_per_step": ncoll / args.sheet_steps,
           "tree_interactions_last_step": int(inter),
           "window": {"steps": window, "steps_per_s": window / wall_window,
                      "collisions_per_step": ncoll_window / window,
                      "note": "steps 2..%d, the CPU baseline's window" % (1 + window)},
           "config": {"N": args.sheet_n, "root_size": args.sheet_root, "n_root": args.sheet_nroot,
                      "ghost_boxes": 2, "opening_angle2": 0.25, "seed": args.sheet_seed}}
    log("sheet: %.2f steps/s (%.1f ms/step), %.0f collisions/step, %d tree interactions/step" % (
        res["value"], res["ms_per_step"], res["collisions_per_step"], inter))
    sim.free()
    if not args.no_cpu:
        exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
        if os.path.exists(exe):
            threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
            cmd = [exe, "sheet", str(args.sheet_n), str(window), repr(float(args.sheet_root)),
                   str(args.sheet_nroot), str(args.sheet_seed)]
            log("sheet cpu baseline:", " ".join(cmd), "threads", threads)
            out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
            if out.returncode == 0:
                r = json.loads(out.stdout.strip().splitlines()[-1])
                res["cpu_baseline"] = {
                    "value": r["steps_per_s"], "unit": "steps/s", "cores": r["threads"], "kind": "reference",
                    "sample": "reference OpenMP build, the same sheet, %d steps after 1 warm-up step, %.1f s "
                              "(%.0f collisions/step resolved, %.0f pairs/step found)" % (
                                  r["steps"], r["seconds"], r["collisions_per_step"],
                                  r.get("resolver_calls_per_step", float("nan")))}
                res["cpu_baseline"]["collisions_per_step"] = r["collisions_per_step"]
                res["speedup_vs_cpu"] = res["window"]["steps_per_s"] / r["steps_per_s"]
                log("sheet cpu: %.3f steps/s on %d threads; GPU on the same steps %.1f steps/s -> %.0fx" % (
                    r["steps_per_s"], r["threads"], res["window"]["steps_per_s"], res["speedup_vs_cpu"]))
            else:
                log(out.stderr)
    return res


def copy_calibration(mib, reps=10):
    """Device-to-device copies of `mib` MiB: the reference reading of the
    FETCH_SIZE / WRITE_SIZE counters for a known byte count."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_calibration_bytes": n, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--softening", type=float, default=0.005)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline mode only")
    ap.add_argument("--acc-rows", type=int, default=512, help="sampled rows of the accuracy check")
    ap.add_argument("--ens-systems", type=int, default=16384)
    ap.add_argument("--ens-steps", type=int, default=100000)
    ap.add_argument("--ens-seed", type=int, default=1)
    ap.add_argument("--ens-probe", type=int, default=500, help="steps per autotuner probe")
    ap.add_argument("--ens-cpu-steps", type=int, default=500)
    ap.add_argument("--ens-full-systems", type=int, default=256,
                    help="systems the reference integrates for the whole request (energy-error check)")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=20.0, help="job arrivals per second")
    ap.add_argument("--serve-seed", type=int, default=7)
    ap.add_argument("--serve-budget", type=int, default=8 << 30, help="device-memory budget in bytes")
    ap.add_argument("--serve-workers", type=int, default=8, help="concurrent N-body jobs")
    ap.add_argument("--serve-tick", type=int, default=2000, help="planetary steps per pool tick")
    ap.add_argument("--serve-cpu-scale", type=float, default=50.0)
    ap.add_argument("--sheet-n", type=int, default=200000)
    ap.add_argument("--sheet-root", type=float, default=136.0)
    ap.add_argument("--sheet-nroot", type=int, default=10)
    ap.add_argument("--sheet-seed", type=int, default=99)
    ap.add_argument("--sheet-steps", type=int, default=500)
    ap.add_argument("--sheet-cpu-steps", type=int, default=3)
    ap.add_argument("--megno-systems", type=int, default=16384)
    ap.add_argument("--megno-steps", type=int, default=5000)
    ap.add_argument("--megno-cpu-steps", type=int, default=200)
    ap.add_argument("--only", choices=["direct", "ensemble", "serve", "sheet", "small", "megno"], default=None)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import rebound_amd as amd
    if amd.device_count() <= 0:
        raise SystemExit("no HIP device")
    log("device:", amd.device_name())

    if args.only == "ensemble":
        print(json.dumps({"ensemble": ensemble_leg(amd, args)}))
        return
    if args.only == "serve":
        print(json.dumps({"serving": serving_leg(amd, args)}))
        return
    if args.only == "sheet":
        print(json.dumps({"sheet": sheet_leg(amd, args)}))
        return
    if args.only == "small":
        print(json.dumps({"small": small_leg(amd, args)}))
        return
    if args.only == "megno":
        print(json.dumps({"megno": megno_leg(amd, args)}))
        return
    head, _ = run_mode(amd, args, "fp64", "fp64", "basic", args.steps, args.warmup)
    modes = {"fp64": head}
    if not args.quick:
        k2 = max(10, args.steps // 4)
        for name, prec, grav in (("compensated", "fp64", "compensated"), ("split", "split", "basic"),
                                 ("exact", "exact", "basic")):
            modes[name], _ = run_mode(amd, args, name, prec, grav, k2, 2)
        acc = accuracy(amd, args)
        for k, v in acc.items():
            modes[k]["accuracy"] = v
    cpu = None if args.no_cpu else cpu_baseline(args)
    ens = None if (args.quick or args.only == "direct") else ensemble_leg(amd, args)
    serving = None if (args.quick or args.only == "direct") else serving_leg(amd, args)
    sheet = None if (args.quick or args.only == "direct") else sheet_leg(amd, args)
    small = None if (args.quick or args.only == "direct") else small_leg(amd, args)
    megno = None if (args.quick or args.only == "direct") else megno_leg(amd, args)

    achieved = head["algorithmic_tflops"]
    tag = "direct_fp64_plummer65536"
    traffic, traffic_src = profiled("traffic_bytes_per_launch", head["kernel"], tag) if args.n == 65536 else (None, None)
    hw_flops, _ = profiled("hw_flops_per_launch", head["kernel"], tag) if args.n == 65536 else (None, None)
    line = {
        "metric": "Gpair-interactions/s at N=65536 (direct gravity, FP64)",
        "value": head["gpairs_per_s"],
        "unit": "Gpair/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Plummer sphere (reb_simulation_add_plummer, seed %d)" % args.seed,
        "config": {"workload": "plummer N=%d softening=%g leapfrog dt=%g, REB_GRAVITY_BASIC FP64" % (
            args.n, args.softening, args.dt), "N": args.n, "softening": args.softening, "dt": args.dt,
            "integrator": "leapfrog", "gravity": "basic", "precision": "fp64"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                     "traffic_source": traffic_src,
                     "hw_tflops": (hw_flops / (1e-3 * head["kernel_ms"]) / 1e12) if hw_flops else None,
                     "hw_tflops_note": "FP64 VALU FLOPs per launch counted by rocprofv3 (SQ_INSTS_VALU_FLOPS_FP64, "
                                       "FMA = 2) over this run's kernel time",
                     "kernel": head["kernel"], "kernel_ms": head["kernel_ms"],
                     "work_unit": head["work_unit"], "flop_per_unit": head["flop_per_unit"],
                     "units_per_launch": head["units_per_launch"]},
        "cpu_baseline": cpu,
        "modes": modes,
        "ensemble": ens,
        "serving": serving,
        "sheet": sheet,
        "small": small,
        "megno": megno,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
