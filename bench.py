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
FLOP_PER_PAIR = {"fp64": 20, "compensated": 28, "split": 28, "exact": 20}


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
    L.reb_amd_set_profiling(sim.ptr, 0)
    L.reb_amd_set_resident(sim.ptr, 0)
    e1 = sim.energy()
    sim.check()
    n = args.n
    pairs = float(n) * float(n - 1)
    wall = t1 - t0
    kernel_ms = ktot.value / max(kn.value, 1)
    res = {
        "precision": precision, "gravity": gravity,
        "steps": steps, "seconds": wall, "ms_per_step": 1e3 * wall / steps,
        "gpairs_per_s": pairs * steps / wall / 1e9,
        "kernel_ms": kernel_ms, "kernel_launches": kn.value,
        "kernel_gpairs_per_s": pairs / (kernel_ms * 1e-3) / 1e9,
        "algorithmic_tflops": FLOP_PER_PAIR[name] * pairs / (kernel_ms * 1e-3) / 1e12,
        "rel_energy_drift": (e1 - e0) / abs(e0),
    }
    log("%-12s %8.2f ms/step  %8.1f Gpair/s (kernel %.3f ms, %.1f TFLOP/s algorithmic)" % (
        name, res["ms_per_step"], res["gpairs_per_s"], kernel_ms, res["algorithmic_tflops"]))
    return res, sim


def accuracy(amd, args):
    """Per-particle error of each mode's initial accelerations against the
    bit-exact FP64 path (identical to the reference's BASIC sum)."""
    def acc_of(prec, gravity="basic"):
        sim = plummer_sim(amd, args.n, args.seed, args.softening, args.dt, prec, gravity)
        a = sim.update_acceleration()
        st = sim.state()
        sim.free()
        return a, st

    ref, st = acc_of("exact")
    x = st[:, :3]
    m = st[:, 9]
    rows = np.random.default_rng(0).choice(args.n, min(args.n, 2048), replace=False)
    d = x[rows, None, :] - x[None, :, :]
    dist2 = (d * d).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = m[None, :] * np.sqrt(dist2) / (dist2 + args.softening ** 2) ** 1.5
    w[~np.isfinite(w)] = 0
    cond = w.sum(1)
    out = {}
    for name, prec, grav in (("fp64", "fp64", "basic"), ("compensated", "fp64", "compensated"),
                             ("split", "split", "basic")):
        a, _ = acc_of(prec, grav)
        da = np.sqrt(((a[rows] - ref[rows]) ** 2).sum(1))
        rel = da / np.sqrt((ref[rows] ** 2).sum(1))
        nrm = da / cond
        out[name] = {"rel_max": float(rel.max()), "rel_p99": float(np.quantile(rel, 0.99)),
                     "rel_median": float(np.median(rel)), "normalised_max": float(nrm.max()),
                     "sampled_rows": int(len(rows))}
        log("accuracy %-12s rel max %.2e p99 %.2e median %.2e | |da|/sum|a_ij| max %.2e" % (
            name, rel.max(), np.quantile(rel, 0.99), np.median(rel), nrm.max()))
    return out


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
            "sample": "reference OpenMP build (gravity.c:215-232 per-target loop), Plummer N=%d, %d leapfrog "
                      "steps after 1 warm-up step, %.1f s" % (args.n, steps, res["seconds"])}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import rebound_amd as amd
    if amd.device_count() <= 0:
        raise SystemExit("no HIP device")
    log("device:", amd.device_name())

    head, _ = run_mode(amd, args, "fp64", "fp64", "basic", args.steps, args.warmup)
    modes = {"fp64": head}
    if not args.quick:
        k2 = max(10, args.steps // 4)
        for name, prec, grav in (("compensated", "fp64", "compensated"), ("split", "split", "basic"),
                                 ("exact", "exact", "basic")):
            modes[name], _ = run_mode(amd, args, name, prec, grav, k2, 2)
        acc = accuracy(amd, args)
        for k, v in acc.items():
            modes[k]["accuracy_vs_exact"] = v
    cpu = None if args.no_cpu else cpu_baseline(args)

    kernel_s = head["kernel_ms"] * 1e-3
    achieved = head["algorithmic_tflops"]
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
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_direct_fast<FM_FP64>", "kernel_ms": head["kernel_ms"],
                     "flop_per_pair": FLOP_PER_PAIR["fp64"], "pairs_per_launch": float(args.n) * (args.n - 1)},
        "cpu_baseline": cpu,
        "modes": modes,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
