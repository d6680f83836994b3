"""Throughput of the NVE Lennard-Jones step loop on one MI355X.

Workload (default): N=256,000 particles (FCC 40^3) at density 0.8442, T=0.72,
cutoff 2.5, dt=0.005, seeded initial condition (seed 5489); K velocity-Verlet
steps with per-step thermodynamic records, in mixed precision (FP32 pair
forces, FP64 positions / velocities / sums).  The same run in plain FP64 is
timed beside it, both on the same GPU, and the CPU restatement of the
reference algorithm is timed on a bounded sample on the host.

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 100]
"""

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "lennard-jonesium_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "atom-steps/s, mixed/batched vs plain FP64 on same GPU at |dE/E|<=1e-4/10^4 st"
PEAK = {"fp32": 157.3, "fp64": 78.6}      # MI355X published vector peaks, TFLOP/s
HBM_PEAK = 8000.0                           # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_flops(c_half, p_half):
    """Algorithmic FLOPs of one force calculation: 8 per candidate pair
    (3 sub, 3 mul, 2 add for r^2) + 30 per in-cutoff pair (r^-6 3, V 8, W 7,
    F 4, accumulation 8), counted on the reference's half-shell pair set."""
    return 8.0 * c_half + 30.0 * p_half


def run_gpu(eng, precision, pos, vel, L, steps, warmup, n, pair_kernel="tiled"):
    e = eng.Engine([eng.SystemSpec(n, L)], precision=precision, pair_kernel=pair_kernel)
    e.set_state(0, pos, vel)
    if warmup:
        e.step(warmup)
    e.set_state(0, pos, vel)        # timed run starts from the initial condition
    t0 = time.perf_counter()
    rec = e.step(steps)
    t1 = time.perf_counter()
    return e, rec[:, 0, :], t1 - t0


def kernel_times(e, steps=40):
    e.set_profiling(True)
    e.step(steps)
    kt = e.kernel_times()
    e.set_profiling(False)
    return kt


def read_traffic(name):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    --pmc summary (a line 'traffic_bytes_per_launch: <bytes>'), or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(n, rho, T, seed, steps):
    """The oracle's timing twin (reference flags -O3 -ffast-math), one core,
    on the same workload for a bounded number of steps."""
    from oracle import ljref

    st, L = ljref.initial_condition(n, rho, T, seed=seed)
    ljref.run(st, L, 0.005, 2, fast=True, records=False)       # warm-up
    t0 = time.perf_counter()
    ljref.run(st, L, 0.005, steps, fast=True)
    dt = time.perf_counter() - t0
    return n * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--pair-kernel", default="tiled", choices=["tiled", "particle"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs every system on a single "
              "MI355X and has no multi-GPU path (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from lennardjonesium import engine as eng

    if eng.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    n, rho, T, seed = 256000, 0.8442, 0.72, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    log(f"workload N={n} L={L:.6f} steps={args.steps} warmup={args.warmup}")

    em, recm, tm = run_gpu(eng, "mixed", pos, vel, L, args.steps, args.warmup, n,
                           args.pair_kernel)
    rate_m = n * args.steps / tm
    Em = recm[:, 3]
    drift_m = float(np.abs(Em - Em[0]).max() / abs(Em[0]))
    log(f"mixed: {tm:.3f} s, {rate_m:.4e} atom-steps/s, max|dE/E0| {drift_m:.3e}")

    res64 = None
    if not args.no_fp64:
        e64, rec64, t64 = run_gpu(eng, "fp64", pos, vel, L, args.steps, args.warmup, n,
                                  args.pair_kernel)
        rate_64 = n * args.steps / t64
        E64 = rec64[:, 3]
        drift_64 = float(np.abs(E64 - E64[0]).max() / abs(E64[0]))
        log(f"fp64: {t64:.3f} s, {rate_64:.4e} atom-steps/s, max|dE/E0| {drift_64:.3e}")
        # force error of the mixed kernel against FP64 on identical positions
        st64 = e64.get_state(0)
        em.set_state(0, st64["pos"], st64["vel"])
        e64.set_state(0, st64["pos"], st64["vel"])
        em.compute_forces()
        e64.compute_forces()
        Fm, F64 = em.get_state(0)["force"], e64.get_state(0)["force"]
        frms = float(np.sqrt(((Fm - F64) ** 2).sum(1).mean() / (F64 ** 2).sum(1).mean()))
        kt64 = kernel_times(e64)
        res64 = dict(value=rate_64, ms_per_step=1e3 * t64 / args.steps, energy_drift=drift_64,
                     kernel_ms=kt64)
        e64.close()
        log(f"force rms rel err mixed vs fp64: {frms:.3e}")
    else:
        frms = None

    # per-kernel durations (HIP events on the engine stream, eager launches)
    ktm = kernel_times(em)
    log("mixed kernel ms:", json.dumps(ktm))
    st = em.get_state(0)
    em.close()
    from oracle import ljref

    c_half, p_half = ljref.pair_counts(st["pos"], L, 2.5)
    flops = pair_flops(c_half, p_half)
    t_pair = ktm["pair"] * 1e-3
    achieved = flops / t_pair / 1e12
    traffic = read_traffic("pair_mixed_256k_traffic.txt")

    cpu = None
    if not args.no_cpu:
        rate_cpu, dt_cpu = cpu_baseline(n, rho, T, seed, args.cpu_steps)
        cpu = {"value": rate_cpu, "unit": "atom-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} FP64 steps of the same N=256,000 workload "
                         f"({dt_cpu:.1f} s) with the oracle's -O3 -ffast-math build, "
                         f"serial cell list + half shell as the reference"}
        log(f"cpu: {rate_cpu:.4e} atom-steps/s over {dt_cpu:.1f} s")

    out = {
        "metric": METRIC,
        "value": rate_m,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tm / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces + f64 state/sums",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities, seed 5489)",
        "config": {"workload": "NVE LJ liquid N=256,000 (FCC 40^3), rho=0.8442, T=0.72, "
                               "rc=2.5, dt=0.005, cell list rebuilt every step, mixed precision",
                   "particles": n, "steps": args.steps, "precision": "mixed",
                   "parallelism": "single GPU"},
        "energy_drift": drift_m,
        "fp64": res64,
        "mixed_over_fp64": (rate_m / res64["value"]) if res64 else None,
        "force_rms_rel_err_vs_fp64": frms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "traffic": traffic,
                     "kernel": ("k_pair_cells<MIXED,LJ,KICK>" if args.pair_kernel == "tiled"
                                else "k_pair<MIXED,LJ,KICK>"),
                     "kernel_ms": ktm["pair"],
                     "flops_per_launch": flops, "c_half": c_half, "p_half": p_half},
        "kernel_ms": ktm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
