"""Throughput of the NVE Lennard-Jones step loop on one MI355X.

Default workload ("liquid"): N=256,000 particles (FCC 40^3) at density 0.8442,
T=0.72, cutoff 2.5, dt=0.005, seeded initial condition (seed 5489); K
velocity-Verlet steps with per-step thermodynamic records, in mixed precision
(FP32 pair forces, FP64 positions / velocities / sums).  The same run in plain
FP64 is timed beside it on the same GPU, and the CPU restatement of the
reference algorithm is timed on a bounded sample on the host.

Other workloads (one JSON line each, for the records in profiles/):
  --workload plain        N=864 (FCC 6^3), rho 0.8, T 1.0, FP64, 1000 steps
  --workload sweep        256 replicas on a 16x16 (T, rho) grid, N = 4 n^3 with
                          n in [5, 20] from the seed; batched in one engine
                          (--no-fp64 / --no-plain drop the FP64 batched leg /
                          the plain per-replica FP64 leg, e.g. under a profiler)
  --workload coexistence  N=2,048,000 at rho 0.30, T 0.80, rc 3.0, dt 0.004
  --workload adjoint      N=4,000, 2000 FP64 steps, adjoint gradients

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps 10000] [--warmup 100] [--workload liquid]
"""

import argparse
import itertools
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


def pair_counts(eng, pos, L):
    """(C_half, P_half) of the reference's half-shell traversal on these
    positions (engine/particle_pair_filter.cpp:141-191): candidate pairs,
    i.e. pairs inside a cell plus the pairs of each cell with its 13
    half-shell neighbour cells, and in-cutoff pairs.  Both come from an FP64
    engine, whose cell lists and cutoff decisions are the reference's."""
    n = len(pos)
    with eng.Engine([eng.SystemSpec(n, L)], precision="fp64") as e:
        e.set_state(0, pos, np.zeros_like(pos))
        start, _ = e.cell_list(0)
        _, nc, _ = e.cell_keys(0)
        p_half = int(e.neighbor_counts(0).sum()) // 2
    m = np.diff(start).astype(np.int64).reshape(tuple(int(x) for x in nc))
    return half_shell_candidates(m), p_half


def half_shell_candidates(m):
    """Candidate pairs of the half-shell traversal over a periodic grid of
    per-cell counts m (x, y, z): pairs inside each cell plus, for each of 13
    neighbour offsets (one of every +/- pair), the products of neighbouring
    counts."""
    assert min(m.shape) >= 3, "the 13 half-shell neighbours are distinct cells from 3 cells a side"
    c = int((m * (m - 1) // 2).sum())
    for d in itertools.product((-1, 0, 1), repeat=3):
        if d > (0, 0, 0):
            c += int((m * np.roll(m, tuple(-x for x in d), axis=(0, 1, 2))).sum())
    return c


def drift_stats(E, window=100, scale=None):
    """(secular drift |<E>_end - <E>_start| / s over `window`-step windows,
    max |E_t - E_0| / s), with s = |E_0| unless a scale is given (a replica
    whose kinetic and potential energies nearly cancel has E_0 ~ 0)."""
    w = max(1, min(window, len(E) // 4))
    s = abs(E[0]) if scale is None else scale
    return (float(abs(E[-w:].mean() - E[:w].mean()) / s), float(np.abs(E - E[0]).max() / s))


def kernel_times(e, steps=40):
    """Mean duration of each step kernel (HIP events on the engine's stream)."""
    e.set_profiling(True)
    e.step(steps)
    kt = e.kernel_times()
    e.set_profiling(False)
    return kt


def read_profile_value(name, key):
    """A 'key: <number>' line of a committed rocprofv3 summary, or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def read_traffic(name):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    --pmc summary (a line 'traffic_bytes_per_launch: <bytes>'), or None."""
    return read_profile_value(name, "traffic_bytes_per_launch")


def cpu_baseline(n, rho, T, seed, steps, rc=2.5, dt=0.005):
    """The oracle's timing twin (reference flags -O3 -ffast-math), one core,
    on the same workload for a bounded number of steps."""
    from oracle import ljref

    st, L = ljref.initial_condition(n, rho, T, seed=seed)
    ljref.run(st, L, dt, 2, rc=rc, fast=True, records=False)       # warm-up
    t0 = time.perf_counter()
    ljref.run(st, L, dt, steps, rc=rc, fast=True)
    dt_ = time.perf_counter() - t0
    return n * steps / dt_, dt_


def timed_run(e, pos, vel, steps, warmup, sys_states=None):
    if sys_states is None:
        sys_states = [(pos, vel)]
    for k, (p, v) in enumerate(sys_states):
        e.set_state(k, p, v)
    if warmup:
        e.step(warmup)
    for k, (p, v) in enumerate(sys_states):
        e.set_state(k, p, v)        # timed run starts from the initial condition
    t0 = time.perf_counter()
    rec = e.step(steps)
    return rec, time.perf_counter() - t0


# ---------------------------------------------------------------------------
def liquid(args, eng):
    n, rho, T, seed = 256000, 0.8442, 0.72, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    log(f"workload N={n} L={L:.6f} steps={args.steps} warmup={args.warmup}")
    spec = [eng.SystemSpec(n, L)]
    em = eng.Engine(spec, precision="mixed", pair_kernel=args.pair_kernel)
    recm, tm = timed_run(em, pos, vel, args.steps, args.warmup)
    rate_m = n * args.steps / tm
    drift_m, fluct_m = drift_stats(recm[:, 0, 3])
    log(f"mixed: {tm:.3f} s, {rate_m:.4e} atom-steps/s, drift {drift_m:.3e}, max dev {fluct_m:.3e}")
    res64, frms = None, None
    if not args.no_fp64:
        e64 = eng.Engine(spec, precision="fp64", pair_kernel=args.pair_kernel)
        rec64, t64 = timed_run(e64, pos, vel, args.steps, args.warmup)
        rate_64 = n * args.steps / t64
        drift_64, fluct_64 = drift_stats(rec64[:, 0, 3])
        log(f"fp64: {t64:.3f} s, {rate_64:.4e} atom-steps/s, drift {drift_64:.3e}")
        st64 = e64.get_state(0)
        em.set_state(0, st64["pos"], st64["vel"])
        e64.set_state(0, st64["pos"], st64["vel"])
        em.compute_forces()
        e64.compute_forces()
        Fm, F64 = em.get_state(0)["force"], e64.get_state(0)["force"]
        frms = float(np.sqrt(((Fm - F64) ** 2).sum(1).mean() / (F64 ** 2).sum(1).mean()))
        res64 = dict(value=rate_64, ms_per_step=1e3 * t64 / args.steps, energy_drift=drift_64,
                     energy_max_rel_dev=fluct_64, kernel_ms=kernel_times(e64))
        e64.close()
        log(f"force rms rel err mixed vs fp64: {frms:.3e}")
    ktm = kernel_times(em)
    log("mixed kernel ms:", json.dumps(ktm))
    st = em.get_state(0)
    em.close()
    c_half, p_half = pair_counts(eng, st["pos"], L)
    flops = pair_flops(c_half, p_half)
    achieved = flops / (ktm["pair"] * 1e-3) / 1e12
    traffic = read_traffic("pair_mixed_256k_traffic.txt")
    # VALU issue: wave-instructions per launch (committed SQ counters) over the
    # CUs' issue capacity in the measured duration (one wave-instruction per CU
    # per clock at 2.4 GHz, 256 CUs)
    valu = read_profile_value("step_kernels_256k_sq_counters.md", "valu_insts_per_launch")
    issue = valu / (ktm["pair"] * 1e-3 * 256 * 2.4e9) if valu else None
    cpu = None
    if not args.no_cpu:
        rate_cpu, dt_cpu = cpu_baseline(n, rho, T, seed, args.cpu_steps)
        cpu = {"value": rate_cpu, "unit": "atom-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} FP64 steps of the same N=256,000 workload "
                         f"({dt_cpu:.1f} s) with the oracle's -O3 -ffast-math build "
                         f"(serial cell list, half shell, Newton 3, as the reference)"}
        log(f"cpu: {rate_cpu:.4e} atom-steps/s over {dt_cpu:.1f} s")
    return {
        "metric": METRIC, "value": rate_m, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * tm / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces + f64 state/sums",
        "data": "synthetic (seeded FCC lattice + Maxwell-Boltzmann velocities, seed 5489)",
        "config": {"workload": "NVE LJ liquid N=256,000 (FCC 40^3), rho=0.8442, T=0.72, "
                               "rc=2.5, dt=0.005, cell list rebuilt every step, mixed precision",
                   "particles": n, "steps": args.steps, "precision": "mixed",
                   "pair_kernel": args.pair_kernel, "parallelism": "single GPU"},
        "energy_drift": drift_m, "energy_max_rel_dev": fluct_m,
        "fp64": res64,
        "mixed_over_fp64": (rate_m / res64["value"]) if res64 else None,
        "force_rms_rel_err_vs_fp64": frms,
        # compute-bound: priced against MI355X's published FP32 compute roof
        # (157.3 TFLOP/s: the matrix cores, or the vector ALUs with packed FP32,
        # which this build leaves off); the kernel runs on the vector ALUs, whose
        # issue fraction is reported beside it
        "roofline": {"bound": "mfma", "executes_on": "VALU (FP32 vector, no MFMA)",
                     "achieved": achieved, "peak": PEAK["fp32"],
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"], "traffic": traffic,
                     "kernel": ("k_pair_cells_mixed<LJ,KICK>" if args.pair_kernel == "tiled"
                                else "k_pair<MIXED,LJ,KICK>"),
                     "kernel_ms": ktm["pair"], "flops_per_launch": flops,
                     "c_half": c_half, "p_half": p_half,
                     "valu_insts_per_launch": valu, "valu_issue_frac": issue},
        "kernel_ms": ktm,
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
def sweep(args, eng):
    """256 replicas, 16 x 16 grid of T in [0.5, 2] and rho in [0.05, 1.2]
    (the reference's linspace arithmetic, temperature-major), N = 4 n^3 with
    n from std::mt19937(seed) in [5, 20]; all replicas in one engine."""
    from lennardjonesium.tools import linspace
    steps = args.steps
    seed = 5489
    grid = list(itertools.product(linspace(0.5, 2.0, 16), linspace(0.05, 1.2, 16)))
    sizes = eng.sweep_sizes(seed, len(grid))
    specs, states, meta = [], [], []
    for k, ((T, rho), n) in enumerate(zip(grid, sizes)):
        N = 4 * int(n) ** 3
        pos, vel, L = eng.initial_condition(N, rho, T, seed=seed + k)
        specs.append(eng.SystemSpec(N, L))
        states.append((pos, vel))
        meta.append((T, rho, N, L))
    total = sum(m[2] for m in meta)
    log(f"sweep: {len(specs)} replicas, {total} particles, steps={steps}")
    out = {}
    for prec in (("mixed",) if args.no_fp64 else ("mixed", "fp64")):
        e = eng.Engine(specs, precision=prec)
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
        variant, ms = e.autotune(10)
        kinds = e.pair_kernels()
        n_tiled = sum(1 for x in kinds if x == "tiled")
        log(f"{prec}: autotuner chose {variant} ({ms:.3f} ms/step; "
            f"{n_tiled} replicas tiled, {len(kinds) - n_tiled} per-particle)")
        rec, t = timed_run(e, None, None, steps, args.warmup, states)
        rate = total * steps / t
        # drift relative to max(|E_0|, KE_0 + |PE_0|)
        drifts = [drift_stats(rec[:, k, 3], scale=max(abs(rec[0, k, 3]),
                                                      rec[0, k, 1] + abs(rec[0, k, 2])))[0]
                  for k in range(len(specs))]
        worst = int(np.argmax(drifts))
        out[prec] = dict(value=rate, seconds=t, pair_kernel=variant,
                         replicas_tiled=n_tiled, replicas_per_particle=len(kinds) - n_tiled,
                         max_replica_drift=float(max(drifts)),
                         median_replica_drift=float(np.median(drifts)),
                         worst_replica={"T": meta[worst][0], "rho": meta[worst][1],
                                        "N": meta[worst][2], "E0": float(rec[0, worst, 3]),
                                        "E_end": float(rec[-1, worst, 3])})
        log(f"{prec}: {t:.2f} s, {rate:.4e} atom-steps/s aggregate, "
            f"max replica drift {max(drifts):.2e}")
        if prec == "mixed":
            # observables on a short sampled continuation: g(r) every 10 steps
            for _ in range(10):
                e.step(10, records=False)
                e.rdf_sample(4.0, 200)
            counts, ns = e.rdf_read(200)
            Ns = np.array([m[2] for m in meta])
            pe = rec[:, :, 2].mean(0) / Ns
            P = np.array([m[1] for m in meta]) * (rec[:, :, 5].mean(0) +
                                                  rec[:, :, 4].mean(0) / (3.0 * Ns))
            out["observables_sample"] = {"pe_per_particle_first4": pe[:4].tolist(),
                                         "pressure_first4": P[:4].tolist(),
                                         "rdf_samples": ns}
        e.close()
    # plain per-replica FP64 runs (one engine per replica, as the reference runs one
    # Simulation per thread), timed on a sample of replicas
    sample = list(range(0, len(specs), 32)) if not args.no_plain else []
    t_plain, work = 0.0, 0
    sst = min(steps, 500)
    for k in sample:
        e = eng.Engine([specs[k]], precision="fp64")
        _, t = timed_run(e, states[k][0], states[k][1], sst, 10)
        t_plain += t
        work += meta[k][2] * sst
        e.close()
    plain_rate = work / t_plain if sample else None
    if sample:
        log(f"plain per-replica fp64: {plain_rate:.4e} atom-steps/s on {len(sample)} sampled replicas")
    return {"metric": METRIC, "value": out["mixed"]["value"], "unit": "atom-steps/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f32 pair forces + f64 state/sums", "data": "synthetic",
            "config": {"workload": "batched sweep: 256 replicas, 16x16 grid T 0.5-2.0 x rho "
                                   "0.05-1.2, N=4n^3, n in [5,20] from seed 5489, rc 2.5, "
                                   "dt 0.005", "particles": total, "replicas": len(specs)},
            "mixed_batched": out["mixed"], "fp64_batched": out.get("fp64"),
            "plain_fp64_per_replica": {"value": plain_rate, "sampled_replicas": sample,
                                       "steps": sst},
            "batched_mixed_over_plain_fp64": (out["mixed"]["value"] / plain_rate
                                              if plain_rate else None),
            "observables": out.get("observables_sample")}


# ---------------------------------------------------------------------------
def coexistence(args, eng):
    n, rho, T, seed, rc, dt = 2048000, 0.30, 0.80, 5489, 3.0, 0.004
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    log(f"coexistence: N={n} L={L:.4f} steps={steps}")
    e = eng.Engine([eng.SystemSpec(n, L, cutoff=rc, time_delta=dt)], precision="mixed")
    e.set_state(0, pos, vel)
    # tune on a state the run spends its time in: the quenched lattice breaks up
    # into liquid and vapour within the first steps, and the fastest kernel on
    # the lattice is not the fastest afterwards (the timed run still starts
    # from the lattice)
    e.step(1000, records=False)
    variant, ms = e.autotune(20)
    rec, t = timed_run(e, pos, vel, steps, args.warmup)
    rate = n * steps / t
    drift, fluct = drift_stats(rec[:, 0, 3])
    start, _ = e.cell_list(0)
    occ = np.diff(start)
    ktm = kernel_times(e, 20)
    e.close()
    bytes_step = 224.0 * n
    log(f"coexistence: {rate:.4e} atom-steps/s, drift {drift:.2e}, occupancy max/mean "
        f"{occ.max() / occ.mean():.2f}")
    return {"metric": METRIC, "value": rate, "unit": "atom-steps/s", "n_gpus": 1, "steps": steps,
            "dtype": "f32 pair forces + f64 state/sums", "data": "synthetic",
            "config": {"workload": "coexistence N=2,048,000 (FCC 80^3) rho 0.30 T 0.80 rc 3.0 "
                                   "dt 0.004, cell list rebuilt every step", "pair_kernel": variant},
            "energy_drift": drift, "energy_max_rel_dev": fluct,
            "load_imbalance": {"cell_occupancy_max_over_mean": float(occ.max() / occ.mean()),
                               "occupancy_histogram": np.bincount(occ).tolist()},
            "algorithmic_hbm_GBps": bytes_step * steps / t / 1e9,
            "kernel_ms": ktm}


# ---------------------------------------------------------------------------
def plain(args, eng):
    n, rho, T, seed = 864, 0.8, 1.0, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    e = eng.Engine([eng.SystemSpec(n, L)], precision="fp64")
    rec, t = timed_run(e, pos, vel, steps, args.warmup)
    e.close()
    rate_cpu, dt_cpu = cpu_baseline(n, rho, T, seed, steps)
    drift, fluct = drift_stats(rec[:, 0, 3])
    return {"metric": METRIC, "value": n * steps / t, "unit": "atom-steps/s", "n_gpus": 1,
            "steps": steps, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "plain FP64 N=864 (FCC 6^3) rho 0.8 T 1.0 rc 2.5 dt 0.005"},
            "energy_drift": drift, "energy_max_rel_dev": fluct,
            "log_every_10": rec[::10, 0, :][:5].tolist(),
            "cpu_baseline": {"value": rate_cpu, "unit": "atom-steps/s", "cores": 1,
                             "kind": "port", "sample": f"{steps} steps, {dt_cpu:.2f} s"}}


def adjoint(args, eng):
    n, rho, T, seed = 4000, 0.8, 1.0, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    e = eng.Engine([eng.SystemSpec(n, L)], precision="fp64")
    e.set_state(0, pos, vel)
    t0 = time.perf_counter()
    loss, ge, gs, gr, gv = e.adjoint(steps, rho, every=100)
    t = time.perf_counter() - t0
    h = 1e-6

    def fwd(eps=1.0, sig=1.0):
        with eng.Engine([eng.SystemSpec(n, L, epsilon=eps, sigma=sig)]) as f:
            f.set_state(0, pos, vel)
            return f.adjoint(steps, rho, every=steps)[0]

    fd_e = (fwd(eps=1 + h) - fwd(eps=1 - h)) / (2 * h)
    fd_s = (fwd(sig=1 + h) - fwd(sig=1 - h)) / (2 * h)
    # agreement with central differences against the horizon: the exact
    # derivative of a chaotic trajectory grows like exp(lambda t)
    horizon = []
    for hs in (50, 100, 200, 400, 800):
        with eng.Engine([eng.SystemSpec(n, L)]) as f:
            f.set_state(0, pos, vel)
            _, ge_h, _, _, _ = f.adjoint(hs, rho, every=50)

        def fh(eps):
            with eng.Engine([eng.SystemSpec(n, L, epsilon=eps)]) as g:
                g.set_state(0, pos, vel)
                return g.adjoint(hs, rho, every=hs)[0]

        fd_h = (fh(1 + h) - fh(1 - h)) / (2 * h)
        horizon.append({"steps": hs, "adjoint_eps": ge_h, "fd_eps": fd_h,
                        "rel_err": abs(ge_h - fd_h) / abs(fd_h)})
        log(f"horizon {hs}: adjoint {ge_h:.6e} fd {fd_h:.6e}")
    e.close()
    return {"metric": "adjoint trajectory seconds", "value": t, "unit": "s", "n_gpus": 1,
            "steps": steps, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "differentiable trajectory N=4000 rho 0.8 T 1.0, checkpoints "
                                   "every 100 steps"},
            "loss": loss, "grad_eps": ge, "grad_sigma": gs, "fd_eps": fd_e, "fd_sigma": fd_s,
            "rel_err_eps": abs(ge - fd_e) / abs(fd_e), "rel_err_sigma": abs(gs - fd_s) / abs(fd_s),
            "grad_r0_norm": float(np.linalg.norm(gr)), "horizon": horizon}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="sweep: skip the plain per-replica FP64 leg")
    ap.add_argument("--pair-kernel", default="tiled", choices=["tiled", "particle"])
    ap.add_argument("--workload", default="liquid",
                    choices=["liquid", "sweep", "coexistence", "plain", "adjoint"])
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build runs every system on a single "
              "MI355X and has no multi-GPU path (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {4 * n} bytes three times")
        return
    if args.steps is None:
        args.steps = {"liquid": 10000, "sweep": 5000, "coexistence": 20000, "plain": 1000,
                      "adjoint": 2000}[args.workload]

    from lennardjonesium import engine as eng

    if eng.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    fn = {"liquid": liquid, "sweep": sweep, "coexistence": coexistence, "plain": plain,
          "adjoint": adjoint}[args.workload]
    print(json.dumps(fn(args, eng)), flush=True)


if __name__ == "__main__":
    main()
