"""Throughput of the NVE Lennard-Jones step loop on one MI355X.

Default workload ("liquid"): N=256,000 particles (FCC 40^3) at density 0.8442,
T=0.72, cutoff 2.5, dt=0.005, seeded initial condition (seed 5489), in mixed
precision (FP32 pair forces, FP64 positions / velocities / sums).  The input
is prepared by 1,000 untimed steps that melt the lattice into the liquid
(--prepare), then W warm-up steps, then K timed steps that continue from them,
each with its thermodynamic record.  The same in plain FP64 is timed beside
it on the same GPU.  Independently of K, an energy-gate leg runs 1,000 + 10^4
steps in each precision and reports |dE/E| over the 10^4 steps (the accuracy
condition the throughput is quoted at).  The CPU restatement of the reference
algorithm is timed on a bounded sample on the host.

Other workloads (one JSON line each; their records are kept in profiles/):
  --workload plain        N=864 (FCC 6^3), rho 0.8, T 1.0, FP64, 1000 steps
  --workload sweep        256 replicas on a 16x16 (T, rho) grid, N = 4 n^3 with
                          n in [5, 20] from the seed; batched in one engine
                          (mixed and FP64), against one FP64 engine per replica;
                          g(r) sampled every 100 steps in every leg
                          (--no-fp64 / --no-plain drop the FP64 batched leg /
                          the per-replica leg, e.g. under a profiler)
  --workload coexistence  N=2,048,000 at rho 0.30, T 0.80, rc 3.0, dt 0.004
  --workload adjoint      N=4,000, FP64 steps, adjoint gradients

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload liquid]
"""

import argparse
import itertools
import json
import os
import pathlib
import platform
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "lennard-jonesium_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "atom-steps/s, mixed/batched vs plain FP64 on same GPU at |dE/E|<=1e-4/10^4 st"
PEAK = {"fp32": 157.3, "fp32_unpacked": 78.6, "fp64": 78.6}   # MI355X vector peaks, TFLOP/s
HBM_PEAK = 8000.0                           # GB/s
GATE = 1e-4                                 # |dE/E| bound over 10^4 steps
GATE_STEPS, GATE_EQUIL = 10000, 1000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and logical CPU count of the measuring host."""
    model = platform.processor() or None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def pair_flops(c_half, p_half):
    """Algorithmic FLOPs of one force calculation: 8 per candidate pair
    (3 sub, 3 mul, 2 add for r^2) + 30 per in-cutoff pair (r^-6 3, V 8, W 7,
    F 4, accumulation 8), counted on the reference's half-shell pair set."""
    return 8.0 * c_half + 30.0 * p_half


def pair_counts(eng, pos, L, rc=2.5):
    """(C_half, P_half) of the reference's half-shell traversal on these
    positions (engine/particle_pair_filter.cpp:141-191): candidate pairs,
    i.e. pairs inside a cell plus the pairs of each cell with its 13
    half-shell neighbour cells, and in-cutoff pairs.  Both come from an FP64
    engine, whose cell lists and cutoff decisions are the reference's."""
    n = len(pos)
    with eng.Engine([eng.SystemSpec(n, L, cutoff=rc)], precision="fp64") as e:
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


def energy_stats(E, scale=None, window=100):
    """Energy conservation of the series E (one value per step):
    endpoint |E_end - E_0| / s, maximum |E_t - E_0| / s, and the windowed
    drift |<E>_last w - <E>_first w| / s with w = min(window, len(E) // 4);
    s = |E_0| unless a scale is given (a system whose kinetic and potential
    energies nearly cancel has E_0 ~ 0)."""
    E = np.asarray(E, dtype=np.float64)
    s = abs(E[0]) if scale is None else scale
    w = max(1, min(window, len(E) // 4))
    return {"steps": int(len(E)),
            "endpoint": float(abs(E[-1] - E[0]) / s),
            "max_dev": float(np.abs(E - E[0]).max() / s),
            "windowed": float(abs(E[-w:].mean() - E[:w].mean()) / s),
            "window": int(w)}


def energy_gate(eng, specs, states, precision, pair_kernel="tiled"):
    """The accuracy condition of the metric, measured on its own run:
    GATE_EQUIL steps from the initial condition (the lattice melts), then
    GATE_STEPS steps.  Reported from the lattice and after the equilibration;
    passes when max_t |E_t - E_s| / |E_s| <= 1e-4 over the GATE_STEPS steps
    after equilibration for every system (E_s: the energy at their start;
    the reference's own conservation figure is also taken on its production
    window, after equilibration).  Also times those steps."""
    e = eng.Engine(specs, precision=precision, pair_kernel=pair_kernel)
    for k, (p, v) in enumerate(states):
        e.set_state(k, p, v)
    e.reserve(GATE_STEPS)
    eq = e.step(GATE_EQUIL)
    t0 = time.perf_counter()
    rec = e.step(GATE_STEPS)
    t = time.perf_counter() - t0
    e.close()
    n_total = sum(s.particle_count for s in specs)
    out = {"definition": "max_t |E_t - E_s| / |E_s| over 10^4 steps after 1,000 "
                         "equilibration steps from the lattice (E_s: energy at step 1,000); "
                         "also: endpoint, windowed (|<E> last 100 - <E> first 100|), and "
                         "the same from the lattice start (steps 1..10^4)",
           "bound": GATE, "rate_10000_steps": n_total * GATE_STEPS / t,
           "ms_per_step_10000_steps": 1e3 * t / GATE_STEPS}
    systems = []
    for k in range(len(specs)):
        series = np.concatenate([eq[:, k, 3], rec[:, k, 3]])
        ke0, pe0 = eq[0, k, 1], eq[0, k, 2]
        alt = ke0 + abs(pe0)
        systems.append({"after_equilibration": energy_stats(rec[:, k, 3]),
                        "from_lattice": energy_stats(series[:GATE_STEPS]),
                        "after_equilibration_vs_ke0_plus_abs_pe0":
                            energy_stats(rec[:, k, 3], scale=alt)})
    out["pass"] = bool(all(s["after_equilibration"]["max_dev"] <= GATE for s in systems))
    if len(systems) == 1:
        out.update(systems[0])
    else:
        out["systems"] = systems
    return out


def step_kernel_ms(kt):
    """Kernel time per step (ms): the kernels launched once per step (a
    resident call's kick-drift and the final record sums run once per call
    and are left out)."""
    return kt["per_step_ms"]


def kernel_times(e, steps=40):
    """Mean duration of each kernel (HIP events on the engine's stream) over
    one call of `steps` steps, its launches, and the per-step kernels' time
    per step."""
    e.set_profiling(True)
    e.step(steps)
    kt = e.kernel_times()
    nl = e.kernel_launches()
    e.set_profiling(False)
    per_step = {k for k, n in nl.items() if n >= steps - 2}
    return {"mean_ms": kt, "launches": nl, "steps": steps,
            "per_step_ms": sum(kt[k] * nl[k] for k in per_step) / steps,
            "per_call_ms": {k: kt[k] * nl[k] for k in kt if k not in per_step}}


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


def cpu_record(rate, seconds, what):
    model, ncpu = host_cpu()
    return {"value": rate, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": what + f" ({seconds:.1f} s) with the oracle's -O3 -ffast-math build "
                             "(serial cell list, half shell, Newton 3, as the reference), one thread",
            "cpu_model": model, "host_cpu_count": ncpu}


def timed_run(e, steps, warmup, sys_states, prepare=0):
    """Loads the states, advances them `prepare` steps (input preparation),
    takes `warmup` steps, then times `steps` more steps that continue from
    them (records included); buffers and step graphs are prepared before the
    timed region."""
    for k, (p, v) in enumerate(sys_states):
        e.set_state(k, p, v)
    e.reserve(max(steps, warmup, prepare))
    if prepare:
        e.step(prepare, records=False)
    if warmup:
        e.step(warmup)
    t0 = time.perf_counter()
    rec = e.step(steps)
    return rec, time.perf_counter() - t0


def step_roofline(flops, bytes_step, ms_step, peak_tflops):
    """T_roof / T_measured for one step: F / P_VALU + B / BW against the
    measured time per step."""
    t_roof = flops / (peak_tflops * 1e12) + bytes_step / (HBM_PEAK * 1e9)
    return {"t_roof_us": 1e6 * t_roof, "t_meas_us": 1e3 * ms_step,
            "frac": t_roof / (ms_step * 1e-3), "flops": flops, "bytes": bytes_step,
            "peak_tflops": peak_tflops, "hbm_gbps": HBM_PEAK}


# ---------------------------------------------------------------------------
def liquid(args, eng):
    n, rho, T, seed = 256000, 0.8442, 0.72, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    log(f"workload N={n} L={L:.6f} steps={args.steps} warmup={args.warmup}")
    spec = [eng.SystemSpec(n, L)]
    em = eng.Engine(spec, precision="mixed", pair_kernel=args.pair_kernel)
    recm, tm = timed_run(em, args.steps, args.warmup, [(pos, vel)], prepare=args.prepare)
    rate_m = n * args.steps / tm
    log(f"mixed: {tm:.4f} s, {rate_m:.4e} atom-steps/s")
    ktm = kernel_times(em)
    log("mixed kernel ms:", json.dumps(ktm))
    resident = em.resident_stats()[0] > 0
    st = em.get_state(0)
    res64, frms, kt64 = None, None, None
    if not args.no_fp64:
        e64 = eng.Engine(spec, precision="fp64", pair_kernel=args.pair_kernel)
        rec64, t64 = timed_run(e64, args.steps, args.warmup, [(pos, vel)], prepare=args.prepare)
        rate_64 = n * args.steps / t64
        log(f"fp64: {t64:.4f} s, {rate_64:.4e} atom-steps/s")
        kt64 = kernel_times(e64)
        # force error of the mixed kernel on the FP64 run's positions
        st64 = e64.get_state(0)
        em.set_state(0, st64["pos"], st64["vel"])
        e64.set_state(0, st64["pos"], st64["vel"])
        em.compute_forces()
        e64.compute_forces()
        Fm, F64 = em.get_state(0)["force"], e64.get_state(0)["force"]
        frms = float(np.sqrt(((Fm - F64) ** 2).sum(1).mean() / (F64 ** 2).sum(1).mean()))
        e64.close()
        res64 = dict(value=rate_64, ms_per_step=1e3 * t64 / args.steps, kernel_ms=kt64)
        log(f"force rms rel err mixed vs fp64: {frms:.3e}")
    em.close()
    gates = {}
    if not args.no_gate:
        for prec in (("mixed",) if args.no_fp64 else ("mixed", "fp64")):
            gates[prec] = energy_gate(eng, spec, [(pos, vel)], prec, args.pair_kernel)
            g = gates[prec]
            log(f"gate {prec}: max dev after equilibration "
                f"{g['after_equilibration']['max_dev']:.3e}, from lattice "
                f"{g['from_lattice']['max_dev']:.3e}, pass={g['pass']}, "
                f"10^4-step rate {g['rate_10000_steps']:.4e}")
    c_half, p_half = pair_counts(eng, st["pos"], L)
    flops = pair_flops(c_half, p_half)
    pair_ms = ktm["mean_ms"]["pair"]
    achieved = flops / (pair_ms * 1e-3) / 1e12
    traffic = read_traffic("pair_mixed_256k_traffic.txt")
    # VALU issue: wave-instructions per launch (committed SQ counters) over the
    # CUs' issue capacity in the measured duration (one wave-instruction per CU
    # per clock at 2.4 GHz, 256 CUs)
    valu = read_profile_value("step_kernels_256k_sq_counters.md", "valu_insts_per_launch")
    issue = valu / (pair_ms * 1e-3 * 256 * 2.4e9) if valu else None
    ms_m = 1e3 * tm / args.steps
    roof = {"bound": "valu", "executes_on": "VALU (FP32 vector; no MFMA: pair forces are "
                                           "not a dense contraction)",
            "achieved": achieved, "peak": PEAK["fp32"], "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp32"], "traffic": traffic,
            "peak_unpacked_fp32": PEAK["fp32_unpacked"],
            "frac_of_unpacked": achieved / PEAK["fp32_unpacked"],
            "kernel": (("k_pair_cells_mixed<LJ,2,4,1,RES,UNIT> (pair forces, kicks, drift of "
                        "resident cells)" if resident else "k_pair_cells_mixed<LJ,KICK>")
                       if args.pair_kernel == "tiled" else "k_pair<MIXED,LJ,KICK>"),
            "kernel_ms": pair_ms, "flops_per_launch": flops,
            "c_half": c_half, "p_half": p_half,
            "valu_insts_per_launch": valu, "valu_issue_frac": issue,
            "step": step_roofline(flops, 224.0 * n, ms_m, PEAK["fp32"])}
    if res64:
        a64 = flops / (kt64["mean_ms"]["pair"] * 1e-3) / 1e12
        res64["roofline"] = {"bound": "valu",
                             "kernel": ("k_pair_cells_fp64<LJ,2,1,RES,UNIT>" if resident
                                        else "k_pair_cells_fp64<LJ,KICK>"),
                             "achieved": a64, "peak": PEAK["fp64"], "unit": "TFLOP/s",
                             "frac": a64 / PEAK["fp64"], "kernel_ms": kt64["mean_ms"]["pair"],
                             "step": step_roofline(flops, 200.0 * n, res64["ms_per_step"],
                                                   PEAK["fp64"])}
        res64["launch_gaps_us_per_step"] = 1e3 * (gates["fp64"]["ms_per_step_10000_steps"] -
                                                  step_kernel_ms(kt64)) if "fp64" in gates else None
    cpu = None
    if not args.no_cpu:
        rate_cpu, dt_cpu = cpu_baseline(n, rho, T, seed, args.cpu_steps)
        cpu = cpu_record(rate_cpu, dt_cpu, f"{args.cpu_steps} FP64 steps of the same "
                                           f"N=256,000 workload")
        log(f"cpu: {rate_cpu:.4e} atom-steps/s over {dt_cpu:.1f} s")
    gate_m = gates.get("mixed")
    return {
        "metric": METRIC, "value": rate_m, "unit": "atom-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_m,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces + f64 state/sums",
        "data": (f"synthetic: seeded FCC lattice + Maxwell-Boltzmann velocities (seed 5489), "
                 f"advanced {args.prepare} untimed NVE steps into the liquid (input preparation, "
                 f"before the warm-up; the lattice melts within them) in each precision"),
        "config": {"workload": "NVE LJ liquid N=256,000 (FCC 40^3), rho=0.8442, T=0.72, "
                               "rc=2.5, dt=0.005, cell lists updated every step, mixed precision",
                   "resident_cells": resident,
                   "prepared_steps": args.prepare,
                   "particles": n, "steps": args.steps, "precision": "mixed",
                   "pair_kernel": args.pair_kernel, "parallelism": "single GPU"},
        "energy_gate": gates or None,
        "energy_gate_pass": (all(g["pass"] for g in gates.values()) if gates else None),
        "fp64": res64,
        "mixed_over_fp64": (rate_m / res64["value"]) if res64 else None,
        "mixed_over_fp64_10000_steps": ((gate_m["rate_10000_steps"] /
                                         gates["fp64"]["rate_10000_steps"])
                                        if gate_m and "fp64" in gates else None),
        "force_rms_rel_err_vs_fp64": frms,
        "roofline": roof,
        "kernel_ms": ktm,
        # time per step of the 10^4-step graph replay less the per-step kernels'
        # event-timed durations (which read 2-5% high): the most a single
        # persistent launch could save
        "launch_gaps_us_per_step": (1e3 * (gate_m["ms_per_step_10000_steps"] - step_kernel_ms(ktm))
                                    if gate_m else None),
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
def sweep_replicas(eng, seed=5489):
    """The 256 replicas: 16 x 16 grid of T in [0.5, 2] and rho in [0.05, 1.2]
    (the reference's linspace arithmetic, temperature-major), N = 4 n^3 with n
    from std::mt19937(seed) in [5, 20], velocity seed seed + k."""
    from lennardjonesium.tools import linspace
    grid = list(itertools.product(linspace(0.5, 2.0, 16), linspace(0.05, 1.2, 16)))
    sizes = eng.sweep_sizes(seed, len(grid))
    specs, states, meta = [], [], []
    for k, ((T, rho), n) in enumerate(zip(grid, sizes)):
        N = 4 * int(n) ** 3
        pos, vel, L = eng.initial_condition(N, rho, T, seed=seed + k)
        specs.append(eng.SystemSpec(N, L))
        states.append((pos, vel))
        meta.append((T, rho, N, L))
    return specs, states, meta


RDF_EVERY = 100


def sweep_leg(e, states, steps, warmup, nsys):
    """Loads, warms up, then times `steps` steps in chunks of RDF_EVERY with a
    g(r) sample (200 bins to 4 sigma) after each chunk.  Returns records
    (steps, nsys, 7), g(r) counts (nsys, 200), samples, seconds."""
    for k, (p, v) in enumerate(states):
        e.set_state(k, p, v)
    e.reserve(max(RDF_EVERY, warmup))
    if warmup:
        e.step(warmup, records=False)
    recs = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(RDF_EVERY, steps - done)
        recs.append(e.step(k))
        e.rdf_sample(4.0, 200)
        done += k
    counts, ns = e.rdf_read(200)
    t = time.perf_counter() - t0
    return np.concatenate(recs), counts, ns, t


def observables(rec, counts, meta, blocks=10):
    """Per-replica time means: PE per particle and the virial pressure
    rho (<T> + <W> / 3N) (physics/analyzers.cpp:73-75), their standard errors
    from `blocks` block averages, and the g(r) counts."""
    Ns = np.array([m[2] for m in meta], dtype=np.float64)
    rhos = np.array([m[1] for m in meta])
    pe_t = rec[:, :, 2] / Ns
    P_t = rhos * (rec[:, :, 5] + rec[:, :, 4] / (3.0 * Ns))
    n = (len(rec) // blocks) * blocks

    def sem(x):
        b = x[:n].reshape(blocks, -1, x.shape[1]).mean(1)
        return b.std(0, ddof=1) / np.sqrt(blocks)

    return {"pe": pe_t.mean(0), "P": P_t.mean(0), "pe_sem": sem(pe_t), "P_sem": sem(P_t),
            "rdf": counts}


def sweep(args, eng):
    steps = args.steps
    specs, states, meta = sweep_replicas(eng)
    total = sum(m[2] for m in meta)
    log(f"sweep: {len(specs)} replicas, {total} particles, steps={steps}")
    out, obs = {}, {}
    for prec in (("mixed",) if args.no_fp64 else ("mixed", "fp64")):
        e = eng.Engine(specs, precision=prec)
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
        variant, ms = e.autotune(10)
        kinds = e.pair_kernels()
        shape = e.launch_shape()
        n_tiled = sum(1 for x in kinds if x == "tiled")
        log(f"{prec}: autotuner chose {variant} ({ms:.3f} ms/step; "
            f"{n_tiled} replicas tiled, {len(kinds) - n_tiled} per-particle; "
            f"{shape[0]} cells per tiled block, sort waves of {shape[1]} particles)")
        rec, counts, ns, t = sweep_leg(e, states, steps, args.warmup, len(specs))
        e.close()
        rate = total * steps / t
        # conservation per replica, relative to max(|E_0|, KE_0 + |PE_0|)
        drifts = []
        for k in range(len(specs)):
            s = max(abs(rec[0, k, 3]), rec[0, k, 1] + abs(rec[0, k, 2]))
            drifts.append(energy_stats(rec[:, k, 3], scale=s)["max_dev"])
        drifts = np.array(drifts)
        ok = drifts <= GATE
        work_ok = sum(meta[k][2] for k in range(len(specs)) if ok[k]) * steps
        worst = int(np.argmax(drifts))
        out[prec] = dict(value=rate, seconds=t, pair_kernel=variant,
                         launch_shape={"cells_per_tiled_block": shape[0],
                                       "sort_wave_particles": shape[1]},
                         replicas_tiled=n_tiled, replicas_per_particle=len(kinds) - n_tiled,
                         energy_definition="max_t |E_t - E_0| / max(|E_0|, KE_0 + |PE_0|) "
                                           "over the run, per replica",
                         replicas_over_gate=int((~ok).sum()),
                         over_gate=[{"k": int(k), "T": meta[k][0], "rho": meta[k][1],
                                     "N": meta[k][2], "dev": float(drifts[k])}
                                    for k in np.argsort(-drifts) if not ok[k]],
                         value_replicas_within_gate=work_ok / t,
                         max_replica_dev=float(drifts.max()),
                         median_replica_dev=float(np.median(drifts)),
                         worst_replica={"T": meta[worst][0], "rho": meta[worst][1],
                                        "N": meta[worst][2], "E0": float(rec[0, worst, 3]),
                                        "E_end": float(rec[-1, worst, 3])},
                         rdf_samples=ns)
        obs[prec] = observables(rec, counts, meta)
        log(f"{prec}: {t:.2f} s, {rate:.4e} atom-steps/s aggregate, "
            f"{int((~ok).sum())} replicas above {GATE:g}")
    # plain per-replica FP64 runs: one engine per replica, as the reference runs
    # one Simulation per worker thread; every replica, the same steps and g(r)
    # samples, buffers and graphs prepared before each replica's timed steps
    plain, accept = None, None
    if not args.no_plain:
        t_plain, recs, cnts = 0.0, [], []
        for k in range(len(specs)):
            e = eng.Engine([specs[k]], precision="fp64", pair_kernel="split")
            rec, counts, _, t = sweep_leg(e, [states[k]], steps, args.warmup, 1)
            e.close()
            t_plain += t
            recs.append(rec[:, 0, :])
            cnts.append(counts[0])
        plain_rate = total * steps / t_plain
        plain = {"value": plain_rate, "seconds": t_plain, "replicas": len(specs),
                 "steps": steps, "pair_kernel": "split (occupancy 6, as the batched runs' rule)"}
        log(f"plain per-replica fp64: {plain_rate:.4e} atom-steps/s over all {len(specs)} replicas")
        po = observables(np.stack(recs, axis=1), np.stack(cnts), meta)
        accept = {}
        for prec, ob in obs.items():
            gr_l1 = np.abs(ob["rdf"] - po["rdf"]).sum(1) / np.maximum(po["rdf"].sum(1), 1)
            z = {}
            for key in ("pe", "P"):
                err = np.sqrt(ob[key + "_sem"] ** 2 + po[key + "_sem"] ** 2)
                z[key] = np.abs(ob[key] - po[key]) / np.maximum(err, 1e-300)
            accept[prec] = {
                "identical": bool(all(np.array_equal(ob[k], po[k]) for k in ("pe", "P", "rdf"))),
                "z_mean_pe": {"median": float(np.median(z["pe"])), "max": float(z["pe"].max()),
                              "replicas_within_3": int((z["pe"] <= 3).sum())},
                "z_pressure": {"median": float(np.median(z["P"])), "max": float(z["P"].max()),
                               "replicas_within_3": int((z["P"] <= 3).sum())},
                "rdf_l1_rel": {"median": float(np.median(gr_l1)), "max": float(gr_l1.max())}}
        accept["criterion"] = (
            "fp64 batched: identical to the per-replica runs (same trajectories, bit for bit); "
            "mixed batched: |mean_batched - mean_plain| / sqrt(sem_b^2 + sem_p^2) per replica "
            "(standard errors from 10 block averages of each run; chaotic trajectories "
            "diverge, so the means agree statistically), and the relative L1 distance "
            "of the g(r) histograms")
    return {"metric": METRIC, "value": out["mixed"]["value"], "unit": "atom-steps/s",
            "n_gpus": 1, "steps": steps, "warmup": args.warmup, "higher_is_better": True,
            "dtype": "f32 pair forces + f64 state/sums", "data": "synthetic",
            "config": {"workload": "batched sweep: 256 replicas, 16x16 grid T 0.5-2.0 x rho "
                                   "0.05-1.2, N=4n^3, n in [5,20] from seed 5489, rc 2.5, "
                                   "dt 0.005, g(r) every 100 steps", "particles": total,
                       "replicas": len(specs)},
            "mixed_batched": out["mixed"], "fp64_batched": out.get("fp64"),
            "plain_fp64_per_replica": plain,
            "batched_mixed_over_plain_fp64": (out["mixed"]["value"] / plain["value"]
                                              if plain else None),
            "batched_fp64_over_plain_fp64": (out["fp64"]["value"] / plain["value"]
                                             if plain and "fp64" in out else None),
            "acceptance_vs_plain": accept}


# ---------------------------------------------------------------------------
def coexistence(args, eng):
    n, rho, T, seed, rc, dt = 2048000, 0.30, 0.80, 5489, 3.0, 0.004
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    log(f"coexistence: N={n} L={L:.4f} steps={steps}")
    spec = [eng.SystemSpec(n, L, cutoff=rc, time_delta=dt)]
    out = {}
    for prec in (("mixed",) if args.no_fp64 else ("mixed", "fp64")):
        e = eng.Engine(spec, precision=prec)
        e.set_state(0, pos, vel)
        # tune on a state the run spends its time in: the quenched lattice breaks
        # up into liquid and vapour within the first steps, and the fastest kernel
        # on the lattice is not the fastest afterwards (the timed run still starts
        # from the lattice)
        e.step(1000, records=False)
        variant, _ = e.autotune(20)
        shape = e.launch_shape()
        nst = steps if prec == "mixed" else min(steps, args.fp64_steps)
        rec, t = timed_run(e, nst, 0, [(pos, vel)])
        E = rec[:, 0, 3]
        alt = rec[0, 0, 1] + abs(rec[0, 0, 2])
        late = None
        if nst >= 2 * GATE_STEPS:
            # after the quench transient: the last 10^4 steps
            late = energy_stats(E[-GATE_STEPS:])
            late["pass"] = late["max_dev"] <= GATE
        start, _ = e.cell_list(0)
        occ = np.diff(start)
        ktm = kernel_times(e, 20)
        e.close()
        out[prec] = {"value": n * nst / t, "steps": nst, "ms_per_step": 1e3 * t / nst,
                     "pair_kernel": variant,
                     "launch_shape": {"cells_per_tiled_block": shape[0],
                                      "sort_wave_particles": shape[1]},
                     "energy_vs_abs_e0": energy_stats(E),
                     "energy_vs_ke0_plus_abs_pe0": energy_stats(E, scale=alt),
                     "energy_last_10000_steps": late,
                     "E0": float(E[0]), "KE0": float(rec[0, 0, 1]), "PE0": float(rec[0, 0, 2]),
                     "load_imbalance": {"cell_occupancy_max_over_mean":
                                        float(occ.max() / occ.mean()),
                                        "occupancy_histogram": np.bincount(occ).tolist()},
                     "kernel_ms": ktm}
        log(f"coexistence {prec}: {n * nst / t:.4e} atom-steps/s over {nst} steps, "
            f"max dev {out[prec]['energy_vs_abs_e0']['max_dev']:.2e} of |E0|, "
            f"{out[prec]['energy_vs_ke0_plus_abs_pe0']['max_dev']:.2e} of KE0+|PE0|")
    m = out["mixed"]
    bytes_step = 224.0 * n
    return {"metric": METRIC, "value": m["value"], "unit": "atom-steps/s", "n_gpus": 1,
            "steps": steps, "dtype": "f32 pair forces + f64 state/sums", "data": "synthetic",
            "config": {"workload": "coexistence N=2,048,000 (FCC 80^3) rho 0.30 T 0.80 rc 3.0 "
                                   "dt 0.004, cell lists updated every step", "pair_kernel":
                       m["pair_kernel"]},
            "mixed": m, "fp64": out.get("fp64"),
            "mixed_over_fp64": (m["value"] / out["fp64"]["value"]) if "fp64" in out else None,
            "algorithmic_hbm_GBps": bytes_step / (m["ms_per_step"] * 1e-3) / 1e9}


# ---------------------------------------------------------------------------
def plain(args, eng):
    n, rho, T, seed = 864, 0.8, 1.0, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    e = eng.Engine([eng.SystemSpec(n, L)], precision="fp64")
    e.set_state(0, pos, vel)
    variant, _ = e.autotune(20)   # untimed; restores the state
    log(f"plain: autotuner chose {variant}")
    rec, t = timed_run(e, steps, args.warmup, [(pos, vel)])
    e.close()
    rate_cpu, dt_cpu = cpu_baseline(n, rho, T, seed, steps)
    gate = energy_gate(eng, [eng.SystemSpec(n, L)], [(pos, vel)], "fp64")
    log10 = rec[::10, 0, :]
    return {"metric": METRIC, "value": n * steps / t, "unit": "atom-steps/s", "n_gpus": 1,
            "steps": steps, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "plain FP64 N=864 (FCC 6^3) rho 0.8 T 1.0 rc 2.5 dt 0.005",
                       "pair_kernel": variant},
            "energy_gate": gate,
            "energy_timed_run": energy_stats(rec[:, 0, 3]),
            "log_every_10_first5": {"columns": ["time", "KE", "PE", "E", "W", "T", "MSD",
                                                "P_virial"],
                                    "rows": [list(r) + [rho * (r[5] + r[4] / (3 * n))]
                                             for r in log10[:5].tolist()]},
            "cpu_baseline": cpu_record(rate_cpu, dt_cpu, f"{steps} FP64 steps of the same "
                                                         f"N=864 workload")}


def adjoint(args, eng):
    n, rho, T, seed = 4000, 0.8, 1.0, 5489
    pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
    steps = args.steps
    e = eng.Engine([eng.SystemSpec(n, L)], precision="fp64")
    e.set_state(0, pos, vel)
    variant, _ = e.autotune(20)   # the trajectory's pair kernel (untimed; restores the state)
    log(f"adjoint: autotuner chose {variant}")
    t0 = time.perf_counter()
    loss, ge, gs, gr, gv = e.adjoint(steps, rho, every=100)
    t = time.perf_counter() - t0
    e.close()
    h = 1e-6

    def fwd(nsteps, eps=1.0, sig=1.0):
        with eng.Engine([eng.SystemSpec(n, L, epsilon=eps, sigma=sig)]) as f:
            f.set_state(0, pos, vel)
            return f.adjoint(nsteps, rho, every=nsteps)[0]

    # agreement with central differences against the horizon: the exact
    # derivative of a chaotic trajectory grows like exp(lambda t)
    horizon = []
    for hs in (50, 100, 200, 400, 800, steps):
        with eng.Engine([eng.SystemSpec(n, L)]) as f:
            f.set_state(0, pos, vel)
            _, ge_h, gs_h, _, _ = f.adjoint(hs, rho, every=min(100, hs))
        fd_e = (fwd(hs, eps=1 + h) - fwd(hs, eps=1 - h)) / (2 * h)
        fd_s = (fwd(hs, sig=1 + h) - fwd(hs, sig=1 - h)) / (2 * h)
        horizon.append({"steps": hs, "adjoint_eps": ge_h, "fd_eps": fd_e,
                        "rel_err_eps": abs(ge_h - fd_e) / abs(fd_e),
                        "adjoint_sigma": gs_h, "fd_sigma": fd_s,
                        "rel_err_sigma": abs(gs_h - fd_s) / abs(fd_s)})
        log(f"horizon {hs}: eps adjoint {ge_h:.6e} fd {fd_e:.6e}; sigma {gs_h:.6e} fd {fd_s:.6e}")
    h200 = next(x for x in horizon if x["steps"] == 200)
    return {"metric": "adjoint trajectory seconds", "value": t, "unit": "s", "n_gpus": 1,
            "steps": steps, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "differentiable trajectory N=4000 rho 0.8 T 1.0, checkpoints "
                                   "every 100 steps", "pair_kernel": variant},
            "loss": loss, "grad_eps": ge, "grad_sigma": gs,
            "rel_err_eps_200_steps": h200["rel_err_eps"],
            "rel_err_sigma_200_steps": h200["rel_err_sigma"],
            "grad_r0_norm": float(np.linalg.norm(gr)), "horizon": horizon,
            "note": "single chaotic trajectory: the exact gradient grows like exp(lambda t), so "
                    "agreement with finite differences holds up to a few hundred steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--prepare", type=int, default=1000,
                    help="liquid: untimed steps that take the lattice into the liquid before "
                         "the warm-up (0: time from the lattice)")
    ap.add_argument("--fp64-steps", type=int, default=2000,
                    help="coexistence: steps of the FP64 leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-gate", action="store_true",
                    help="liquid: skip the 10^4-step energy-gate legs")
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
