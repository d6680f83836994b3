"""``run_sweep()``: a temperature x density grid of simulations
(signature and file layout of src/python/lennardjonesium/orchestration/run_sweep.py:37-336).

All simulations of the chunk run together as one batched GPU engine (the
SimulationPool); ``thread_count`` is kept for compatibility and reported in
the summary, which, like the reference's, prices time per thread."""

import os
import pathlib
import time
from copy import deepcopy
from types import BuiltinFunctionType, FunctionType
from typing import Optional, Union

from lennardjonesium.orchestration.sweep_configuration import SweepConfiguration
from lennardjonesium.orchestration.sweep_result import SweepResult
from lennardjonesium.simulation import Configuration, Simulation, SimulationPool


def run_sweep(sweep_config_file: Union[str, pathlib.Path], echo_status: bool = True,
              polling_interval: float = 0.5, thread_count: int = 4,
              random_seed: Union[None, int, FunctionType, BuiltinFunctionType] = None,
              chunk_count: int = 1, chunk_index: int = 0,
              sweep_config_object: Optional[SweepConfiguration] = None,
              precision: str = "fp64") -> SweepResult:
    path = pathlib.Path(sweep_config_file).resolve()
    if sweep_config_object is None and not path.is_file():
        raise ValueError(
            'sweep_config_object must be given if sweep_config_file is not an existing file')
    if sweep_config_object is None:
        cfg = SweepConfiguration.from_file(path)
    else:
        cfg = deepcopy(sweep_config_object)
        path.parent.mkdir(parents=True, exist_ok=True)
        cfg.write(path)
    cwd = os.getcwd()
    os.chdir(path.parent)
    try:
        sims = _create_simulations(cfg, random_seed, chunk_count, chunk_index, precision)
        pool = SimulationPool(thread_count)
        if echo_status:
            _preamble(cfg, len(sims), thread_count, chunk_count, chunk_index)
        t0 = time.perf_counter()
        for sim in sims:
            pool.push(sim)
        if echo_status:
            _report(pool, len(sims), t0, polling_interval)
        pool.wait()
        elapsed = time.perf_counter() - t0
    finally:
        os.chdir(cwd)
    result = SweepResult(path, chunk_count, chunk_index)
    if echo_status:
        _postamble(result, thread_count, elapsed)
    return result


def _create_simulations(cfg, random_seed, chunk_count, chunk_index, precision):
    sims = []
    for T, rho in cfg.sweep_range(chunk_count, chunk_index):
        sim_dir = cfg.simulation_dir(T, rho)
        run_ini = sim_dir / cfg.templates.run_config_file
        run_cfg = _run_configuration(cfg, T, rho)
        if random_seed is None and run_ini.is_file():
            run_cfg.system.random_seed = Configuration.from_file(run_ini).system.random_seed
        elif isinstance(random_seed, int):
            run_cfg.system.random_seed = random_seed
        elif random_seed is not None:
            run_cfg.system.random_seed = random_seed()
        run_ini.parent.mkdir(parents=True, exist_ok=True)
        run_cfg.write(run_ini)
        fp = run_cfg.filepaths
        fp.event_log = str(sim_dir / fp.event_log)
        fp.thermodynamic_log = str(sim_dir / fp.thermodynamic_log)
        fp.observation_log = str(sim_dir / fp.observation_log)
        fp.snapshot_log = str(sim_dir / fp.snapshot_log)
        sims.append(Simulation(run_cfg, precision=precision))
    return sims


def _run_configuration(cfg, T, rho):
    rc = Configuration()
    rc.system.temperature = T
    rc.system.density = rho
    rc.system.particle_count = cfg.system.particle_count
    rc.system.cutoff_distance = cfg.system.cutoff_distance
    rc.system.time_delta = cfg.system.time_delta
    for src, dst in ((cfg.equilibration, rc.equilibration), (cfg.observation, rc.observation)):
        for key in vars(src):
            setattr(dst, key, getattr(src, key))
        dst.name = cfg.templates.phase_name.format(temperature=T, density=rho, name=src.name)
    for key in vars(cfg.filenames):
        setattr(rc.filepaths, key, getattr(cfg.filenames, key))
    return rc


def _preamble(cfg, jobs, threads, chunk_count, chunk_index):
    s = cfg.system
    print("\n".join([
        "================= Lennard-Jones Sweep Simulation ==================", "",
        f"Temperature: {s.temperature_steps} points in the range "
        f"[{s.temperature_start}, {s.temperature_stop}]",
        f"Density: {s.density_steps} points in the range [{s.density_start}, {s.density_stop}]",
        "", f"Number of Particles: {s.particle_count}", f"Time Step: {s.time_delta}", "",
        f"Chunk {chunk_index + 1} of {chunk_count} (index {chunk_index})",
        f"Running {jobs} jobs as one batched GPU engine", "", "Begin simulation sweep..."]),
        flush=True)


def _report(pool, jobs, t0, interval):
    while True:
        st = pool.status()
        print(f"Jobs queued: {st.waiting}, Running: {st.running}, Completed: {st.completed}, "
              f"Elapsed time: {time.perf_counter() - t0:.2f} seconds     ", end="\r", flush=True)
        if st.completed == jobs:
            break
        time.sleep(interval)
    print()


def _postamble(result, threads, elapsed):
    done = result.completed + result.equilibration_aborted + result.observation_aborted
    steps = sum(r.event_data.total_time_steps for r in done)
    per_thread = steps / threads
    sps = per_thread / elapsed if elapsed > 0 else float("nan")
    print("\n".join([
        "End simulation sweep", "", "Job status:", f"Completed: {len(result.completed)}",
        f"Aborted during Equilibration: {len(result.equilibration_aborted)}",
        f"Aborted during Observation: {len(result.observation_aborted)}", "",
        f"{steps} time steps computed using {threads} threads in {elapsed:.3f} seconds",
        f"{elapsed / per_thread if per_thread else float('nan'):.3f} thread-seconds per step, "
        f"or {1000 * elapsed / per_thread if per_thread else float('nan'):.3f} milliseconds",
        f"Average framerate per thread: {sps:.3f} fps"]), flush=True)
