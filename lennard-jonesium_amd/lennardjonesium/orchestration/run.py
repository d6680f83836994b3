"""``run()``: one simulation from a config file
(signature and file semantics of src/python/lennardjonesium/orchestration/run.py:35-107)."""

import os
import pathlib
import time
from copy import deepcopy
from types import BuiltinFunctionType, FunctionType
from typing import Optional, Union

from lennardjonesium.orchestration.run_result import RunResult
from lennardjonesium.simulation import Configuration, Simulation


def _resolve_seed(random_seed, current):
    if isinstance(random_seed, int):
        return random_seed
    if isinstance(random_seed, (FunctionType, BuiltinFunctionType)) or callable(random_seed):
        return random_seed()
    return current


def run(config_file: Union[str, pathlib.Path], echo_status: bool = True,
        random_seed: Union[None, int, FunctionType, BuiltinFunctionType] = None,
        config_object: Optional[Configuration] = None, precision: str = "fp64") -> RunResult:
    """Runs one simulation on the GPU and returns its RunResult.

    The config file is (re)written with the seed actually used; output files
    are placed relative to the config file's directory.  ``precision`` selects
    the engine arithmetic ('fp64' as the reference, or 'mixed')."""
    path = pathlib.Path(config_file).resolve()
    if config_object is None and not path.is_file():
        raise ValueError('config_object must be given if config_file is not an existing file')
    cfg = Configuration.from_file(path) if config_object is None else deepcopy(config_object)
    cfg.system.random_seed = _resolve_seed(random_seed, cfg.system.random_seed)
    path.parent.mkdir(parents=True, exist_ok=True)
    cfg.write(path)
    if echo_status:
        _preamble(cfg)
    cwd = os.getcwd()
    os.chdir(path.parent)
    try:
        sim = Simulation(cfg, precision=precision)
        t0 = time.perf_counter()
        sim.run(echo=echo_status)
        elapsed = time.perf_counter() - t0
    finally:
        os.chdir(cwd)
    result = RunResult(path)
    if echo_status:
        _postamble(result, elapsed)
    return result


def _preamble(cfg):
    s, f = cfg.system, cfg.filepaths
    print("\n".join([
        "==================== Lennard-Jones Simulation ====================", "",
        f"Temperature: {s.temperature}", f"Density: {s.density}",
        f"Number of Particles: {s.particle_count}", f"Time Step: {s.time_delta}",
        f"Random Seed: {s.random_seed}", "",
        "Simulation will run in 2 phases:", f"Phase 1: {cfg.equilibration.name}",
        f"Phase 2: {cfg.observation.name}", "",
        "Results to be output to the following files:", f"Events: {f.event_log}",
        f"Thermodynamics: {f.thermodynamic_log}", f"Observations: {f.observation_log}",
        f"Snapshots: {f.snapshot_log}", "", "Begin simulation..."]), flush=True)


def _postamble(result, elapsed):
    steps = result.total_time_steps
    per = elapsed / steps if steps else float("nan")
    fps = steps / elapsed if elapsed > 0 else float("nan")
    print("\n".join([
        "End simulation", "",
        f"{steps} time steps computed in {elapsed:.3f} seconds",
        f"{per:.3f} seconds per step, or {per * 1000:.3f} milliseconds",
        f"Framerate: {fps:.3f} fps"]), flush=True)
