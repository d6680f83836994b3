"""CPU tests of the Python / host API that mirrors the reference's
(configuration round trips, sweep grid and naming, run-result parsing)."""

import csv
import pathlib

import numpy as np
import pytest

from tests.conftest import GOLDEN


def test_configuration_from_dict():
    # tests/python/lennardjonesium/simulation/test_configuration.py
    from lennardjonesium.simulation import Configuration
    d = {'system': {'temperature': 0.7, 'density': 0.3, 'particle_count': 10},
         'equilibration': {'name': "T=0.7, d=0.3 Equilibration Phase", 'timeout': 2000},
         'observation': {'name': "T=0.7, d=0.3 Observation Phase", 'observation_count': 5},
         'filepaths': {'event_log': 'data/events.log'}}
    c = Configuration.from_dict(d)
    assert c.system.particle_count == 10 and c.equilibration.timeout == 2000
    assert c.to_dict()['filepaths']['event_log'] == 'data/events.log'


def test_configuration_ini_round_trip(tmp_path):
    from lennardjonesium.simulation import Configuration
    src = Configuration()
    src.write(tmp_path / 'test.ini')
    assert Configuration.from_file(tmp_path / 'test.ini') == src
    # the reference's own test input and the fixture's run.ini parse with their types
    c = Configuration.from_file(GOLDEN / 'test_simulation.ini')
    assert c.system.random_seed == 17 and c.observation.observation_count == 19
    c = Configuration.from_file(GOLDEN / 'quick_single_run/run.ini')
    assert (c.system.random_seed, c.equilibration.adjustment_interval) == (117208264, 50)
    assert isinstance(c.system.temperature, float) and c.system.particle_count == 500


def test_sweep_grid_matches_reference_sweep():
    """notebooks/quick_sweep_run: 540 runs; T-major order, linspace arithmetic,
    directory and phase-name templates."""
    from lennardjonesium import SweepConfiguration
    from lennardjonesium.orchestration.run_sweep import _run_configuration
    cfg = SweepConfiguration.from_file(GOLDEN / 'quick_sweep_run/sweep.ini')
    rows = list(csv.DictReader(open(GOLDEN / 'quick_sweep_run/runs.csv')))
    pts = list(cfg.sweep_range())
    assert len(pts) == len(rows) == 540
    for (T, rho), row in zip(pts, rows):
        assert str(cfg.simulation_dir(T, rho)) == row['directory']
        rc = _run_configuration(cfg, T, rho)
        assert str(rc.system.temperature) == row['temperature']
        assert str(rc.system.density) == row['density']
        assert rc.equilibration.name == row['equilibration_name']
        assert rc.observation.name == row['observation_name']
    assert sum(int(r['total_time_steps']) for r in rows) == 6586600


def test_sweep_chunks_partition_the_grid():
    from lennardjonesium import SweepConfiguration
    cfg = SweepConfiguration()
    whole = list(cfg.sweep_range())
    parts = [list(cfg.sweep_range(7, k)) for k in range(7)]
    assert sum(parts, []) == whole
    assert [len(p) for p in parts] == [15] * 2 + [14] * 5


def test_run_result_parses_fixture_events(tmp_path):
    from lennardjonesium.orchestration import RunResult, SimulationStatus
    import shutil
    shutil.copy(GOLDEN / 'quick_single_run/run.ini', tmp_path / 'run.ini')
    shutil.copy(GOLDEN / 'quick_single_run/events.log', tmp_path / 'events.log')
    r = RunResult(tmp_path / 'run.ini')
    assert r.simulation_status == SimulationStatus.completed
    assert (r.equilibration_started, r.equilibration_completed, r.observation_completed) == (0, 600, 1600)
    assert r.temperature_adjustments == [50, 100] and len(r.observations_recorded) == 20
    assert r.total_time_steps == 1600


def test_simulation_pair_functions_match_oracle():
    from oracle import ljref
    from lennardjonesium.simulation import Configuration, Simulation
    s = Simulation(Configuration())
    for r in (0.95, 1.12, 1.5, 2.2, 2.6):
        v, w, fz = ljref.lj_along_z(r, 2.5)
        assert s.potential(r) == v and s.virial(r) == w and s.force(r) == pytest.approx(fz, rel=1e-15)
