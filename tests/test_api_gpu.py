"""GPU tests of lj.run / lj.run_sweep against the reference's own outputs."""

import csv
import shutil

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _rows(path):
    rows = list(csv.reader(open(path)))
    return rows[0], np.array([[float(x) for x in r] for r in rows[1:]])


def test_run_reproduces_reference_single_run(tmp_path):
    """notebooks/quick_single_run: same config and seed as the reference binary."""
    import lennardjonesium as lj
    shutil.copy(GOLDEN / 'quick_single_run/run.ini', tmp_path / 'run.ini')
    res = lj.run(tmp_path / 'run.ini', echo_status=False)
    assert (tmp_path / 'events.log').read_text() == (GOLDEN / 'quick_single_run/events.log').read_text()
    assert res.total_time_steps == 1600 and res.temperature_adjustments == [50, 100]
    hdr, got = _rows(tmp_path / 'thermodynamics.csv')
    rhdr, ref = _rows(GOLDEN / 'quick_single_run/thermodynamics_head.csv')
    assert hdr == rhdr and len(got) == 1600
    rel = np.abs(got[:200, 1:] - ref[:, 1:]) / np.abs(ref[:, 1:])
    assert rel.max() < 1e-10
    ohdr, obs = _rows(tmp_path / 'observations.csv')
    rohdr, robs = _rows(GOLDEN / 'quick_single_run/observations.csv')
    assert ohdr == rohdr and obs[:, 0].tolist() == robs[:, 0].tolist()
    np.testing.assert_allclose(obs[:, 1:], robs[:, 1:], rtol=1e-3, atol=1e-4)
    snap = (tmp_path / 'snapshots.csv').read_text().splitlines()
    assert len(snap) == 502 and snap[2].startswith('1600,0,')


def test_reference_python_test_simulation(tmp_path):
    # tests/python/lennardjonesium/simulation/test_simulation.py: the observation
    # log of a completed run has observation_count + 1 lines
    import lennardjonesium as lj
    cfg = lj.Configuration.from_file(GOLDEN / 'test_simulation.ini')
    for name in ('event_log', 'thermodynamic_log', 'observation_log', 'snapshot_log'):
        p = tmp_path / getattr(cfg.filepaths, name)
        p.parent.mkdir(parents=True, exist_ok=True)
        setattr(cfg.filepaths, name, str(p))
    lj.Simulation(cfg).run(echo=False)
    n = sum(1 for _ in open(cfg.filepaths.observation_log))
    assert n == cfg.observation.observation_count + 1


def _small_sweep():
    import lennardjonesium as lj
    s = lj.SweepConfiguration()
    s.system.temperature_start, s.system.temperature_stop, s.system.temperature_steps = 0.6, 1.2, 2
    s.system.density_start, s.system.density_stop, s.system.density_steps = 0.5, 0.9, 2
    s.system.particle_count = 108
    s.equilibration.adjustment_interval = 50
    s.equilibration.steady_state_time = 200
    s.equilibration.timeout = 1000
    s.equilibration.sample_size = 20
    s.observation.observation_interval = 50
    s.observation.observation_count = 4
    s.observation.sample_size = 20
    return s


def test_batched_sweep_equals_individual_runs(tmp_path):
    """All replicas in one engine give byte-identical files to separate runs."""
    import lennardjonesium as lj
    s = _small_sweep()
    res = lj.run_sweep(tmp_path / 'sweep' / 'sweep.ini', echo_status=False, random_seed=11,
                       sweep_config_object=s)
    assert len(res.completed) + len(res.equilibration_aborted) + len(res.observation_aborted) == 4
    for T, rho in s.sweep_range():
        d = s.simulation_dir(T, rho)
        single = tmp_path / 'single' / d
        single.mkdir(parents=True)
        shutil.copy(tmp_path / 'sweep' / d / 'run.ini', single / 'run.ini')
        lj.run(single / 'run.ini', echo_status=False)
        for f in ('events.log', 'thermodynamics.csv', 'observations.csv', 'snapshots.csv'):
            assert (single / f).read_bytes() == (tmp_path / 'sweep' / d / f).read_bytes(), (d, f)
