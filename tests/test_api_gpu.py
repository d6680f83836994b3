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


# ---- abort paths of the controller (control/simulation_phase.cpp:76-80, 111-118;
#      control/simulation_controller.cpp:124-135) ----------------------------------

def _config(tmp_path, **sections):
    import lennardjonesium as lj
    cfg = lj.Configuration()
    cfg.system.temperature, cfg.system.density = 1.0, 0.8
    cfg.system.particle_count, cfg.system.random_seed = 108, 4242
    for sec, values in sections.items():
        for k, v in values.items():
            setattr(getattr(cfg, sec), k, v)
    ini = tmp_path / 'run.ini'
    cfg.write(ini)
    return cfg, ini


def _oracle_events(cfg):
    """The oracle's controller over the oracle's stepper from the same initial
    condition as the engine (the host's initial condition, loaded as is)."""
    from lennardjonesium import engine as eng
    from oracle import controller as C
    from oracle import ljref
    s = cfg.system
    pos, vel, L = eng.initial_condition(s.particle_count, s.density, s.temperature,
                                        seed=s.random_seed)
    st = ljref.State(pos=pos, vel=vel)
    e, o = cfg.equilibration, cfg.observation
    eq = C.EquilibrationPhase(e.name, s.temperature, C.EquilibrationParameters(
        e.tolerance, e.sample_size, e.adjustment_interval, e.steady_state_time, e.timeout))
    ob = C.ObservationPhase(o.name, s.temperature, s.density, s.particle_count,
                            C.ObservationParameters(o.tolerance, o.sample_size,
                                                    o.observation_interval, o.observation_count))
    events, rows, _ = C.run_controller(st, L, s.time_delta, s.cutoff_distance, [eq, ob])
    return events, np.array(rows), st


def _snapshot_steps(path):
    lines = path.read_text().splitlines()
    return len(lines) - 2, {int(x.split(',')[0]) for x in lines[2:]}


def test_equilibration_abort_could_not_equilibrate(tmp_path):
    """test_system_equilibration.cpp:178-249: a 1e-6 equilibration tolerance
    cannot be met; the phase times out with "Could not equilibrate", the
    controller logs the abort and a snapshot, and the run reads as aborted in
    equilibration.  events.log equals the oracle controller's line for line."""
    import lennardjonesium as lj
    from lennardjonesium.orchestration.run_result import RunResult, SimulationStatus
    cfg, ini = _config(tmp_path, equilibration=dict(
        tolerance=1e-6, sample_size=50, adjustment_interval=100, steady_state_time=5000,
        timeout=600))
    res = lj.run(ini, echo_status=False)
    events = (tmp_path / 'events.log').read_text().splitlines()
    want, rows, st = _oracle_events(cfg)
    assert events == want
    assert events[-1] == '600: Simulation aborted: Could not equilibrate'
    assert sum('adjusted to' in x for x in events) == 6
    r = RunResult(ini)
    assert r.simulation_status == SimulationStatus.equilibration_aborted
    assert r.total_time_steps == 600 and r.equilibration_completed is None
    assert res.simulation_status == SimulationStatus.equilibration_aborted
    count, steps = _snapshot_steps(tmp_path / 'snapshots.csv')
    assert count == 108 and steps == {600}
    hdr, got = _rows(tmp_path / 'thermodynamics.csv')
    assert len(got) == 600
    rel = np.abs(got[:100, 1:] - rows[:100, 1:]) / np.abs(rows[:100, 1:])
    assert rel.max() < 1e-10
    # the reference's check: the temperature stays outside the tolerance
    T_end = got[-50:, 6].mean()
    assert abs(T_end / 1.0 - 1.0) > 1e-6


def test_observation_abort_temperature_drift(tmp_path):
    """control/simulation_phase.cpp:111-118: an observation whose mean
    temperature is off by more than the (here 1e-6) tolerance aborts the run
    with "Temperature drifted outside acceptable range" and a snapshot; the run
    reads as aborted in observation.  events.log equals the oracle's."""
    import lennardjonesium as lj
    from lennardjonesium.orchestration.run_result import RunResult, SimulationStatus
    cfg, ini = _config(tmp_path,
                       equilibration=dict(tolerance=0.05, sample_size=20, adjustment_interval=50,
                                          steady_state_time=200, timeout=2000),
                       observation=dict(tolerance=1e-6, sample_size=20, observation_interval=50,
                                        observation_count=4))
    lj.run(ini, echo_status=False)
    events = (tmp_path / 'events.log').read_text().splitlines()
    want, _, _ = _oracle_events(cfg)
    assert events == want
    assert events[-1].endswith(': Simulation aborted: Temperature drifted outside acceptable range')
    r = RunResult(ini)
    assert r.simulation_status == SimulationStatus.observation_aborted
    step = int(events[-1].split(':')[0])
    assert step == r.observation_started + 50
    count, steps = _snapshot_steps(tmp_path / 'snapshots.csv')
    assert count == 108 and steps == {step}
    assert len((tmp_path / 'observations.csv').read_text().splitlines()) == 1


def test_observations_equal_the_analyzers_on_the_logged_records(tmp_path):
    """Every observations.csv line of a completed run, recomputed from the
    run's own thermodynamics.csv (printed as shortest round-trip doubles, so
    the records are exact): the host analyzer reproduces it bit for bit and
    the oracle's analyzer within 1e-12."""
    import lennardjonesium as lj
    from lennardjonesium import engine as eng
    from oracle import ljref
    cfg, ini = _config(tmp_path,
                       equilibration=dict(tolerance=0.05, sample_size=20, adjustment_interval=50,
                                          steady_state_time=200, timeout=2000),
                       observation=dict(tolerance=0.5, sample_size=40, observation_interval=60,
                                        observation_count=5))
    lj.run(ini, echo_status=False)
    _, th = _rows(tmp_path / 'thermodynamics.csv')
    _, obs = _rows(tmp_path / 'observations.csv')
    events = (tmp_path / 'events.log').read_text().splitlines()
    start = int(next(x for x in events if 'Phase started: Observation' in x).split(':')[0])
    assert len(obs) == 5
    n, rho, size = cfg.system.particle_count, cfg.system.density, cfg.observation.sample_size
    for row in obs:
        t = int(row[0])
        # records the phase collected: TimeStep start .. t - 1, the last `size` of them
        w = th[(th[:, 0] >= max(start, t - size)) & (th[:, 0] <= t - 1), 1:]
        assert len(w) == size
        host = eng.thermo_observation(w, size, rho, n)
        orc = ljref.observation(w[:, 5], w[:, 3], w[:, 4], w[:, 0], w[:, 6], rho, n)
        keys = ("temperature", "density", "total_energy", "pressure", "specific_heat",
                "diffusion_coefficient")
        assert [host[k] for k in keys] == row[1:].tolist()
        for k, v in zip(keys, row[1:]):
            assert abs(orc[k] - v) <= 1e-12 * abs(v), k
