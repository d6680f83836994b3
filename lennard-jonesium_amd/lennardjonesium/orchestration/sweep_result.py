"""Classification of a sweep's runs by status
(src/python/lennardjonesium/orchestration/sweep_result.py:33-83)."""

import pathlib
from dataclasses import dataclass

from lennardjonesium.orchestration.run_result import RunResult, SimulationStatus
from lennardjonesium.orchestration.sweep_configuration import SweepConfiguration


class SweepResult:
    @dataclass
    class _SimulationResult:
        simulation_dir: pathlib.Path
        event_data: RunResult

    def __init__(self, sweep_config_file, chunk_count: int = 1, chunk_index: int = 0):
        self.completed = []
        self.equilibration_aborted = []
        self.observation_aborted = []
        sweep_config_file = pathlib.Path(sweep_config_file)
        cfg = SweepConfiguration.from_file(sweep_config_file)
        bins = {SimulationStatus.completed: self.completed,
                SimulationStatus.equilibration_aborted: self.equilibration_aborted,
                SimulationStatus.observation_aborted: self.observation_aborted}
        for sim_dir in cfg.simulation_dir_range(chunk_count, chunk_index):
            rr = RunResult(sweep_config_file.parent / sim_dir / cfg.templates.run_config_file)
            bins[rr.simulation_status].append(self._SimulationResult(sim_dir, rr))
