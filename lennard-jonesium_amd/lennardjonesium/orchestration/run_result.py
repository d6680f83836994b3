"""Summary of one run read back from its events.log
(semantics of src/python/lennardjonesium/orchestration/run_result.py:31-127)."""

import pathlib
from enum import Enum, auto
from typing import Optional

from lennardjonesium.simulation import Configuration


class SimulationStatus(Enum):
    completed = auto()
    equilibration_aborted = auto()
    observation_aborted = auto()


class RunResult:
    simulation_status: Optional[SimulationStatus] = None
    equilibration_started: Optional[int] = None
    equilibration_completed: Optional[int] = None
    equilibration_time_steps: Optional[int] = None
    observation_started: Optional[int] = None
    observation_completed: Optional[int] = None
    observation_time_steps: Optional[int] = None
    total_time_steps: int = 0

    def __init__(self, config_filepath):
        config_filepath = pathlib.Path(config_filepath)
        cfg = Configuration.from_file(config_filepath)
        log = pathlib.Path(cfg.filepaths.event_log)
        if not log.is_absolute():
            log = config_filepath.parent / log
        self.temperature_adjustments = []
        self.observations_recorded = []
        with open(log) as fh:
            self._parse(fh.readlines())

    def _parse(self, lines):
        for line in lines:
            head, text = line.split(': ', 1)
            step = int(head)
            self.total_time_steps = step
            if text.startswith('Phase started'):
                if self.equilibration_started is None:
                    self.equilibration_started = step
                else:
                    self.observation_started = step
            elif text.startswith('Phase complete'):
                if self.equilibration_completed is None:
                    self.equilibration_completed = step
                    self.equilibration_time_steps = step - self.equilibration_started
                else:
                    self.observation_completed = step
                    self.observation_time_steps = step - self.observation_started
                    self.simulation_status = SimulationStatus.completed
            elif text.startswith('Temperature'):
                self.temperature_adjustments.append(step)
            elif text.startswith('Observation'):
                self.observations_recorded.append(step)
            elif text.startswith('Simulation aborted'):
                self.simulation_status = (SimulationStatus.equilibration_aborted
                                          if self.equilibration_completed is None
                                          else SimulationStatus.observation_aborted)
