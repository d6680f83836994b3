"""Entry points (the reference's lennardjonesium.orchestration package)."""

from lennardjonesium.orchestration.run import run  # noqa: F401
from lennardjonesium.orchestration.sweep_configuration import SweepConfiguration  # noqa: F401
from lennardjonesium.orchestration.sweep_result import SweepResult  # noqa: F401
from lennardjonesium.orchestration.run_sweep import run_sweep  # noqa: F401
from lennardjonesium.orchestration.run_result import RunResult, SimulationStatus  # noqa: F401
