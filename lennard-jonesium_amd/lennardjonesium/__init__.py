"""MI355X-native engine for lennard-jonesium's microcanonical LJ step loop.

The public names mirror the reference package
(src/python/lennardjonesium/__init__.py): ``run``, ``run_sweep``,
``SweepConfiguration``, ``SeedGenerator``, ``Configuration``, ``Simulation``,
``SimulationPool``.  ``Engine`` / ``SystemSpec`` expose the GPU step engine
directly.
"""

from lennardjonesium.orchestration import run, SweepConfiguration, run_sweep  # noqa: F401
from lennardjonesium.simulation import (SeedGenerator, Configuration, Simulation,  # noqa: F401
                                        SimulationPool)
from lennardjonesium.engine import Engine, SystemSpec, initial_condition  # noqa: F401
