"""Simulation objects (the reference's lennardjonesium.simulation package,
src/python/lennardjonesium/simulation/__init__.py) backed by the GPU engine."""

from lennardjonesium.simulation.seed_generator import SeedGenerator  # noqa: F401
from lennardjonesium.simulation.configuration import Configuration  # noqa: F401
from lennardjonesium.simulation.simulation import Simulation, SimulationPool  # noqa: F401
