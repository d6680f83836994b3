"""MI355X-native engine for lennard-jonesium's microcanonical LJ step loop.

The public names mirror the reference package (src/python/lennardjonesium/__init__.py).
"""

from lennardjonesium.engine import Engine, SystemSpec, initial_condition  # noqa: F401
