"""Simulation and SimulationPool over the host C++ layer (include/lj_host.h).

``Simulation`` mirrors the reference's Cython class
(src/python/lennardjonesium/simulation/_simulation.pyx:42-117): construct from
a Configuration, ``run(echo=...)`` writes events.log / thermodynamics.csv /
observations.csv / snapshots.csv, and potential / virial / force evaluate the
pair law for plots.

``SimulationPool`` keeps the reference's push / close / wait / status API
(src/python/lennardjonesium/simulation/_simulation_pool.pyx:36-108) but runs
the queued simulations as ONE batched GPU engine instead of one CPU thread
per simulation; ``thread_count`` is accepted for compatibility and ignored.
"""

import ctypes
import threading
from dataclasses import dataclass

from lennardjonesium import _native
from lennardjonesium.simulation.configuration import Configuration

PRECISIONS = {"fp64": 0, "mixed": 1}


class lj_sim_config(ctypes.Structure):
    _fields_ = [
        ("temperature", ctypes.c_double), ("density", ctypes.c_double),
        ("particle_count", ctypes.c_int), ("random_seed", ctypes.c_uint),
        ("cutoff_distance", ctypes.c_double), ("time_delta", ctypes.c_double),
        ("equilibration_name", ctypes.c_char_p), ("equilibration_tolerance", ctypes.c_double),
        ("equilibration_sample_size", ctypes.c_int), ("adjustment_interval", ctypes.c_int),
        ("steady_state_time", ctypes.c_int), ("equilibration_timeout", ctypes.c_int),
        ("observation_name", ctypes.c_char_p), ("observation_tolerance", ctypes.c_double),
        ("observation_sample_size", ctypes.c_int), ("observation_interval", ctypes.c_int),
        ("observation_count", ctypes.c_int),
        ("event_log", ctypes.c_char_p), ("thermodynamic_log", ctypes.c_char_p),
        ("observation_log", ctypes.c_char_p), ("snapshot_log", ctypes.c_char_p),
        ("precision", ctypes.c_int),
    ]


def _host():
    lib = _native.host()
    if not getattr(lib, "_sim_bound", False):
        V = ctypes.c_void_p
        lib.lj_simulation_create.restype = ctypes.c_int
        lib.lj_simulation_create.argtypes = [ctypes.POINTER(lj_sim_config),
                                             ctypes.POINTER(ctypes.c_void_p)]
        lib.lj_simulation_destroy.argtypes = [V]
        lib.lj_simulation_run.argtypes = [V, ctypes.c_int]
        lib.lj_simulation_run_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                                ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        for fn in ("lj_simulation_potential", "lj_simulation_virial", "lj_simulation_force"):
            getattr(lib, fn).restype = ctypes.c_double
            getattr(lib, fn).argtypes = [V, ctypes.c_double]
        lib.lj_host_last_error.restype = ctypes.c_char_p
        lib._sim_bound = True
    return lib


def _check(status):
    if status < 0:
        raise _native.LJError((_host().lj_host_last_error() or b"").decode() or f"status {status}")


class Simulation:
    """One NVE simulation (equilibration then observation) on the GPU."""

    def __init__(self, configuration: Configuration = None, precision: str = "fp64"):
        if configuration is None:
            configuration = Configuration()
        if not isinstance(configuration, Configuration):
            raise TypeError('Argument must be of type Configuration')
        self.configuration = configuration
        c = configuration
        self._strings = [s.encode() for s in (c.equilibration.name, c.observation.name,
                                              c.filepaths.event_log, c.filepaths.thermodynamic_log,
                                              c.filepaths.observation_log,
                                              c.filepaths.snapshot_log)]
        cfg = lj_sim_config(
            c.system.temperature, c.system.density, c.system.particle_count,
            c.system.random_seed & 0xFFFFFFFF, c.system.cutoff_distance, c.system.time_delta,
            self._strings[0], c.equilibration.tolerance, c.equilibration.sample_size,
            c.equilibration.adjustment_interval, c.equilibration.steady_state_time,
            c.equilibration.timeout,
            self._strings[1], c.observation.tolerance, c.observation.sample_size,
            c.observation.observation_interval, c.observation.observation_count,
            self._strings[2], self._strings[3], self._strings[4], self._strings[5],
            PRECISIONS[precision])
        h = ctypes.c_void_p()
        _check(_host().lj_simulation_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _host().lj_simulation_destroy(h)
            self._h = None

    def run(self, *, echo: bool = True, delay: float = 0.05):
        """Runs the schedule; re-running overwrites the output files.  ``delay``
        is accepted for compatibility (lines are printed as they are logged)."""
        _check(_host().lj_simulation_run(self._h, 1 if echo else 0))

    def potential(self, separation: float) -> float:
        return _host().lj_simulation_potential(self._h, float(separation))

    def virial(self, separation: float) -> float:
        return _host().lj_simulation_virial(self._h, float(separation))

    def force(self, separation: float) -> float:
        return _host().lj_simulation_force(self._h, float(separation))


@dataclass
class PoolStatus:
    waiting: int
    running: int
    completed: int


class SimulationPool:
    """Runs pushed simulations together as one batched GPU engine."""

    def __init__(self, thread_count: int = 1):
        self.thread_count = thread_count
        self._queue = []
        self._batches = []          # (sims, status array, thread)
        self._lock = threading.Lock()
        self._closed = False
        self._error = None

    def push(self, simulation: Simulation):
        if self._closed:
            raise RuntimeError("SimulationPool is closed")
        with self._lock:
            self._queue.append(simulation)

    def _launch(self):
        with self._lock:
            sims, self._queue = self._queue, []
        if not sims:
            return
        status = (ctypes.c_int * len(sims))(*([3] * len(sims)))
        handles = (ctypes.c_void_p * len(sims))(*[s._h for s in sims])

        def work():
            try:
                _check(_host().lj_simulation_run_batch(handles, len(sims), 0, status))
            except Exception as exc:   # surfaced by wait()
                self._error = exc

        t = threading.Thread(target=work, daemon=True)
        self._batches.append((sims, status, t, handles))
        t.start()

    def status(self) -> PoolStatus:
        self._launch()
        waiting = len(self._queue)
        running = completed = 0
        for sims, status, t, _ in self._batches:
            for k in range(len(sims)):
                if status[k] == 3 and t.is_alive():
                    running += 1
                else:
                    completed += 1
        return PoolStatus(waiting, running, completed)

    def close(self):
        self._closed = True

    def wait(self):
        self._launch()
        for _, _, t, _ in self._batches:
            t.join()
        if self._error is not None:
            err, self._error = self._error, None
            raise err
