"""Restatement of the reference's control loop, phases and analyzers.

TEST INFRASTRUCTURE ONLY (see oracle/ljref.py): drives the oracle's
velocity-Verlet step one step at a time and reproduces

* control/simulation_controller.cpp:39-146   (command loop, logging order)
* control/simulation_phase.cpp:40-135         (equilibration / observation decisions)
* control/simulation_phase.hpp:120-244        (phase clocks)
* physics/analyzers.cpp:44-159                (observation statistics)
* tools/moving_sample.hpp:81-189              (fixed-size window, Bessel variance)
* tools/math.hpp:33-34                        (relative_error)
* output/sinks.cpp:40-98                      (events.log line texts)

The thermodynamic rows are returned as floats (time step, then the seven
ThermodynamicMeasurement values); formatting is checked separately.
"""

from collections import deque
from dataclasses import dataclass

from oracle import ljref


def relative_error(measured, desired):
    return abs(measured / desired - 1.0)


@dataclass
class EquilibrationParameters:
    tolerance: float = 0.05
    sample_size: int = 50
    adjustment_interval: int = 200
    steady_state_time: int = 1000
    timeout: int = 5000


@dataclass
class ObservationParameters:
    tolerance: float = 0.10
    sample_size: int = 50
    observation_interval: int = 200
    observation_count: int = 20


def _mean(xs):
    s = 0.0
    for x in xs:
        s += x
    return s / len(xs)


class EquilibrationPhase:
    def __init__(self, name, temperature, params: EquilibrationParameters):
        self.name = name
        self.temperature = temperature
        self.p = params
        self.sample = deque(maxlen=params.sample_size)
        self.set_start_time(0)

    def set_start_time(self, t):
        self.start_time = t
        self.last_check = t
        self.last_adjust = t

    def evaluate(self, queue, time_step, rec):
        self.sample.append(rec[5])
        if time_step - self.last_check >= self.p.adjustment_interval:
            self.last_check = time_step
            measured = _mean(self.sample)
            if relative_error(measured, self.temperature) >= self.p.tolerance:
                self.last_adjust = time_step
                queue.append(("adjust", measured, self.temperature))
        if time_step - self.last_adjust >= self.p.steady_state_time:
            queue.append(("complete",))
            return
        if time_step - self.start_time >= self.p.timeout:
            queue.append(("abort", "Could not equilibrate"))
            return
        queue.append(("advance",))


class ObservationPhase:
    def __init__(self, name, temperature, density, n, params: ObservationParameters):
        self.name = name
        self.temperature = temperature
        self.density = density
        self.n = n
        self.p = params
        self.window = deque(maxlen=params.sample_size)
        self.count = 0
        self.set_start_time(0)

    def set_start_time(self, t):
        self.start_time = t
        self.last_obs = t

    def evaluate(self, queue, time_step, rec):
        self.window.append((rec[5], rec[3], rec[4], rec[0], rec[6]))
        if time_step - self.last_obs >= self.p.observation_interval:
            self.last_obs = time_step
            T, E, W, t, m = zip(*self.window)
            obs = ljref.observation(T, E, W, t, m, self.density, self.n)
            if relative_error(obs["temperature"], self.temperature) >= self.p.tolerance:
                queue.append(("abort", "Temperature drifted outside acceptable range"))
                return
            self.count += 1
            queue.append(("observe", obs))
        if self.count >= self.p.observation_count:
            queue.append(("complete",))
            return
        queue.append(("advance",))


def run_controller(state, L, dt, rc, phases, max_steps=None):
    """SimulationController::operator() over the oracle stepper.

    Returns (events, thermo_rows, observations) where events are the text
    lines of events.log, thermo_rows are [time_step, time, KE, PE, E, W, T, MSD]
    and observations are (time_step, dict).
    """
    phases = list(phases)
    events, rows, observations = [], [], []
    time_step = 0
    phases[0].set_start_time(time_step)
    events.append(f"{time_step}: Phase started: {phases[0].name}")
    queue = deque([("advance",)])
    while queue:
        cmd = queue.popleft()
        kind = cmd[0]
        if kind == "advance":
            rec = ljref.run(state, L, dt, 1, rc=rc)[0]
            rows.append([time_step] + list(rec))
            time_step += 1
            phases[0].evaluate(queue, time_step, rec)
            if max_steps is not None and time_step >= max_steps:
                break
        elif kind == "observe":
            observations.append((time_step, cmd[1]))
            events.append(f"{time_step}: Observation recorded")
        elif kind == "adjust":
            state.vel = ljref.set_temperature(state.vel, cmd[2])
            events.append(f"{time_step}: Temperature measured at: {cmd[1]:.4g}, "
                          f"adjusted to: {cmd[2]:.4g}")
        elif kind == "complete":
            events.append(f"{time_step}: Phase complete: {phases[0].name}")
            phases.pop(0)
            if phases:
                phases[0].set_start_time(time_step)
                queue.append(("advance",))
                events.append(f"{time_step}: Phase started: {phases[0].name}")
        elif kind == "abort":
            events.append(f"{time_step}: Simulation aborted: {cmd[1]}")
    return events, rows, observations
