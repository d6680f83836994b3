"""Evenly spaced floats without numpy.

Same arithmetic as the reference's tools/linspace.py:24-35 (each point is
start + (k / denominator) * (stop - start)), which the sweep directory names
and run.ini values depend on bit for bit (e.g. 0.44999999999999996)."""


def linspace(start: float, stop: float, points: int, endpoint: bool = True) -> list:
    span = stop - start
    steps = points - 1 if endpoint else points
    out = []
    for k in range(points):
        out.append(start + (k / steps) * span)
    return out
