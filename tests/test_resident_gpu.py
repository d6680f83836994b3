"""Resident cells: step calls that keep every cell in a fixed block of slots and
move only the particles that change cell (and stay on them between calls)
must leave records and state equal, bit for bit, to the same calls with the
full cell build every step (the reference's per-step cell lists,
engine/particle_pair_filter.cpp:141-156).  Steps that exceed a capacity are
redone with the full build: same results."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "disp", "force")


def _eng():
    from lennardjonesium import engine as eng
    assert eng.device_count() > 0, "no HIP device"
    return eng


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _systems(eng, specs):
    """(SystemSpec, pos, vel) per (n, rho, T, seed)."""
    out = []
    for n, rho, T, seed, rc in specs:
        pos, vel, L = eng.initial_condition(n, rho, T, seed=seed)
        out.append((eng.SystemSpec(n, L, cutoff=rc), pos, vel))
    return out


def _run(eng, systems, precision, calls, resident, graphs=True, kernel="tiled"):
    e = eng.Engine([s for s, _, _ in systems], precision=precision, pair_kernel=kernel)
    e.set_graphs(graphs)
    e.set_resident_cells(resident)
    for k, (_, p, v) in enumerate(systems):
        e.set_state(k, p, v)
    recs = [e.step(n) for n in calls]
    states = [e.get_state(k) for k in range(len(systems))]
    stats = e.resident_stats()
    e.close()
    return np.concatenate(recs), states, stats


def _same(a, b):
    ra, sa, _ = a
    rb, sb, _ = b
    assert ra.shape == rb.shape
    assert np.array_equal(ra, rb), f"records differ: max {np.abs(ra - rb).max():.3e}"
    for x, y in zip(sa, sb):
        for f in FIELDS:
            assert np.array_equal(x[f], y[f]), f"{f} differs: max {np.abs(x[f] - y[f]).max():.3e}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_equals_full_build_batch(precision):
    """Three systems in one engine (a dense liquid, a hot fluid with many
    cell changes per step, a small box): one call of 300 steps."""
    eng = _eng()
    systems = _systems(eng, [(32000, 0.8442, 0.72, 5489, 2.5), (13500, 0.6, 2.0, 7, 2.5),
                             (4000, 0.8, 1.0, 11, 2.5)])
    _say("resident vs full build, batch of 3,", precision)
    a = _run(eng, systems, precision, [300], True)
    b = _run(eng, systems, precision, [300], False)
    assert a[2] == (1, 0), f"resident calls / fallbacks: {a[2]}"
    assert b[2] == (0, 0)
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_small_systems_multiwave(precision):
    """Small systems on the tiled kernel with four waves per home cell (N = 864
    and 4,000 in one engine): 200 steps."""
    eng = _eng()
    systems = _systems(eng, [(864, 0.8, 1.0, 5489, 2.5), (4000, 0.8, 1.0, 17, 2.5)])
    _say("resident multiwave,", precision)
    a = _run(eng, systems, precision, [200], True, kernel="multiwave")
    b = _run(eng, systems, precision, [200], False, kernel="multiwave")
    assert a[2] == (1, 0), f"resident calls / fallbacks: {a[2]}"
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_split_calls_and_eager(precision):
    """Calls of 5, 37, 4, 3, 2 and 100 steps (the first enters resident
    cells, the others continue on them), replayed from graphs and launched
    eagerly, equal one call of 151 steps with the full build."""
    eng = _eng()
    systems = _systems(eng, [(32000, 0.8442, 0.72, 5489, 2.5)])
    _say("resident split calls,", precision)
    ref = _run(eng, systems, precision, [151], False)
    for graphs in (True, False):
        a = _run(eng, systems, precision, [5, 37, 4, 3, 2, 100], True, graphs=graphs)
        assert a[2] == (6, 0), f"resident calls / fallbacks: {a[2]}"
        _same(a, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_split_calls_multiwave(precision):
    """The four-wave small-system kernels on live cells (N = 864 and 4,000 in
    one engine): calls of 5, 3, 17, 1 and 40 steps equal one call of 66 steps
    with the full build."""
    eng = _eng()
    systems = _systems(eng, [(864, 0.8, 1.0, 5489, 2.5), (4000, 0.8, 1.0, 17, 2.5)])
    _say("resident split calls multiwave,", precision)
    ref = _run(eng, systems, precision, [66], False, kernel="multiwave")
    a = _run(eng, systems, precision, [5, 3, 17, 1, 40], True, kernel="multiwave")
    assert a[2] == (5, 0), f"resident calls / fallbacks: {a[2]}"
    _same(a, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_live_between_calls(precision):
    """Between resident calls: reading the state, measuring, a velocity
    rescale, a call of 1 step, a g(r) sample, the cell lists and the
    neighbour counts.  Each reads or changes the packed state (the live cells
    are packed back first) and the steps after it continue exactly as on the
    full-build path."""
    eng = _eng()
    systems = _systems(eng, [(32000, 0.8442, 0.72, 5489, 2.5), (4000, 0.8, 1.0, 11, 2.5)])
    _say("resident live between calls,", precision)
    out = []
    for resident in (True, False):
        e = eng.Engine([s for s, _, _ in systems], precision=precision)
        e.set_resident_cells(resident)
        for k, (_, p, v) in enumerate(systems):
            e.set_state(k, p, v)
        got = [e.step(7)]
        got.append(e.get_state(1))
        got.append(e.step(1))
        got.append(e.step(9))
        got.append(e.measure())
        got.append(e.step(4))
        e.set_temperature(0, 0.9)
        got.append(e.step(12))
        e.rdf_sample(4.0, 200)
        got.append(e.rdf_read(200)[0])
        got.append(e.step(6))
        got.extend(e.cell_list(1))
        got.append(e.step(5))
        got.append(e.neighbor_counts(0))
        got.append(e.get_state(0))
        got.append(e.get_state(1))
        got.append(e.resident_stats())
        e.close()
        out.append(got)
    a, b = out
    assert a[-1] == (6, 0) and b[-1] == (0, 0), (a[-1], b[-1])
    for x, y in zip(a[:-1], b[:-1]):
        if isinstance(x, dict):
            for f in FIELDS:
                assert np.array_equal(x[f], y[f]), f
            assert (x["time"], x["pe"], x["vir"]) == (y["time"], y["pe"], y["vir"])
        else:
            assert np.array_equal(x, y)


def test_resident_snapshot_refresh():
    """After 16,384 steps on live cells the next call packs them back and
    enters again: calls of 9,000, 9,000 and 100 steps (the third re-enters)
    equal the full-build run."""
    eng = _eng()
    systems = _systems(eng, [(4000, 0.8, 1.0, 11, 2.5)])
    _say("resident snapshot refresh")
    a = _run(eng, systems, "mixed", [9000, 9000, 100], True)
    b = _run(eng, systems, "mixed", [9000, 9000, 100], False)
    assert a[2] == (3, 0), f"resident calls / fallbacks: {a[2]}"
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_fallback_after_live_calls(precision):
    """Nine particles on a 3 x 3 grid at one x, moving along +x at the same
    speed, all enter the next cell in the sixth step: more arrivals than one
    cell takes in a step.  Calls of 4, 4 and 4 steps: the second call, on live
    cells, is redone with the previous call's steps on the full-build path;
    the third enters resident cells again.  Records and state equal the
    full-build run's."""
    eng = _eng()
    L = 7.8   # 3 cells of 2.6 per axis; the crossing is at x = 2.6
    ys = (2.8, 3.9, 5.0)
    pos = np.array([[2.545, y, z] for y in ys for z in ys])
    vel = np.tile([2.0, 0.0, 0.0], (9, 1))
    systems = [(eng.SystemSpec(9, L), pos, vel)]
    _say("resident fallback after live calls,", precision)
    a = _run(eng, systems, precision, [4, 4, 4], True)
    b = _run(eng, systems, precision, [4, 4, 4], False)
    assert a[2] == (3, 1), f"resident calls / fallbacks: {a[2]}"
    _same(a, b)
    assert a[1][0]["pos"][0, 0] > 2.6   # the crossing happened


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_liquid_256k(precision):
    """The N = 256,000 liquid (rho 0.8442, T 0.72): 200 steps."""
    eng = _eng()
    systems = _systems(eng, [(256000, 0.8442, 0.72, 5489, 2.5)])
    _say("resident liquid 256k,", precision)
    a = _run(eng, systems, precision, [200], True)
    b = _run(eng, systems, precision, [200], False)
    assert a[2] == (1, 0)
    _same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resident_coexistence_2048k(precision):
    """The N = 2,048,000 coexistence workload (rho 0.30, T 0.80, rc 3.0, dt
    0.004; 250,047 cells, so the packing scan runs in two launches): 30 steps."""
    eng = _eng()
    n = 2048000
    pos, vel, L = eng.initial_condition(n, 0.30, 0.80, seed=5489)
    systems = [(eng.SystemSpec(n, L, cutoff=3.0, time_delta=0.004), pos, vel)]
    _say("resident coexistence 2048k,", precision)
    a = _run(eng, systems, precision, [30], True)
    b = _run(eng, systems, precision, [30], False)
    assert a[2] == (1, 0)
    _same(a, b)


def test_resident_fallback_when_cells_overflow():
    """A cutoff of 5 sigma at rho 0.8442 puts about 148 particles in a cell,
    beyond the 48 resident slots: the call is redone with the full build and
    its results equal a run with resident cells off."""
    eng = _eng()
    systems = _systems(eng, [(4000, 0.8442, 0.72, 3, 5.0)])
    _say("resident fallback")
    a = _run(eng, systems, "mixed", [20, 20], True)
    b = _run(eng, systems, "mixed", [20, 20], False)
    assert a[2] == (2, 2), f"resident calls / fallbacks: {a[2]}"
    _same(a, b)


def test_resident_noalloc_after_allocation():
    """lj_engine_step_noalloc uses resident cells once an allocating call has
    created their buffers: its records and state equal lj_engine_step's."""
    import ctypes

    from lennardjonesium import _native

    eng = _eng()
    systems = _systems(eng, [(32000, 0.8442, 0.72, 5489, 2.5)])
    _say("resident noalloc")
    es = []
    for _ in range(2):
        e = eng.Engine([s for s, _, _ in systems], precision="mixed")
        e.set_state(0, systems[0][1], systems[0][2])
        e.step(8)   # allocates the resident buffers
        es.append(e)
    a, b = es
    want = a.step(40)
    got = np.full((40, 1, 7), np.nan)
    _native.check(_native.engine().lj_engine_step_noalloc(
        b.h, 40, got.ctypes.data_as(ctypes.POINTER(_native.lj_thermo))))
    assert np.array_equal(got, want)
    sa, sb = a.get_state(0), b.get_state(0)
    for f in FIELDS:
        assert np.array_equal(sa[f], sb[f]), f
    assert a.resident_stats() == (2, 0) and b.resident_stats() == (2, 0)
    a.close()
    b.close()
