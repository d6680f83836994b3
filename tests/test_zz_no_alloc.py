"""lj_engine_step_noalloc (the step entry point that allocates nothing,
records staged in a fixed device array) against lj_engine_step, both called
through the C ABI on engines holding the same state: records and state bit
for bit, on two calls in a row, the second shorter than the first."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(precision):
    from lennardjonesium import _native
    from lennardjonesium import engine as eng
    assert eng.device_count() > 0, "no HIP device"
    specs, states = [], []
    for k, (n, rho) in enumerate([(4000, 0.8), (864, 0.5), (500, 1.1)]):
        pos, vel, L = eng.initial_condition(n, rho, 1.0, seed=40 + k)
        specs.append(eng.SystemSpec(n, L))
        states.append((pos, vel))
    engines = [eng.Engine(specs, precision=precision) for _ in range(2)]
    for e in engines:
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
    return _native, engines


def _noalloc(native, e, nsteps):
    rec = np.full((nsteps, e.nsys, 7), np.nan)
    rc = native.engine().lj_engine_step_noalloc(
        e.h, nsteps, rec.ctypes.data_as(ctypes.POINTER(native.lj_thermo)))
    native.check(rc)
    return rec


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_step_equals_step(precision):
    native, (a, b) = _setup(precision)
    print("noalloc", precision, flush=True)
    for nsteps in (61, 9):
        want = a.step(nsteps)
        got = _noalloc(native, b, nsteps)
        assert np.array_equal(got, want), nsteps
        for k in range(a.nsys):
            sa, sb = a.get_state(k), b.get_state(k)
            for key in ("pos", "vel", "disp", "force", "time", "pe", "vir"):
                assert np.array_equal(np.asarray(sa[key]), np.asarray(sb[key])), (nsteps, k, key)
    # a frozen system's rows read zero in both entry points
    a.set_active(1, False)
    b.set_active(1, False)
    want = a.step(5)
    got = _noalloc(native, b, 5)
    assert np.array_equal(got, want) and np.all(got[:, 1, :] == 0)
    a.close()
    b.close()


def test_noalloc_rejects_more_rows_than_its_array():
    native, (a, b) = _setup("fp64")
    lib = native.engine()
    cap = 16384 // a.nsys
    rc = lib.lj_engine_step_noalloc(a.h, cap + 1, None)
    assert rc == -1 and b"exceeds the 16384 record rows" in lib.lj_last_error()
    assert lib.lj_engine_step_noalloc(a.h, cap, None) == 0
    a.close()
    b.close()
