"""CPU-side checks of the native libraries: they load, export every symbol
the C headers declare, and the host-layer routines match the oracle."""

import re

import numpy as np
import pytest

from tests.conftest import ROOT


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lj_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("lj_engine.h", "engine"), ("lj_host.h", "host")])
def test_library_exports_every_declared_symbol(header, lib):
    from lennardjonesium import _native
    so = getattr(_native, lib)()
    names = _declared(header)
    assert names
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_engine_create_without_device_fails_cleanly():
    from lennardjonesium import _native
    lib = _native.engine()
    if lib.lj_device_count() > 0:
        pytest.skip("a device is visible")
    from lennardjonesium.engine import Engine, SystemSpec
    with pytest.raises(_native.LJError):
        Engine([SystemSpec(10, 5.0)])


def _desc(n=100, box=10.0, cutoff=2.5, dt=0.005, law=0):
    from lennardjonesium import _native
    d = _native.lj_system_desc()
    d.particle_count = n
    d.box[0] = d.box[1] = d.box[2] = box
    d.cutoff, d.time_delta, d.epsilon, d.sigma, d.strength = cutoff, dt, 1.0, 1.0, 1.0
    d.force_law, d.periodic = law, 1
    return d


@pytest.mark.parametrize("descs,precision,message", [
    ([_desc()], 7, "unknown precision"),
    ([_desc(n=0)], 0, "particle_count < 1"),
    ([_desc(n=(1 << 30) + 1)], 0, "particle_count exceeds 2^30"),
    # N rounds up to 2^30 slots, so a second system does not fit
    ([_desc(n=(1 << 30) - 100), _desc(n=1000)], 1, "exceed 2^30"),
    ([_desc(dt=0.0)], 0, "time_delta"),
    ([_desc(cutoff=1.1)], 0, "Cutoff distance is too short"),          # 1.21 < 2^(1/3)
    ([_desc(box=2.0)], 0, "less than the cutoff distance"),
    ([_desc(box=2600.0)], 1, "more than 1023 cells"),
    ([_desc(), _desc(law=1)], 0, "same force law"),
    ([_desc(), _desc(law=2)], 0, "LJ_LAW_NONE"),
])
def test_engine_create_rejects_invalid_descriptors(descs, precision, message):
    # the descriptors are checked before any device call, so this runs with
    # or without a GPU and allocates nothing
    import ctypes
    from lennardjonesium import _native
    lib = _native.engine()
    arr = (_native.lj_system_desc * len(descs))(*descs)
    out = ctypes.c_void_p(1)
    rc = lib.lj_engine_create(arr, len(descs), precision, ctypes.byref(out))
    assert rc == -1                      # LJ_E_INVALID
    assert out.value is None
    assert message in lib.lj_last_error().decode()


@pytest.mark.parametrize("n,rho,T,seed,lat", [(500, 0.8, 0.5, 117208264, 2), (864, 0.8, 1.0, 5489, 2),
                                              (100, 0.8, 0.3, 5489, 2), (7, 7 / 8, 0.5, 1, 0),
                                              (54, 0.5, 1.2, 99, 1)])
def test_host_initial_condition_matches_oracle(n, rho, T, seed, lat):
    from oracle import ljref
    from lennardjonesium.engine import initial_condition
    st, L0 = ljref.initial_condition(n, rho, T, seed=seed, cell_type=lat)
    pos, vel, L = initial_condition(n, rho, T, seed=seed, lattice=lat)
    assert L == L0
    assert np.array_equal(pos, st.pos)
    np.testing.assert_allclose(vel, st.vel, rtol=0, atol=1e-13 * np.abs(st.vel).max())


def test_format_double_matches_reference_sinks():
    # tests/cpp/lennardjonesium/output/test_sinks.cpp:100-225 and the reference's
    # notebook CSVs (fmt "{}" output)
    from lennardjonesium.engine import format_double
    cases = {3.5: "3.5", 2.25: "2.25", 0.5: "0.5", 7.25: "7.25", 0.0: "0", 2.0: "2",
             9.0: "9", -4.0: "-4", 3.7500000000000003e-05: "3.7500000000000003e-05",
             2.9999999999999583: "2.9999999999999583", -10810.337076759344: "-10810.337076759344",
             0.0001498634676421689: "0.0001498634676421689", 1e16: "1e+16",
             1234567890123456.0: "1234567890123456", 0.005: "0.005", 1e-5: "1e-05",
             -2.5e-300: "-2.5e-300", 0.2500000000000001: "0.2500000000000001"}
    for x, want in cases.items():
        assert format_double(x) == want, (x, format_double(x), want)


def test_format_double_reproduces_fixture_lines(golden):
    from lennardjonesium.engine import format_double
    lines = (golden / "quick_single_run/thermodynamics_head.csv").read_text().splitlines()[1:]
    for line in lines:
        fields = line.split(",")
        assert ",".join([fields[0]] + [format_double(float(x)) for x in fields[1:]]) == line


def test_sweep_sizes_deterministic_and_in_range():
    from lennardjonesium.engine import sweep_sizes
    a = sweep_sizes(5489, 4096)
    assert np.array_equal(a, sweep_sizes(5489, 4096))
    assert a.min() == 5 and a.max() == 20
    counts = np.bincount(a - 5, minlength=16)
    assert counts.min() > 180 and counts.max() < 340      # roughly uniform over 16 values
    # std::mt19937(5489): first output 3499211612 -> 3499211612 % 16 = 12 -> 17
    assert a[0] == 17


@pytest.mark.parametrize("n,rho", [(864, 0.8), (4000, 0.3), (32000, 0.05)])
def test_bench_candidate_count_matches_oracle(n, rho):
    # bench.py prices the pair kernel's flops on the half-shell candidate
    # count it forms from the engine's per-cell counts; the same formula on
    # the oracle's cells must give the oracle's count
    import sys
    sys.path.insert(0, str(ROOT))
    import bench
    from oracle import ljref
    st, L = ljref.initial_condition(n, rho, 1.0, seed=11)
    pos = ljref.wrap(st.pos + np.random.default_rng(n).uniform(-0.4, 0.4, st.pos.shape), L)
    nc = int(np.floor(L / 2.5))
    k = np.clip(np.floor(pos * (nc / L)).astype(int), 0, nc - 1)
    m = np.zeros((nc, nc, nc), dtype=np.int64)
    np.add.at(m, (k[:, 0], k[:, 1], k[:, 2]), 1)
    c_ref, _ = ljref.pair_counts(pos, L, 2.5)
    assert bench.half_shell_candidates(m) == c_ref


def test_bench_energy_stats_known_answer():
    """bench.py's conservation figures on a hand-made series."""
    import bench
    E = np.array([-10.0, -10.001, -9.999, -10.0005])
    got = bench.energy_stats(E)
    assert got["steps"] == 4 and got["window"] == 1
    assert got["endpoint"] == pytest.approx(5e-5, rel=1e-9)
    assert got["max_dev"] == pytest.approx(1e-4, rel=1e-9)
    assert got["windowed"] == pytest.approx(5e-5, rel=1e-9)
    assert bench.energy_stats(E, scale=20.0)["max_dev"] == pytest.approx(5e-5, rel=1e-9)


# ---- the observation statistics (physics/analyzers.cpp:44-159) -------------------

def _records(T=None, tm=None, n=None):
    """Records (k, 7) with the given temperatures and (time, MSD) pairs."""
    k = len(T) if T is not None else len(tm)
    rec = np.zeros((k, 7))
    rec[:, 5] = T if T is not None else 1.0 + np.arange(k)
    rec[:, 3] = -3.0
    if tm is not None:
        rec[:, 0] = [a for a, _ in tm]
        rec[:, 6] = [b for _, b in tm]
    else:
        rec[:, 0] = np.arange(k)
        rec[:, 6] = 2.0 * np.arange(k)
    return rec


@pytest.mark.parametrize("values,mean,var", [([2, 3], 2.5, 0.5), ([2, 3, 10], 5.0, 19.0),
                                             ([2, 3, 8, 1], 4.0, 13.0)])
def test_moving_sample_known_answers_scalar(values, mean, var):
    """tests/cpp/lennardjonesium/tools/test_moving_sample.cpp:10-56 (a window
    of 3 values; the fourth pushes the first out), read back through the
    heat capacity of the host analyzer: Cv = 1.5 / (1 - 1.5 N var / <T>^2)."""
    from lennardjonesium import engine as eng
    n = 1
    o = eng.thermo_observation(_records(T=values), 3, 1.0, n)
    assert o["temperature"] == pytest.approx(mean, rel=1e-15)
    var_back = (1.0 - 1.5 / o["specific_heat"]) * mean * mean / (1.5 * n)
    assert var_back == pytest.approx(var, rel=1e-12)


@pytest.mark.parametrize("pairs,slope", [([(1, 0), (0, 1)], -0.5 / 0.5),
                                         ([(2, 0), (0, 3), (1, 0)], -1.5 / 1.0),
                                         ([(2, 0), (0, 3), (1, 0), (5, 6)], 6.0 / 7.0)])
def test_moving_sample_known_answers_covariance(pairs, slope):
    """tests/cpp/lennardjonesium/tools/test_moving_sample.cpp:58-140 (vector
    samples: covariance(0, 1) / covariance(0, 0) of the first two components
    in a window of 3) as the diffusion coefficient D = slope / 6 of (time, MSD)."""
    from lennardjonesium import engine as eng
    o = eng.thermo_observation(_records(T=[1.0] * len(pairs), tm=pairs), 3, 1.0, 10)
    assert o["diffusion_coefficient"] == pytest.approx(slope / 6.0, rel=1e-14)


@pytest.mark.parametrize("k,size,seed", [(50, 50, 1), (120, 50, 2), (7, 5, 3)])
def test_host_analyzer_equals_oracle(k, size, seed):
    """The simulation's own analyzer against the oracle's restatement on
    identical records: every observation field within 1e-12 relative."""
    from lennardjonesium import engine as eng
    from oracle import ljref
    rng = np.random.default_rng(seed)
    rec = np.zeros((k, 7))
    rec[:, 0] = 3.0 + 0.005 * np.arange(k)
    rec[:, 5] = 0.8 + 0.01 * rng.standard_normal(k)
    rec[:, 3] = -2500.0 + rng.standard_normal(k)
    rec[:, 4] = -800.0 + 40 * rng.standard_normal(k)
    rec[:, 6] = 0.05 * np.arange(k) + 0.01 * rng.random(k)
    n, rho = 500, 0.8
    got = eng.thermo_observation(rec, size, rho, n)
    w = rec[-size:]
    want = ljref.observation(w[:, 5], w[:, 3], w[:, 4], w[:, 0], w[:, 6], rho, n)
    for key, v in want.items():
        assert got[key] == pytest.approx(v, rel=1e-12, abs=0), key
