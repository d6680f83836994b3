"""Parity of the HIP engine (through the C ABI) against the CPU oracle.

Bars (written where they are checked):
  * integer work -- cell keys, cell lists, pair sets: bit-exact;
  * FP64 forces / PE / virial on identical positions: 1e-12 relative
    (force norm-wise: max_i |F_i - F_i^ref| / max(|F^ref|_rms, |F_i^ref|));
  * trajectories: track the oracle / the reference fixture until chaos;
  * mixed precision: stated error bounds against the FP64 oracle.
"""

import csv
import sys

import numpy as np
import pytest

from oracle import ljref

pytestmark = pytest.mark.gpu

FTOL = 1e-12


def _engine_mod():
    from lennardjonesium import engine as eng
    assert eng.device_count() > 0, "no HIP device"
    return eng


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def perturbed_lattice(n, rho, amp, seed):
    pos, L = ljref.cubic_lattice(n, rho)
    rng = np.random.default_rng(seed)
    pos = ljref.wrap(pos + rng.uniform(-amp, amp, size=pos.shape), L)
    return pos, L


def force_err(F, Fref):
    rms = np.sqrt((Fref ** 2).sum(1).mean())
    den = np.maximum(np.linalg.norm(Fref, axis=1), rms)
    return (np.linalg.norm(F - Fref, axis=1) / den).max()


def make(eng, n, L, precision="fp64", pair_kernel="tiled", **kw):
    e = eng.Engine([eng.SystemSpec(n, L, **kw)], precision=precision, pair_kernel=pair_kernel)
    return e


# ---- cell keys / cell lists / pair sets: bit-exact -------------------------------

@pytest.mark.parametrize("n,rho,amp", [(864, 0.8, 0.3), (4000, 0.8, 0.5), (500, 1.2, 0.4),
                                       (32000, 0.05, 1.0)])
def test_cell_keys_and_lists_bit_exact(n, rho, amp):
    eng = _engine_mod()
    _say("cell keys", n, rho)
    pos, L = perturbed_lattice(n, rho, amp, n)
    # exact boundary values: 0, just below L, and the wrap's rounding cases
    pos[0] = [0.0, np.nextafter(L, 0), L / 2]
    pos[1] = [np.nextafter(0.0, 1), L * (1 - 1e-16), 0.0]
    e = make(eng, n, L)
    e.set_state(0, pos, np.zeros_like(pos))
    keys, nc, clamped = e.cell_keys(0)
    rkeys, rclamped = ljref.cell_keys(pos, L, 2.5)
    assert np.array_equal(keys, rkeys) and clamped == rclamped
    assert list(nc) == list(ljref.cell_shape(L, 2.5))
    start, atoms = e.cell_list(0)
    rstart, ratoms = ljref.cell_list(pos, L, 2.5)
    assert np.array_equal(start, rstart) and np.array_equal(atoms, ratoms)
    e.close()


def _pairset(idx, sep):
    """Orientation-free pair set {(i, j, image)} with i < j (bit-exact keys)."""
    out = set()
    for (i, j, ix, iy, iz), s in zip(idx.tolist(), sep.tolist()):
        if i > j:
            i, j, ix, iy, iz = j, i, -ix, -iy, -iz
            s = [-x for x in s]
        out.add((i, j, ix, iy, iz, *(float(x) for x in s)))
    return out


@pytest.mark.parametrize("n,L,rc", [(10, 10.0, 2.0), (864, None, 2.5), (60, 5.0, 2.5),
                                    (40, 3.0, 2.5), (500, None, 2.5)])
def test_pair_set_bit_exact(n, L, rc):
    eng = _engine_mod()
    _say("pairs", n, L, rc)
    if n == 10:   # tests/cpp/lennardjonesium/engine/test_particle_pair_filter.cpp:59-125
        pos = np.array([[0, 0, 0], [1, 1, 1], [1, 9, 9], [9, 1, 9], [9, 9, 1], [5, 5, 5],
                        [4, 4, 4], [4, 6, 6], [6, 4, 6], [6, 6, 4]], dtype=float)
    elif L is None:
        pos, L = perturbed_lattice(n, 0.8 if n == 864 else 1.2, 0.3, 7)
    else:
        pos = np.random.default_rng(n).uniform(0, L, size=(n, 3))
    e = eng.Engine([eng.SystemSpec(n, L, cutoff=rc)])
    e.set_state(0, pos, np.zeros_like(pos))
    got = _pairset(*e.pairs(0))
    want = _pairset(*ljref.pairs_cell(pos, L, rc))
    assert got == want and len(want) > 0
    e.close()


def _oracle_counts(pos, L, rc=2.5):
    idx, _ = ljref.pairs_cell(pos, L, rc)
    n = len(pos)
    return np.bincount(idx[:, 0], minlength=n) + np.bincount(idx[:, 1], minlength=n)


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
@pytest.mark.parametrize("n,rho,amp", [(864, 0.8, 0.3), (500, 1.2, 0.4), (40, 0.05, 0.5),
                                       (4000, 0.3, 1.0), (256000, 0.8442, 0.2),
                                       (32, 1.185, 0.1)])
def test_production_kernel_cutoff_decisions_bit_exact(n, rho, amp, kernel):
    """Per-particle in-cutoff neighbour counts of the production pair kernels
    (FP64) equal the reference filter's, particle by particle (N = 32 at
    rho 1.185: L = 3.0, one cell per axis, all 27 neighbour cells are images
    of the home cell; N = 500 at rho 1.2: two per axis)."""
    eng = _engine_mod()
    _say("neighbour counts", n, rho, kernel)
    pos, L = perturbed_lattice(n, rho, amp, 17)
    e = make(eng, n, L, pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    got = e.neighbor_counts(0)
    assert np.array_equal(got, _oracle_counts(pos, L))
    e.close()


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_mixed_kernel_cutoff_decisions(kernel):
    """FP32 separations may flip only pairs within ~1e-5 of the cutoff."""
    eng = _engine_mod()
    pos, L = perturbed_lattice(32000, 0.8442, 0.2, 19)
    e = make(eng, 32000, L, precision="mixed", pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    got = e.neighbor_counts(0)
    ref = _oracle_counts(pos, L)
    idx, sep = ljref.pairs_cell(pos, L, 2.5)
    near = (np.abs(np.sqrt((sep ** 2).sum(1)) - 2.5) < 1e-5).sum()
    assert np.abs(got - ref).sum() <= 2 * near + 4
    e.close()


# ---- forces, PE, virial: FP64 within 1e-12 ------------------------------------------

@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
@pytest.mark.parametrize("n,rho,amp", [(864, 0.8, 0.2), (4000, 0.8, 0.3), (500, 1.2, 0.2),
                                       (256000, 0.8442, 0.15), (32, 1.185, 0.1)])
def test_forces_fp64(n, rho, amp, kernel):
    eng = _engine_mod()
    _say("forces", n, kernel)
    pos, L = perturbed_lattice(n, rho, amp, 11)
    e = make(eng, n, L, pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    pe, w = e.compute_forces()
    st = e.get_state(0)
    Fr, per, wr = ljref.forces(pos, L)
    assert force_err(st["force"], Fr) < FTOL
    assert abs(pe[0] - per) / abs(per) < FTOL
    assert abs(w[0] - wr) / abs(wr) < FTOL
    assert np.array_equal(st["pos"], pos)          # forces do not move particles
    e.close()


@pytest.mark.parametrize("z", [(0.2, 0.6), (0.8, 1.2), (2.8, 0.2)])
def test_mock_force_known_answers(z):
    # tests/cpp/lennardjonesium/engine/test_short_range_force_calculation.cpp:39-109
    eng = _engine_mod()
    pos = np.array([[0.2, 0.2, z[0]], [0.2, 0.2, z[1]]])
    e = eng.Engine([eng.SystemSpec(2, 3.0, cutoff=1.0, force_law="constant", strength=-10.0)])
    e.set_state(0, pos, np.zeros_like(pos))
    pe, w = e.compute_forces()
    np.testing.assert_allclose(e.get_state(0)["force"], [[0, 0, 10], [0, 0, -10]], atol=1e-12)
    assert pe[0] == pytest.approx(-6.0) and w[0] == pytest.approx(-4.0)
    e.close()


def test_mock_force_l_shape():
    eng = _engine_mod()
    pos = np.array([[0.6, 0.6, 0.2], [1.4, 0.6, 0.2], [0.6, 1.4, 0.2]])
    e = eng.Engine([eng.SystemSpec(3, 3.0, cutoff=1.0, force_law="constant", strength=-10.0)])
    e.set_state(0, pos, np.zeros_like(pos))
    pe, w = e.compute_forces()
    np.testing.assert_allclose(e.get_state(0)["force"], [[10, 10, 0], [-10, 0, 0], [0, -10, 0]],
                               atol=1e-12)
    assert pe[0] == pytest.approx(-4.0) and w[0] == pytest.approx(-16.0)
    e.close()


# ---- integrator known answers (dt = 1): test_velocity_verlet_integrator.cpp -------------

def test_vv_known_answers():
    eng = _engine_mod()
    vel = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    e = eng.Engine([eng.SystemSpec(2, 1.0, time_delta=1.0, force_law="none", periodic=False)])
    e.set_state(0, np.zeros((2, 3)), vel)
    rec = e.step(4)
    st = e.get_state(0)
    assert st["pos"].tolist() == [[4, 0, 0], [0, 4, 0]] and st["disp"].tolist() == st["pos"].tolist()
    assert st["vel"].tolist() == vel.tolist() and st["time"] == 4.0
    assert rec[:, 0, 0].tolist() == [1.0, 2.0, 3.0, 4.0]
    e.set_state(0, np.zeros((2, 3)), vel, force=[[0, 0, -1.0]] * 2)
    e.step(4)
    st = e.get_state(0)
    z = -(0.5 + 1.5 + 2.5 + 3.5)
    assert st["pos"].tolist() == [[4, 0, z], [0, 4, z]]
    assert st["vel"].tolist() == [[1, 0, -4], [0, 1, -4]]
    e.close()
    e = eng.Engine([eng.SystemSpec(2, 3.0, time_delta=1.0, force_law="none", periodic=True)])
    e.set_state(0, np.zeros((2, 3)), vel)
    e.step(4)
    st = e.get_state(0)
    assert st["pos"].tolist() == [[1, 0, 0], [0, 1, 0]]
    assert st["disp"].tolist() == [[4, 0, 0], [0, 4, 0]]
    e.close()


# ---- trajectories ---------------------------------------------------------------------------

def _fixture_rows(golden):
    rows = list(csv.reader(open(golden / "quick_single_run/thermodynamics_head.csv")))
    return np.array([[float(x) for x in r] for r in rows[1:]])


def test_trajectory_tracks_reference_fixture(golden):
    """notebooks/quick_single_run (reference binary): 200 steps incl. two rescales."""
    eng = _engine_mod()
    ref = _fixture_rows(golden)
    pos, vel, L = eng.initial_condition(500, 0.8, 0.5, seed=117208264)
    e = make(eng, 500, L)
    e.set_state(0, pos, vel)
    got = []
    for chunk in range(4):
        rec = e.step(50)[:, 0, :]
        got.append(rec)
        if chunk < 2:   # events.log: rescale after steps 50 and 100
            T = rec[:, 5].mean()
            e.set_temperature(0, 0.5)
    got = np.concatenate(got)
    rel = np.abs(got - ref[:, 1:]) / np.abs(ref[:, 1:])
    _say("fixture max rel", rel.max())
    assert rel[:50].max() < 1e-12 and rel.max() < 1e-10
    e.close()


def test_trajectory_tracks_oracle_and_graphs_are_exact():
    eng = _engine_mod()
    st, L = ljref.initial_condition(864, 0.8, 1.0, seed=5489)
    ref = ljref.run(st.copy(), L, 0.005, 400)
    out = []
    for graphs in (True, False):
        e = eng.Engine([eng.SystemSpec(864, L)], graphs=graphs)
        e.set_state(0, st.pos, st.vel)
        out.append(e.step(400)[:, 0, :])
        e.close()
    assert np.array_equal(out[0], out[1])       # graph replay == eager launches, bit for bit
    rel = np.abs(out[0] - ref) / np.abs(ref)
    _say("oracle max rel", rel[:100].max(), rel.max())
    assert rel[:100].max() < 1e-11 and rel.max() < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_single_particle_drifts_through_the_wrap(precision, kernel):
    """N = 1 in a box of 3 cells a side: its own images lie beyond the cutoff,
    so no pair exists and F = PE = W = 0 exactly; the particle drifts at
    constant velocity and wraps across two faces (x up, y down).  Positions,
    displacements and times equal the oracle's bit for bit over 40 steps
    (a 32-step graph and 8 eager steps)."""
    eng = _engine_mod()
    _say("single particle", precision, kernel)
    L = 8.0
    pos = np.array([[7.9, 0.05, 4.0]])
    vel = np.array([[1.0, -2.0, 0.5]])
    e = make(eng, 1, L, precision=precision, pair_kernel=kernel)
    e.set_state(0, pos, vel)
    rec = e.step(40)[:, 0, :]
    out = e.get_state(0)
    e.close()
    st = ljref.State(pos=pos, vel=vel)
    ref = ljref.run(st, L, 0.005, 40)
    assert st.pos[0, 0] < 1.0 and st.pos[0, 1] > 7.0          # both wraps happened
    assert np.array_equal(out["pos"], st.pos)
    assert np.array_equal(out["disp"], st.disp)
    assert np.array_equal(out["vel"], vel)
    assert not out["force"].any() and out["pe"] == 0.0 and out["vir"] == 0.0
    assert not rec[:, 2].any() and not rec[:, 4].any()       # PE, W
    assert np.array_equal(rec[:, 0], ref[:, 0])               # time
    np.testing.assert_allclose(rec, ref, rtol=1e-15, atol=0)


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave", "split"])
def test_batched_replicas_match_single_runs(kernel):
    """One engine for three uneven replicas gives each replica's single-engine
    records bit for bit; 'split' runs the tiled kernel for two replicas and
    the per-particle kernel for the dilute one (threshold 10 per cell) in the
    same step."""
    eng = _engine_mod()
    cases = [(500, 0.8, 0.7, 3), (256, 0.3, 1.5, 4), (864, 1.1, 0.9, 5)]
    specs, states = [], []
    for n, rho, T, seed in cases:
        st, L = ljref.initial_condition(n, rho, T, seed=seed)
        specs.append(eng.SystemSpec(n, L))
        states.append((st, L))

    def engine(sp):
        e = eng.Engine(sp, pair_kernel="tiled" if kernel == "split" else kernel)
        if kernel == "split":
            e.set_pair_split(10.0)
        return e

    e = engine(specs)
    if kernel == "split":
        assert e.pair_kernels() == ["tiled", "particle", "tiled"]
    for k, (st, _) in enumerate(states):
        e.set_state(k, st.pos, st.vel)
    rec = e.step(100)
    for k, (st, L) in enumerate(states):
        single = engine([specs[k]])
        single.set_state(0, st.pos, st.vel)
        r1 = single.step(100)[:, 0, :]
        assert np.array_equal(rec[:, k, :], r1)      # batching does not change the arithmetic
        ref = ljref.run(st.copy(), L, 0.005, 100)
        assert (np.abs(r1 - ref) / np.abs(ref)).max() < 1e-10
        single.close()
    e.close()


def test_split_kernels_neighbour_counts_exact():
    """FP64 in-cutoff counts per particle are the oracle's with both kernels
    active in one engine (a dense and a dilute replica)."""
    eng = _engine_mod()
    cases = [(4000, 0.8, 1.0, 7), (2048, 0.1, 1.5, 8)]
    specs, states = [], []
    for n, rho, T, seed in cases:
        st, L = ljref.initial_condition(n, rho, T, seed=seed)
        st.pos = ljref.wrap(st.pos + np.random.default_rng(seed).normal(0, 0.05, st.pos.shape), L)
        specs.append(eng.SystemSpec(n, L))
        states.append((st, L))
    e = eng.Engine(specs)
    e.set_pair_split(4.0)
    assert e.pair_kernels() == ["tiled", "particle"]
    for k, (st, _) in enumerate(states):
        e.set_state(k, st.pos, st.vel)
    for k, (st, L) in enumerate(states):
        pos = st.pos
        assert np.array_equal(e.neighbor_counts(k), _oracle_counts(pos, L))
        F = e.get_state(k)["force"]
        Fr, _, _ = ljref.forces(pos, L)
        assert force_err(F, Fr) < FTOL
    e.close()


def test_set_temperature_and_measure():
    eng = _engine_mod()
    st, L = ljref.initial_condition(864, 0.8, 1.0, seed=5489)
    e = make(eng, 864, L)
    e.set_state(0, st.pos, st.vel * 1.3)
    measured = e.set_temperature(0, 0.7)
    assert measured == pytest.approx(1.69, rel=1e-12)
    m = e.measure()[0]
    assert m[5] == pytest.approx(0.7, rel=1e-13)
    e.close()


# ---- mixed precision: error stated against the FP64 oracle -------------------------------

@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
@pytest.mark.parametrize("n,rho", [(4000, 0.8), (256000, 0.8442)])
def test_mixed_forces_error(n, rho, kernel):
    eng = _engine_mod()
    pos, L = perturbed_lattice(n, rho, 0.15, 3)
    e = make(eng, n, L, precision="mixed", pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    pe, w = e.compute_forces()
    F = e.get_state(0)["force"]
    Fr, per, wr = ljref.forces(pos, L)
    rms_rel = np.sqrt(((F - Fr) ** 2).sum(1).mean() / (Fr ** 2).sum(1).mean())
    _say("mixed force rms rel", rms_rel, "pe rel", abs(pe[0] - per) / abs(per))
    assert rms_rel < 1e-5
    assert abs(pe[0] - per) / abs(per) < 1e-5 and abs(w[0] - wr) / abs(wr) < 1e-5
    e.close()


def energy_drift(E, window=100):
    """Secular drift |<E>_end - <E>_start| / |E_0| over windows of `window` steps
    (separates drift from the O(dt^2) fluctuation of velocity Verlet)."""
    return abs(E[-window:].mean() - E[:window].mean()) / abs(E[0])


def test_mixed_energy_drift_matches_fp64():
    eng = _engine_mod()
    pos, vel, L = eng.initial_condition(4000, 0.8, 1.0, seed=5489)
    out = {}
    for prec in ("fp64", "mixed"):
        e = make(eng, 4000, L, precision=prec)
        e.set_state(0, pos, vel)
        E = e.step(2000)[:, 0, 3]
        out[prec] = (energy_drift(E), np.abs(E - E[0]).max() / abs(E[0]))
        e.close()
    _say("drift / max fluctuation", out)
    # the fluctuation is the integrator's (same in both precisions); drift < 1e-4
    assert out["mixed"][0] < 1e-4 and out["fp64"][0] < 1e-4
    assert out["mixed"][1] < 1.2 * out["fp64"][1] + 1e-5


# ---- g(r) ------------------------------------------------------------------------------------

@pytest.mark.parametrize("n,rho", [(500, 1.2), (256, 0.3), (32, 1.185)])
def test_rdf_histogram_matches_brute_force(n, rho):
    """Bit-exact against the brute-force oracle, including N = 32 at rho 1.185
    (L = 3.0, one cell per axis: every stencil cell is an image of the home
    cell, and each particle meets images of itself)."""
    eng = _engine_mod()
    pos, L = perturbed_lattice(n, rho, 0.3, 5)
    e = make(eng, n, L)
    e.set_state(0, pos, np.zeros_like(pos))
    e.rdf_sample(4.0, 200)
    counts, ns = e.rdf_read(200)
    ref = ljref.rdf_hist(pos, L, 4.0, 200)
    assert ns == 1
    assert np.array_equal(counts[0], ref)
    e.close()


def test_autotune_leaves_no_trace():
    eng = _engine_mod()
    specs, states = [], []
    for k, (n, rho) in enumerate([(500, 0.9), (2048, 0.1), (864, 0.7)]):
        pos, vel, L = eng.initial_condition(n, rho, 1.0, seed=k)
        specs.append(eng.SystemSpec(n, L))
        states.append((pos, vel))
    ref = eng.Engine(specs)
    tuned = eng.Engine(specs)
    for e in (ref, tuned):
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
    variant, ms = tuned.autotune(4)
    assert variant in ("tiled", "particle", "multiwave", "split") and ms > 0
    ref.set_pair_kernel(variant)
    assert np.array_equal(ref.step(30), tuned.step(30))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_launch_shapes_give_identical_records(precision):
    """Waves per workgroup of the mixed tiled kernel and particles per sort
    wave change how work is grouped, not what is computed: every shape gives
    the default shape's records and state bit for bit (a dense and a dilute
    system in one engine, both pair kernels)."""
    eng = _engine_mod()
    specs, states = [], []
    for k, (n, rho) in enumerate([(13500, 0.8), (4000, 0.1)]):
        pos, vel, L = eng.initial_condition(n, rho, 1.0, seed=20 + k)
        specs.append(eng.SystemSpec(n, L))
        states.append((pos, vel))
    ref = None
    for cpb in (4, 2, 8):
        for sw in (44, 32, 56):
            with eng.Engine(specs, precision=precision, pair_kernel="split") as e:
                e.set_launch_shape(cpb, sw)
                assert e.launch_shape() == (cpb, sw)
                for k, st in enumerate(states):
                    e.set_state(k, *st)
                rec = e.step(25)
                fin = e.get_state(0)["vel"]
            if ref is None:
                ref = (rec, fin)
            else:
                assert np.array_equal(rec, ref[0]), (cpb, sw)
                assert np.array_equal(fin, ref[1]), (cpb, sw)


def test_autotune_after_steps_leaves_valid_cell_lists():
    """Tuning after steps (the slots already sorted) must not leave the cell
    ranges of a trial's last sort behind: the pair set and g(r) read
    afterwards equal those of a fresh engine on the same state."""
    eng = _engine_mod()
    pos, vel, L = eng.initial_condition(2048, 0.7, 1.0, seed=3)
    spec = [eng.SystemSpec(2048, L)]
    e = eng.Engine(spec)
    e.set_state(0, pos, vel)
    e.step(7)
    st = e.get_state(0)
    e.autotune(4)
    got_pairs = _pairset(*e.pairs(0))
    e.rdf_sample(4.0, 200)
    got_rdf, _ = e.rdf_read(200)
    after = e.get_state(0)
    e.close()
    for key in ("pos", "vel", "disp", "force"):
        assert np.array_equal(after[key], st[key]), key
    with eng.Engine(spec) as f:
        f.set_state(0, st["pos"], st["vel"], st["disp"], st["force"], st["time"])
        assert got_pairs == _pairset(*f.pairs(0))
        f.rdf_sample(4.0, 200)
        want_rdf, _ = f.rdf_read(200)
    assert np.array_equal(got_rdf, want_rdf)
    assert np.array_equal(got_rdf[0], ljref.rdf_hist(st["pos"], L, 4.0, 200))


def test_records_buffer_growth_keeps_graphs_valid():
    """A call longer than any before grows the record buffer; the step graphs
    captured on the old buffer are dropped first and the records stay exact
    (one engine growing 10 -> 70 -> 300 steps against one reserved up front)."""
    eng = _engine_mod()
    pos, vel, L = eng.initial_condition(864, 0.8, 1.0, seed=9)
    a = eng.Engine([eng.SystemSpec(864, L)])
    b = eng.Engine([eng.SystemSpec(864, L)])
    b.reserve(380)
    for e in (a, b):
        e.set_state(0, pos, vel)
    ra = np.concatenate([a.step(k) for k in (10, 70, 300)])
    rb = np.concatenate([b.step(k) for k in (10, 70, 300)])
    a.close()
    b.close()
    assert np.array_equal(ra, rb)
    st = ljref.State(pos=pos, vel=vel)
    ref = ljref.run(st, L, 0.005, 380)
    assert (np.abs(ra[:, 0, :] - ref) / np.abs(ref))[:200].max() < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_replica_beside_a_large_one_matches_single_run(precision):
    """The record sums of a system are split into first-level parts by its
    own block count: a 500-particle replica batched with a 256,000-particle
    one (64 parts) keeps the records of its single run (one part) bit for bit."""
    eng = _engine_mod()
    specs, states = [], []
    for k, (n, rho) in enumerate([(256000, 0.8442), (500, 0.8)]):
        pos, vel, L = eng.initial_condition(n, rho, 0.72, seed=k)
        specs.append(eng.SystemSpec(n, L))
        states.append((pos, vel))
    with eng.Engine(specs, precision=precision) as e:
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
        rec = e.step(20)
    for k in range(2):
        with eng.Engine([specs[k]], precision=precision) as s:
            s.set_state(0, *states[k])
            assert np.array_equal(rec[:, k, :], s.step(20)[:, 0, :]), k


def test_frozen_replica_records_read_zero():
    """Rows of a frozen system are not written by the steps; the call clears
    them, so they read as zeros (not the previous call's rows)."""
    eng = _engine_mod()
    specs, states = [], []
    for k, n in enumerate((500, 864)):
        pos, vel, L = eng.initial_condition(n, 0.8, 1.0, seed=k)
        specs.append(eng.SystemSpec(n, L))
        states.append((pos, vel))
    with eng.Engine(specs) as e:
        for k, (p, v) in enumerate(states):
            e.set_state(k, p, v)
        first = e.step(6)
        e.set_active(1, False)
        second = e.step(6)
    assert np.all(first[:, 1, :] != 0)
    assert np.all(second[:, 1, :] == 0) and np.all(second[:, 0, 0] > first[-1, 0, 0])


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_clumped_cell_chunks_and_stage_splits(kernel):
    """300 particles inside one cell of a 4^3 grid: the tiled kernel walks
    several 64-particle chunks and splits its neighbour stages; results must
    still match the oracle (FP64 counts bit-exact, forces 1e-12)."""
    eng = _engine_mod()
    rng = np.random.default_rng(23)
    n = 300
    pos = rng.uniform(0.1, 2.4, size=(n, 3))
    pos[:40] = rng.uniform(0, 10.0, size=(40, 3))      # some spread-out particles too
    e = make(eng, n, 10.0, pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    assert np.array_equal(e.neighbor_counts(0), _oracle_counts(pos, 10.0))
    F = e.get_state(0)["force"]
    Fr, per, wr = ljref.forces(pos, 10.0)
    assert force_err(F, Fr) < FTOL
    e.close()
    e = make(eng, n, 10.0, precision="mixed", pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    got = e.neighbor_counts(0)
    idx, sep = ljref.pairs_cell(pos, 10.0, 2.5)
    near = (np.abs(np.sqrt((sep ** 2).sum(1)) - 2.5) < 1e-5).sum()
    assert np.abs(got - _oracle_counts(pos, 10.0)).sum() <= 2 * near + 4
    e.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_records_independent_of_call_boundaries(graphs):
    """Each step's record is finished inside the next step (or at the end of
    the call): two calls of 50 steps must reproduce one call of 100, bit for
    bit, for every system and both graph sizes."""
    eng = _engine_mod()
    cases = [(500, 0.8, 0.7, 3), (864, 1.1, 0.9, 5)]
    specs, states = [], []
    for n, rho, T, seed in cases:
        st, L = ljref.initial_condition(n, rho, T, seed=seed)
        specs.append(eng.SystemSpec(n, L))
        states.append(st)
    runs = []
    for split in ((100,), (50, 50)):
        e = eng.Engine(specs, precision="mixed", graphs=graphs)
        for k, st in enumerate(states):
            e.set_state(k, st.pos, st.vel)
        runs.append(np.concatenate([e.step(n) for n in split], axis=0))
        e.close()
    assert runs[0].shape == runs[1].shape == (100, 2, 7)
    assert np.array_equal(runs[0], runs[1])
    assert np.all(np.diff(runs[0][:, :, 0], axis=0) > 0)   # time advances every step


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_chained_steps_leave_the_reference_state(precision, kernel):
    """Inside one call the pair kernel also applies the next step's first half
    kick (forces are then not stored); the last step of a call must leave
    full-step velocities and the forces, bit for bit as one step per call
    does, which applies each half kick from the stored force.  Eager
    launches, so that the chain spans all 37 steps."""
    eng = _engine_mod()
    st, L = ljref.initial_condition(864, 0.8, 1.0, seed=11)
    out = []
    for split in ((37,), (1,) * 37):
        e = eng.Engine([eng.SystemSpec(864, L)], precision=precision, graphs=False,
                       pair_kernel=kernel)
        e.set_state(0, st.pos, st.vel)
        rec = np.concatenate([e.step(k) for k in split], axis=0)
        out.append((rec, e.get_state(0)))
        e.close()
    (r0, s0), (r1, s1) = out
    assert np.array_equal(r0, r1)
    for key in ("pos", "vel", "disp", "force"):
        assert np.array_equal(s0[key], s1[key]), key
    assert s0["time"] == s1["time"] and s0["pe"] == s1["pe"] and s0["vir"] == s1["vir"]
    # and the chained run still tracks the oracle's first steps
    if precision == "fp64":
        ref = ljref.run(st.copy(), L, 0.005, 37)
        rel = np.abs(r0[:, 0, :] - ref) / np.abs(ref)
        assert rel.max() < 1e-11


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_inactive_replica_inside_chained_graphs(kernel):
    """Step graphs are captured while every replica is active; a replica
    switched off later must stay where it is (its particles still enter the
    next cell builds) while the other one runs on exactly as alone."""
    eng = _engine_mod()
    cases = [(500, 0.8, 0.7, 3), (864, 0.8, 0.9, 5)]
    specs, states = [], []
    for n, rho, T, seed in cases:
        st, L = ljref.initial_condition(n, rho, T, seed=seed)
        specs.append(eng.SystemSpec(n, L))
        states.append(st)
    e = eng.Engine(specs, precision="mixed", pair_kernel=kernel)
    for k, st in enumerate(states):
        e.set_state(k, st.pos, st.vel)
    e.step(64)                       # graphs built and replayed, both active
    held = e.get_state(1)
    e.set_active(1, False)
    rec = e.step(70)[:, 0, :]        # 32-step and 2-step graph replays
    got0, got1 = e.get_state(0), e.get_state(1)
    e.close()
    alone = eng.Engine([specs[0]], precision="mixed", pair_kernel=kernel)
    alone.set_state(0, states[0].pos, states[0].vel)
    alone.step(64)
    ref = alone.step(70)[:, 0, :]
    ref0 = alone.get_state(0)
    alone.close()
    assert np.array_equal(rec, ref)
    for key in ("pos", "vel", "disp", "force"):
        assert np.array_equal(got0[key], ref0[key]), key
        assert np.array_equal(got1[key], held[key]), key


def test_large_cell_grid_records_track_oracle():
    """More than 32768 cells (dilute box, 37^3 cells): the two-launch cell scan
    and the two-level record sums; FP64 records follow the oracle, and split
    calls reproduce a single call bit for bit."""
    eng = _engine_mod()
    st, L = ljref.initial_condition(40000, 0.05, 1.0, seed=11)
    assert int(L // 2.5) ** 3 > 32768
    ref = ljref.run(st.copy(), L, 0.005, 40)
    runs = []
    for split in ((40,), (13, 27)):
        e = eng.Engine([eng.SystemSpec(40000, L)])
        e.set_state(0, st.pos, st.vel)
        runs.append(np.concatenate([e.step(n) for n in split], axis=0)[:, 0, :])
        e.close()
    assert np.array_equal(runs[0], runs[1])
    err = np.abs(runs[0] - ref)
    _say("large grid max abs err", err.max(), "max |ref|", np.abs(ref).max())
    assert np.all(err <= 1e-10 * np.abs(ref) + 1e-12)   # PE and W are ~0 in this gas


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sort_groups_split_in_passes(precision):
    """A slab of 4096 particles (simple cubic, spacing 1.25) in two cell layers
    of a 16^3 grid: the mean occupancy (1 per cell) gives each sort wave 44
    cells, which here hold ~350 particles at 8 per cell, so every occupied
    group is sorted in several 64-lane passes.  Cell lists equal the oracle's
    (bit-exact) and 20 steps track the oracle's records."""
    eng = _engine_mod()
    L = 40.0
    g = np.arange(32) * 1.25 + 0.3
    ii, jj, kk = np.meshgrid(np.arange(4) * 1.25 + 0.3, g, g, indexing="ij")
    pos = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1)
    rng = np.random.default_rng(4)
    vel = rng.normal(0.0, 0.3, size=pos.shape)
    n = len(pos)
    st = ljref.State(n, pos.copy(), vel.copy())
    ref = ljref.run(st, L, 0.005, 20)
    e = eng.Engine([eng.SystemSpec(n, L)], precision=precision)
    e.set_state(0, pos, vel)
    start, atoms = e.cell_list(0)
    rstart, ratoms = ljref.cell_list(pos, L, 2.5)
    assert np.array_equal(start, rstart) and np.array_equal(atoms, ratoms)
    assert np.diff(start).max() == 8
    rec = e.step(20)[:, 0, :]
    start, atoms = e.cell_list(0)
    p_end = e.get_state(0)["pos"]
    e.close()
    rstart, ratoms = ljref.cell_list(p_end, L, 2.5)
    assert np.array_equal(start, rstart) and np.array_equal(atoms, ratoms)
    err = np.abs(rec - ref) / np.maximum(np.abs(ref), 1e-300)
    _say("slab passes", precision, "max rel err", err.max())
    assert err.max() < (1e-11 if precision == "fp64" else 1e-5)

@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_crowded_cells_track_oracle(precision):
    """Cutoff 3.9 sigma on a 500-particle liquid (L = 8.40): a 2^3 cell grid of
    about 62 particles per cell, so cells often receive more arrivals than
    their 64 member records and the sort ranks the rest from the overflow
    list.  After 30 steps the cell list equals the oracle's for the engine's
    own positions (bit-exact), and the records track the oracle's run."""
    eng = _engine_mod()
    st, L = ljref.initial_condition(500, 0.8442, 0.72, seed=5)
    assert int(L // 3.9) == 2
    ref_st = st.copy()
    ref = ljref.run(ref_st, L, 0.005, 30, rc=3.9)
    e = eng.Engine([eng.SystemSpec(500, L, cutoff=3.9)], precision=precision)
    e.set_state(0, st.pos, st.vel)
    rec = e.step(30)[:, 0, :]
    start, atoms = e.cell_list(0)
    pos = e.get_state(0)["pos"]
    e.close()
    rstart, ratoms = ljref.cell_list(pos, L, 3.9)
    _say("crowded cells", precision, np.diff(start))
    assert np.diff(start).max() > 64          # the overflow path ran
    assert np.array_equal(start, rstart) and np.array_equal(atoms, ratoms)
    err = np.abs(rec - ref) / np.abs(ref)
    _say("crowded cells max rel err", err.max())
    assert err.max() < (1e-11 if precision == "fp64" else 1e-5)
