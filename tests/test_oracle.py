"""Pins the CPU oracle (oracle/ljref.cpp) before anything is compared with it.

Each check is one the reference's own tests make, a fixture the reference
holds, or a property the mathematics fixes -- never a retyped formula.
"""

import csv
import itertools
import math

import numpy as np
import pytest

from oracle import ljref
from oracle import controller as C


# ---- LJ pair law: physics/lennard_jones_force.cpp ---------------------------

def test_lj_vanishes_at_cutoff():
    # The spline makes V(rc) = W(rc) = 0 (physics/lennard_jones_force.cpp:47-66)
    for rc in (2.0, 2.5, 3.0, 4.5):
        a, b = ljref.lj_spline(rc)
        s = rc ** -6
        # at r = rc: 4 s (s-1) + alpha + beta*0 and 24 s (2s-1) - 2 beta
        assert abs(4 * s * (s - 1) + a) < 1e-15
        assert abs(24 * s * (2 * s - 1) - 2 * b) < 1e-15
        just_in = rc * (1 - 1e-9)
        v, w, fz = ljref.lj_along_z(just_in, rc)
        assert abs(v) < 1e-9 and abs(w) < 1e-8 and abs(fz) < 1e-8


def test_lj_minimum_and_signs():
    # tests/cpp/lennardjonesium/physics/test_lennard_jones_force.cpp:38-116
    v, w, fz = ljref.lj_along_z(2 ** (1 / 6), 4.5)
    assert abs(v + 1.0) < 0.002 and abs(w) < 0.002 and abs(fz) < 0.002
    v, w, fz = ljref.lj_along_z(0.95, 2.5)        # repulsive core
    assert v > 0 and w > 0 and fz > 0
    v, w, fz = ljref.lj_along_z(1.5, 2.5)         # attractive well
    assert v < 0 and w < 0 and fz < 0


def test_lj_zero_beyond_cutoff():
    # test_lennard_jones_force.cpp:138-156
    for r in (2.5, 2.6, 10.0):
        assert ljref.lj_along_z(r, 2.5) == (0.0, 0.0, 0.0)


def test_lj_virial_is_minus_r_dvdr():
    # W = -r dV/dr, checked by central differences of the oracle's own V
    for r in (0.97, 1.1, 1.5, 2.0, 2.4):
        h = 1e-6
        vp = ljref.lj_along_z(r + h, 2.5)[0]
        vm = ljref.lj_along_z(r - h, 2.5)[0]
        w = ljref.lj_along_z(r, 2.5)[1]
        assert abs(-r * (vp - vm) / (2 * h) - w) < 1e-6 * max(1, abs(w))


def test_lj_force_direction_and_eps_sigma_scaling():
    rng = np.random.default_rng(3)
    for _ in range(20):
        sep = rng.normal(size=3)
        sep *= rng.uniform(0.95, 2.45) / np.linalg.norm(sep)
        f, v, w, inside = ljref.lj_pair(sep)
        assert inside
        r2 = sep @ sep
        np.testing.assert_allclose(f, w * sep / r2, rtol=1e-14)
        # V = eps v(r / sigma) with rc = 2.5 sigma
        eps, sig = 1.7, 1.3
        f2, v2, w2, _ = ljref.lj_pair(sep * sig, 2.5, eps, sig)
        assert abs(v2 - eps * v) < 1e-12 * max(1, abs(v))
        assert abs(w2 - eps * w) < 1e-12 * max(1, abs(w))


# ---- force accumulation: tests/cpp/lennardjonesium/engine/test_short_range_force_calculation.cpp

@pytest.mark.parametrize("z", [(0.2, 0.6), (0.8, 1.2), (2.8, 0.2)])
def test_mock_force_two_particles(z):
    pos = np.array([[0.2, 0.2, z[0]], [0.2, 0.2, z[1]]])
    F, pe, w = ljref.forces(pos, 3.0, rc=1.0, law=ljref.LAW_CONST, strength=-10.0)
    np.testing.assert_allclose(F, [[0, 0, 10.0], [0, 0, -10.0]], atol=1e-12)
    assert pe == pytest.approx(-6.0) and w == pytest.approx(-4.0)


def test_mock_force_l_shape():
    pos = np.array([[0.6, 0.6, 0.2], [1.4, 0.6, 0.2], [0.6, 1.4, 0.2]])
    F, pe, w = ljref.forces(pos, 3.0, rc=1.0, law=ljref.LAW_CONST, strength=-10.0)
    np.testing.assert_allclose(F, [[10, 10, 0], [-10, 0, 0], [0, -10, 0]], atol=1e-12)
    assert pe == pytest.approx(-4.0) and w == pytest.approx(-16.0)


# ---- pair filters: tests/cpp/lennardjonesium/engine/test_particle_pair_filter.cpp:59-125

KAT_POS = np.array([[0, 0, 0], [1, 1, 1], [1, 9, 9], [9, 1, 9], [9, 9, 1], [5, 5, 5],
                    [4, 4, 4], [4, 6, 6], [6, 4, 6], [6, 6, 4]], dtype=float)
KAT_PAIRS = {(0, 1): (-1, -1, -1), (0, 2): (-1, 1, 1), (0, 3): (1, -1, 1), (0, 4): (1, 1, -1),
             (5, 6): (1, 1, 1), (5, 7): (1, -1, -1), (5, 8): (-1, 1, -1), (5, 9): (-1, -1, 1)}


def _pairset(idx, sep):
    out = {}
    for (i, j, *_), s in zip(idx.tolist(), sep.tolist()):
        if i > j:
            i, j, s = j, i, [-x for x in s]
        key = (i, j, tuple(round(x, 9) for x in s))
        assert key not in out
        out[key] = True
    return set(out)


@pytest.mark.parametrize("fn", [ljref.pairs_naive, ljref.pairs_cell])
def test_pair_filter_known_answer(fn):
    idx, sep = fn(KAT_POS, 10.0, 2.0)
    got = _pairset(idx, sep)
    want = {(i, j, tuple(float(x) for x in s)) for (i, j), s in KAT_PAIRS.items()}
    assert got == want


@pytest.mark.parametrize("n,L,rc", [(60, 5.0, 2.5), (80, 7.3, 2.4), (40, 3.0, 2.5),
                                    (120, [6.0, 4.1, 8.7], 2.0)])
def test_cell_filter_equals_naive(n, L, rc):
    # nc = 1 and 2 included: explicit lattice images, not minimum image
    rng = np.random.default_rng(n)
    Lv = np.broadcast_to(np.asarray(L, float), 3)
    pos = rng.uniform(0, 1, size=(n, 3)) * Lv
    a = _pairset(*ljref.pairs_naive(pos, Lv, rc))
    b = _pairset(*ljref.pairs_cell(pos, Lv, rc))
    assert a == b and len(a) > 0


# ---- cell grid: tests/cpp/lennardjonesium/tools/test_cell_list_array.cpp

@pytest.mark.parametrize("rc,n", [(0.3, 3), (0.48, 2), (0.67, 1)])
def test_cell_shape(rc, n):
    assert list(ljref.cell_shape(1.0, rc)) == [n, n, n]


@pytest.mark.parametrize("nc", [[3, 3, 3], [4, 3, 5], [2, 2, 2], [1, 1, 1]])
def test_adjacent_cells_cover_all_26_neighbours_once(nc):
    nb, img = ljref.adjacent_cells(nc)
    ncell = int(np.prod(nc))
    coord = lambda c: ((c // (nc[1] * nc[2])), (c // nc[2]) % nc[1], c % nc[2])
    seen = {}
    for c in range(ncell):
        for s in range(13):
            o = nb[c, s]
            # unwrapped displacement from c to o (image * nc + (o - c))
            d = tuple(coord(o)[a] + img[c, s, a] * nc[a] - coord(c)[a] for a in range(3))
            seen.setdefault(c, []).append(d)
            seen.setdefault(int(o), []).append(tuple(-x for x in d))
    for c in range(ncell):
        assert sorted(seen[c]) == sorted(d for d in itertools.product((-1, 0, 1), repeat=3)
                                         if d != (0, 0, 0))


def test_cell_list_ascending_and_consistent():
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 9.0, size=(500, 3))
    start, atoms = ljref.cell_list(pos, 9.0, 2.5)
    keys, clamped = ljref.cell_keys(pos, 9.0, 2.5)
    assert clamped == 0
    lin = (keys[:, 0] * 3 + keys[:, 1]) * 3 + keys[:, 2]
    for c in range(27):
        members = atoms[start[c]:start[c + 1]]
        assert list(members) == sorted(members)
        assert set(members.tolist()) == set(np.nonzero(lin == c)[0].tolist())
    # key = floor(r * (nc / L)) with exact (integer) expectations on grid points
    k, _ = ljref.cell_keys(np.array([[0.0, 3.0, 6.0], [8.999, 5.999, 2.999]]), 9.0, 2.5)
    assert k.tolist() == [[0, 1, 2], [2, 1, 0]]


def test_cell_key_clamps_reference_out_of_range_case():
    # A coordinate equal to L maps to nc in the reference (an out-of-range
    # write); the restatement clamps it to nc - 1 and counts it.
    keys, clamped = ljref.cell_keys(np.array([[9.0, 0.0, 0.0]]), 9.0, 2.5)
    assert keys.tolist() == [[2, 0, 0]] and clamped == 1


# ---- boundary condition: test_periodic_boundary_condition.cpp:12-42

def test_wrap_known_answer():
    out = ljref.wrap(np.array([[1.7, 0, 0], [0, 3.2, 0], [0, 0, -4.3]]), 1.5)
    assert out[0, 0] == pytest.approx(0.2) and out[1, 1] == pytest.approx(0.2)
    assert out[2, 2] == pytest.approx(0.2)


# ---- integrator: test_velocity_verlet_integrator.cpp:25-239 (dt = 1)

def test_vv_inertial():
    st = ljref.State(pos=np.zeros((2, 3)), vel=[[1, 0, 0], [0, 1, 0]])
    ljref.run(st, None, 1.0, 4, law=ljref.LAW_NONE)
    assert st.pos.tolist() == [[4, 0, 0], [0, 4, 0]] and st.disp.tolist() == st.pos.tolist()
    assert st.vel.tolist() == [[1, 0, 0], [0, 1, 0]] and st.time == 4.0


def test_vv_gravity_uses_stored_force_on_first_kick():
    st = ljref.State(pos=np.zeros((2, 3)), vel=[[1, 0, 0], [0, 1, 0]])
    st.F[:] = [0, 0, -1.0]
    ljref.run(st, None, 1.0, 4, law=ljref.LAW_NONE)
    z = -(0.5 + 1.5 + 2.5 + 3.5)
    assert st.pos.tolist() == [[4, 0, z], [0, 4, z]]
    assert st.vel.tolist() == [[1, 0, -4], [0, 1, -4]]


def test_vv_box_three():
    st = ljref.State(pos=np.zeros((2, 3)), vel=[[1, 0, 0], [0, 1, 0]])
    ljref.run(st, 3.0, 1.0, 4, law=ljref.LAW_NONE)
    assert st.pos.tolist() == [[1, 0, 0], [0, 1, 0]]
    assert st.disp.tolist() == [[4, 0, 0], [0, 4, 0]]


# ---- lattice / initial condition: test_cubic_lattice.cpp, test_initial_condition.cpp

@pytest.mark.parametrize("cell_type,sites", [
    (0, [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0)]),
    (1, [(0, 0, 0), (.5, .5, .5), (0, 0, 1), (.5, .5, 1.5), (0, 1, 0), (.5, 1.5, .5), (0, 1, 1)]),
    (2, [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5), (0, 0, 1), (.5, .5, 1), (.5, 0, 1.5)]),
])
def test_cubic_lattice_sites(cell_type, sites):
    pos, L = ljref.cubic_lattice(7, 7 / 8, cell_type)
    assert sorted(map(tuple, pos.tolist())) == sorted(sites)


@pytest.mark.parametrize("n,rho", [(864, 0.8), (500, 1.2), (100, 0.8), (4000, 0.8)])
def test_lattice_density_and_box(n, rho):
    pos, L = ljref.cubic_lattice(n, rho)
    assert n / L ** 3 == pytest.approx(rho, rel=1e-13)
    assert pos.min() >= 0 and pos.max() < L


@pytest.mark.parametrize("n,T", [(864, 1.0), (100, 0.3)])
def test_initial_condition_invariants(n, T):
    st, L = ljref.initial_condition(n, 0.8, T, seed=5489)
    assert np.abs(st.vel.sum(0)).max() < 1e-10
    com = st.pos.mean(0)
    Lm = np.cross(st.pos - com, st.vel).sum(0)
    assert np.abs(Lm).max() < 1e-9
    ke = 0.5 * (st.vel ** 2).sum()
    assert (2 / 3) * ke / n == pytest.approx(T, rel=1e-14)
    assert np.all(st.F == 0)


# Four particles at (0,0,0), (0,0,2), (2,2,0), (2,2,2): centre of mass (1,1,1)
# and the two velocity sets of test_measurements.cpp:80-188 (rotation about z,
# then about x) with their known kinetic energy 4, temperature 2/3 and angular
# momenta (0,0,8) and (-8,4,0) about the centre.
_FOUR = np.array([[0, 0, 0], [0, 0, 2], [2, 2, 0], [2, 2, 2]], dtype=np.float64)


@pytest.mark.parametrize("vel,ang", [
    ([[1, -1, 0], [1, -1, 0], [-1, 1, 0], [-1, 1, 0]], [0, 0, 8]),
    ([[0, -1, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1]], [-8, 4, 0]),
])
def test_four_particle_measurements_and_zeroing(vel, ang):
    vel = np.array(vel, dtype=np.float64)
    st = ljref.State(pos=_FOUR.copy(), vel=vel.copy())
    rec = ljref.measure(st)                       # time, KE, PE, E, W, T, MSD
    assert rec[1] == 4.0 and rec[5] == pytest.approx(2.0 / 3.0, rel=1e-15)
    np.testing.assert_array_equal(np.cross(_FOUR - 1.0, vel).sum(0), ang)
    # physics/transformations.cpp:35-92: zeroing P and L about the centre
    v = vel.copy()
    ljref.lib().ljref_zero_momenta(4, ljref._d(np.ascontiguousarray(_FOUR)), ljref._d(v))
    assert np.abs(v.sum(0)).max() < 1e-14
    assert np.abs(np.cross(_FOUR - 1.0, v).sum(0)).max() < 1e-14


def test_zeroing_removes_a_rigid_motion_exactly():
    """v = V + w x (r - com) is a rigid translation plus rotation: removing the
    momentum and then the angular momentum (I dw = -L) must leave v = 0.  A
    wrong sign or a transposed cross product would double it instead."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 5, (50, 3))
    com = pos.mean(0)
    V, w = np.array([0.3, -1.1, 0.7]), np.array([0.9, 0.2, -0.5])
    v = V + np.cross(w, pos - com)
    ljref.lib().ljref_zero_momenta(50, ljref._d(np.ascontiguousarray(pos)), ljref._d(v))
    assert np.abs(v).max() < 1e-13


# ---- end-to-end fixture: notebooks/quick_single_run/data (reference binary output)

def _csv(path):
    rows = list(csv.reader(open(path)))
    return np.array([[float(x) for x in r] for r in rows[1:]])


def test_trajectory_fixture_first_200_steps(golden):
    ref = _csv(golden / "quick_single_run/thermodynamics_head.csv")
    st, L = ljref.initial_condition(500, 0.8, 0.5, seed=117208264)
    eq = C.EquilibrationPhase("Equilibration Phase", 0.5,
                              C.EquilibrationParameters(0.1, 50, 50, 500, 2000))
    ob = C.ObservationPhase("Observation Phase", 0.5, 0.8, 500,
                            C.ObservationParameters(0.2, 50, 50, 20))
    events, rows, _ = C.run_controller(st, L, 0.005, 2.5, [eq, ob], max_steps=200)
    rows = np.array(rows)
    assert rows[:, 0].tolist() == ref[:, 0].tolist()
    rel = np.abs(rows[:, 1:] - ref[:, 1:]) / np.abs(ref[:, 1:])
    assert rel.max() < 1e-13, rel.max()
    assert events[:3] == (golden / "quick_single_run/events.log").read_text().splitlines()[:3]


@pytest.mark.slow
def test_full_run_fixture_events_and_observations(golden):
    st, L = ljref.initial_condition(500, 0.8, 0.5, seed=117208264)
    eq = C.EquilibrationPhase("Equilibration Phase", 0.5,
                              C.EquilibrationParameters(0.1, 50, 50, 500, 2000))
    ob = C.ObservationPhase("Observation Phase", 0.5, 0.8, 500,
                            C.ObservationParameters(0.2, 50, 50, 20))
    events, rows, obs = C.run_controller(st, L, 0.005, 2.5, [eq, ob])
    assert events == (golden / "quick_single_run/events.log").read_text().splitlines()
    every = _csv(golden / "quick_single_run/thermodynamics_every50.csv")
    rows = np.array(rows)[::50]
    rel = np.abs(rows[:, 1:] - every[:, 1:]) / np.abs(every[:, 1:])
    # chaotic growth from ulp-level differences: tight early, loose at the end
    assert rel[:17].max() < 1e-10 and rel.max() < 1e-4
    ref_obs = _csv(golden / "quick_single_run/observations.csv")
    got = np.array([[t, o["temperature"], o["density"], o["total_energy"], o["pressure"],
                     o["specific_heat"], o["diffusion_coefficient"]] for t, o in obs])
    assert got[:, 0].tolist() == ref_obs[:, 0].tolist()
    np.testing.assert_allclose(got[:, 1:], ref_obs[:, 1:], rtol=1e-3, atol=1e-4)


# ---- g(r) histogram (no reference counterpart): FCC shell counts

def test_rdf_fcc_shells():
    pos, L = ljref.cubic_lattice(4 * 4 ** 3, 1.0)
    a = L / 4
    hist = ljref.rdf_hist(pos, L, rmax=1.3 * a, bins=1300)
    r = (np.arange(1300) + 0.5) * (1.3 * a / 1300)
    n = len(pos)
    for shell, count in [(a / math.sqrt(2), 12), (a, 6), (a * math.sqrt(1.5), 24)]:
        k = int(shell / (1.3 * a / 1300))
        assert hist[max(k - 2, 0):k + 3].sum() == n * count
    assert hist.sum() == n * (12 + 6 + 24)


@pytest.mark.parametrize("n,L,rmax", [(400, 14.0, 4.0), (300, [12.5, 16.0, 13.1], 4.0),
                                      (200, 9.0, 2.9), (150, 10.0, 4.0)])
def test_rdf_cells_equals_brute_force(n, L, rmax):
    """The cell-grid histogram (used for replicas too large for the brute
    force) counts exactly what the brute force counts, with particles placed
    on and next to cell faces (L = 10, rmax = 4: two cells per axis, so the
    brute force itself runs)."""
    rng = np.random.default_rng(n)
    Lv = np.broadcast_to(np.asarray(L, float), 3)
    pos = rng.uniform(0, 1, size=(n, 3)) * Lv
    a = Lv / np.floor(Lv / rmax)
    pos[:6] = [a, np.nextafter(a, 0), np.nextafter(a, 2 * a), 2 * a, np.zeros(3),
               np.nextafter(Lv, 0)]
    ref = ljref.rdf_hist(pos, Lv, rmax, 100)
    got = ljref.rdf_cells(pos, Lv, rmax, 100)
    assert np.array_equal(got, ref) and ref.sum() > 0


@pytest.mark.parametrize("n,L,rc", [(60, 5.0, 2.5), (40, 3.0, 2.5), (500, 9.0, 2.5)])
def test_neighbor_counts_equal_pair_set(n, L, rc):
    """Per-particle counts equal the pair set's, each pair counted for both."""
    pos = np.random.default_rng(n).uniform(0, L, size=(n, 3))
    idx, _ = ljref.pairs_naive(pos, L, rc)
    want = np.bincount(idx[:, 0], minlength=n) + np.bincount(idx[:, 1], minlength=n)
    assert np.array_equal(ljref.neighbor_counts(pos, L, rc), want)


# ---- the torch restatement used as the independent reference gradient ----------

@pytest.mark.parametrize("eps,sig", [(1.0, 1.0), (1.3, 0.95)])
def test_torch_ref_forward_equals_ljref(eps, sig):
    """oracle/torch_ref.py's trajectory (brute force over all 27 images, no
    wrap, F(r0) before the first half kick) against ljref.run started with
    the stored force F(r0): per-step PE, W and T, and the loss, within 1e-12
    relative for 20 steps at N = 108."""
    import torch
    from oracle import torch_ref
    st, L = ljref.initial_condition(108, 0.8, 1.0, seed=7)
    L = L * sig
    pos, vel = st.pos * sig, st.vel
    s = ljref.State(pos=pos, vel=vel)
    s.F, s.pe, s.vir = ljref.forces(pos, L, rc=2.5, eps=eps, sigma=sig)
    ref = ljref.run(s, L, 0.005, 20, rc=2.5, eps=eps, sigma=sig)
    with torch.no_grad():
        got = [(pe.item(), w.item(), T.item()) for pe, w, T in torch_ref.trajectory(
            torch.tensor(eps, dtype=torch.float64), torch.tensor(sig, dtype=torch.float64),
            torch.tensor(pos), torch.tensor(vel), L, 20, 0.005, 2.5)]
        loss = torch_ref.trajectory_loss(
            torch.tensor(eps, dtype=torch.float64), torch.tensor(sig, dtype=torch.float64),
            torch.tensor(pos), torch.tensor(vel), L, 20, 0.005, 2.5, 0.8).item()
    got = np.array(got)
    for col, k in ((0, 2), (1, 4), (2, 5)):
        rel = np.abs(got[:, col] - ref[:, k]) / np.abs(ref[:, k])
        assert rel.max() < 1e-12, (col, rel.max())
    want = (ref[:, 2] + 0.8 * (ref[:, 5] + ref[:, 4] / (3 * 108))).mean()
    assert abs(loss - want) < 1e-12 * abs(want)
