"""The benchmark workloads at their own sizes, through the C ABI, against the
CPU oracle:

  * coexistence system: N = 2,048,000 (FCC 80^3) at density 0.30, cutoff 3.0,
    dt 0.004 -- FP64 cutoff decisions bit-exact per particle for both pair
    kernels, FP64 forces / PE / virial within 1e-12, the mixed-precision force
    error stated, 20 FP64 steps tracking the oracle;
  * batched sweep: the 256-replica engine of the sweep benchmark (16 x 16 grid
    of T in [0.5, 2] and density in [0.05, 1.2], N = 4 n^3 with n from the
    seed in [5, 20]) -- replicas spanning all 16 densities, the 34-cells-a-side
    dilute replicas and the nc = 2 dense one included: records bit-equal to
    single-replica engines and within 1e-10 of the oracle over 50 steps,
    g(r) bit-exact;
  * differentiable trajectory: N = 4,000 at density 0.8, checkpoints every
    100 steps -- adjoint gradients within 1e-4 of central differences, and the
    closed forms dU/deps = U / eps, dU/dsigma = W / sigma at the shortest
    horizon.
"""

import itertools
import sys

import numpy as np
import pytest

from oracle import ljref

pytestmark = pytest.mark.gpu

FTOL = 1e-12


def _eng():
    from lennardjonesium import engine as eng
    assert eng.device_count() > 0, "no HIP device"
    return eng


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def force_err(F, Fref):
    rms = np.sqrt((Fref ** 2).sum(1).mean())
    den = np.maximum(np.linalg.norm(Fref, axis=1), rms)
    return (np.linalg.norm(F - Fref, axis=1) / den).max()


# ---- coexistence: N = 2,048,000, rho 0.30, rc 3.0, dt 0.004 --------------------------------

N5, RHO5, T5, RC5, DT5 = 2048000, 0.30, 0.80, 3.0, 0.004


@pytest.fixture(scope="module")
def coexistence_positions():
    """The lattice of the coexistence workload with every particle displaced
    by up to 0.4 sigma per axis (uneven cell occupancy, no symmetry; nearest
    pairs stay above 0.29 sigma, where float r^-12 is finite)."""
    pos, L = ljref.cubic_lattice(N5, RHO5)
    rng = np.random.default_rng(55)
    pos = ljref.wrap(pos + rng.uniform(-0.4, 0.4, size=pos.shape), L)
    F, pe, w = ljref.forces(pos, L, rc=RC5)
    counts = ljref.neighbor_counts(pos, L, RC5)
    return pos, L, F, pe, w, counts


@pytest.mark.parametrize("kernel", ["tiled", "particle", "multiwave"])
def test_coexistence_fp64_cutoff_decisions_and_forces(coexistence_positions, kernel):
    eng = _eng()
    pos, L, F, pe, w, counts = coexistence_positions
    _say("coexistence fp64", kernel, N5)
    e = eng.Engine([eng.SystemSpec(N5, L, cutoff=RC5, time_delta=DT5)], precision="fp64",
                   pair_kernel=kernel)
    e.set_state(0, pos, np.zeros_like(pos))
    got = e.neighbor_counts(0)
    assert np.array_equal(got, counts), int((got != counts).sum())
    st = e.get_state(0)
    e.close()
    err = force_err(st["force"], F)
    _say(f"  force err {err:.3e}, pe {st['pe']} vs {pe}, w {st['vir']} vs {w}")
    assert err < FTOL
    assert abs(st["pe"] - pe) <= FTOL * abs(pe)
    assert abs(st["vir"] - w) <= FTOL * abs(w)


def test_coexistence_mixed_force_error(coexistence_positions):
    """Mixed precision (float pair math on cell-local coordinates): relative
    RMS force error against the FP64 oracle below 1e-5, PE and virial within
    1e-6 relative."""
    eng = _eng()
    pos, L, F, pe, w, _ = coexistence_positions
    e = eng.Engine([eng.SystemSpec(N5, L, cutoff=RC5, time_delta=DT5)], precision="mixed")
    e.set_state(0, pos, np.zeros_like(pos))
    pe_m, w_m = e.compute_forces()
    Fm = e.get_state(0)["force"]
    e.close()
    rms = np.sqrt(((Fm - F) ** 2).sum(1).mean() / (F ** 2).sum(1).mean())
    _say(f"coexistence mixed: force rms rel err {rms:.3e}, pe {pe_m[0]} vs {pe}, w {w_m[0]} vs {w}")
    assert rms < 1e-5
    assert abs(pe_m[0] - pe) < 1e-6 * abs(pe)
    assert abs(w_m[0] - w) < 1e-6 * abs(w)


def test_coexistence_fp64_trajectory_tracks_oracle():
    """20 FP64 steps of the seeded coexistence workload (the quenched lattice
    at T 0.80) against the oracle: every record within 1e-10 relative."""
    eng = _eng()
    st, L = ljref.initial_condition(N5, RHO5, T5, seed=5489)
    e = eng.Engine([eng.SystemSpec(N5, L, cutoff=RC5, time_delta=DT5)], precision="fp64")
    e.set_state(0, st.pos, st.vel)
    rec = e.step(20)[:, 0, :]
    gpu = e.get_state(0)
    e.close()
    ref = ljref.run(st, L, DT5, 20, rc=RC5)
    rel = np.abs(rec - ref) / np.abs(ref)
    _say(f"coexistence 20 steps: max rel record err {rel.max():.3e}")
    assert rel.max() < 1e-10
    assert np.abs(gpu["pos"] - st.pos).max() < 1e-10 * L


# ---- batched sweep: the benchmark's 256 replicas -----------------------------------------

def sweep_replicas(eng, seed=5489):
    """(specs, states, meta) of the sweep benchmark: temperature-major 16 x 16
    grid (the reference's linspace arithmetic), N = 4 n^3, velocity seed
    seed + k."""
    from lennardjonesium.tools import linspace
    grid = list(itertools.product(linspace(0.5, 2.0, 16), linspace(0.05, 1.2, 16)))
    sizes = eng.sweep_sizes(seed, len(grid))
    specs, states, meta = [], [], []
    for k, ((T, rho), n) in enumerate(zip(grid, sizes)):
        N = 4 * int(n) ** 3
        pos, vel, L = eng.initial_condition(N, rho, T, seed=seed + k)
        specs.append(eng.SystemSpec(N, L))
        states.append((pos, vel))
        meta.append((T, rho, N, L))
    return specs, states, meta


def sweep_sample(meta):
    """One replica of every density column (rows spread over temperature),
    the largest dilute replica (density 0.05, 34 cells a side), the smallest
    dense one (N = 500 at density 1.2: 2 cells a side) and the compressed
    N = 864 start at T 1.2, density 1.2."""
    ks = {((3 * j + 9) % 16) * 16 + j for j in range(16)}
    dilute = [k for k in range(256) if k % 16 == 0]
    ks.add(max(dilute, key=lambda k: meta[k][2]))
    dense = [k for k in range(256) if k % 16 == 15]
    ks.add(min(dense, key=lambda k: meta[k][2]))
    ks.add(7 * 16 + 15)
    return sorted(ks)


@pytest.fixture(scope="module")
def sweep_setup():
    eng = _eng()
    specs, states, meta = sweep_replicas(eng)
    sample = sweep_sample(meta)
    return specs, states, meta, sample


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_256_replicas_batched_equals_single(sweep_setup, precision):
    eng = _eng()
    specs, states, meta, sample = sweep_setup
    _say(f"sweep {precision}: 256 replicas, {sum(m[2] for m in meta)} particles, "
         f"sample {sample}")
    ncell = [int(np.prod(ljref.cell_shape(m[3], 2.5))) for m in meta]
    assert max(ncell[k] for k in sample) == 34 ** 3
    assert min(ljref.cell_shape(meta[k][3], 2.5)[0] for k in sample) == 2
    assert len({meta[k][1] for k in sample}) == 16
    nsteps = 50
    e = eng.Engine(specs, precision=precision, pair_kernel="split")
    for k, (p, v) in enumerate(states):
        e.set_state(k, p, v)
    kinds = e.pair_kernels()
    assert "tiled" in kinds and "particle" in kinds
    rec = e.step(nsteps)
    finals = {k: e.get_state(k) for k in sample}
    e.rdf_sample(4.0, 200)
    counts, ns = e.rdf_read(200)
    e.close()
    assert ns == 1
    for k in sample:
        T, rho, N, L = meta[k]
        with eng.Engine([specs[k]], precision=precision, pair_kernel="split") as s:
            assert s.pair_kernels() == [kinds[k]]
            s.set_state(0, *states[k])
            single = s.step(nsteps)[:, 0, :]
        assert np.array_equal(rec[:, k, :], single), k
        want = ljref.rdf_cells(finals[k]["pos"], L, 4.0, 200)
        assert np.array_equal(counts[k], want), k
        if precision == "fp64":
            st = ljref.State(pos=states[k][0], vel=states[k][1])
            ref = ljref.run(st, L, 0.005, nsteps)
            # relative per record column; a column that is zero in the oracle
            # (PE and virial of a dilute lattice with no pair inside the
            # cutoff) must be zero here too
            scale = np.maximum(np.abs(ref), 1e-300)
            rel = np.abs(rec[:, k, :] - ref) / scale
            _say(f"  replica {k}: T {T:.2f} rho {rho:.3f} N {N} ({kinds[k]}) "
                 f"max rel err {rel.max():.2e}")
            assert rel.max() < 1e-10, k


# ---- differentiable trajectory: N = 4,000, density 0.8 ------------------------------------

def _adjoint(eng, pos, vel, L, nsteps, rho, every, eps=1.0, sig=1.0, dt=0.005):
    with eng.Engine([eng.SystemSpec(len(pos), L, time_delta=dt, epsilon=eps,
                                    sigma=sig)]) as e:
        e.set_state(0, pos, vel)
        return e.adjoint(nsteps, rho, every=every)


def test_adjoint_n4000_matches_central_differences():
    """200 steps (two checkpoint segments of 100): eps, sigma and sampled
    initial-position gradients within 1e-4 of central differences (h = 1e-6)."""
    eng = _eng()
    pos, vel, L = eng.initial_condition(4000, 0.8, 1.0, seed=5489)
    rho, nsteps, h = 0.8, 200, 1e-6
    loss, ge, gs, gr, gv = _adjoint(eng, pos, vel, L, nsteps, rho, every=100)

    def fwd(eps=1.0, sig=1.0, p=pos):
        return _adjoint(eng, p, vel, L, nsteps, rho, every=nsteps, eps=eps, sig=sig)[0]

    assert fwd() == loss
    fd_e = (fwd(eps=1 + h) - fwd(eps=1 - h)) / (2 * h)
    fd_s = (fwd(sig=1 + h) - fwd(sig=1 - h)) / (2 * h)
    _say(f"N=4000 adjoint: eps {ge:.10e} fd {fd_e:.10e}; sigma {gs:.10e} fd {fd_s:.10e}")
    assert abs(ge - fd_e) < 1e-4 * abs(fd_e)
    assert abs(gs - fd_s) < 1e-4 * abs(fd_s)
    rng = np.random.default_rng(1)
    scale = np.abs(gr).max()
    for i, a in zip(rng.integers(0, 4000, 4), rng.integers(0, 3, 4)):
        pp, pm = pos.copy(), pos.copy()
        pp[i, a] += h
        pm[i, a] -= h
        fd = (fwd(p=pp) - fwd(p=pm)) / (2 * h)
        _say(f"  d/dr0[{i},{a}]: {gr[i, a]:.8e} fd {fd:.8e}")
        assert abs(gr[i, a] - fd) < 1e-4 * max(abs(fd), 1e-3 * scale)


@pytest.mark.parametrize("eps,sig,rho", [(1.0, 1.0, 0.0), (1.3, 0.95, 0.0), (1.3, 0.95, 0.8)])
def test_adjoint_closed_forms_at_shortest_horizon(eps, sig, rho):
    """One step of dt = 1e-9: the trajectory does not move (r_1 - r_0 ~ 1e-9,
    d r_1 / d theta ~ dt^2), so the gradients reduce to the static closed forms
    of V = eps v(r / sigma) with rc = 2.5 sigma:
        dU/deps = U / eps,  dU/dsigma = W / sigma,
    and with the pressure term, d/deps [U + rho W / 3N] = (U + rho W / 3N) / eps."""
    eng = _eng()
    st, L = ljref.initial_condition(4000, 0.8, 1.0, seed=11)
    rng = np.random.default_rng(2)
    pos = ljref.wrap(st.pos + rng.uniform(-0.1, 0.1, size=st.pos.shape), L) * sig
    L = L * sig
    loss, ge, gs, _, _ = _adjoint(eng, pos, st.vel, L, 1, rho, every=1, eps=eps, sig=sig, dt=1e-9)
    F, U, W = ljref.forces(pos, L, rc=2.5, eps=eps, sigma=sig)
    n = len(pos)
    T = (2.0 / 3.0) * 0.5 * (st.vel ** 2).sum() / n
    assert abs(loss - (U + rho * (T + W / (3 * n)))) < 1e-7 * abs(loss)
    _say(f"closed forms eps={eps} sig={sig} rho={rho}: d/deps {ge:.12e} vs "
         f"{(U + rho * W / (3 * n)) / eps:.12e}; d/dsig {gs:.12e} vs U-part {W / sig:.12e}")
    assert abs(ge - (U + rho * W / (3 * n)) / eps) < 1e-7 * abs(ge)
    if rho == 0.0:
        assert abs(gs - W / sig) < 1e-7 * abs(gs)
