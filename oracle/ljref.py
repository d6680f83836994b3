"""ctypes front end of the CPU oracle ``ljref`` (oracle/ljref.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, where it is the checker (or the timed CPU
baseline), never the thing measured or shipped.  The HIP engine does not import
this module.

Arrays are numpy ``float64`` with shape (N, 3) (x, y, z per particle).
Each wrapper names the reference routine it restates (paths relative to the
reference repository root).
"""

import ctypes
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
_LIBS = {}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle (parity build and timing twin) with its Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)


def lib(fast: bool = False):
    """Load the parity build (``fast=False``) or the -ffast-math timing twin."""
    name = "libljref_fast.so" if fast else "libljref.so"
    if name in _LIBS:
        return _LIBS[name]
    path = HERE / "build" / name
    if not path.exists():
        build()
    so = ctypes.CDLL(str(path))
    so.ljref_cubic_lattice.restype = ctypes.c_double
    so.ljref_initial_condition.restype = ctypes.c_double
    so.ljref_pairs_cell.restype = ctypes.c_longlong
    so.ljref_pairs_naive.restype = ctypes.c_longlong
    so.ljref_cell_list.restype = ctypes.c_int
    so.ljref_cell_keys.restype = ctypes.c_int
    so.ljref_lj_pair.restype = ctypes.c_int
    _LIBS[name] = so
    return so


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _box(L):
    L = np.asarray(L, dtype=np.float64)
    if L.ndim == 0:
        L = np.full(3, float(L))
    return np.ascontiguousarray(L)


def _pos(p):
    return np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(-1, 3))


# ---- pair laws --------------------------------------------------------------

def lj_spline(rc):
    """physics/lennard_jones_force.cpp:68-73 -> (alpha, beta)."""
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().ljref_lj_spline(ctypes.c_double(rc), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def lj_pair(sep, rc=2.5, eps=1.0, sigma=1.0):
    """physics/lennard_jones_force.cpp:76-128 -> (F[3], V, W, inside)."""
    sep = np.ascontiguousarray(sep, dtype=np.float64)
    f = np.zeros(3)
    v, w = ctypes.c_double(), ctypes.c_double()
    inside = lib().ljref_lj_pair(_d(sep), ctypes.c_double(rc), ctypes.c_double(eps),
                                 ctypes.c_double(sigma), _d(f), ctypes.byref(v), ctypes.byref(w))
    return f, v.value, w.value, bool(inside)


def lj_along_z(distance, rc=2.5):
    """physics/forces.hpp:80-94 -> (potential, virial, force_z)."""
    v, w, fz = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().ljref_lj_along_z(ctypes.c_double(distance), ctypes.c_double(rc), ctypes.byref(v),
                           ctypes.byref(w), ctypes.byref(fz))
    return v.value, w.value, fz.value


# ---- cells ------------------------------------------------------------------

def cell_shape(L, rc):
    """tools/cell_list_array.cpp:45-51."""
    L = _box(L)
    nc = np.zeros(3, dtype=np.int32)
    lib().ljref_cell_shape(_d(L), ctypes.c_double(rc), _i(nc))
    return nc


def cell_keys(pos, L, rc):
    """engine/particle_pair_filter.cpp:144-147 -> (keys (N,3) int32, clamp count)."""
    pos, L = _pos(pos), _box(L)
    keys = np.zeros((len(pos), 3), dtype=np.int32)
    clamped = lib().ljref_cell_keys(len(pos), _d(pos), _d(L), ctypes.c_double(rc), _i(keys))
    return keys, clamped


def cell_list(pos, L, rc):
    """Cell list as CSR (start[ncell+1], atoms[N]); cells x-slowest, ascending index."""
    pos, L = _pos(pos), _box(L)
    nc = cell_shape(L, rc)
    ncell = int(np.prod(nc))
    start = np.zeros(ncell + 1, dtype=np.int32)
    atoms = np.zeros(len(pos), dtype=np.int32)
    lib().ljref_cell_list(len(pos), _d(pos), _d(L), ctypes.c_double(rc), _i(start), _i(atoms))
    return start, atoms


def adjacent_cells(nc):
    """tools/cell_list_array.cpp:83-157 -> (neighbour[ncell,13], image[ncell,13,3])."""
    nc = np.ascontiguousarray(nc, dtype=np.int32)
    ncell = int(np.prod(nc))
    nb = np.zeros((ncell, 13), dtype=np.int32)
    img = np.zeros((ncell, 13, 3), dtype=np.int32)
    lib().ljref_adjacent_cells(_i(nc), _i(nb), _i(img))
    return nb, img


def _pairs(fn, pos, L, rc):
    pos, L = _pos(pos), _box(L)
    n = len(pos)
    total = fn(n, _d(pos), _d(L), ctypes.c_double(rc), ctypes.c_longlong(0), None, None)
    idx = np.zeros((max(total, 1), 5), dtype=np.int32)
    sep = np.zeros((max(total, 1), 3))
    fn(n, _d(pos), _d(L), ctypes.c_double(rc), ctypes.c_longlong(total), _i(idx), _d(sep))
    return idx[:total], sep[:total]


def pairs_cell(pos, L, rc):
    """CellListParticlePairFilter::pairs (engine/particle_pair_filter.cpp:131-192).

    Returns (idx (P,5) = first, second, image xyz; sep (P,3)) in generator order.
    """
    return _pairs(lib().ljref_pairs_cell, pos, L, rc)


def pairs_naive(pos, L, rc):
    """NaiveParticlePairFilter::pairs (engine/particle_pair_filter.cpp:67-123)."""
    return _pairs(lib().ljref_pairs_naive, pos, L, rc)


def pair_counts(pos, L, rc):
    """(C_half, P_half): the reference filter's candidate and in-cutoff pair counts."""
    pos, L = _pos(pos), _box(L)
    c, p = ctypes.c_longlong(), ctypes.c_longlong()
    lib().ljref_pair_counts(len(pos), _d(pos), _d(L), ctypes.c_double(rc), ctypes.byref(c),
                            ctypes.byref(p))
    return c.value, p.value


# ---- forces / integration ---------------------------------------------------

LAW_LJ, LAW_CONST, LAW_NONE = 0, 1, 2


def forces(pos, L, rc=2.5, law=LAW_LJ, eps=1.0, sigma=1.0, strength=0.0, fast=False):
    """ShortRangeForceCalculation (engine/force_calculation.cpp:46-71) -> (F, PE, W)."""
    pos, L = _pos(pos), _box(L)
    F = np.zeros_like(pos)
    pe, w = ctypes.c_double(), ctypes.c_double()
    lib(fast).ljref_forces(len(pos), _d(pos), _d(L), law, ctypes.c_double(rc),
                           ctypes.c_double(eps), ctypes.c_double(sigma),
                           ctypes.c_double(strength), _d(F), ctypes.byref(pe), ctypes.byref(w))
    return F, pe.value, w.value


def wrap(pos, L):
    """engine/boundary_condition.cpp:37-51."""
    pos, L = _pos(pos).copy(), _box(L)
    lib().ljref_wrap(len(pos), _d(pos), _d(L))
    return pos


class State:
    """physics::SystemState (physics/system_state.hpp:35-102), N x 3 arrays."""

    def __init__(self, n=0, pos=None, vel=None):
        self.pos = np.zeros((n, 3)) if pos is None else _pos(pos).copy()
        n = len(self.pos)
        self.vel = np.zeros((n, 3)) if vel is None else _pos(vel).copy()
        self.disp = np.zeros((n, 3))
        self.F = np.zeros((n, 3))
        self.pe = 0.0
        self.vir = 0.0
        self.time = 0.0

    @property
    def n(self):
        return len(self.pos)

    def copy(self):
        s = State(pos=self.pos, vel=self.vel)
        s.disp = self.disp.copy()
        s.F = self.F.copy()
        s.pe, s.vir, s.time = self.pe, self.vir, self.time
        return s


def run(state: State, L, dt, nsteps, rc=2.5, law=LAW_LJ, eps=1.0, sigma=1.0, strength=0.0,
        fast=False, records=True):
    """VelocityVerletIntegrator (engine/integrator.cpp:51-88), nsteps steps in place.

    ``L=None`` runs without the boundary condition.  Returns the per-step
    records (nsteps, 7) = time, KE, PE, E, W, T, MSD (after each step).
    """
    n = state.n
    rec = np.zeros((nsteps, 7)) if records else None
    pe, w, t = ctypes.c_double(state.pe), ctypes.c_double(state.vir), ctypes.c_double(state.time)
    Lp = _d(_box(L)) if L is not None else None
    Lkeep = _box(L) if L is not None else None
    if Lkeep is not None:
        Lp = _d(Lkeep)
    lib(fast).ljref_run(n, _d(state.pos), _d(state.vel), _d(state.disp), _d(state.F),
                        ctypes.byref(pe), ctypes.byref(w), ctypes.byref(t), Lp, law,
                        ctypes.c_double(rc), ctypes.c_double(eps), ctypes.c_double(sigma),
                        ctypes.c_double(strength), ctypes.c_double(dt), nsteps,
                        _d(rec) if rec is not None else None)
    state.pe, state.vir, state.time = pe.value, w.value, t.value
    return rec


def measure(state: State):
    """ThermodynamicMeasurement (physics/measurements.hpp:59-70) -> 7 values."""
    rec = np.zeros(7)
    lib().ljref_measure(state.n, _d(state.vel), _d(state.disp), ctypes.c_double(state.time),
                        ctypes.c_double(state.pe), ctypes.c_double(state.vir), _d(rec))
    return rec


def set_temperature(vel, target):
    """physics/transformations.cpp:94-117."""
    vel = _pos(vel).copy()
    lib().ljref_set_temperature(len(vel), _d(vel), ctypes.c_double(target))
    return vel


def cubic_lattice(n, density, cell_type=2):
    """tools/cubic_lattice.cpp:38-86 -> (pos (N,3), box side)."""
    pos = np.zeros((n, 3))
    L = lib().ljref_cubic_lattice(n, ctypes.c_double(density), cell_type, _d(pos))
    return pos, L


def initial_condition(n, density, temperature, seed=5489, cell_type=2):
    """engine/initial_condition.cpp:49-104 -> (State, box side)."""
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    L = lib().ljref_initial_condition(n, ctypes.c_double(density), ctypes.c_double(temperature),
                                      ctypes.c_uint(seed & 0xFFFFFFFF), cell_type, _d(pos),
                                      _d(vel))
    return State(pos=pos, vel=vel), L


def rdf_hist(pos, L, rmax=4.0, bins=200):
    """Brute-force ordered-pair distance histogram (no reference counterpart)."""
    pos, L = _pos(pos), _box(L)
    hist = np.zeros(bins)
    lib().ljref_rdf_hist(len(pos), _d(pos), _d(L), ctypes.c_double(rmax), bins, _d(hist))
    return hist


def rdf_cells(pos, L, rmax=4.0, bins=200):
    """rdf_hist's histogram over a cell grid of side >= rmax (for large N;
    the brute force when a box holds fewer than three such cells per axis)."""
    pos, L = _pos(pos), _box(L)
    hist = np.zeros(bins)
    lib().ljref_rdf_cells(len(pos), _d(pos), _d(L), ctypes.c_double(rmax), bins, _d(hist))
    return hist


def neighbor_counts(pos, L, rc):
    """In-cutoff neighbours of each particle under the reference's cell filter
    (each pair counted for both particles, every image separately)."""
    pos, L = _pos(pos), _box(L)
    out = np.zeros(len(pos), dtype=np.int32)
    lib().ljref_neighbor_counts(len(pos), _d(pos), _d(L), ctypes.c_double(rc), _i(out))
    return out


# ---- analyzers (physics/analyzers.cpp:44-159, tools/moving_sample.hpp:81-189) ----

def observation(temperatures, energies, virials, times, msds, density, n):
    """ThermodynamicAnalyzer::result over one sample window.

    Mean T and E; P = rho (<T> + <W>/(3N)); Cv = 1.5 / (1 - 1.5 N var(T)/<T>^2)
    with Bessel-corrected variance; D = cov(t, MSD) / var(t) / 6.
    """
    T = np.asarray(temperatures, dtype=np.float64)
    E = np.asarray(energies, dtype=np.float64)
    W = np.asarray(virials, dtype=np.float64)
    t = np.asarray(times, dtype=np.float64)
    m = np.asarray(msds, dtype=np.float64)
    k = len(T)
    tm = sum(T.tolist()) / k
    em = sum(E.tolist()) / k
    wm = sum(W.tolist()) / k
    tvar = sum(((x - tm) * (x - tm) for x in T.tolist())) / (k - 1)
    P = density * (tm + wm / (3.0 * n))
    cv = (3. / 2.) / (1 - (3. / 2.) * n * (tvar / (tm * tm)))
    mt = sum(t.tolist()) / k
    mm = sum(m.tolist()) / k
    ctt = sum(((a - mt) * (a - mt) for a in t.tolist())) / (k - 1)
    ctm = sum(((a - mt) * (b - mm) for a, b in zip(t.tolist(), m.tolist()))) / (k - 1)
    D = (1. / 6.) * (ctm / ctt)
    return dict(temperature=tm, density=density, total_energy=em, pressure=P,
                specific_heat=cv, diffusion_coefficient=D)


if os.environ.get("LJREF_BUILD_ON_IMPORT"):
    build()
