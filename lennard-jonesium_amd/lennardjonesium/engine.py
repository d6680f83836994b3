"""Python front end of the GPU step engine (include/lj_engine.h).

``Engine`` owns one native engine handle: one or more independent periodic
systems resident on the GPU.  It is the object the operator-level API
(``lennardjonesium.operators``) and the orchestration (``run``,
``run_sweep``) drive; tests use it to reach every C-ABI entry point.

Host arrays are numpy float64 of shape (N, 3), indexed by particle id.
"""

import ctypes
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from lennardjonesium import _native
from lennardjonesium._native import check

LAWS = {"lj": 0, "constant": 1, "none": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}
PAIR_KERNELS = {"particle": 0, "tiled": 1, "split": 2, "multiwave": 3}
THERMO_FIELDS = ("time", "kinetic_energy", "potential_energy", "total_energy", "virial",
                 "temperature", "mean_square_displacement")


@dataclass
class SystemSpec:
    """One periodic system (the reference's SystemParameters + force parameters)."""

    particle_count: int
    box: Sequence[float]
    cutoff: float = 2.5
    time_delta: float = 0.005
    epsilon: float = 1.0
    sigma: float = 1.0
    strength: float = 0.0
    force_law: str = "lj"
    periodic: bool = True

    def desc(self):
        d = _native.lj_system_desc()
        d.particle_count = int(self.particle_count)
        box = np.broadcast_to(np.asarray(self.box, dtype=np.float64), 3)
        for a in range(3):
            d.box[a] = float(box[a])
        d.cutoff = float(self.cutoff)
        d.time_delta = float(self.time_delta)
        d.epsilon = float(self.epsilon)
        d.sigma = float(self.sigma)
        d.strength = float(self.strength)
        d.force_law = LAWS[self.force_law]
        d.periodic = 1 if self.periodic else 0
        return d


def _dptr(a):
    return a.ctypes.data_as(_native.P_d) if a is not None else None


def _arr(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n, 3))
    return a


class Engine:
    """GPU engine for ``systems`` (list of SystemSpec), precision 'fp64' or 'mixed'."""

    def __init__(self, systems, precision="fp64", graphs=True, pair_kernel="tiled"):
        self.lib = _native.engine()
        self.specs = list(systems)
        self.precision = precision
        arr = (_native.lj_system_desc * len(self.specs))(*[s.desc() for s in self.specs])
        h = ctypes.c_void_p()
        check(self.lib.lj_engine_create(arr, len(self.specs), PRECISIONS[precision],
                                        ctypes.byref(h)))
        self.h = h
        if not graphs:
            check(self.lib.lj_engine_set_graphs(self.h, 0))
        if pair_kernel != "tiled":
            self.set_pair_kernel(pair_kernel)

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.lj_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nsys(self):
        return len(self.specs)

    # -- state --------------------------------------------------------------
    def set_state(self, sys, pos, vel, disp=None, force=None, time=0.0):
        n = self.specs[sys].particle_count
        pos, vel, disp, force = (_arr(x, n) for x in (pos, vel, disp, force))
        check(self.lib.lj_engine_set_state(self.h, sys, _dptr(pos), _dptr(vel), _dptr(disp),
                                           _dptr(force), float(time)))

    def get_state(self, sys):
        n = self.specs[sys].particle_count
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "disp", "force")}
        t, pe, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.lj_engine_get_state(self.h, sys, _dptr(out["pos"]), _dptr(out["vel"]),
                                           _dptr(out["disp"]), _dptr(out["force"]),
                                           ctypes.byref(t), ctypes.byref(pe), ctypes.byref(w)))
        out.update(time=t.value, pe=pe.value, vir=w.value)
        return out

    # -- stepping -----------------------------------------------------------
    def step(self, nsteps, records=True):
        """Advances nsteps; returns records of shape (nsteps, nsys, 7) or None."""
        if records:
            rec = np.zeros((nsteps, self.nsys, 7))
            ptr = rec.ctypes.data_as(ctypes.POINTER(_native.lj_thermo))
        else:
            rec, ptr = None, None
        check(self.lib.lj_engine_step(self.h, int(nsteps), ptr))
        return rec

    def reserve(self, max_steps):
        """Sizes the record buffer and captures the step graphs for calls of
        up to max_steps steps, so such calls allocate and capture nothing."""
        check(self.lib.lj_engine_reserve(self.h, int(max_steps)))

    def compute_forces(self):
        pe = np.zeros(self.nsys)
        w = np.zeros(self.nsys)
        check(self.lib.lj_engine_compute_forces(self.h, _dptr(pe), _dptr(w)))
        return pe, w

    def measure(self):
        out = np.zeros((self.nsys, 7))
        check(self.lib.lj_engine_measure(self.h, out.ctypes.data_as(
            ctypes.POINTER(_native.lj_thermo))))
        return out

    def set_temperature(self, sys, target):
        m = ctypes.c_double()
        check(self.lib.lj_engine_set_temperature(self.h, sys, float(target), ctypes.byref(m)))
        return m.value

    def set_active(self, sys, active):
        check(self.lib.lj_engine_set_active(self.h, sys, 1 if active else 0))

    def adjoint(self, nsteps, density, every=100):
        """Loss (1/n) sum_t [PE_t + rho (T_t + W_t / 3N)] over nsteps FP64 steps
        from the current state and its gradients: returns
        (loss, d/deps, d/dsigma, d/dr0 (N, 3), d/dv0 (N, 3))."""
        n = self.specs[0].particle_count
        L, ge, gs = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        gr = np.zeros((n, 3))
        gv = np.zeros((n, 3))
        check(self.lib.lj_engine_adjoint(self.h, int(nsteps), int(every), float(density),
                                         ctypes.byref(L), ctypes.byref(ge), ctypes.byref(gs),
                                         _dptr(gr), _dptr(gv)))
        return L.value, ge.value, gs.value, gr, gv

    # -- parity hooks ---------------------------------------------------------
    def cell_keys(self, sys):
        n = self.specs[sys].particle_count
        keys = np.zeros((n, 3), dtype=np.int32)
        nc = np.zeros(3, dtype=np.int32)
        cl = ctypes.c_int()
        check(self.lib.lj_engine_cell_keys(self.h, sys, keys.ctypes.data_as(_native.P_i),
                                           nc.ctypes.data_as(_native.P_i), ctypes.byref(cl)))
        return keys, nc, cl.value

    def cell_list(self, sys):
        n = self.specs[sys].particle_count
        keys, nc, _ = self.cell_keys(sys)
        start = np.zeros(int(np.prod(nc)) + 1, dtype=np.int32)
        atoms = np.zeros(n, dtype=np.int32)
        check(self.lib.lj_engine_cell_list(self.h, sys, start.ctypes.data_as(_native.P_i),
                                           atoms.ctypes.data_as(_native.P_i)))
        return start, atoms

    def pairs(self, sys, cap=None):
        n = self.specs[sys].particle_count
        cap = int(cap if cap is not None else 64 * n)
        while True:
            idx = np.zeros((max(cap, 1), 5), dtype=np.int32)
            sep = np.zeros((max(cap, 1), 3))
            cnt = ctypes.c_longlong()
            check(self.lib.lj_engine_pairs(self.h, sys, cap, idx.ctypes.data_as(_native.P_i),
                                           _dptr(sep), ctypes.byref(cnt)))
            if cnt.value <= cap:
                return idx[:cnt.value], sep[:cnt.value]
            cap = cnt.value

    def neighbor_counts(self, sys):
        n = self.specs[sys].particle_count
        cnt = np.zeros(n, dtype=np.int32)
        check(self.lib.lj_engine_neighbor_counts(self.h, sys, cnt.ctypes.data_as(_native.P_i)))
        return cnt

    # -- observers ------------------------------------------------------------
    def rdf_sample(self, rmax=4.0, bins=200):
        check(self.lib.lj_engine_rdf_sample(self.h, float(rmax), int(bins)))

    def rdf_read(self, bins=200):
        counts = np.zeros((self.nsys, bins))
        ns = ctypes.c_int()
        check(self.lib.lj_engine_rdf_read(self.h, _dptr(counts), ctypes.byref(ns)))
        return counts, ns.value

    # -- tuning / timing ---------------------------------------------------------
    def set_graphs(self, enabled):
        check(self.lib.lj_engine_set_graphs(self.h, 1 if enabled else 0))

    def set_resident_cells(self, mode):
        """Resident cells, entered by a step call of 4 or more steps: the cell
        lists are built once and afterwards, across later step calls too, only
        the particles that change cell are moved; results are the same bit for
        bit.  mode: False / 'off', 'auto' (default: at 10 or more particles
        per cell), True / 'on'."""
        modes = {False: 0, "off": 0, "auto": 1, True: 2, "on": 2}
        check(self.lib.lj_engine_set_resident_cells(self.h, modes[mode]))

    def resident_cells(self):
        return {0: "off", 1: "auto", 2: "on"}[self.lib.lj_engine_resident_cells(self.h)]

    def resident_stats(self):
        """(calls run on resident cells, calls redone with the full cell build)."""
        c, f = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.lj_engine_resident_stats(self.h, ctypes.byref(c), ctypes.byref(f)))
        return c.value, f.value

    def set_pair_kernel(self, variant):
        """'tiled' (one wave per home cell, LDS-staged neighbours), 'particle'
        (one thread per particle), 'multiwave' (the tiled kernel with four
        waves per home cell; small systems) or 'split' (per system, by cell
        occupancy)."""
        check(self.lib.lj_engine_set_pair_kernel(self.h, PAIR_KERNELS[variant]))

    def set_pair_split(self, min_occupancy):
        """Per-system kernels: tiled where N / cells >= min_occupancy."""
        check(self.lib.lj_engine_set_pair_split(self.h, float(min_occupancy)))

    def pair_kernels(self):
        """The kernel each system uses now ('tiled', 'multiwave' or 'particle')."""
        out = np.zeros(self.nsys, dtype=np.int32)
        check(self.lib.lj_engine_pair_kernels(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        names = {k: n for n, k in PAIR_KERNELS.items()}
        return [names[int(v)] for v in out]

    def set_launch_shape(self, cells_per_block=4, sort_wave_particles=44):
        """Waves per workgroup of the mixed tiled kernel (2, 4, 8) and particles
        per sort wave (8..64); results do not depend on either."""
        check(self.lib.lj_engine_set_launch_shape(self.h, int(cells_per_block),
                                                  int(sort_wave_particles)))

    def launch_shape(self):
        cpb, sw = ctypes.c_int(), ctypes.c_int()
        check(self.lib.lj_engine_launch_shape(self.h, ctypes.byref(cpb), ctypes.byref(sw)))
        return cpb.value, sw.value

    def autotune(self, steps=20):
        """Times each pair-kernel variant on this workload and keeps the fastest;
        returns (variant name, ms per step)."""
        v = ctypes.c_int()
        ms = ctypes.c_double()
        check(self.lib.lj_engine_autotune(self.h, int(steps), ctypes.byref(v), ctypes.byref(ms)))
        name = {k: n for n, k in PAIR_KERNELS.items()}[v.value]
        return name, ms.value

    def set_profiling(self, enabled):
        check(self.lib.lj_engine_set_profiling(self.h, 1 if enabled else 0))

    def kernel_times(self):
        names = (ctypes.c_char_p * 32)()
        ms = np.zeros(32)
        cnt = ctypes.c_int()
        check(self.lib.lj_engine_kernel_times(self.h, 32, names, _dptr(ms), ctypes.byref(cnt)))
        return {names[k].decode(): float(ms[k]) for k in range(min(cnt.value, 32))}

    def kernel_launches(self):
        """Launches of each kernel of kernel_times() (same keys)."""
        names = (ctypes.c_char_p * 32)()
        ms = np.zeros(32)
        cnt = ctypes.c_int()
        check(self.lib.lj_engine_kernel_times(self.h, 32, names, _dptr(ms), ctypes.byref(cnt)))
        n = (ctypes.c_longlong * 32)()
        check(self.lib.lj_engine_kernel_launches(self.h, 32, n, ctypes.byref(cnt)))
        return {names[k].decode(): int(n[k]) for k in range(min(cnt.value, 32))}


def device_count():
    return _native.engine().lj_device_count()


def initial_condition(n, density, temperature, seed=5489, lattice=2):
    """engine::InitialCondition on the host (csrc/host/initial_condition.cpp).

    Returns (pos, vel, box_side)."""
    lib = _native.host()
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    L = lib.lj_initial_condition(int(n), float(density), float(temperature),
                                 int(seed) & 0xFFFFFFFF, int(lattice), _dptr(pos), _dptr(vel))
    if L < 0:
        raise ValueError("invalid initial-condition parameters")
    return pos, vel, L


def cubic_lattice(n, density, lattice=2):
    lib = _native.host()
    pos = np.zeros((n, 3))
    L = lib.lj_cubic_lattice(int(n), float(density), int(lattice), _dptr(pos))
    if L < 0:
        raise ValueError("invalid lattice parameters")
    return pos, L


def sweep_sizes(seed, count, lo=5, hi=20):
    """Replica lattice sizes n_k in [lo, hi] from std::mt19937(seed) with
    rejection sampling (csrc/host/initial_condition.cpp)."""
    out = np.zeros(count, dtype=np.int32)
    if _native.host().lj_sweep_sizes(int(seed) & 0xFFFFFFFF, int(count), int(lo), int(hi),
                                     out.ctypes.data_as(_native.P_i)) < 0:
        raise ValueError("invalid range")
    return out


def thermo_observation(records, sample_size, density, particle_count):
    """ThermodynamicAnalyzer::result over the last sample_size rows of
    records ((k, 7): time, KE, PE, E, W, T, MSD) -> dict of the six
    observation fields (csrc/host/simulation.cpp, the phase's own code)."""
    rec = np.ascontiguousarray(np.asarray(records, dtype=np.float64).reshape(-1, 7))
    out = np.zeros(6)
    rc = _native.host().lj_thermo_observation(
        rec.ctypes.data_as(ctypes.POINTER(_native.lj_thermo)), len(rec), int(sample_size),
        float(density), int(particle_count), _dptr(out))
    if rc < 0:
        raise ValueError("invalid observation window")
    return dict(zip(("temperature", "density", "total_energy", "pressure", "specific_heat",
                     "diffusion_coefficient"), out.tolist()))


def format_double(x):
    buf = ctypes.create_string_buffer(64)
    _native.host().lj_format_double(float(x), buf, 64)
    return buf.value.decode()
