/*
 * lj_engine.h -- C ABI of the MI355X (gfx950) NVE Lennard-Jones step engine.
 *
 * The engine replaces, for one or more independent periodic systems
 * ("replicas") resident on one GPU, the reference's per-step operator chain
 *
 *   engine::VelocityVerletIntegrator::operator()   engine/integrator.cpp:62-88
 *   engine::PeriodicBoundaryCondition::operator()  engine/boundary_condition.cpp:37-51
 *   engine::ShortRangeForceCalculation::operator() engine/force_calculation.cpp:46-71
 *   engine::CellListParticlePairFilter::pairs      engine/particle_pair_filter.cpp:131-192
 *   physics::LennardJonesForce::compute            physics/lennard_jones_force.cpp:76-128
 *   physics::ThermodynamicMeasurement::operator()  physics/measurements.hpp:59-70
 *   physics::set_temperature                       physics/transformations.cpp:94-117
 *
 * (paths relative to the reference repository root).  Plain C types only;
 * no exception crosses this boundary.  Every function returns LJ_OK (0) or a
 * negative LJ_E* code; lj_last_error() describes the most recent failure on
 * the calling thread.  One engine handle owns its device buffers and one HIP
 * stream; use a handle from one host thread at a time.
 *
 * Host arrays are row-major N x 3 doubles (x, y, z per particle), indexed by
 * particle id -- the columns of the reference's 4 x N Eigen matrices
 * (physics/system_state.hpp:85-90) without the always-zero 4th row.
 */
#ifndef LJ_ENGINE_H
#define LJ_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

#define LJ_OK 0
#define LJ_E_INVALID -1   /* bad argument or configuration */
#define LJ_E_HIP -2       /* HIP runtime failure */
#define LJ_E_STATE -3     /* call not valid in the engine's current state */

/* Arithmetic of the pair kernel.  LJ_FP64: everything in double.  LJ_MIXED:
 * pair separations, r^-6 and forces in float on cell-local coordinates;
 * positions, velocities, displacements and the PE / virial / KE / MSD sums
 * stay in double. */
#define LJ_PRECISION_FP64 0
#define LJ_PRECISION_MIXED 1

/* Pair law.  LJ_LAW_LJ is the reference's spline-shifted LJ
 * (physics/lennard_jones_force.cpp:33-128) generalised to V = eps v(r/sigma)
 * with the cutoff rc = cutoff * sigma; eps = sigma = 1 is the reference.
 * LJ_LAW_CONSTANT is the reference's test double
 * tests/cpp/mock/constant_short_range_force.cpp:31-59 (V = k (rc - r),
 * W = k r).  LJ_LAW_NONE performs no force calculation: forces keep their
 * stored values (the reference's Integrator built without a short-range
 * force, engine/integrator.cpp:79). */
#define LJ_LAW_LJ 0
#define LJ_LAW_CONSTANT 1
#define LJ_LAW_NONE 2

typedef struct lj_system_desc {
    int particle_count;       /* 1 <= N <= 2^30 */
    double box[3];            /* periodic box sides (tools/bounding_box.hpp:30-61) */
    double cutoff;            /* LJ: reduced cutoff rc/sigma; CONSTANT: absolute cutoff */
    double time_delta;        /* dt */
    double epsilon;           /* LJ energy scale (1 in the reference) */
    double sigma;             /* LJ length scale (1 in the reference) */
    double strength;          /* CONSTANT law strength k */
    int force_law;            /* LJ_LAW_* */
    int periodic;             /* 1: apply the periodic boundary condition each step;
                                 0 only allowed with LJ_LAW_NONE */
} lj_system_desc;

/* One row of the reference's thermodynamics log
 * (physics/measurements.hpp:42-56): the state after a step. */
typedef struct lj_thermo {
    double time;
    double kinetic_energy;
    double potential_energy;
    double total_energy;
    double virial;
    double temperature;
    double mean_square_displacement;
} lj_thermo;

typedef struct lj_engine lj_engine;

/* Message of the most recent failing call on this thread ("" if none). */
const char* lj_last_error(void);

/* Number of visible HIP devices (0 if none); never fails. */
int lj_device_count(void);

/* Creates an engine for nsys independent systems on the current HIP
 * device.  Replaces api::Simulation's construction of the integrator
 * (api/simulation.cpp:119-125, Integrator::Builder(dt).bounding_box(box)
 * .short_range_force(lj).build(), engine/integrator_builder.hpp:36-170) and,
 * for nsys > 1, the per-replica integrators of api::SimulationPool
 * (api/simulation_pool.cpp:93-101).  The reference's preconditions are
 * errors here: box >= cutoff on every axis (engine/particle_pair_filter.cpp:61-64),
 * LJ cutoff^2 > 2^(1/3) (physics/lennard_jones_force.cpp:45).  Limits of
 * this build: at most 1023 cells along a box side, and all systems of one
 * engine together at most 2^30 particle slots (N rounded up to a multiple of
 * 256 each) and 2^30 cells.  The descriptors are checked before the device
 * is looked for, so LJ_E_INVALID comes back on a machine without a GPU too. */
int lj_engine_create(const lj_system_desc* systems, int nsys, int precision, lj_engine** out);

int lj_engine_destroy(lj_engine* e);

int lj_engine_system_count(const lj_engine* e);

/* Loads the state of system `sys` (the reference's SystemState,
 * physics/system_state.hpp:35-102): pos, vel, disp, force as N x 3 host
 * arrays (disp or force may be NULL = zeros) plus the elapsed time. */
int lj_engine_set_state(lj_engine* e, int sys, const double* pos, const double* vel,
                        const double* disp, const double* force, double time);

/* Copies the state of system `sys` back, in particle-id order.  Any output
 * pointer may be NULL.  potential_energy / virial are those of the last
 * force calculation. */
int lj_engine_get_state(lj_engine* e, int sys, double* pos, double* vel, double* disp,
                        double* force, double* time, double* potential_energy, double* virial);

/* Advances every system by nsteps velocity-Verlet steps
 * (engine/integrator.cpp:51-88 applied nsteps times).  If records is not
 * NULL it receives nsteps * nsys lj_thermo rows, step-major
 * (records[step * nsys + sys]), each the measurement of
 * physics/measurements.hpp:59-70 after that step.  The loop runs from a
 * captured HIP graph unless graphs were disabled.  After a failed call the
 * systems' states are unspecified until lj_engine_set_state has set each of
 * them again. */
int lj_engine_step(lj_engine* e, int nsteps, lj_thermo* records);

/* lj_engine_step without any allocation: no hipMalloc / hipFree, no graph
 * capture or instantiation, no library routine that allocates.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture elsewhere in the process,
 * tight serving loops, harnesses that forbid allocation).  The records are
 * staged in a fixed device array of LJ_NOALLOC_MAX_ROWS rows (one lj_thermo
 * per step and system) that every engine of a device shares: nsteps *
 * nsys beyond that is LJ_E_INVALID; profiling must be off (LJ_E_STATE).
 * Results, records and state equal lj_engine_step's bit for bit.  Calls on
 * engines of the same device that overlap in time race on that array. */
#define LJ_NOALLOC_MAX_ROWS 16384
int lj_engine_step_noalloc(lj_engine* e, int nsteps, lj_thermo* records);

/* Prepares calls of lj_engine_step of up to max_steps steps: sizes the
 * device record buffer and captures the step graphs now, so that such a call
 * neither allocates device memory nor captures graphs inside a caller's timed
 * region.  No reference counterpart (the reference allocates nothing per
 * step; its per-step measurement goes straight to a message queue,
 * control/simulation_controller.cpp:60-69). */
int lj_engine_reserve(lj_engine* e, int max_steps);

/* Recomputes forces, potential energy and virial on the current positions
 * without moving the particles (engine/force_calculation.cpp:46-71).
 * pe / vir receive nsys values each when not NULL. */
int lj_engine_compute_forces(lj_engine* e, double* pe, double* vir);

/* physics::set_temperature (physics/transformations.cpp:94-117) on system
 * `sys`: v *= sqrt(target / T_instantaneous).  measured (optional) receives
 * the temperature before scaling. */
int lj_engine_set_temperature(lj_engine* e, int sys, double target, double* measured);

/* ThermodynamicMeasurement of the current state (physics/measurements.hpp:59-70)
 * for every system, without stepping; out receives nsys rows. */
int lj_engine_measure(lj_engine* e, lj_thermo* out);

/* Freezes (active = 0) or resumes system `sys`: a frozen system is not
 * integrated, produces no records and keeps its time, so replicas of a
 * batch that have finished their schedule stop consuming steps (the role of
 * a finished worker in api::SimulationPool, api/simulation_pool.cpp:93-101). */
int lj_engine_set_active(lj_engine* e, int sys, int active);

/* ---- parity hooks ---------------------------------------------------- */

/* Integer cell coordinates of every particle of system `sys` on the current
 * positions, as engine/particle_pair_filter.cpp:144-147 computes them
 * (keys: N x 3 ints, particle-id order).  nc (optional) receives the grid
 * shape of tools/cell_list_array.cpp:45-51.  Returns LJ_OK; *clamped
 * (optional) receives the number of coordinates clamped into [0, nc). */
int lj_engine_cell_keys(lj_engine* e, int sys, int* keys, int* nc, int* clamped);

/* Cell list of system `sys` built by the engine's counting sort on the
 * current positions, in CSR form: start[ncell + 1] (offsets from 0),
 * atoms[N] (particle ids, ascending within each cell; cells x-slowest as in
 * tools/cell_list_array.cpp:63-81). */
int lj_engine_cell_list(lj_engine* e, int sys, int* start, int* atoms);

/* In-cutoff pairs of system `sys` found by the engine's pair traversal on
 * the current positions.  Each unordered pair appears once as
 * (first < second, image x, image y, image z) in idx[5*k..] with separation
 * r_first - r_second - image * L in sep[3*k..] (the reference's ParticlePair,
 * engine/particle_pair_filter.hpp:35-50, oriented by id).  *count receives
 * the number of pairs; at most cap are written (order unspecified). */
int lj_engine_pairs(lj_engine* e, int sys, long long cap, int* idx, double* sep,
                    long long* count);

/* Number of in-cutoff neighbours (pairs with each image counted) of every
 * particle of system `sys` as decided by the production pair kernel (the
 * current precision and pair-kernel variant) on the current positions;
 * counts: N ints in particle-id order.  Also recomputes the forces. */
int lj_engine_neighbor_counts(lj_engine* e, int sys, int* counts);

/* Ordered-pair distance histogram of every system on the current positions
 * (bins of width rmax / bins over [0, rmax), images included); accumulates
 * into the engine's g(r) buffer.  lj_engine_rdf_read copies the counts
 * (nsys x bins doubles) and the number of accumulated samples, then clears. */
int lj_engine_rdf_sample(lj_engine* e, double rmax, int bins);
int lj_engine_rdf_read(lj_engine* e, double* counts, int* samples);

/* ---- differentiation ---------------------------------------------------- */

/* Hand-written adjoint of the FP64 velocity-Verlet trajectory of a
 * single-system LJ engine (no reference counterpart).  Starting from the
 * engine's current state, with forces F(r0) computed first, it integrates
 * nsteps steps and returns
 *   loss = (1/nsteps) sum_{t=1..nsteps} [PE_t + rho (T_t + W_t / (3N))]
 * (rho = density) and its exact gradients with respect to epsilon, sigma
 * (rc = cutoff * sigma moves with sigma), the initial positions and
 * velocities (N x 3, particle-id order).  States are checkpointed every
 * `every` steps and recomputed segment by segment in the reverse sweep.
 * The engine is left holding the state it had on entry -- positions,
 * velocities, displacements and elapsed time -- with the forces, potential
 * energy and virial recomputed on those positions. */
int lj_engine_adjoint(lj_engine* e, int nsteps, int every, double density, double* loss,
                      double* grad_eps, double* grad_sigma, double* grad_pos, double* grad_vel);

/* ---- tuning ----------------------------------------------------------- */

/* 0 disables HIP graph capture of the step loop (launches each kernel). */
int lj_engine_set_graphs(lj_engine* e, int enabled);

/* Resident cells: a step call of at least 4 steps on an engine whose systems
 * all use the tiled pair kernel (all LJ_PAIR_CELL_TILED or all
 * LJ_PAIR_CELL_MULTIWAVE) builds its cell lists once, keeps every cell in a
 * fixed block of slots and, in each later step, moves only the particles that
 * changed cell (in place of the key pass, scan and sort).  The cells stay live
 * for later step calls of any length; every other entry point that reads or
 * changes the state packs them back first.  Records and state are the same
 * bit for bit with any mode.  When a cell would exceed its slots (48 members)
 * or receive more than 8 particles in one step, the steps taken since the
 * cells were entered are redone with the full cell build (same results, more
 * time; records already returned are not touched).  The first resident call
 * allocates the resident buffers (about 175 bytes per cell slot, 48 slots per
 * cell).  Modes: LJ_RESIDENT_OFF; LJ_RESIDENT_AUTO (default): when the engine
 * holds at least 10 particles per cell (liquid densities; dilute systems gain
 * nothing); LJ_RESIDENT_ON: whenever the kernels allow it.  lj_engine_autotune
 * also chooses between on and off.  No counterpart in the reference, which
 * rebuilds its cell lists every step (engine/particle_pair_filter.cpp:141-156). */
#define LJ_RESIDENT_OFF 0
#define LJ_RESIDENT_AUTO 1
#define LJ_RESIDENT_ON 2
int lj_engine_set_resident_cells(lj_engine* e, int mode);

/* The resident-cell mode (LJ_RESIDENT_OFF for a null engine). */
int lj_engine_resident_cells(const lj_engine* e);

/* Step calls run on resident cells since creation, and how many of them
 * were redone with the full cell build. */
int lj_engine_resident_stats(const lj_engine* e, long long* calls, long long* fallbacks);

/* Pair-kernel variant: LJ_PAIR_CELL_TILED (default) runs one wave64 per home
 * cell with the neighbour cells staged in LDS; LJ_PAIR_PER_PARTICLE runs one
 * thread per particle over the 27 neighbour cells; LJ_PAIR_BY_OCCUPANCY picks
 * per system: tiled where the mean cell occupancy N / cells is at least the
 * threshold of lj_engine_set_pair_split (default 6), per-particle elsewhere
 * (dilute replicas of a sweep), both kernels running in the same step;
 * LJ_PAIR_CELL_MULTIWAVE runs the tiled kernel with four waves per home cell,
 * each over a quarter of the neighbour cells (small systems, where a wave per
 * cell leaves most of the GPU idle).  All give the same pair set and, in
 * FP64, the same cutoff decisions.  No
 * counterpart in the reference (its single pair filter,
 * engine/particle_pair_filter.cpp:131-192, has no variants). */
#define LJ_PAIR_PER_PARTICLE 0
#define LJ_PAIR_CELL_TILED 1
#define LJ_PAIR_BY_OCCUPANCY 2
#define LJ_PAIR_CELL_MULTIWAVE 3
int lj_engine_set_pair_kernel(lj_engine* e, int variant);

/* Selects LJ_PAIR_BY_OCCUPANCY with the given threshold (particles per cell). */
int lj_engine_set_pair_split(lj_engine* e, double min_occupancy);

/* The kernel each system uses now: per_system[k] = LJ_PAIR_CELL_TILED,
 * LJ_PAIR_CELL_MULTIWAVE or LJ_PAIR_PER_PARTICLE. */
int lj_engine_pair_kernels(lj_engine* e, int* per_system);

/* Launch shape (no reference counterpart; neither knob changes any result):
 * cells_per_block = home cells (one wave64 each) per workgroup of the mixed
 * tiled pair kernel, 2, 4 (default) or 8; sort_wave_particles = particles the
 * cell sort gives each wave (cells per wave = that / mean occupancy), 8..64
 * (default 44). */
int lj_engine_set_launch_shape(lj_engine* e, int cells_per_block, int sort_wave_particles);
int lj_engine_launch_shape(lj_engine* e, int* cells_per_block, int* sort_wave_particles);

/* Project autotuner, for the reference's one-Simulation-per-thread pool
 * (api/simulation_pool.cpp:93-101) that this engine's batches replace.  It
 * times calls of `steps` and 2 `steps` steps of each pair-kernel assignment
 * (their difference: the time per step of long calls, without what a call
 * costs once) -- all tiled, all
 * per-particle, and per-system splits at cell occupancies 1.5, 3, 6 and 12
 * that give distinct assignments -- then, with the fastest assignment, each
 * launch shape (cells per block 2 / 4 / 8 for mixed tiled systems, sort waves
 * of 32 / 44 / 56 particles), on the engine's own systems from the current
 * state, which is saved and restored, so tuning leaves no trace; it keeps the
 * fastest.  chosen / ms_per_step (optional) receive the winning variant and
 * its time per step; lj_engine_pair_kernels and lj_engine_launch_shape tell
 * the choices. */
int lj_engine_autotune(lj_engine* e, int steps, int* chosen, double* ms_per_step);

/* Mean duration in milliseconds of each step-loop kernel over the last
 * lj_engine_step call with profiling enabled (HIP events on the engine's
 * stream).  names/ms receive up to cap entries; returns the entry count. */
int lj_engine_set_profiling(lj_engine* e, int enabled);
int lj_engine_kernel_times(lj_engine* e, int cap, const char** names, double* ms, int* count);

/* Launch count of each kernel of lj_engine_kernel_times, in the same order
 * (per-step kernels launch once per step; a resident call's kick-drift and
 * record finish once per call). */
int lj_engine_kernel_launches(lj_engine* e, int cap, long long* launches, int* count);

#ifdef __cplusplus
}
#endif

#endif /* LJ_ENGINE_H */
