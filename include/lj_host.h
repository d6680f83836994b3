/*
 * lj_host.h -- C ABI of the host-side C++ layer above the GPU engine.
 *
 * These entry points replace the reference's api / control / output layers
 * as far as the NVE step loop needs them (paths relative to the reference
 * repository root).  Plain C types; status codes as in lj_engine.h.
 */
#ifndef LJ_HOST_H
#define LJ_HOST_H

#include "lj_engine.h"   /* lj_thermo, status codes */

#ifdef __cplusplus
extern "C" {
#endif

/* Lattice types of tools/cubic_lattice.hpp:53-65. */
#define LJ_LATTICE_SIMPLE 0
#define LJ_LATTICE_BODY_CENTERED 1
#define LJ_LATTICE_FACE_CENTERED 2

/* tools::CubicLattice (tools/cubic_lattice.cpp:38-86): writes N x 3 lattice
 * sites to pos and returns the cubic box side (or -1 on bad input). */
double lj_cubic_lattice(int n, double density, int lattice, double* pos);

/* engine::InitialCondition (engine/initial_condition.cpp:49-104): lattice
 * positions, seeded Maxwell-Boltzmann velocities with zero linear and angular
 * momentum, rescaled exactly to the temperature.  Returns the box side (or
 * -1 on bad input).  Forces of the initial state are zero. */
double lj_initial_condition(int n, double density, double temperature, unsigned int seed,
                            int lattice, double* pos, double* vel);

/* Replica sizes for batched sweeps: count integers in [lo, hi] drawn from
 * std::mt19937(seed) by rejection sampling (u < floor(2^32 / R) * R,
 * value lo + u mod R).  No reference counterpart: the reference's sweeps use
 * one particle count (orchestration/sweep_configuration.py:55).  Returns 0,
 * or -1 on bad arguments. */
int lj_sweep_sizes(unsigned int seed, int count, int lo, int hi, int* out);

/* Formats a double the way the reference's fmt "{}" sinks print it
 * (output/sinks.cpp:100-204): shortest round-trip digits, fixed notation
 * when the decimal exponent of the leading digit is in [-4, 16), otherwise
 * d.ddde+XX (at least two exponent digits); integral values without a
 * decimal point.  Writes at most cap bytes incl. the terminator;
 * returns the length. */
int lj_format_double(double value, char* buf, int cap);


/* api::Simulation (api/simulation.hpp:81-161) over the GPU engine.  The
 * configuration mirrors api::Configuration (api/configuration.hpp:35-102);
 * NULL strings take the reference's defaults. */
typedef struct lj_sim_config {
    double temperature;
    double density;
    int particle_count;
    unsigned int random_seed;
    double cutoff_distance;
    double time_delta;
    /* equilibration phase (control/simulation_phase.hpp:120-160) */
    const char* equilibration_name;
    double equilibration_tolerance;
    int equilibration_sample_size;
    int adjustment_interval;
    int steady_state_time;
    int equilibration_timeout;
    /* observation phase (control/simulation_phase.hpp:200-244) */
    const char* observation_name;
    double observation_tolerance;
    int observation_sample_size;
    int observation_interval;
    int observation_count;
    /* output files */
    const char* event_log;
    const char* thermodynamic_log;
    const char* observation_log;
    const char* snapshot_log;
    /* engine precision: LJ_PRECISION_* of lj_engine.h */
    int precision;
} lj_sim_config;

typedef struct lj_simulation lj_simulation;

/* api::make_simulation (api/configuration.cpp:40-92): builds the initial
 * condition and the force parameters; the GPU is not touched until run. */
int lj_simulation_create(const lj_sim_config* cfg, lj_simulation** out);
int lj_simulation_destroy(lj_simulation* sim);

/* api::Simulation::run (api/simulation.cpp:72-117): equilibration then
 * observation, writing the four output files; echo != 0 also prints each
 * events.log line to stdout.  Re-runnable; files are overwritten. */
int lj_simulation_run(lj_simulation* sim, int echo);

/* Runs nsim simulations as one batched GPU engine (one launch per kernel for
 * all replicas), each with its own schedule and files -- the role of
 * api::SimulationPool (api/simulation_pool.cpp:35-101).  Replicas that finish
 * are frozen.  status (optional, nsim ints) reads 3 while running, then
 * 0 = completed, 1 = aborted in equilibration, 2 = aborted in observation. */
int lj_simulation_run_batch(lj_simulation** sims, int nsim, int echo, int* status);

/* physics::ShortRangeForce::potential / virial / force along z
 * (physics/forces.hpp:80-94), exposed for plots as api::Simulation does
 * (api/simulation.hpp:148-150). */
double lj_simulation_potential(const lj_simulation* sim, double distance);
double lj_simulation_virial(const lj_simulation* sim, double distance);
double lj_simulation_force(const lj_simulation* sim, double distance);

/* physics::ThermodynamicAnalyzer::result (physics/analyzers.cpp:44-159) over
 * a tools::MovingSample window (tools/moving_sample.hpp:81-189) of the last
 * sample_size of the count records given, oldest first -- the statistics an
 * observation phase records: out[6] = temperature, density, total energy,
 * pressure, specific heat, diffusion coefficient.  The same code the
 * simulation's observation phase runs. */
int lj_thermo_observation(const lj_thermo* records, int count, int sample_size, double density,
                          int particle_count, double* out);

/* Message of the most recent failing host-layer call on this thread. */
const char* lj_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LJ_HOST_H */
