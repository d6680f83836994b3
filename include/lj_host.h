/*
 * lj_host.h -- C ABI of the host-side C++ layer above the GPU engine.
 *
 * These entry points replace the reference's api / control / output layers
 * as far as the NVE step loop needs them (paths relative to the reference
 * repository root).  Plain C types; status codes as in lj_engine.h.
 */
#ifndef LJ_HOST_H
#define LJ_HOST_H

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

/* Formats a double the way the reference's fmt "{}" sinks print it
 * (output/sinks.cpp:100-204): shortest round-trip digits, fixed notation
 * when the decimal exponent of the leading digit is in [-4, 16), otherwise
 * d.ddde+XX (at least two exponent digits); integral values without a
 * decimal point.  Writes at most cap bytes incl. the terminator;
 * returns the length. */
int lj_format_double(double value, char* buf, int cap);


#ifdef __cplusplus
}
#endif

#endif /* LJ_HOST_H */
