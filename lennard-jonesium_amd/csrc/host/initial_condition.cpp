// initial_condition.cpp -- host construction of the initial SystemState.
//
// Follows engine/initial_condition.cpp:49-104 and tools/cubic_lattice.cpp:38-86
// (paths relative to the reference repository root): lattice sites of a
// simple / body-centred / face-centred cubic lattice scaled to the density,
// Maxwell-Boltzmann velocities from std::mt19937(seed) and
// std::normal_distribution<>(0, sqrt(T)) drawn in the order vx0, vy0, vz0,
// vx1, ..., then zero linear momentum, zero angular momentum about the
// centre of mass and an exact rescale to T (physics/transformations.cpp:35-117).
// Forces start at zero: the reference never computes F(0) before the first
// step.

#include <cmath>
#include <random>
#include <vector>

#include "lj_host.h"

namespace {

struct UnitCell {
    int m;
    double site[4][3];
};

const UnitCell kCells[3] = {
    {1, {{0, 0, 0}}},
    {2, {{0, 0, 0}, {0.5, 0.5, 0.5}}},
    {4, {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}}},
};

// Solves the 3x3 system A x = b by Gaussian elimination with partial pivoting;
// singular directions (a lattice with fewer than two distinct sites on an
// axis) receive zero.
void solve3(double A[3][3], const double b[3], double x[3])
{
    double M[3][4];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) M[r][c] = A[r][c];
        M[r][3] = b[r];
    }
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (piv != c)
            for (int k = 0; k < 4; ++k) std::swap(M[c][k], M[piv][k]);
        if (M[c][c] == 0.0) continue;
        for (int r = c + 1; r < 3; ++r) {
            const double f = M[r][c] / M[c][c];
            for (int k = c; k < 4; ++k) M[r][k] -= f * M[c][k];
        }
    }
    for (int r = 2; r >= 0; --r) {
        if (M[r][r] == 0.0) { x[r] = 0.0; continue; }
        double s = M[r][3];
        for (int k = r + 1; k < 3; ++k) s -= M[r][k] * x[k];
        x[r] = s / M[r][r];
    }
}

}  // namespace

extern "C" {

double lj_cubic_lattice(int n, double density, int lattice, double* pos)
{
    if (n < 1 || !(density > 0) || lattice < 0 || lattice > 2 || !pos) return -1.0;
    const UnitCell& uc = kCells[lattice];
    const int nonempty = (int)std::ceil((double)n / (double)uc.m);
    const int cps = (int)std::ceil(std::cbrt((double)nonempty));
    const double proto = (double)n / (double)(cps * cps * cps);
    const double scale = std::cbrt(proto / density);
    for (int idx = 0; idx < n; ++idx) {
        // index = ((x * cps + y) * cps + z) * m + s
        const int s = idx % uc.m, q = idx / uc.m;
        const int z = q % cps, y = (q / cps) % cps, x = q / (cps * cps);
        pos[3 * (size_t)idx] = ((double)x + uc.site[s][0]) * scale;
        pos[3 * (size_t)idx + 1] = ((double)y + uc.site[s][1]) * scale;
        pos[3 * (size_t)idx + 2] = ((double)z + uc.site[s][2]) * scale;
    }
    return (double)cps * scale;
}

double lj_initial_condition(int n, double density, double temperature, unsigned int seed,
                            int lattice, double* pos, double* vel)
{
    const double L = lj_cubic_lattice(n, density, lattice, pos);
    if (!(L > 0) || !vel || !(temperature > 0)) return -1.0;
    std::mt19937 gen(seed);
    std::normal_distribution<> mb(0, std::sqrt(temperature));
    for (size_t k = 0; k < 3 * (size_t)n; ++k) vel[k] = mb(gen);

    // centre of mass before any velocity change (engine/initial_condition.cpp:96)
    double com[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) com[a] += pos[3 * (size_t)i + a];
    for (int a = 0; a < 3; ++a) com[a] /= (double)n;

    // zero linear momentum: v += (0 - P) / N
    double P[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) P[a] += vel[3 * (size_t)i + a];
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) vel[3 * (size_t)i + a] += (0.0 - P[a]) / (double)n;

    // zero angular momentum about the centre of mass: I dw = -L, v += dw x r
    double Lm[3] = {0, 0, 0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < n; ++i) {
        const double r[3] = {pos[3 * (size_t)i] - com[0], pos[3 * (size_t)i + 1] - com[1],
                             pos[3 * (size_t)i + 2] - com[2]};
        const double* v = vel + 3 * (size_t)i;
        Lm[0] += r[1] * v[2] - r[2] * v[1];
        Lm[1] += r[2] * v[0] - r[0] * v[2];
        Lm[2] += r[0] * v[1] - r[1] * v[0];
        const double r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) I[a][b] += (a == b ? r2 : 0.0) - r[a] * r[b];
    }
    const double rhs[3] = {0.0 - Lm[0], 0.0 - Lm[1], 0.0 - Lm[2]};
    double dw[3];
    solve3(I, rhs, dw);
    for (int i = 0; i < n; ++i) {
        const double r[3] = {pos[3 * (size_t)i] - com[0], pos[3 * (size_t)i + 1] - com[1],
                             pos[3 * (size_t)i + 2] - com[2]};
        double* v = vel + 3 * (size_t)i;
        v[0] += dw[1] * r[2] - dw[2] * r[1];
        v[1] += dw[2] * r[0] - dw[0] * r[2];
        v[2] += dw[0] * r[1] - dw[1] * r[0];
    }

    // set_temperature(T): v *= sqrt(T / T_inst), T_inst = (2/3) KE / N
    double sv = 0;
    for (size_t k = 0; k < 3 * (size_t)n; k += 3)
        sv += (vel[k] * vel[k] + vel[k + 1] * vel[k + 1]) + vel[k + 2] * vel[k + 2];
    const double tinst = (2. / 3.) * (0.5 * sv) / (double)n;
    const double scale = std::sqrt(temperature / tinst);
    for (size_t k = 0; k < 3 * (size_t)n; ++k) vel[k] *= scale;
    return L;
}

int lj_sweep_sizes(unsigned int seed, int count, int lo, int hi, int* out)
{
    if (count < 0 || hi < lo || !out) return -1;
    std::mt19937 gen(seed);
    const unsigned long long range = (unsigned long long)(hi - lo) + 1ull;
    const unsigned long long limit = (0x100000000ull / range) * range;   // accept u < limit
    for (int k = 0; k < count; ++k) {
        unsigned long long u;
        do {
            u = gen();
        } while (u >= limit);
        out[k] = lo + (int)(u % range);
    }
    return 0;
}

}  // extern "C"
