// ljref -- CPU restatement of bniehoff/lennard-jonesium's NVE step loop.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP engine is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker (or, for the timing
// twin built with the reference's -O3 -march=native -ffast-math flags, as
// the reported CPU baseline).  Nothing in the product links or calls it.
//
// Every routine follows one reference routine; the citation is given above
// each one (paths relative to the reference repository root).  The
// reference stores positions, velocities, displacements and forces as
// Eigen 4xN column-major matrices with the 4th row zero
// (src/cpp/lennardjonesium/physics/system_state.hpp:85-90).  Here the same
// data are N x 3 row-major double arrays (x, y, z per particle); the 4th
// component is always zero in the reference and never contributes.
//
// Evaluation order is fixed and documented so that the parity build
// (-O2 -ffp-contract=off) is bit-reproducible:
//   squared norm      (x*x + y*y) + z*z
//   separation        (r_first - r_second) - image*L     (image in {-1,0,1})
//   cell key          floor(r * (nc / L))                (quotient first)
//   wrap              r - floor(r / L) * L
//   half kick         v + (0.5 * F) * dt
//   drift             r + v * dt ;  d + v * dt          (same increment)
//
// Parity pins (what checks each function; the tests live under tests/):
//   lj_pair               closed forms V(rc)=W(rc)=0, V(2^(1/6)) ~ -1, sign
//                         table, zero beyond rc, F = W sep / r^2
//                         (tests/cpp/lennardjonesium/physics/test_lennard_jones_force.cpp)
//   const_pair            mock ConstantShortRangeForce known answers
//                         (tests/cpp/lennardjonesium/engine/test_short_range_force_calculation.cpp)
//   pairs_naive/pairs_cell 10-particle known pair set
//                         (tests/cpp/lennardjonesium/engine/test_particle_pair_filter.cpp:59-125)
//                         and naive == cell on random inputs (brute force)
//   cell_shape/adjacent   shapes and 26-neighbour coverage
//                         (tests/cpp/lennardjonesium/tools/test_cell_list_array.cpp)
//   wrap                  1.7/3.2/-4.3 -> 0.2 in box 1.5
//                         (tests/cpp/lennardjonesium/engine/test_periodic_boundary_condition.cpp)
//   step                  dt=1 inertial / gravity / box-3 known answers
//                         (tests/cpp/lennardjonesium/engine/test_velocity_verlet_integrator.cpp)
//   cubic_lattice         FCC/BCC/SC site lists, density
//                         (tests/cpp/lennardjonesium/tools/test_cubic_lattice.cpp)
//   initial_condition     zero P, zero L, exact T; trajectory fixture
//                         notebooks/quick_single_run/data/thermodynamics.csv
//   thermo                same fixture (KE, PE, E, W, T, MSD per step)
//   rdf                   FCC shell counts and brute force on tiny N (no
//                         reference counterpart)

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include <algorithm>

namespace {

// ---------------------------------------------------------------------------
// Pair laws
// ---------------------------------------------------------------------------

// physics/lennard_jones_force.cpp:68-73 (spline coefficients).  The (eps,
// sigma) generalisation V = eps * v(r / sigma) keeps rc / sigma fixed, so
// alpha and beta depend only on the reduced cutoff; eps = sigma = 1 gives
// the reference's arithmetic exactly.
struct LJ {
    double rc;        // absolute cutoff (rc_reduced * sigma)
    double rc2;       // rc * rc (reference square_cutoff_distance_)
    double rc2_red;   // reduced squared cutoff (rc / sigma)^2
    double alpha, beta;
    double eps, sigma2;
};

LJ make_lj(double rc_reduced, double eps, double sigma)
{
    LJ p;
    p.rc = rc_reduced * sigma;
    p.rc2 = p.rc * p.rc;
    p.rc2_red = rc_reduced * rc_reduced;
    double rcm6 = 1 / (p.rc2_red * p.rc2_red * p.rc2_red);
    p.alpha = -4.0 * rcm6 * (rcm6 - 1.0);
    p.beta = 12.0 * rcm6 * (2.0 * rcm6 - 1.0);
    p.eps = eps;
    p.sigma2 = sigma * sigma;
    return p;
}

inline double sqnorm(double x, double y, double z) { return (x * x + y * y) + z * z; }

// physics/lennard_jones_force.cpp:76-128.  Returns false (zero contribution)
// outside the cutoff.  With sigma == 1 the reduced r^2 is r^2 itself, and
// with eps == 1 the scaling is exact, so this is the reference arithmetic.
inline bool lj_pair(const LJ& p, const double sep[3], double f[3], double& v, double& w)
{
    double r2 = sqnorm(sep[0], sep[1], sep[2]);
    if (!(r2 < p.rc2)) { f[0] = f[1] = f[2] = 0; v = w = 0; return false; }
    double r2r = (p.sigma2 == 1.0) ? r2 : r2 / p.sigma2;
    double s = 1.0 / (r2r * r2r * r2r);
    double vv = 4.0 * s * (s - 1.0) + p.alpha + p.beta * (r2r / p.rc2_red - 1.0);
    double ww = 24.0 * s * (2.0 * s - 1.0) - 2.0 * p.beta * (r2r / p.rc2_red);
    if (p.eps != 1.0) { vv = p.eps * vv; ww = p.eps * ww; }
    v = vv; w = ww;
    f[0] = ww * sep[0] / r2; f[1] = ww * sep[1] / r2; f[2] = ww * sep[2] / r2;
    return true;
}

// tests/cpp/mock/constant_short_range_force.cpp:31-59 (mock force used by
// the reference's force-accumulation known-answer tests).
inline bool const_pair(double strength, double rc, const double sep[3], double f[3],
                       double& v, double& w)
{
    double norm = std::sqrt(sqnorm(sep[0], sep[1], sep[2]));
    if (!(norm < rc)) { f[0] = f[1] = f[2] = 0; v = w = 0; return false; }
    v = strength * (rc - norm);
    w = strength * norm;
    double n2 = norm * norm;
    f[0] = w * sep[0] / n2; f[1] = w * sep[1] / n2; f[2] = w * sep[2] / n2;
    return true;
}

// ---------------------------------------------------------------------------
// Cell grid
// ---------------------------------------------------------------------------

// tools/cell_list_array.cpp:45-51: nc_a = floor(L_a / rc)
void cell_shape(const double L[3], double rc, int nc[3])
{
    for (int a = 0; a < 3; ++a) nc[a] = (int)std::floor(L[a] / rc);
}

// engine/particle_pair_filter.cpp:144-147: key = floor(r * (nc / L)), the
// quotient nc/L is formed first.  The reference indexes out of bounds when
// the key reaches nc (r == L after wrap) or -1; the restatement clamps and
// reports how many clamps happened.
inline int cell_key_axis(double r, double q, int nc, int* clamped)
{
    int k = (int)std::floor(r * q);
    if (k < 0) { k = 0; if (clamped) ++*clamped; }
    if (k >= nc) { k = nc - 1; if (clamped) ++*clamped; }
    return k;
}

struct CellList {
    int nc[3];
    int ncell;
    std::vector<int> start;   // ncell + 1
    std::vector<int> atoms;   // N, ascending particle index inside each cell
};

// engine/particle_pair_filter.cpp:141-154 and tools/cell_list_array.cpp:63-81.
// Cells are laid out x-slowest / z-fastest, the order of the reference's
// (i, j, k) iteration over boost::multi_array extents; each list holds
// particle indices in ascending order (push_back in index order).
void build_cells(int n, const double* pos, const double L[3], double rc, CellList& cl,
                 int* keys_out = nullptr, int* clamped = nullptr)
{
    cell_shape(L, rc, cl.nc);
    cl.ncell = cl.nc[0] * cl.nc[1] * cl.nc[2];
    double q[3];
    for (int a = 0; a < 3; ++a) q[a] = (double)cl.nc[a] / L[a];
    std::vector<int> key(n);
    std::vector<int> count(cl.ncell, 0);
    for (int i = 0; i < n; ++i) {
        int k[3];
        for (int a = 0; a < 3; ++a) k[a] = cell_key_axis(pos[3 * i + a], q[a], cl.nc[a], clamped);
        if (keys_out) { keys_out[3 * i] = k[0]; keys_out[3 * i + 1] = k[1]; keys_out[3 * i + 2] = k[2]; }
        key[i] = (k[0] * cl.nc[1] + k[1]) * cl.nc[2] + k[2];
        count[key[i]]++;
    }
    cl.start.assign(cl.ncell + 1, 0);
    for (int c = 0; c < cl.ncell; ++c) cl.start[c + 1] = cl.start[c] + count[c];
    cl.atoms.assign(n, 0);
    std::vector<int> fill(cl.start.begin(), cl.start.end() - 1);
    for (int i = 0; i < n; ++i) cl.atoms[fill[key[i]]++] = i;
}

// tools/cell_list_array.cpp:108-114 (half-shell steps, leading nonzero +1)
const int kHalfSteps[13][3] = {
    {1, 1, 1}, {1, 1, 0}, {1, 1, -1}, {1, 0, 1}, {1, 0, 0}, {1, 0, -1}, {1, -1, 1},
    {1, -1, 0}, {1, -1, -1}, {0, 1, 1}, {0, 1, 0}, {0, 1, -1}, {0, 0, 1}};

// tools/cell_list_array.cpp:118-155: neighbour index and lattice image.  The
// image coordinate equals the step where the neighbour index leaves
// [0, nc), and the neighbour is mapped back by image * nc.
inline void neighbour(const int nc[3], const int c[3], const int step[3], int nb[3], int img[3])
{
    for (int a = 0; a < 3; ++a) {
        int t = c[a] + step[a];
        img[a] = (t < 0 || t >= nc[a]) ? step[a] : 0;
        nb[a] = t - img[a] * nc[a];
    }
}

// A pair visitor receives (first, second, separation, image) in the
// reference's generator order.
template <class Visit>
void for_each_pair_cell(int n, const double* pos, const double L[3], double rc, Visit&& visit,
                        CellList* keep = nullptr)
{
    CellList local;
    CellList& cl = keep ? *keep : local;
    build_cells(n, pos, L, rc, cl);
    const double rc2 = rc * rc;
    // engine/particle_pair_filter.cpp:156-171: intra-cell pairs, i < j
    for (int c = 0; c < cl.ncell; ++c) {
        for (int a = cl.start[c]; a < cl.start[c + 1]; ++a) {
            for (int b = a + 1; b < cl.start[c + 1]; ++b) {
                int i = cl.atoms[a], j = cl.atoms[b];
                double sep[3] = {pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                                 pos[3 * i + 2] - pos[3 * j + 2]};
                if (sqnorm(sep[0], sep[1], sep[2]) < rc2) {
                    const int img[3] = {0, 0, 0};
                    visit(i, j, sep, img);
                }
            }
        }
    }
    // engine/particle_pair_filter.cpp:173-191 over tools/cell_list_array.cpp:118-155
    for (int cx = 0; cx < cl.nc[0]; ++cx)
        for (int cy = 0; cy < cl.nc[1]; ++cy)
            for (int cz = 0; cz < cl.nc[2]; ++cz) {
                int c[3] = {cx, cy, cz};
                int home = (cx * cl.nc[1] + cy) * cl.nc[2] + cz;
                for (int s = 0; s < 13; ++s) {
                    int nb[3], img[3];
                    neighbour(cl.nc, c, kHalfSteps[s], nb, img);
                    int other = (nb[0] * cl.nc[1] + nb[1]) * cl.nc[2] + nb[2];
                    double sh[3] = {img[0] * L[0], img[1] * L[1], img[2] * L[2]};
                    for (int a = cl.start[home]; a < cl.start[home + 1]; ++a) {
                        int i = cl.atoms[a];
                        for (int b = cl.start[other]; b < cl.start[other + 1]; ++b) {
                            int j = cl.atoms[b];
                            double sep[3] = {(pos[3 * i] - pos[3 * j]) - sh[0],
                                             (pos[3 * i + 1] - pos[3 * j + 1]) - sh[1],
                                             (pos[3 * i + 2] - pos[3 * j + 2]) - sh[2]};
                            if (sqnorm(sep[0], sep[1], sep[2]) < rc2) visit(i, j, sep, img);
                        }
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// Force calculation: engine/force_calculation.cpp:46-71 (clear, then
// Newton-3 scatter and PE/W accumulation in generator order).
// ---------------------------------------------------------------------------
enum Law { LAW_LJ = 0, LAW_CONST = 1, LAW_NONE = 2 };

struct ForceModel {
    int law;
    LJ lj;
    double strength;   // LAW_CONST
    double rc;         // cutoff used by the pair filter
};

void compute_forces(const ForceModel& m, int n, const double* pos, const double L[3],
                    double* F, double& pe, double& vir)
{
    if (m.law == LAW_NONE) return;   // no force calculation: forces kept as they are
    std::fill(F, F + 3 * n, 0.0);    // physics/system_state.cpp:40-47 (clear_dynamics)
    pe = 0; vir = 0;
    for_each_pair_cell(n, pos, L, m.rc, [&](int i, int j, const double* sep, const int*) {
        double f[3], v, w;
        if (m.law == LAW_LJ) lj_pair(m.lj, sep, f, v, w);
        else const_pair(m.strength, m.rc, sep, f, v, w);
        for (int a = 0; a < 3; ++a) { F[3 * i + a] += f[a]; F[3 * j + a] -= f[a]; }
        pe += v;
        vir += w;
    });
}

// engine/boundary_condition.cpp:45-48: r -= floor(r / L) * L
inline double wrap(double r, double L) { return r - std::floor(r / L) * L; }

// physics/derived_properties.cpp:37-51 and physics/derived_properties.hpp:62-86
struct Thermo { double time, ke, pe, e, w, t, msd; };

Thermo measure(int n, const double* v, const double* d, double time, double pe, double w)
{
    double sv = 0, sd = 0;
    for (int i = 0; i < n; ++i) {
        sv += sqnorm(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        sd += sqnorm(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    }
    Thermo t;
    t.time = time;
    t.ke = (1. / 2.) * sv;
    t.pe = pe;
    t.e = pe + t.ke;
    t.w = w;
    t.t = (2. / 3.) * t.ke / (double)n;
    t.msd = sd / (double)n;
    return t;
}

}  // namespace

extern "C" {

// ---- pair laws -----------------------------------------------------------

void ljref_lj_spline(double rc, double* alpha, double* beta)
{
    LJ p = make_lj(rc, 1.0, 1.0);
    *alpha = p.alpha;
    *beta = p.beta;
}

// Single pair with the LJ law (eps, sigma generalisation; rc is the reduced
// cutoff).  Returns 1 inside the cutoff.
int ljref_lj_pair(const double* sep, double rc, double eps, double sigma, double* f, double* v,
                  double* w)
{
    LJ p = make_lj(rc, eps, sigma);
    return lj_pair(p, sep, f, *v, *w) ? 1 : 0;
}

// physics/forces.hpp:80-94: potential/virial/force along z at a distance
void ljref_lj_along_z(double distance, double rc, double* v, double* w, double* fz)
{
    LJ p = make_lj(rc, 1.0, 1.0);
    double sep[3] = {0, 0, distance}, f[3];
    lj_pair(p, sep, f, *v, *w);
    *fz = f[2];
}

// ---- cell grid -----------------------------------------------------------

void ljref_cell_shape(const double* L, double rc, int* nc) { cell_shape(L, rc, nc); }

// Per-particle integer cell coordinates (N x 3) plus the count of clamped
// coordinates (reference UB cases).
int ljref_cell_keys(int n, const double* pos, const double* L, double rc, int* keys)
{
    CellList cl;
    int clamped = 0;
    build_cells(n, pos, L, rc, cl, keys, &clamped);
    return clamped;
}

// Cell list as CSR: start[ncell+1], atoms[n].  Returns ncell.
int ljref_cell_list(int n, const double* pos, const double* L, double rc, int* start, int* atoms)
{
    CellList cl;
    build_cells(n, pos, L, rc, cl);
    std::copy(cl.start.begin(), cl.start.end(), start);
    std::copy(cl.atoms.begin(), cl.atoms.end(), atoms);
    return cl.ncell;
}

// tools/cell_list_array.cpp:83-157: for every cell (x-slowest) and each of
// the 13 half-shell steps, the neighbour cell's linear index and image.
// out_nb[ncell*13], out_img[ncell*13*3].
void ljref_adjacent_cells(const int* nc, int* out_nb, int* out_img)
{
    int k = 0;
    for (int cx = 0; cx < nc[0]; ++cx)
        for (int cy = 0; cy < nc[1]; ++cy)
            for (int cz = 0; cz < nc[2]; ++cz) {
                int c[3] = {cx, cy, cz};
                for (int s = 0; s < 13; ++s, ++k) {
                    int nb[3], img[3];
                    neighbour(nc, c, kHalfSteps[s], nb, img);
                    out_nb[k] = (nb[0] * nc[1] + nb[1]) * nc[2] + nb[2];
                    for (int a = 0; a < 3; ++a) out_img[3 * k + a] = img[a];
                }
            }
}

// ---- pair enumeration ----------------------------------------------------

// Cell-list filter, reference generator order.  Writes up to cap pairs as
// (first, second, imgx, imgy, imgz) into out_idx[5*cap] and separations
// into out_sep[3*cap].  Returns the total number of pairs (may exceed cap).
long long ljref_pairs_cell(int n, const double* pos, const double* L, double rc, long long cap,
                           int* out_idx, double* out_sep)
{
    long long k = 0;
    for_each_pair_cell(n, pos, L, rc, [&](int i, int j, const double* sep, const int* img) {
        if (k < cap) {
            out_idx[5 * k] = i; out_idx[5 * k + 1] = j;
            out_idx[5 * k + 2] = img[0]; out_idx[5 * k + 3] = img[1]; out_idx[5 * k + 4] = img[2];
            out_sep[3 * k] = sep[0]; out_sep[3 * k + 1] = sep[1]; out_sep[3 * k + 2] = sep[2];
        }
        ++k;
    });
    return k;
}

// engine/particle_pair_filter.cpp:67-123: all i<j and all 27 images, image
// order x-slowest; the separation is (r_i - r_j) - image * L.
long long ljref_pairs_naive(int n, const double* pos, const double* L, double rc, long long cap,
                            int* out_idx, double* out_sep)
{
    long long k = 0;
    const double rc2 = rc * rc;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            for (int ix = -1; ix <= 1; ++ix)
                for (int iy = -1; iy <= 1; ++iy)
                    for (int iz = -1; iz <= 1; ++iz) {
                        double sep[3] = {(pos[3 * i] - pos[3 * j]) - ix * L[0],
                                         (pos[3 * i + 1] - pos[3 * j + 1]) - iy * L[1],
                                         (pos[3 * i + 2] - pos[3 * j + 2]) - iz * L[2]};
                        if (sqnorm(sep[0], sep[1], sep[2]) < rc2) {
                            if (k < cap) {
                                out_idx[5 * k] = i; out_idx[5 * k + 1] = j;
                                out_idx[5 * k + 2] = ix; out_idx[5 * k + 3] = iy;
                                out_idx[5 * k + 4] = iz;
                                for (int a = 0; a < 3; ++a) out_sep[3 * k + a] = sep[a];
                            }
                            ++k;
                        }
                    }
    return k;
}

// Half-shell candidate and in-cutoff pair counts (C_half, P_half) of the
// reference's cell filter on these positions: the work units that the
// throughput roofline is priced in.
void ljref_pair_counts(int n, const double* pos, const double* L, double rc, long long* cand,
                       long long* inrange)
{
    CellList cl;
    build_cells(n, pos, L, rc, cl);
    const double rc2 = rc * rc;
    long long c = 0, p = 0;
    for (int h = 0; h < cl.ncell; ++h) {
        long long m = cl.start[h + 1] - cl.start[h];
        c += m * (m - 1) / 2;
    }
    for (int cx = 0; cx < cl.nc[0]; ++cx)
        for (int cy = 0; cy < cl.nc[1]; ++cy)
            for (int cz = 0; cz < cl.nc[2]; ++cz) {
                int cc[3] = {cx, cy, cz};
                int home = (cx * cl.nc[1] + cy) * cl.nc[2] + cz;
                for (int s = 0; s < 13; ++s) {
                    int nb[3], img[3];
                    neighbour(cl.nc, cc, kHalfSteps[s], nb, img);
                    int other = (nb[0] * cl.nc[1] + nb[1]) * cl.nc[2] + nb[2];
                    c += (long long)(cl.start[home + 1] - cl.start[home]) *
                         (cl.start[other + 1] - cl.start[other]);
                }
            }
    for_each_pair_cell(n, pos, L, rc, [&](int, int, const double*, const int*) { ++p; });
    *cand = c;
    *inrange = p;
}

// In-cutoff neighbours of every particle (each pair counted for both of its
// particles, every image separately) of the reference's cell filter.
void ljref_neighbor_counts(int n, const double* pos, const double* L, double rc, int* counts)
{
    std::fill(counts, counts + n, 0);
    for_each_pair_cell(n, pos, L, rc, [&](int i, int j, const double*, const int*) {
        ++counts[i];
        ++counts[j];
    });
}

// ---- forces --------------------------------------------------------------

// law: 0 = LJ (rc reduced, eps, sigma), 1 = constant mock (strength, rc),
// 2 = none.  F is N x 3, overwritten (except law 2).
void ljref_forces(int n, const double* pos, const double* L, int law, double rc, double eps,
                  double sigma, double strength, double* F, double* pe, double* vir)
{
    ForceModel m;
    m.law = law;
    m.lj = make_lj(rc, eps, sigma);
    m.strength = strength;
    m.rc = (law == LAW_LJ) ? m.lj.rc : rc;
    double p = 0, w = 0;
    compute_forces(m, n, pos, L, F, p, w);
    *pe = p;
    *vir = w;
}

// engine/boundary_condition.cpp:37-51
void ljref_wrap(int n, double* pos, const double* L)
{
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) pos[3 * i + a] = wrap(pos[3 * i + a], L[a]);
}

// ---- integration ---------------------------------------------------------

// engine/integrator.cpp:51-88 (VelocityVerletIntegrator), one record per
// step as control/simulation_controller.cpp:60-69 logs it (the state after
// the step).  L == NULL disables the boundary condition (the reference's
// Integrator::Builder without bounding_box).  state_time is in/out.
// rec[7*nsteps] receives {time, KE, PE, E, W, T, MSD} when non-null.
void ljref_run(int n, double* pos, double* vel, double* disp, double* F, double* pe, double* vir,
               double* state_time, const double* L, int law, double rc, double eps, double sigma,
               double strength, double dt, int nsteps, double* rec)
{
    ForceModel m;
    m.law = law;
    m.lj = make_lj(rc, eps, sigma);
    m.strength = strength;
    m.rc = (law == LAW_LJ) ? m.lj.rc : rc;
    double t = *state_time, p = *pe, w = *vir;
    for (int s = 0; s < nsteps; ++s) {
        for (int k = 0; k < 3 * n; ++k) vel[k] = vel[k] + ((1. / 2.) * F[k]) * dt;
        for (int k = 0; k < 3 * n; ++k) {
            double inc = vel[k] * dt;
            pos[k] = pos[k] + inc;
            disp[k] = disp[k] + inc;
        }
        if (L) ljref_wrap(n, pos, L);
        compute_forces(m, n, pos, L ? L : nullptr, F, p, w);
        for (int k = 0; k < 3 * n; ++k) vel[k] = vel[k] + ((1. / 2.) * F[k]) * dt;
        t = t + dt;
        if (rec) {
            Thermo th = measure(n, vel, disp, t, p, w);
            double* r = rec + 7 * (size_t)s;
            r[0] = th.time; r[1] = th.ke; r[2] = th.pe; r[3] = th.e; r[4] = th.w; r[5] = th.t;
            r[6] = th.msd;
        }
    }
    *state_time = t;
    *pe = p;
    *vir = w;
}

// physics/measurements.hpp:59-70
void ljref_measure(int n, const double* vel, const double* disp, double time, double pe,
                   double vir, double* rec)
{
    Thermo th = measure(n, vel, disp, time, pe, vir);
    rec[0] = th.time; rec[1] = th.ke; rec[2] = th.pe; rec[3] = th.e; rec[4] = th.w;
    rec[5] = th.t; rec[6] = th.msd;
}

// physics/transformations.cpp:94-117: v *= sqrt(T_target / T_inst)
void ljref_set_temperature(int n, double* vel, double target)
{
    double sv = 0;
    for (int i = 0; i < n; ++i) sv += sqnorm(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    double ke = 0.5 * sv;
    double tinst = (2. / 3.) * ke / (double)n;
    double s = std::sqrt(target / tinst);
    for (int k = 0; k < 3 * n; ++k) vel[k] *= s;
}

// ---- initial state -------------------------------------------------------

// tools/cubic_lattice.cpp:38-86 with the unit cells of
// tools/cubic_lattice.hpp:53-65 (0 = simple, 1 = body-centred, 2 = face-centred).
// Returns the cubic box side (cells_per_side * scale).
double ljref_cubic_lattice(int n, double density, int cell_type, double* pos)
{
    static const double sc[1][3] = {{0, 0, 0}};
    static const double bcc[2][3] = {{0, 0, 0}, {0.5, 0.5, 0.5}};
    static const double fcc[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
    const double(*uc)[3] = cell_type == 0 ? sc : (cell_type == 1 ? bcc : fcc);
    int m = cell_type == 0 ? 1 : (cell_type == 1 ? 2 : 4);
    int nonempty = (int)std::ceil((double)n / (double)m);
    int cps = (int)std::ceil(std::cbrt((double)nonempty));
    double proto = (double)n / (double)(cps * cps * cps);
    double scale = std::cbrt(proto / density);
    for (int idx = 0; idx < n; ++idx) {
        int s = idx % m, q = idx / m;
        int z = q % cps, q2 = q / cps;
        int y = q2 % cps, x = q2 / cps;
        pos[3 * idx] = ((double)x + uc[s][0]) * scale;
        pos[3 * idx + 1] = ((double)y + uc[s][1]) * scale;
        pos[3 * idx + 2] = ((double)z + uc[s][2]) * scale;
    }
    return (double)cps * scale;
}

// physics/transformations.cpp:35-92 applied as in engine/initial_condition.cpp:91-101:
// zero the momentum, then the angular momentum about the centre of mass
// (taken before the momentum shift), then set the temperature.
void ljref_zero_momenta(int n, const double* pos, double* vel)
{
    double com[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) com[a] += pos[3 * i + a];
    for (int a = 0; a < 3; ++a) com[a] /= (double)n;
    // set_momentum(0): v += (0 - P) / N
    double P[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) P[a] += vel[3 * i + a];
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) vel[3 * i + a] += (0.0 - P[a]) / (double)n;
    // set_angular_momentum(0, com): solve I * dw = (0 - L)
    double Lm[3] = {0, 0, 0}, I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < n; ++i) {
        double r[3] = {pos[3 * i] - com[0], pos[3 * i + 1] - com[1], pos[3 * i + 2] - com[2]};
        const double* v = vel + 3 * i;
        Lm[0] += r[1] * v[2] - r[2] * v[1];
        Lm[1] += r[2] * v[0] - r[0] * v[2];
        Lm[2] += r[0] * v[1] - r[1] * v[0];
        double r2 = sqnorm(r[0], r[1], r[2]);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) I[a][b] += (a == b ? r2 : 0.0) - r[a] * r[b];
    }
    // Gaussian elimination with partial pivoting on the 3x3 block (the 4th
    // row/column of the reference's 4x4 tensor decouples and gives dw_w = 0).
    double A[3][4];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) A[a][b] = I[a][b];
        A[a][3] = 0.0 - Lm[a];
    }
    for (int col = 0; col < 3; ++col) {
        int piv = col;
        for (int r = col + 1; r < 3; ++r)
            if (std::fabs(A[r][col]) > std::fabs(A[piv][col])) piv = r;
        if (piv != col)
            for (int k = 0; k < 4; ++k) std::swap(A[col][k], A[piv][k]);
        if (A[col][col] == 0.0) continue;
        for (int r = col + 1; r < 3; ++r) {
            double f = A[r][col] / A[col][col];
            for (int k = col; k < 4; ++k) A[r][k] -= f * A[col][k];
        }
    }
    double dw[3] = {0, 0, 0};
    for (int r = 2; r >= 0; --r) {
        if (A[r][r] == 0.0) { dw[r] = 0; continue; }
        double s = A[r][3];
        for (int k = r + 1; k < 3; ++k) s -= A[r][k] * dw[k];
        dw[r] = s / A[r][r];
    }
    for (int i = 0; i < n; ++i) {
        double r[3] = {pos[3 * i] - com[0], pos[3 * i + 1] - com[1], pos[3 * i + 2] - com[2]};
        vel[3 * i] += dw[1] * r[2] - dw[2] * r[1];
        vel[3 * i + 1] += dw[2] * r[0] - dw[0] * r[2];
        vel[3 * i + 2] += dw[0] * r[1] - dw[1] * r[0];
    }
}

// engine/initial_condition.cpp:49-104.  std::mt19937(seed) +
// std::normal_distribution<>(0, sqrt(T)) (libstdc++), filled vx0, vy0, vz0,
// vx1, ...; then zero momenta and set the temperature.  Forces are zero.
// Returns the box side.
double ljref_initial_condition(int n, double density, double temperature, unsigned int seed,
                               int cell_type, double* pos, double* vel)
{
    double L = ljref_cubic_lattice(n, density, cell_type, pos);
    std::mt19937 gen(seed);
    std::normal_distribution<> mb(0, std::sqrt(temperature));
    for (int k = 0; k < 3 * n; ++k) vel[k] = mb(gen);
    ljref_zero_momenta(n, pos, vel);
    ljref_set_temperature(n, vel, temperature);
    return L;
}

// ---- g(r) (no reference counterpart) -------------------------------------

// Brute force over all ordered pairs i != j and all images within rmax;
// hist[bins] counts ordered pairs with r in [k*dr, (k+1)*dr).
void ljref_rdf_hist(int n, const double* pos, const double* L, double rmax, int bins,
                    double* hist)
{
    // bin = floor(r * (bins / rmax)), the GPU kernel's formula
    const double inv_dr = (double)bins / rmax;
    int im[3];
    for (int a = 0; a < 3; ++a) im[a] = (int)std::ceil(rmax / L[a]) + 1;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int ix = -im[0]; ix <= im[0]; ++ix)
                for (int iy = -im[1]; iy <= im[1]; ++iy)
                    for (int iz = -im[2]; iz <= im[2]; ++iz) {
                        if (i == j && ix == 0 && iy == 0 && iz == 0) continue;
                        double dx = (pos[3 * i] - pos[3 * j]) - ix * L[0];
                        double dy = (pos[3 * i + 1] - pos[3 * j + 1]) - iy * L[1];
                        double dz = (pos[3 * i + 2] - pos[3 * j + 2]) - iz * L[2];
                        double r2 = sqnorm(dx, dy, dz);
                        if (r2 < rmax * rmax) {
                            int k = (int)(std::sqrt(r2) * inv_dr);
                            if (k < bins) hist[k] += 1.0;
                        }
                    }
}

// The histogram of ljref_rdf_hist computed over a cell grid of side >= rmax
// (three or more cells per axis, else the brute force above).  Each (j,
// image) within rmax of particle i lies in one of the 27 cells around i's
// cell, and with three or more cells per axis those 27 (cell, image)
// combinations are distinct, so every ordered pair and image is counted once
// with the brute force's arithmetic, (r_i - r_j) - image * L.  Positions must
// lie in [0, L).
void ljref_rdf_cells(int n, const double* pos, const double* L, double rmax, int bins,
                     double* hist)
{
    int nc[3];
    for (int a = 0; a < 3; ++a) nc[a] = (int)std::floor(L[a] / rmax);
    if (nc[0] < 3 || nc[1] < 3 || nc[2] < 3) {
        ljref_rdf_hist(n, pos, L, rmax, bins, hist);
        return;
    }
    CellList cl;
    build_cells(n, pos, L, rmax, cl);
    const double inv_dr = (double)bins / rmax;
    for (int cx = 0; cx < nc[0]; ++cx)
        for (int cy = 0; cy < nc[1]; ++cy)
            for (int cz = 0; cz < nc[2]; ++cz) {
                const int home = (cx * nc[1] + cy) * nc[2] + cz;
                for (int ox = -1; ox <= 1; ++ox)
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int oz = -1; oz <= 1; ++oz) {
                            const int t[3] = {cx + ox, cy + oy, cz + oz};
                            int img[3], nb[3];
                            for (int a = 0; a < 3; ++a) {
                                img[a] = t[a] < 0 ? -1 : (t[a] >= nc[a] ? 1 : 0);
                                nb[a] = t[a] - img[a] * nc[a];
                            }
                            const int other = (nb[0] * nc[1] + nb[1]) * nc[2] + nb[2];
                            for (int a = cl.start[home]; a < cl.start[home + 1]; ++a) {
                                const int i = cl.atoms[a];
                                for (int b = cl.start[other]; b < cl.start[other + 1]; ++b) {
                                    const int j = cl.atoms[b];
                                    if (i == j && img[0] == 0 && img[1] == 0 && img[2] == 0)
                                        continue;
                                    double dx = (pos[3 * i] - pos[3 * j]) - img[0] * L[0];
                                    double dy = (pos[3 * i + 1] - pos[3 * j + 1]) - img[1] * L[1];
                                    double dz = (pos[3 * i + 2] - pos[3 * j + 2]) - img[2] * L[2];
                                    double r2 = sqnorm(dx, dy, dz);
                                    if (r2 < rmax * rmax) {
                                        int k = (int)(std::sqrt(r2) * inv_dr);
                                        if (k < bins) hist[k] += 1.0;
                                    }
                                }
                            }
                        }
            }
}

}  // extern "C"
