// lj_kernels.hpp -- device code of the gfx950 NVE Lennard-Jones step engine.
//
// Layout in HBM (one engine = one or more independent periodic systems):
//   * particles live in "slots"; system k owns slots [atom_off, atom_off + N)
//     of a range padded to a multiple of kBlock, so every kBlock-wide block
//     of slots belongs to exactly one system (blk_sys[] maps block -> system);
//   * per-particle state is structure-of-arrays double x / y / z planes
//     (plane stride = total slot count): positions, velocities, displacements
//     and forces, plus the particle id held in each slot;
//   * after every cell build the slots of a system are sorted by cell
//     (cells x-slowest, z-fastest, as tools/cell_list_array.cpp:63-81 visits
//     them) and, inside a cell, by ascending particle id -- exactly the
//     membership and order of the reference's std::vector cell lists
//     (engine/particle_pair_filter.cpp:150-154);
//   * cells of all systems are numbered consecutively; each system is
//     followed by one "tail" cell holding its padding slots, so that one
//     exclusive scan over all cell counts yields every system's slot range.
//
// Arithmetic that decides integers (cell keys, the r^2 < rc^2 test) and the
// periodic wrap follows the reference's formulas in one fixed evaluation order
// (engine/particle_pair_filter.cpp:143-148 keys, :172-187 the r^2 test;
// engine/boundary_condition.cpp:36-48 the wrap): r^2 = (x^2 + y^2) + z^2,
// the order of the CPU oracle, against which cell keys and cutoff decisions
// are bit-exact.  The reference binary itself was built with -ffast-math and
// Eigen's packet reductions, whose summation order is not fixed by its
// source, so parity with the reference's own outputs is pinned by tolerances
// on its recorded trajectory, not bit for bit.  The whole file is compiled
// with -ffp-contract=off, and fused multiply-adds appear only where they are
// written explicitly (mixed-precision force math).

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ljk {

constexpr int kBlock = 256;        // slots per block in particle-space kernels
constexpr int kScanTile = 1024;    // cells per block in the scan kernels
constexpr int kCellsPerBlockMixed = 4;  // default home cells (one wave64 each) per block, mixed tiled kernel
constexpr int kCellsPerBlockFp64 = 1;   // FP64 tiled kernel: single-wave blocks retire independently
constexpr int kWavesPerCellSmall = 4;   // tiled kernels for small systems: waves per home cell
constexpr int kHistBins = 1024;    // LDS histogram window of the cell count (cells)
constexpr int kMaxParts = 64;      // first-level parts per system of the record sums
constexpr int kPartSums = 1024;    // partial sums per first-level part (below kMaxParts parts)

enum Law { LAW_LJ = 0, LAW_CONSTANT = 1, LAW_NONE = 2 };
enum Prec { FP64 = 0, MIXED = 1 };

// Per-system constants (device resident, read through uniform loads).
struct SysDev {
    int atom_off;      // first slot
    int n;             // particle count
    int cell_off;      // first cell (tail cell at cell_off + ncell)
    int ncell;         // real cells nc0 * nc1 * nc2
    int nc[3];
    int law;
    int periodic;
    int blk_first;     // first particle-space block
    int nblk;          // blocks of this system
    int cblk_first;    // first cell-space block (the tiled kernel's cells per block each)
    int ncblk;         // cell-space blocks of this system
    int active;        // 0: frozen (no integration, forces or records; cells still rebuilt)
    int tiled;         // pair forces from the tiled kernel with one wave (1) or several waves (2) per
                       // home cell, or from the per-particle kernel (0)
    int sort_k;        // cells per wave in the sort (about 44 particles per wave)
    int sblk_first;    // first sort block (kBlock / 64 waves of sort_k cells each)
    double L[3];       // box
    double q[3];       // nc / L (the reference forms this quotient first)
    double a[3];       // L / nc (cell side, used for cell-local coordinates)
    double rc2;        // absolute filter cutoff squared
    double rc2_red;    // LJ reduced cutoff squared
    double inv_rc2s;   // 1 / (sigma^2 rc2_red): (r / rc)^2 = r^2 * inv_rc2s
    double alpha, beta;
    double eps, sigma2;
    double strength, rc;   // constant law
    double dt;
    // mixed-precision forms of the above (filled by the host)
    float rc2f;        // (float) rc2
    float inv_s2f;     // (float) (1 / sigma2)
    float cwf;         // (float) (-2 beta inv_rc2s)
    float cw24f;       // (float) (-2 beta inv_rc2s / 24)
    // image shift of a neighbour cell, per axis: shf[a][3 (o + 1) + (i + 1)] =
    // (float) ((i nc - o) a - i L) for step o and lattice image i in {-1, 0, 1}
    float shf[3][9];
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// 1 / x for 0 < x < ~1e300 from the hardware reciprocal refined by two
// Newton steps (error within an ulp or two; no IEEE special cases needed).
__device__ __forceinline__ double rcp_f64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    return __builtin_fma(y, e, y);
}

// One DPP lane move of a double (two 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Wave-wide sum in a fixed order: in-row butterflies over DPP (quad swaps,
// half-row and row mirrors), then the four row sums.  Uniform result.
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_f64<0xB1>(v);    // quad_perm 1,0,3,2
    v += dpp_f64<0x4E>(v);    // quad_perm 2,3,0,1
    v += dpp_f64<0x141>(v);   // row_half_mirror
    v += dpp_f64<0x140>(v);   // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Wave-wide minimum (MAX = false) or maximum of an int: four in-row DPP steps
// (quad swaps, half-row and row mirrors), then the two cross-row exchanges.
template <bool MAX>
__device__ __forceinline__ int wave_extreme(int v)
{
    auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));    // quad_perm 1,0,3,2
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));    // quad_perm 2,3,0,1
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));   // row_half_mirror
    v = op(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true));   // row_mirror
    v = op(v, __shfl_xor(v, 16, 64));
    v = op(v, __shfl_xor(v, 32, 64));
    return v;
}

// Sums four per-thread doubles over an NT-thread block; thread 0 gets them.
template <int NT = kBlock>
__device__ __forceinline__ void block_sum4(double v[4], double out[4])
{
    constexpr int NW = NT / 64;
    __shared__ double red[4][NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double s = wave_sum_dpp(v[k]);
        if (lane == 0) red[k][wid] = s;
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = wave_sum_dpp(lane < NW ? red[k][lane] : 0.0);
    }
}

// Integer cell coordinate along one axis: floor(r * (nc / L)), clamped into
// [0, nc) (the reference indexes out of range for r == L or a rounded-up
// quotient; clamping keeps such particles in the boundary cell).
__device__ __forceinline__ int cell_axis(double r, double q, int nc)
{
    int k = (int)floor(r * q);
    k = k < 0 ? 0 : k;
    return k >= nc ? nc - 1 : k;
}

// Integer coordinates of global cell c from the per-cell table.
__device__ __forceinline__ void cell_xyz_of(const int* tab, int c, int& cx, int& cy, int& cz)
{
    const int v = tab[c];
    cx = v & 1023;
    cy = (v >> 10) & 1023;
    cz = v >> 20;
}

// Neighbour cell along one axis for step o in {-1, 0, 1}: the reference's
// lattice-image rule (tools/cell_list_array.cpp:142-151).
__device__ __forceinline__ void nb_axis(int c, int o, int nc, int& nb, int& img)
{
    int t = c + o;
    img = (t < 0) ? -1 : (t >= nc ? 1 : 0);
    nb = t - img * nc;
}

// ---------------------------------------------------------------------------
// kernel parameter blocks
// ---------------------------------------------------------------------------

struct Planes {       // three coordinate planes of one per-slot vector field
    double* x;
    double* y;
    double* z;
};

struct StepArgs {
    const SysDev* sys;
    const int* blk_sys;
    const int* cblk_sys;                // cell-space block -> system
    const int* sblk_sys;                // sort block -> system
    Planes pos, vel, disp, force;       // current buffers
    Planes pos_n, vel_n, disp_n;        // next buffers (sort destination)
    int* id;                            // current ids
    int* id_n;                          // next ids
    int* count;                         // per cell, incl. tail cells
    const int* count_init;             // tail cells hold their padding count
    int* start;                         // exclusive scan, ncell_total + 1
    int* tile_sum;
    int* members;                       // [cell][cell_cap]: slots of a cell's first arrivals
    int2* ovf;                          // (slot, cell) of the later arrivals (nslot entries)
    int* novf;                          // [0] overflow entries of this pass, [1] of the last scanned pass
    int cell_cap;                       // member records per cell
    int* scell;                         // cell of each sorted slot
    float4* lpos;                       // mixed: cell-local float coordinates
    double* partial;                    // per particle-space block: sum v^2, PE, W, sum d^2
    double* partial_c;                  // the same per cell (tiled pair kernels), indexed like cells
    int skip_tiled;                     // per-particle pair kernel: leave tiled systems alone
    const int* cblk_list;               // tiled pair kernels: cell-space blocks to run (null: all)
    const int* blk_list;                // per-particle pair kernel: blocks to run (null: all)
    int* pending;                       // per system: a step's block sums await the finalize
    const int* cell_xyz;                // per cell: x | y << 10 | z << 20 (cell coordinates)
    int* ncount;                        // optional: in-cutoff neighbours per slot
    int ncell_total;
    int nslot;
    const int* lblk_list;               // several waves per cell (small systems): blocks to run (null: all)
    int cpb;                            // cells per cell-space block
    // End of each cell's slots: start + 1 in the packed layout (start[c + 1]),
    // its own array for resident cells (section 4b).
    const int* cend;
    // Resident cells (section 4b): every real cell owns kResCap slots from
    // rc_base[c]; a step's pair kernel moves each particle into the next
    // buffer set itself and hands the few that change cell to the fix-up.
    float4* lpos_n;                     // next buffers' cell-local float coordinates
    const int* rc_base;                 // first slot of each cell
    int* rc_count;                      // members per cell (tail cells: their padding count)
    int* rc_end;                        // rc_base + rc_count
    int* rc_dirty;                      // per cell: to be fixed up this step
    int* rc_narr;                       // per cell: arrivals this step
    struct ResMover* rc_arr;            // [cell][kResArr]: the arrivals' integrated states
    int* rc_flag;                       // nonzero: a capacity was exceeded (the call is redone packed)
    const int* cell_sys;                // system of each cell
    int rc_par;                         // this step's parity (record parts and pending flags)
};

// Resident cells: slots per cell (a cell with more members makes the call
// fall back to the packed layout) and arrivals per cell and step.
constexpr int kResCap = 48;
constexpr int kResArr = 8;
constexpr int kResFixCells = 1;   // cells per fix-up wave (1: 5.6 us per step, 2: 6.5, 4: 8.7 on the liquid)
static_assert(kResCap + kResArr <= 64, "a cell's members and arrivals share one wave");

// A particle that left its cell during a resident step: its integrated state
// and its new cell.
struct ResMover {
    double x, y, z, vx, vy, vz, dx, dy, dz;
    int id, cell;
};

// ---------------------------------------------------------------------------
// per-system reduction parameters (the reduction itself is section 7)
//   mode 0: after a step; 1: after a force-only calculation; 2: measure only
// ---------------------------------------------------------------------------
struct FinalArgs {
    const SysDev* sys;
    const double* partial;    // particle-space block sums
    const double* partial_c;  // per-cell sums (systems with tiled pair forces)
    double* time;      // per system
    double* pe;        // per system (last force calculation)
    double* vir;
    int* step;         // per system record counter
    double* records;   // [cap][nsys][7]
    double* measure;   // [nsys][7] (mode 2)
    int nsys;
    int cap;
    int mode;
    int cell_blocks;   // 1: a system with tiled pair forces has cell-space sums; 0: all particle-space
    int* pending;      // per system: set by a step's pair kernel, cleared here
    double* partial2;  // [nsys][nf] first-level sums (see finalize_part)
    int nf;            // first-level part slots per system: the most any system uses
};

template <int NT>
__device__ void finalize_range(const FinalArgs& f, int k, const double* part, int nb);
__device__ void finalize_part(const FinalArgs& f, int k, int r);
__device__ int parts_of(const FinalArgs& f, int k);

// ---------------------------------------------------------------------------
// 1. kick + drift + wrap + cell key + cell count
//    engine/integrator.cpp:69-78 and engine/particle_pair_filter.cpp:141-154
// ---------------------------------------------------------------------------
// r - floor(r / L) L (engine/boundary_condition.cpp:45-48).  Inside the box,
// 0 < r < L (1 - 2^-50), the rounded quotient is below 1, the floor is 0 and
// the result is r itself bit for bit, so the division is skipped there (a
// wave divides only when one of its particles left the box).
__device__ __forceinline__ double wrap_axis(double r, double L)
{
    if (r > 0.0 && r < L - L * 0x1p-50) return r;
    return r - floor(r / L) * L;
}

// First half kick, drift and periodic wrap of one particle (velocity Verlet,
// engine/integrator.cpp:69-78, engine/boundary_condition.cpp:45-48):
// v += (1/2) F dt with the stored force (zero before the first calculation),
// then r += v dt and d += v dt (the same increment).  The cell-key pass and
// the sort evaluate it identically.  KICK1 false: the velocities already
// carry the half kick (a chained step, whose previous pair kernel applied it
// with this arithmetic and the same force), and fx, fy, fz are unused.
template <bool KICK1 = true>
__device__ __forceinline__ void advance(const SysDev& S, double& x, double& y, double& z,
                                        double& vx, double& vy, double& vz, double fx, double fy,
                                        double fz, double& ix, double& iy, double& iz)
{
    const double dt = S.dt;
    if (KICK1) {
        vx = vx + ((1. / 2.) * fx) * dt;
        vy = vy + ((1. / 2.) * fy) * dt;
        vz = vz + ((1. / 2.) * fz) * dt;
    }
    ix = vx * dt; iy = vy * dt; iz = vz * dt;
    x = x + ix; y = y + iy; z = z + iz;
    if (S.periodic) {
        x = wrap_axis(x, S.L[0]);
        y = wrap_axis(y, S.L[1]);
        z = wrap_axis(z, S.L[2]);
    }
}

// The g-th arrival (0-based) of slot s in cell c: the first cell_cap of a
// cell go to its member record, later ones to the overflow list.
__device__ __forceinline__ void record_member(const StepArgs& p, int s, int c, int g)
{
    if (g < p.cell_cap) {
        p.members[(size_t)c * p.cell_cap + g] = s;
    } else {
        // at most one entry per slot and pass (the scan empties the list)
        const int o = atomicAdd(p.novf, 1);
        if (o < p.nslot) p.ovf[o] = make_int2(s, c);
    }
}

// MODE 0: keys of the current positions; 1: integrate in place (and key);
// 2: key of the integrated position only -- the sort applies the step while
// it moves the particle, so the state is read once and written once; 3: as
// 2 for velocities that already carry the half kick (no force is read).
// The first nfin blocks compute the first level of the previous step's
// record sums instead (f.nsys * f.nf parts, section 7); the scan's extra
// blocks finish them.
template <int MODE, bool KEYS>
__global__ __launch_bounds__(kBlock) void k_integrate_key(StepArgs p, FinalArgs f, int nfin)
{
    if ((int)blockIdx.x < nfin) {
        finalize_part(f, blockIdx.x / f.nf, blockIdx.x % f.nf);
        return;
    }
    const int b = blockIdx.x - nfin;
    const SysDev& S = p.sys[p.blk_sys[b]];
    const int s = b * kBlock + threadIdx.x;
    const bool active = s - S.atom_off < S.n;
    int c = -1;
    if (active) {
        double x = p.pos.x[s], y = p.pos.y[s], z = p.pos.z[s];
        if (MODE != 0 && S.active) {
            double vx = p.vel.x[s], vy = p.vel.y[s], vz = p.vel.z[s];
            double fx = 0., fy = 0., fz = 0.;
            if (MODE != 3) { fx = p.force.x[s]; fy = p.force.y[s]; fz = p.force.z[s]; }
            double ix, iy, iz;
            advance<MODE != 3>(S, x, y, z, vx, vy, vz, fx, fy, fz, ix, iy, iz);
            if (MODE == 1) {
                p.disp.x[s] = p.disp.x[s] + ix;
                p.disp.y[s] = p.disp.y[s] + iy;
                p.disp.z[s] = p.disp.z[s] + iz;
                p.vel.x[s] = vx; p.vel.y[s] = vy; p.vel.z[s] = vz;
                p.pos.x[s] = x; p.pos.y[s] = y; p.pos.z[s] = z;
            }
        }
        if (KEYS) {
            const int kx = cell_axis(x, S.q[0], S.nc[0]);
            const int ky = cell_axis(y, S.q[1], S.nc[1]);
            const int kz = cell_axis(z, S.q[2], S.nc[2]);
            c = S.cell_off + (kx * S.nc[1] + ky) * S.nc[2] + kz;
        }
    }
    if (KEYS) {
        // Slots arrive nearly cell-sorted (the previous step's order), so a
        // block's keys fall in a narrow window of cells: count them in an LDS
        // histogram (LDS atomics give each particle its rank inside the block),
        // then reserve each occupied bin's ranks with one global atomic.
        __shared__ int hist[kHistBins];
        __shared__ int win[2][kBlock / 64];
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        const int lo = wave_extreme<false>(c >= 0 ? c : 0x7fffffff);
        const int hi = wave_extreme<true>(c);
        if (lane == 0) { win[0][wid] = lo; win[1][wid] = hi; }
        __syncthreads();
        int kmin = 0x7fffffff, kmax = -1;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) {
            kmin = min(kmin, win[0][w]);
            kmax = max(kmax, win[1][w]);
        }
        if (kmax < 0) return;   // no particles in this block
        const int width = kmax - kmin + 1;
        if (width <= kHistBins) {
            for (int t = threadIdx.x; t < width; t += kBlock) hist[t] = 0;
            __syncthreads();
            const int r = (c >= 0) ? atomicAdd(&hist[c - kmin], 1) : 0;
            __syncthreads();
            for (int t = threadIdx.x; t < width; t += kBlock) {
                const int n = hist[t];
                if (n) hist[t] = atomicAdd(&p.count[kmin + t], n);
            }
            __syncthreads();
            if (c >= 0) record_member(p, s, c, hist[c - kmin] + r);
            return;
        }
        // Wide windows (slots not yet in cell order): one atomic per run of
        // equal keys in the wave reserves the run's arrival ranks, handed to
        // its members by position.
        const int prev = __shfl_up(c, 1, 64);
        const bool head = (lane == 0) || (c != prev);
        const unsigned long long heads = __ballot(head);
        const unsigned long long upto =
            (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);   // bits 0..lane
        const int h = 63 - __builtin_clzll(heads & upto);
        const unsigned long long after = heads & ~upto;
        const int end = after ? __builtin_ctzll(after) : 64;
        int base = 0;
        if (head && c >= 0) base = atomicAdd(&p.count[c], end - lane);
        base = __shfl(base, h, 64);
        if (c >= 0) record_member(p, s, c, base + (lane - h));
    }
}

// ---------------------------------------------------------------------------
// 2-3. exclusive scan of the cell counts (tile sums, then per-tile wave64
//      scans offset by the preceding tiles); counts are reset afterwards.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// The same inclusive scan from DPP row shifts and row broadcasts (no LDS).
__device__ __forceinline__ int wave_incl_scan_dpp(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 -> rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 -> rows 2, 3
    return v;
}

// Exclusive block scan of one int per thread (kBlock threads); *total gets the sum.
__device__ __forceinline__ int block_excl_scan(int v, int* total)
{
    __shared__ int wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

//    k_scan_apply's blocks past the tiles finish the previous step's record
//    sums (the second level; section 7).
__global__ __launch_bounds__(kBlock) void k_scan_tiles(StepArgs p)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    int v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < p.ncell_total) v += p.count[base + k];
    int tot;
    block_excl_scan(v, &tot);
    if (threadIdx.x == 0) p.tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(StepArgs p, FinalArgs f)
{
    const int ntile = (p.ncell_total + kScanTile - 1) / kScanTile;
    if ((int)blockIdx.x >= ntile) {
        const int k = blockIdx.x - ntile;
        finalize_range<kBlock>(f, k, f.partial2 + 4 * (size_t)k * f.nf, parts_of(f, k));
        return;
    }
    __shared__ int s_off;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the key pass's overflow count, for the sort; the next pass starts empty
        p.novf[1] = p.novf[0];
        p.novf[0] = 0;
    }
    // offset of this tile: sum of the preceding tile sums
    int acc = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kBlock) acc += p.tile_sum[t];
    int tot;
    block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) s_off = tot;
    __syncthreads();
    const int base = blockIdx.x * kScanTile + threadIdx.x * 4;
    int c[4], v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = (base + k < p.ncell_total) ? p.count[base + k] : 0;
        v += c[k];
    }
    int tile_tot;
    int ex = block_excl_scan(v, &tile_tot) + s_off;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < p.ncell_total) {
            p.start[base + k] = ex;
            p.count[base + k] = p.count_init[base + k];
            if (base + k == p.ncell_total - 1) p.start[p.ncell_total] = ex + c[k];
        }
        ex += c[k];
    }
}

// Single-block form for up to kScanOneMax * 4096 cells: 1024 threads load
// all their int4 tiles first (one memory latency), scan every tile's waves
// side by side, and one wave combines the wave sums into offsets: two block
// barriers in all.  The counts are zeroed later, by k_sort_cells.
// Blocks 1.. of the launch finish the previous step's records instead (the
// second level of the record sums, one system each; section 7).
constexpr int kScanOneThreads = 1024;
constexpr int kScanOneMax = 8;

__global__ __launch_bounds__(kScanOneThreads) void k_scan_one(StepArgs p, FinalArgs f)
{
    if (blockIdx.x > 0) {
        // at most 64 parts: all in wave 0, the same sums as a kBlock reduction
        const int k = blockIdx.x - 1;
        finalize_range<kScanOneThreads>(f, k, f.partial2 + 4 * (size_t)k * f.nf, parts_of(f, k));
        return;
    }
    constexpr int NW = kScanOneThreads / 64;
    __shared__ int wsum[kScanOneMax][NW];
    const int n = p.ncell_total;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int c[kScanOneMax][4], v[kScanOneMax], inc[kScanOneMax];
#pragma unroll
    for (int t = 0; t < kScanOneMax; ++t) {
        // one 16-byte load per tile (the buffers are 16-byte aligned); the
        // ragged end falls back to single words
        const int base = t * 4 * kScanOneThreads + 4 * threadIdx.x;
        if (base + 3 < n) {
            const int4 q = *reinterpret_cast<const int4*>(p.count + base);
            c[t][0] = q.x; c[t][1] = q.y; c[t][2] = q.z; c[t][3] = q.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) c[t][k] = (base + k < n) ? p.count[base + k] : 0;
        }
    }
#pragma unroll
    for (int t = 0; t < kScanOneMax; ++t) {
        v[t] = c[t][0] + c[t][1] + c[t][2] + c[t][3];
        inc[t] = wave_incl_scan_dpp(v[t]);
        if (lane == 63) wsum[t][wid] = inc[t];
    }
    __syncthreads();
    if (wid == 0) {
        // lane w < NW: exclusive offsets of wave w in every tile, plus the
        // running total of the preceding tiles
        int carry = 0;
#pragma unroll
        for (int t = 0; t < kScanOneMax; ++t) {
            const int x = lane < NW ? wsum[t][lane] : 0;
            const int xi = wave_incl_scan_dpp(x);
            const int tile_tot = __builtin_amdgcn_readlane(xi, NW - 1);
            if (lane < NW) wsum[t][lane] = carry + xi - x;
            carry += tile_tot;
        }
        if (lane == 0) {
            p.start[n] = carry;
            // the key pass's overflow count, for the sort; the next pass starts empty
            p.novf[1] = p.novf[0];
            p.novf[0] = 0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kScanOneMax; ++t) {
        const int ex = wsum[t][wid] + inc[t] - v[t];
        const int base = t * 4 * kScanOneThreads + 4 * threadIdx.x;
        const int4 q = make_int4(ex, ex + c[t][0], ex + c[t][0] + c[t][1],
                                 ex + c[t][0] + c[t][1] + c[t][2]);
        if (base + 3 < n) {
            *reinterpret_cast<int4*>(p.start + base) = q;
        } else {
            const int e4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (base + k < n) p.start[base + k] = e4[k];
        }
    }
}

// ---------------------------------------------------------------------------
// 4. sort: one wave per cell orders the cell's members by particle id and
//    moves each into its sorted slot start[c] + rank (next buffers), with the
//    first half kick, drift and wrap when integrating; mixed precision also
//    writes cell-local float coordinates.  The sorted order is a function of
//    the positions alone, whatever order the key pass's atomics took.
// ---------------------------------------------------------------------------
template <bool MIXED_, int INTEGRATE>
__device__ __forceinline__ void move_particle(const StepArgs& p, const SysDev& S, int c, int cx,
                                              int cy, int cz, int s, int me, int dst)
{
    double x = p.pos.x[s], y = p.pos.y[s], z = p.pos.z[s];
    double vx = p.vel.x[s], vy = p.vel.y[s], vz = p.vel.z[s];
    double dx = p.disp.x[s], dy = p.disp.y[s], dz = p.disp.z[s];
    if (INTEGRATE && S.active) {
        double fx = 0., fy = 0., fz = 0.;
        if (INTEGRATE == 1) { fx = p.force.x[s]; fy = p.force.y[s]; fz = p.force.z[s]; }
        double ix, iy, iz;
        advance<INTEGRATE == 1>(S, x, y, z, vx, vy, vz, fx, fy, fz, ix, iy, iz);
        dx = dx + ix; dy = dy + iy; dz = dz + iz;
    }
    float4 lp = make_float4(0.f, 0.f, 0.f, 0.f);
    if (MIXED_)
        lp = make_float4((float)(x - cx * S.a[0]), (float)(y - cy * S.a[1]),
                         (float)(z - cz * S.a[2]), 0.0f);
    p.pos_n.x[dst] = x; p.pos_n.y[dst] = y; p.pos_n.z[dst] = z;
    p.vel_n.x[dst] = vx; p.vel_n.y[dst] = vy; p.vel_n.z[dst] = vz;
    p.disp_n.x[dst] = dx; p.disp_n.y[dst] = dy; p.disp_n.z[dst] = dz;
    p.id_n[dst] = me;
    p.scell[dst] = c;
    if (MIXED_) p.lpos[dst] = lp;
}

// One cell (local index lc) of n members from slot c0, one wave: the path for
// cells that do not fit a shared wave.  Lane j holds the j-th arrival; cells
// with more members than records take the rest from the overflow list.
template <bool MIXED_, int INTEGRATE>
__device__ __forceinline__ void sort_one_cell(const StepArgs& p, const SysDev& S, int lc, int c0,
                                              int n)
{
    const int lane = threadIdx.x & 63;
    const int c = S.cell_off + lc;
    const int cz = lc % S.nc[2], cxy = lc / S.nc[2];
    const int cy = cxy % S.nc[1], cx = cxy / S.nc[1];
    const int cap = p.cell_cap;
    const int* mem = p.members + (size_t)c * cap;
    if (n <= cap) {
        const bool act = lane < n;
        const int s = mem[act ? lane : 0];
        const int me = act ? p.id[s] : 0x7fffffff;
        int rank = 0;
        for (int j = 0; j < n; ++j) rank += (__builtin_amdgcn_readlane(me, j) < me) ? 1 : 0;
        if (act) move_particle<MIXED_, INTEGRATE>(p, S, c, cx, cy, cz, s, me, c0 + rank);
        return;
    }
    // more members than records: the others are in the overflow list of the
    // pass (unordered), searched once per 64 entries for each group of
    // members whose ranks are wanted
    const int novf = __builtin_amdgcn_readfirstlane(min(p.novf[1], p.nslot));
    const bool actA = lane < cap;
    const int sA = mem[actA ? lane : 0];
    const int idA = actA ? p.id[sA] : 0x7fffffff;
    auto rank_of = [&](int v) {   // members of c with an id below v
        int r = 0;
        for (int j = 0; j < cap; ++j) r += (__builtin_amdgcn_readlane(idA, j) < v) ? 1 : 0;
        for (int w0 = 0; w0 < novf; w0 += 64) {
            const int o = w0 + lane;
            int idB = 0x7fffffff;
            if (o < novf) {
                const int2 en = p.ovf[o];
                if (en.y == c) idB = p.id[en.x];
            }
            const int m = min(64, novf - w0);
            for (int j = 0; j < m; ++j) r += (__builtin_amdgcn_readlane(idB, j) < v) ? 1 : 0;
        }
        return r;
    };
    const int rA = rank_of(idA);
    if (actA) move_particle<MIXED_, INTEGRATE>(p, S, c, cx, cy, cz, sA, idA, c0 + rA);
    for (int w0 = 0; w0 < novf; w0 += 64) {
        const int o = w0 + lane;
        int sB = 0, idB = 0x7fffffff;
        bool mine = false;
        if (o < novf) {
            const int2 en = p.ovf[o];
            if (en.y == c) { mine = true; sB = en.x; idB = p.id[sB]; }
        }
        if (__ballot(mine) == 0ull) continue;
        const int rB = rank_of(idB);
        if (mine) move_particle<MIXED_, INTEGRATE>(p, S, c, cx, cy, cz, sB, idB, c0 + rB);
    }
}

// Cells k0 .. k1 - 1 of a wave's group (their members fit 64 lanes and every
// one is recorded): lane l holds member l - off(k) of cell k (off: the cell's
// start relative to cell k0's), and ranks by id are counted among the lanes
// of the same cell.  st: lane k holds the start of the group's cell k.
template <bool MIXED_, int INTEGRATE>
__device__ __forceinline__ void sort_group(const StepArgs& p, const SysDev& S, int cf, int st,
                                           int k0, int k1, int maxcnt)
{
    const int lane = threadIdx.x & 63;
    const int g0 = __builtin_amdgcn_readlane(st, k0);
    const int total = __builtin_amdgcn_readlane(st, k1) - g0;
    if (total == 0) return;
    // my cell kk and its members' lanes [jb, je)
    int kk = k0, jb = 0, je = total;
    for (int k = k0; k < k1; ++k) {
        const int o = __builtin_amdgcn_readlane(st, k) - g0;
        if (o <= lane) { kk = k; jb = o; } else { je = min(je, o); }
    }
    const bool act = lane < total;
    const int c = cf + kk;
    const int rec = p.members[(size_t)c * p.cell_cap + (act ? lane - jb : 0)];
    const int s = act ? rec : S.atom_off;   // idle lanes read a valid slot
    const int me = act ? p.id[s] : 0x7fffffff;
    double x = p.pos.x[s], y = p.pos.y[s], z = p.pos.z[s];
    double vx = p.vel.x[s], vy = p.vel.y[s], vz = p.vel.z[s];
    double dx = p.disp.x[s], dy = p.disp.y[s], dz = p.disp.z[s];
    double fx = 0., fy = 0., fz = 0.;
    if (INTEGRATE == 1) { fx = p.force.x[s]; fy = p.force.y[s]; fz = p.force.z[s]; }
    int cx = 0, cy = 0, cz = 0;
    if (MIXED_) cell_xyz_of(p.cell_xyz, c, cx, cy, cz);
    // rank: members of my cell (lanes jb .. je - 1) with a smaller id
    // (maxcnt is uniform, so every lane runs every shuffle)
    const int m = je - jb;
    int rank = 0;
    for (int t = 0; t < maxcnt; ++t) {
        const int v = __shfl(me, jb + t, 64);
        rank += (t < m && v < me) ? 1 : 0;
    }
    if (act) {
        const int dst = g0 + jb + rank;
        if (INTEGRATE && S.active) {
            double ix, iy, iz;
            advance<INTEGRATE == 1>(S, x, y, z, vx, vy, vz, fx, fy, fz, ix, iy, iz);
            dx = dx + ix; dy = dy + iy; dz = dz + iz;
        }
        float4 lp = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MIXED_)
            lp = make_float4((float)(x - cx * S.a[0]), (float)(y - cy * S.a[1]),
                             (float)(z - cz * S.a[2]), 0.0f);
        p.pos_n.x[dst] = x; p.pos_n.y[dst] = y; p.pos_n.z[dst] = z;
        p.vel_n.x[dst] = vx; p.vel_n.y[dst] = vy; p.vel_n.z[dst] = vz;
        p.disp_n.x[dst] = dx; p.disp_n.y[dst] = dy; p.disp_n.z[dst] = dz;
        p.id_n[dst] = me;
        p.scell[dst] = c;
        if (MIXED_) p.lpos[dst] = lp;
    }
}

// A wave takes sort_k consecutive cells.  Runs of consecutive cells whose
// members fit 64 lanes are sorted together (sort_group; normally the whole
// group in one pass); when a cell has more members than records, the wave
// walks the cells one by one (sort_one_cell).  INTEGRATE 0: move only; 1:
// first half kick from the stored force, drift and wrap; 2: drift and wrap
// (a chained step: the velocities already carry the half kick).
template <bool MIXED_, int INTEGRATE>
__global__ __launch_bounds__(kBlock) void k_sort_cells(StepArgs p)
{
    const int b = blockIdx.x;
    const SysDev& S = p.sys[p.sblk_sys[b]];
    const int K = S.sort_k;
    const int lane = threadIdx.x & 63;
    const int wv = (b - S.sblk_first) * (kBlock / 64) +
                   __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lc0 = wv * K;
    if (lc0 >= S.ncell) return;
    const int ng = min(K, S.ncell - lc0);    // cells of this wave (at most 63)
    const int cf = S.cell_off + lc0;
    // lane k <= ng: start of the group's cell k (k = ng: the group's end)
    const int st = p.start[cf + min(lane, ng)];
    const int total = __builtin_amdgcn_readlane(st, ng) - __builtin_amdgcn_readfirstlane(st);
    if (total == 0) return;   // the counts are zero already
    // the shuffle runs on every lane (a lane reading an inactive one gets 0)
    const int nxt = __shfl_down(st, 1, 64);
    const int cnt = (lane < ng) ? nxt - st : 0;
    const int maxcnt = __builtin_amdgcn_readfirstlane(wave_extreme<true>(cnt));
    if (maxcnt <= p.cell_cap) {
        // passes over runs of cells [k0, k1) with at most 64 members (one
        // cell has at most cell_cap <= 64, so every pass takes a cell)
        for (int k0 = 0; k0 < ng;) {
            const int base = __builtin_amdgcn_readlane(st, k0);
            int k1 = k0 + 1;
            while (k1 < ng && __builtin_amdgcn_readlane(st, k1 + 1) - base <= 64) ++k1;
            sort_group<MIXED_, INTEGRATE>(p, S, cf, st, k0, k1, maxcnt);
            k0 = k1;
        }
    } else {
        for (int k = 0; k < ng; ++k) {
            const int c0 = __builtin_amdgcn_readlane(st, k);
            const int n = __builtin_amdgcn_readlane(st, k + 1) - c0;
            if (n) sort_one_cell<MIXED_, INTEGRATE>(p, S, lc0 + k, c0, n);
        }
    }
    // ready for the next counting pass (tail cells keep theirs); stored last,
    // after every read of the system's parameters
    if (cnt > 0) p.count[cf + lane] = 0;
}

// ---------------------------------------------------------------------------
// 4b. resident cells: chained steps without a cell build.
//
//   Every real cell owns kResCap slots from rc_base[c] in the resident
//   buffers; its members occupy the first rc_count[c] of them in
//   particle-id order, exactly the order of the packed layout's cell lists.
//   A resident step is two kernels: the tiled pair kernel (RES), whose owner
//   lanes also drift and wrap their particles and hand the few that left
//   their cell to the fix-up as mover records (res_kick_drift), and the
//   fix-up, which rebuilds only the cells that lost or gained members (again
//   in id order).  Positions and cell-local coordinates have two buffer sets
//   (neighbours read the current one while owners write the next);
//   velocities, displacements and ids are updated in place.  The pair kernels therefore see the same candidates in the
//   same order as after a full cell build, and results are the same bit for
//   bit.  k_res_enter moves a first cell build into the resident slots; a
//   call's last step stores its forces in the resident slots, and the next
//   call starts with k_res_kick_drift; k_res_leave packs the state back (after
//   a scan of rc_count) when anything else needs the packed layout.
//   Capacities exceeded anywhere (a cell beyond kResCap, more than kResArr
//   arrivals in a step) set rc_flag; every kernel stays within its buffers
//   regardless, and the host then redoes the steps on the packed path from
//   the packed state the resident steps started from.
// ---------------------------------------------------------------------------

// The first cell build of a call into the resident slots: one wave per cell takes
// the key pass's member records, ranks them by id and moves each (with the
// integration of k_sort_cells) to rc_base[c] + rank.  Resets the key pass's
// counts.
template <bool MIXED_, int INTEGRATE>
__global__ __launch_bounds__(kBlock) void k_res_enter(StepArgs p)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (c >= p.ncell_total) return;
    const SysDev& S = p.sys[p.cell_sys[c]];
    const int lc = c - S.cell_off;
    if (lc >= S.ncell) return;   // tail cell: rc_count keeps its padding count
    const int n = __builtin_amdgcn_readfirstlane(p.count[c]);
    const int base = p.rc_base[c];
    int kept = n;
    if (n > min(p.cell_cap, kResCap)) {   // beyond the member records or the resident slots
        if (lane == 0) atomicOr(p.rc_flag, 1);
        kept = 0;
    }
    if (kept > 0) {
        const bool act = lane < kept;
        const int s = p.members[(size_t)c * p.cell_cap + (act ? lane : 0)];
        const int me = act ? p.id[s] : 0x7fffffff;
        int rank = 0;
        for (int j = 0; j < kept; ++j) rank += (__builtin_amdgcn_readlane(me, j) < me) ? 1 : 0;
        if (act) {
            double x = p.pos.x[s], y = p.pos.y[s], z = p.pos.z[s];
            double vx = p.vel.x[s], vy = p.vel.y[s], vz = p.vel.z[s];
            double dx = p.disp.x[s], dy = p.disp.y[s], dz = p.disp.z[s];
            if (INTEGRATE && S.active) {
                double fx = 0., fy = 0., fz = 0.;
                if (INTEGRATE == 1) { fx = p.force.x[s]; fy = p.force.y[s]; fz = p.force.z[s]; }
                double ix, iy, iz;
                advance<INTEGRATE == 1>(S, x, y, z, vx, vy, vz, fx, fy, fz, ix, iy, iz);
                dx = dx + ix; dy = dy + iy; dz = dz + iz;
            }
            const int dst = base + rank;
            p.pos_n.x[dst] = x; p.pos_n.y[dst] = y; p.pos_n.z[dst] = z;
            p.vel_n.x[dst] = vx; p.vel_n.y[dst] = vy; p.vel_n.z[dst] = vz;
            p.disp_n.x[dst] = dx; p.disp_n.y[dst] = dy; p.disp_n.z[dst] = dz;
            p.id_n[dst] = me;
            if (MIXED_) {
                int cx, cy, cz;
                cell_xyz_of(p.cell_xyz, c, cx, cy, cz);
                p.lpos_n[dst] = make_float4((float)(x - cx * S.a[0]), (float)(y - cy * S.a[1]),
                                            (float)(z - cz * S.a[2]), 0.0f);
            }
        }
    }
    if (lane == 0) {
        p.rc_count[c] = kept;
        p.rc_end[c] = base + kept;
        p.count[c] = 0;   // ready for the next key pass
    }
}

// Marks cell c for this step's fix-up (a plain store: every writer stores 1;
// no counter that all movers would share).
__device__ __forceinline__ void res_mark(const StepArgs& p, int c) { p.rc_dirty[c] = 1; }

// Rebuilds one marked cell c of the resident buffers of the next step (p.pos
// and p.lpos: the sets the pair kernel wrote): its remaining members
// (vacated slots hold id -1) and its arrivals, in id order from rc_base[c];
// clears the cell's step marks.  One wave.
template <bool MIXED_>
__device__ __forceinline__ void res_fix_cell(const StepArgs& p, int c, int k, int base, int n0,
                                             int na, int lane)
{
    const SysDev& S = p.sys[k];
    const int tot = n0 + na;   // <= kResCap + kResArr <= 64
    double x = 0., y = 0., z = 0., vx = 0., vy = 0., vz = 0., dx = 0., dy = 0., dz = 0.;
    int me = 0x7fffffff;       // no particle
    if (lane < n0) {
        const int s = base + lane;
        me = p.id[s];
        x = p.pos.x[s]; y = p.pos.y[s]; z = p.pos.z[s];
        vx = p.vel.x[s]; vy = p.vel.y[s]; vz = p.vel.z[s];
        dx = p.disp.x[s]; dy = p.disp.y[s]; dz = p.disp.z[s];
    } else if (lane < tot) {
        const ResMover& mv = p.rc_arr[(size_t)c * kResArr + (lane - n0)];
        me = mv.id;
        x = mv.x; y = mv.y; z = mv.z;
        vx = mv.vx; vy = mv.vy; vz = mv.vz;
        dx = mv.dx; dy = mv.dy; dz = mv.dz;
    }
    if (me < 0) me = 0x7fffffff;   // a vacated slot
    const bool valid = me != 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < tot; ++j) rank += (__builtin_amdgcn_readlane(me, j) < me) ? 1 : 0;
    const int nn = __popcll(__ballot(valid));
    const int keep = min(nn, kResCap);
    // every lane's loads are consumed above, before any slot is written
    if (valid && rank < kResCap) {
        const int dst = base + rank;
        p.pos.x[dst] = x; p.pos.y[dst] = y; p.pos.z[dst] = z;
        p.vel.x[dst] = vx; p.vel.y[dst] = vy; p.vel.z[dst] = vz;
        p.disp.x[dst] = dx; p.disp.y[dst] = dy; p.disp.z[dst] = dz;
        p.id[dst] = me;
        if (MIXED_) {
            int cx, cy, cz;
            cell_xyz_of(p.cell_xyz, c, cx, cy, cz);
            p.lpos[dst] = make_float4((float)(x - cx * S.a[0]), (float)(y - cy * S.a[1]),
                                      (float)(z - cz * S.a[2]), 0.0f);
        }
    }
    if (lane == 0) {
        if (nn > kResCap) atomicOr(p.rc_flag, 1);
        p.rc_count[c] = keep;
        p.rc_end[c] = base + keep;
        p.rc_narr[c] = 0;
        p.rc_dirty[c] = 0;
    }
}

// The fix-up after a resident step's pair kernel.  Blocks [0, nred): the second
// level of the previous step's record (f2: its parity's parts and pending
// flags); [nred, nred + nfin): the first level of this step's record (f1);
// the rest: waves of kResFixCells consecutive cells, each fixing its marked
// cells one after another.
template <bool MIXED_>
__global__ __launch_bounds__(kBlock) void k_res_fixup(StepArgs p, FinalArgs f1, FinalArgs f2,
                                                      int nred, int nfin)
{
    int b = blockIdx.x;
    if (b < nred) {
        finalize_range<kBlock>(f2, b, f2.partial2 + 4 * (size_t)b * f2.nf, parts_of(f2, b));
        return;
    }
    b -= nred;
    if (b < nfin) {
        finalize_part(f1, b / f1.nf, b % f1.nf);
        return;
    }
    b -= nfin;
    const int lane = threadIdx.x & 63;
    const int c0 = (b * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * kResFixCells;
    // lane t < kResFixCells: cell c0 + t's mark, members, arrivals, system and
    // first slot, all in one round trip
    const bool mine = lane < kResFixCells && c0 + lane < p.ncell_total;
    const int cl = mine ? c0 + lane : 0;
    const int dirty = mine ? p.rc_dirty[cl] : 0;
    const int n0 = p.rc_count[cl], nar = p.rc_narr[cl], ks = p.cell_sys[cl], rb = p.rc_base[cl];
    unsigned long long marked = __ballot(dirty != 0);
    while (marked) {
        const int t = __builtin_ctzll(marked);
        marked &= marked - 1;
        res_fix_cell<MIXED_>(p, c0 + t, __builtin_amdgcn_readlane(ks, t),
                             __builtin_amdgcn_readlane(rb, t), __builtin_amdgcn_readlane(n0, t),
                             min(__builtin_amdgcn_readlane(nar, t), kResArr), lane);
    }
}

// Packs resident set p.pos etc. into the packed layout of set p.pos_n etc.:
// cell c's members go to start[c] + rank (start: the scan of rc_count), with
// their cell-local coordinates and cells, as a sort leaves them.
// FORCE: also the resident forces p.force into fdst (a call that ended on
// resident cells stored its last step's forces there).
template <bool MIXED_, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_res_leave(StepArgs p, Planes fdst)
{
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (c >= p.ncell_total) return;
    const SysDev& S = p.sys[p.cell_sys[c]];
    if (c - S.cell_off >= S.ncell) return;   // tail cell: padding slots
    const int n = p.rc_count[c];
    if (lane >= n) return;
    const int src = p.rc_base[c] + lane, dst = p.start[c] + lane;
    p.pos_n.x[dst] = p.pos.x[src]; p.pos_n.y[dst] = p.pos.y[src]; p.pos_n.z[dst] = p.pos.z[src];
    p.vel_n.x[dst] = p.vel.x[src]; p.vel_n.y[dst] = p.vel.y[src]; p.vel_n.z[dst] = p.vel.z[src];
    p.disp_n.x[dst] = p.disp.x[src]; p.disp_n.y[dst] = p.disp.y[src];
    p.disp_n.z[dst] = p.disp.z[src];
    p.id_n[dst] = p.id[src];
    p.scell[dst] = c;
    if (MIXED_) p.lpos_n[dst] = p.lpos[src];
    if (FORCE) {
        fdst.x[dst] = p.force.x[src]; fdst.y[dst] = p.force.y[src];
        fdst.z[dst] = p.force.z[src];
    }
}

// ---------------------------------------------------------------------------
// 6. pair forces (full shell, one thread per sorted particle), fused with the
//    second half kick and the per-block sums of |v|^2, PE, W, |d|^2.  The
//    pair kernels' KICK: 0 forces only; 1 second half kick (the force is
//    stored); 2 also the next step's first half kick (a chained step,
//    kick_store; the force is not stored).
//    physics/lennard_jones_force.cpp:76-128, engine/force_calculation.cpp:46-71,
//    engine/integrator.cpp:81-82, physics/derived_properties.cpp:37-51.
// ---------------------------------------------------------------------------

// Candidates whose positions the one-thread-per-particle walks load together
// (one memory round trip per batch): mixed, FP64.  Measured on the dilute
// sweep replicas against one candidate ahead: 4 (mixed) and 2 (FP64) are
// 12% and 4% faster; 16 and 8 lose the gain to register pressure.
constexpr int kPairBatch = 4;
constexpr int kPairBatch64 = 2;

// Visits every in-cutoff neighbour j (with its image) of sorted slot s in
// double precision, with the reference pair filter's formula in the oracle's
// fixed order (engine/particle_pair_filter.cpp:172-187): sep = (r_i - r_j) -
// image * L, r^2 = (x^2 + y^2) + z^2, keep if r^2 < rc^2.
template <class Visit>
__device__ __forceinline__ void for_each_neighbour_fp64(const SysDev& S, const StepArgs& p,
                                                        int s, Visit&& visit)
{
    int cx, cy, cz;
    cell_xyz_of(p.cell_xyz, p.scell[s], cx, cy, cz);
    const double xi = p.pos.x[s], yi = p.pos.y[s], zi = p.pos.z[s];
    const double rc2 = S.rc2;
    for (int ox = -1; ox <= 1; ++ox) {
        int nbx, ix;
        nb_axis(cx, ox, S.nc[0], nbx, ix);
        const double shx = ix * S.L[0];
        for (int oy = -1; oy <= 1; ++oy) {
            int nby, iy;
            nb_axis(cy, oy, S.nc[1], nby, iy);
            const double shy = iy * S.L[1];
            // the row's three cells (oz = -1, 0, 1) walked as one run of nt
            // candidates in the same order as three separate loops, the
            // positions loaded kPairBatch64 at a time
            int jb[3], je[3], izs[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                int nbz;
                nb_axis(cz, k - 1, S.nc[2], nbz, izs[k]);
                const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
                jb[k] = p.start[nb];
                je[k] = p.start[nb + 1];
            }
            const int n0 = je[0] - jb[0];
            const int n01 = n0 + (je[1] - jb[1]);
            const int nt = n01 + (je[2] - jb[2]);
            auto cand = [&](int t) {
                return t < n0 ? jb[0] + t : (t < n01 ? jb[1] + (t - n0) : jb[2] + (t - n01));
            };
            const int last = p.nslot - 1;
            for (int t0 = 0; t0 < nt; t0 += kPairBatch64) {
                int jbuf[kPairBatch64];
                double xb[kPairBatch64], yb[kPairBatch64], zb[kPairBatch64];
#pragma unroll
                for (int u = 0; u < kPairBatch64; ++u) {
                    jbuf[u] = cand(t0 + u);
                    const int jc = min(jbuf[u], last);
                    xb[u] = p.pos.x[jc]; yb[u] = p.pos.y[jc]; zb[u] = p.pos.z[jc];
                }
#pragma unroll
                for (int u = 0; u < kPairBatch64; ++u) {
                    const int t = t0 + u;
                    const int j = jbuf[u];
                    if (t < nt && j != s) {
                        const int iz = t < n0 ? izs[0] : (t < n01 ? izs[1] : izs[2]);
                        const double shz = iz * S.L[2];
                        const double dx = (xi - xb[u]) - shx;
                        const double dy = (yi - yb[u]) - shy;
                        const double dz = (zi - zb[u]) - shz;
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        if (r2 < rc2) visit(j, dx, dy, dz, r2, ix, iy, iz);
                    }
                }
            }
        }
    }
}

// FP64 pair law.  One IEEE division (1 / r^2); r^-6, the spline term and
// F = W sep / r^2 are formed from it by multiplication, within a few ulp of
// the reference's divide-per-term arithmetic.
template <int LAW>
__device__ __forceinline__ void pair_fp64(const SysDev& S, double dx, double dy, double dz,
                                          double r2, double& fx, double& fy, double& fz,
                                          double& pe, double& w)
{
    double vv, ww, fm;
    if (LAW == LAW_LJ) {
        const double inv = 1.0 / r2;
        const double t = (S.sigma2 == 1.0) ? inv : S.sigma2 * inv;   // (sigma / r)^2
        const double s6 = t * t * t;
        const double u = r2 * S.inv_rc2s;                            // (r / rc)^2
        vv = 4.0 * s6 * (s6 - 1.0) + S.alpha + S.beta * (u - 1.0);
        ww = 24.0 * s6 * (2.0 * s6 - 1.0) - 2.0 * S.beta * u;
        if (S.eps != 1.0) { vv = S.eps * vv; ww = S.eps * ww; }
        fm = ww * inv;
    } else {   // constant-magnitude test law
        const double nrm = sqrt(r2);
        vv = S.strength * (S.rc - nrm);
        ww = S.strength * nrm;
        fm = ww / (nrm * nrm);
    }
    fx += fm * dx; fy += fm * dy; fz += fm * dz;
    pe += vv;
    w += ww;
}

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), and each XCD has its own L2.  Renumbering so that XCD x walks one
// contiguous range of logical blocks keeps each XCD on one slab of
// neighbouring cells, so a cell's neighbours are fetched into one L2 instead
// of all eight.  Bijective on [0, gridDim.x).  (A pair launch over a subset
// of the systems maps this index through its ordered block list.)
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_slab_block()
{
    const int b = blockIdx.x, n = gridDim.x;
    const int x = b % kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Second half kick of slot s (engine/integrator.cpp:80-81) with the force just
// computed: v += (1/2) F dt, left in vx, vy, vz for the kinetic energy.  For a
// chained step (chain: another step follows in the same launch sequence) the
// stored velocity also takes the next step's first half kick with the same
// force and the same arithmetic, v + (1/2) F dt, so that step's key pass and
// sort read no force; the force is then not stored at all.
__device__ __forceinline__ void kick_store(const StepArgs& p, const SysDev& S, int s, double fx,
                                           double fy, double fz, double& vx, double& vy,
                                           double& vz, bool chain)
{
    const double dt = S.dt;
    const double hx = ((1. / 2.) * fx) * dt, hy = ((1. / 2.) * fy) * dt,
                 hz = ((1. / 2.) * fz) * dt;
    vx = vx + hx;
    vy = vy + hy;
    vz = vz + hz;
    if (chain) {
        p.vel.x[s] = vx + hx; p.vel.y[s] = vy + hy; p.vel.z[s] = vz + hz;
    } else {
        p.vel.x[s] = vx; p.vel.y[s] = vy; p.vel.z[s] = vz;
    }
}

// The owner of slot si of home cell c in a resident step's pair kernel, with
// its new force: the second half kick (vx, vy, vz become the full-step
// velocity, for the kinetic energy), then the next step's first half kick,
// drift and wrap with the arithmetic of kick_store, the key pass and the sort
// (chained steps).  Velocity and displacement are updated in place; the new
// position goes to the next position set (FP64 neighbours still read the
// current one) and, in mixed precision, the cell-local coordinates to the
// next lpos set.  A particle whose new cell differs is handed to the fix-up
// as an arrival record of its new cell (its slot is vacated: id -1); both
// cells are marked.  d: the displacement before the drift.
template <bool MIXED_>
__device__ __forceinline__ void res_place(const StepArgs& p, const SysDev& S, int c, int si,
                                          double x, double y, double z, double wx, double wy,
                                          double wz, double nx, double ny, double nz);

template <bool MIXED_>
__device__ __forceinline__ void res_kick_drift(const StepArgs& p, const SysDev& S, int c, int si,
                                               double fx, double fy, double fz, double& vx,
                                               double& vy, double& vz, double ddx, double ddy,
                                               double ddz)
{
    const double dt = S.dt;
    const double hx = ((1. / 2.) * fx) * dt, hy = ((1. / 2.) * fy) * dt,
                 hz = ((1. / 2.) * fz) * dt;
    vx = vx + hx;
    vy = vy + hy;
    vz = vz + hz;
    double wx = vx + hx, wy = vy + hy, wz = vz + hz;
    double x = p.pos.x[si], y = p.pos.y[si], z = p.pos.z[si];
    double ix, iy, iz;
    advance<false>(S, x, y, z, wx, wy, wz, 0., 0., 0., ix, iy, iz);
    res_place<MIXED_>(p, S, c, si, x, y, z, wx, wy, wz, ddx + ix, ddy + iy, ddz + iz);
}

// Slot si of cell c after its drift: the new (wrapped) position x, y, z, the
// velocity w carried into the next step and the displacement n.  A particle
// that stays in c keeps its slot (position and cell-local coordinates into
// the next sets, velocity and displacement in place); one that left becomes
// an arrival record of its new cell (the slot is vacated: id -1), and both
// cells are marked for the fix-up.
template <bool MIXED_>
__device__ __forceinline__ void res_place(const StepArgs& p, const SysDev& S, int c, int si,
                                          double x, double y, double z, double wx, double wy,
                                          double wz, double nx, double ny, double nz)
{
    const int kx = cell_axis(x, S.q[0], S.nc[0]);
    const int ky = cell_axis(y, S.q[1], S.nc[1]);
    const int kz = cell_axis(z, S.q[2], S.nc[2]);
    const int cn = S.cell_off + (kx * S.nc[1] + ky) * S.nc[2] + kz;
    p.vel.x[si] = wx; p.vel.y[si] = wy; p.vel.z[si] = wz;
    if (cn == c) {
        p.disp.x[si] = nx; p.disp.y[si] = ny; p.disp.z[si] = nz;
        p.pos_n.x[si] = x; p.pos_n.y[si] = y; p.pos_n.z[si] = z;
        if (MIXED_)
            p.lpos_n[si] = make_float4((float)(x - kx * S.a[0]), (float)(y - ky * S.a[1]),
                                       (float)(z - kz * S.a[2]), 0.0f);
        return;
    }
    const int a = atomicAdd(p.rc_narr + cn, 1);
    if (a < kResArr) {
        ResMover& mv = p.rc_arr[(size_t)cn * kResArr + a];
        mv.x = x; mv.y = y; mv.z = z;
        mv.vx = wx; mv.vy = wy; mv.vz = wz;
        mv.dx = nx; mv.dy = ny; mv.dz = nz;
        mv.id = p.id[si];
        mv.cell = cn;
    } else {
        atomicOr(p.rc_flag, 1);
    }
    p.id[si] = -1;   // the slot is vacated
    res_mark(p, c);
    res_mark(p, cn);
}

// The first half kick and drift of a call that starts on live resident cells
// (the previous call's last step stored its forces in the resident slots),
// with the arithmetic of the key pass and the sort (advance), then res_place.
// kResDriftCells cells per wave, 16 lanes each (lane t: slots t, t + 16, ...
// of its cell).  The fix-up follows.  As the call's first kernel it also does
// k_begin_call's work: clears the call's nrec record doubles and resets every
// system's record counter.
constexpr int kResDriftCells = 4;

template <bool MIXED_>
__global__ __launch_bounds__(kBlock) void k_res_kick_drift(StepArgs p, double* rec, long long nrec,
                                                           int* step, int nsys)
{
    {
        const long long t0 = (long long)blockIdx.x * kBlock + threadIdx.x;
        const long long stride = (long long)gridDim.x * kBlock;
        for (long long k = t0; k < nrec; k += stride) rec[k] = 0.0;
        for (long long k = t0; k < nsys; k += stride) step[k] = 0;
    }
    const int lane = threadIdx.x & 63;
    const int c = (blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6)) * kResDriftCells + (lane >> 4);
    if (c >= p.ncell_total) return;
    const SysDev& S = p.sys[p.cell_sys[c]];
    if (c - S.cell_off >= S.ncell) return;   // tail cell: padding slots
    const int n = p.rc_count[c], base = p.rc_base[c];
    for (int t = lane & 15; t < n; t += 16) {
        const int si = base + t;
        double x = p.pos.x[si], y = p.pos.y[si], z = p.pos.z[si];
        double vx = p.vel.x[si], vy = p.vel.y[si], vz = p.vel.z[si];
        const double fx = p.force.x[si], fy = p.force.y[si], fz = p.force.z[si];
        const double dx = p.disp.x[si], dy = p.disp.y[si], dz = p.disp.z[si];
        double ix, iy, iz;
        advance<true>(S, x, y, z, vx, vy, vz, fx, fy, fz, ix, iy, iz);
        res_place<MIXED_>(p, S, c, si, x, y, z, vx, vy, vz, dx + ix, dy + iy, dz + iz);
    }
}



template <int PREC, int LAW, int KICK>
__global__ __launch_bounds__(kBlock) void k_pair(StepArgs p)
{
    const int b = p.blk_list ? p.blk_list[xcd_slab_block()] : xcd_slab_block();
    const SysDev& S = p.sys[p.blk_sys[b]];
    if (p.skip_tiled && S.tiled) return;   // this system's forces come from a tiled kernel
    const int s = b * kBlock + threadIdx.x;
    const bool active = (s - S.atom_off) < S.n && (!KICK || S.active);
    double fx = 0, fy = 0, fz = 0, pe = 0, w = 0;
    int nct = 0;
    if (active) {
        if (LAW == LAW_NONE) {
            fx = p.force.x[s]; fy = p.force.y[s]; fz = p.force.z[s];
        } else if (PREC == FP64) {
            for_each_neighbour_fp64(S, p, s, [&](int, double dx, double dy, double dz, double r2,
                                                 int, int, int) {
                pair_fp64<LAW>(S, dx, dy, dz, r2, fx, fy, fz, pe, w);
                ++nct;
            });
        } else {
            // cell-local float coordinates; sep = (l_i - l_j) + shift, where
            // shift = (image * nc - step) * a - image * L for the neighbour cell
            // The LJ law as in the tiled mixed kernel: F / 24 accumulated from
            // g / 24 = s^6 (2 s^6 - 1), with per-particle float sums of g, s^6
            // and r^2 from which V and W follow in double.
            int cx, cy, cz;
            cell_xyz_of(p.cell_xyz, p.scell[s], cx, cy, cz);
            const float4 li = p.lpos[s];
            const float rc2f = S.rc2f;
            const float inv_s2 = S.inv_s2f;
            const float cw24 = S.cw24f;
            float ax = 0.f, ay = 0.f, az = 0.f;
            float gs = 0.f, s6s = 0.f, r2s = 0.f;   // LJ sums
            float spe = 0.f, sww = 0.f;             // constant law
            auto visit = [&](int j, const float4 lj, float shx, float shy, float shz) {
                if (j == s) return;
                const float dx = (li.x - lj.x) + shx;
                const float dy = (li.y - lj.y) + shy;
                const float dz = (li.z - lj.z) + shz;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                if (r2 < rc2f) {
                    ++nct;
                    if (LAW == LAW_LJ) {
                        const float inv = __builtin_amdgcn_rcpf(r2);
                        const float t = inv * inv_s2;
                        const float s6 = t * t * t;
                        const float g = s6 * __builtin_fmaf(2.0f, s6, -1.0f);
                        gs += g;
                        s6s += s6;
                        r2s += r2;
                        const float fm = __builtin_fmaf(g, inv, cw24);
                        ax = __builtin_fmaf(fm, dx, ax);
                        ay = __builtin_fmaf(fm, dy, ay);
                        az = __builtin_fmaf(fm, dz, az);
                    } else {
                        const float nrm = sqrtf(r2);
                        const float vv = (float)S.strength * ((float)S.rc - nrm);
                        const float ww = (float)S.strength * nrm;
                        const float fm = ww / r2;
                        ax += fm * dx; ay += fm * dy; az += fm * dz;
                        spe += vv;
                        sww += ww;
                    }
                }
            };
            const int last = p.nslot - 1;
            for (int ox = -1; ox <= 1; ++ox) {
                int nbx, ix;
                nb_axis(cx, ox, S.nc[0], nbx, ix);
                const float shx = S.shf[0][3 * (ox + 1) + ix + 1];
                for (int oy = -1; oy <= 1; ++oy) {
                    int nby, iy;
                    nb_axis(cy, oy, S.nc[1], nby, iy);
                    const float shy = S.shf[1][3 * (oy + 1) + iy + 1];
                    // the row's three cells, walked as one run of nt
                    // candidates (cell 0's members, then cell 1's, then cell
                    // 2's: the order of three separate loops), so a wave
                    // waits for the longest row of its lanes rather than for
                    // the longest cell of each of the three; positions are
                    // loaded kPairBatch at a time (dilute systems hold a
                    // particle or two per cell)
                    int jb[3], je[3];
                    float shz[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        int nbz, iz;
                        nb_axis(cz, k - 1, S.nc[2], nbz, iz);
                        shz[k] = S.shf[2][3 * k + iz + 1];
                        const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
                        jb[k] = p.start[nb];
                        je[k] = p.start[nb + 1];
                    }
                    const int n0 = je[0] - jb[0];
                    const int n01 = n0 + (je[1] - jb[1]);
                    const int nt = n01 + (je[2] - jb[2]);
                    auto cand = [&](int t) {   // slot of the row's t-th candidate
                        return t < n0 ? jb[0] + t : (t < n01 ? jb[1] + (t - n0) : jb[2] + (t - n01));
                    };
                    for (int t0 = 0; t0 < nt; t0 += kPairBatch) {
                        int jbuf[kPairBatch];
                        float4 lb[kPairBatch];
#pragma unroll
                        for (int u = 0; u < kPairBatch; ++u) {
                            jbuf[u] = cand(t0 + u);
                            lb[u] = p.lpos[min(jbuf[u], last)];
                        }
#pragma unroll
                        for (int u = 0; u < kPairBatch; ++u) {
                            const int t = t0 + u;
                            if (t < nt) {
                                const float sz = t < n0 ? shz[0] : (t < n01 ? shz[1] : shz[2]);
                                visit(jbuf[u], lb[u], shx, shy, sz);
                            }
                        }
                    }
                }
            }
            if (LAW == LAW_LJ) {
                // V = 4 s^12 - 4 s^6 + (alpha - beta) + beta u = 2 g' - 2 s^6 + ...,
                // W = 24 g' - 2 beta u (g' = g / 24), F = 24 (F / 24), times epsilon
                const double u_sum = (double)r2s * S.inv_rc2s;
                pe = S.eps * (2.0 * (double)gs - 2.0 * (double)s6s +
                              (S.alpha - S.beta) * (double)nct + S.beta * u_sum);
                w = S.eps * (24.0 * (double)gs - 2.0 * S.beta * u_sum);
                const double fs = 24.0 * S.eps;
                fx = fs * (double)ax; fy = fs * (double)ay; fz = fs * (double)az;
            } else {
                pe = spe;
                w = sww;
                fx = ax; fy = ay; fz = az;
            }
        }
    }
    double acc[4] = {0, 0, 0, 0};
    if (active) {
        if (p.ncount) p.ncount[s] = nct;
        if (LAW != LAW_NONE && KICK != 2) {
            p.force.x[s] = fx; p.force.y[s] = fy; p.force.z[s] = fz;
        }
        double vx = p.vel.x[s], vy = p.vel.y[s], vz = p.vel.z[s];
        if (KICK) kick_store(p, S, s, fx, fy, fz, vx, vy, vz, KICK == 2);
        const double ddx = p.disp.x[s], ddy = p.disp.y[s], ddz = p.disp.z[s];
        acc[0] = (vx * vx + vy * vy) + vz * vz;
        acc[1] = (1. / 2.) * pe;
        acc[2] = (1. / 2.) * w;
        acc[3] = (ddx * ddx + ddy * ddy) + ddz * ddz;
    }
    double out[4];
    block_sum4(acc, out);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) p.partial[4 * b + k] = out[k];
        if (KICK && b == S.blk_first) p.pending[p.blk_sys[b]] = 1;   // the step awaits its finalize
    }
}


__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block sums (kBlock threads) of four per-lane doubles through an LDS
// transpose: sc is 256 doubles of LDS owned by the calling wave.  Lane l adds
// the four entries 4 (l % 16) .. + 3 of quantity l / 16, a 16-lane DPP
// reduction completes each quantity within one row, and wave 0 adds the
// waves' results with a quad DPP reduction.  Thread 0 gets the sums.
template <int NW = kBlock / 64>
__device__ __forceinline__ void block_sum4_lds(const double v[4], double* sc, double out[4])
{
    static_assert(NW == 1 || NW == 2 || NW == 4, "one, two or four waves");
    __shared__ double red[4][NW];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    wave_lds_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) sc[q * 64 + lane] = v[q];
    wave_lds_sync();
    const double* row = sc + (lane >> 4) * 64 + 4 * (lane & 15);
    double a = (row[0] + row[1]) + (row[2] + row[3]);
    a += dpp_f64<0xB1>(a);    // quad_perm 1,0,3,2
    a += dpp_f64<0x4E>(a);    // quad_perm 2,3,0,1
    a += dpp_f64<0x141>(a);   // row_half_mirror
    a += dpp_f64<0x140>(a);   // row_mirror
    double w4[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w4[q] = readlane_f64(a, 16 * q);
    if constexpr (NW == 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = w4[q];
        return;
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) red[q][wid] = w4[q];
    }
    __syncthreads();
    if (wid == 0) {
        // lane NW q + w holds wave w's quantity q
        double b = lane < 4 * NW ? red[lane / NW][lane % NW] : 0.0;
        b += dpp_f64<0xB1>(b);
        if (NW == 4) b += dpp_f64<0x4E>(b);
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = readlane_f64(b, NW * q);
    }
}

// ---------------------------------------------------------------------------
// 6b. tiled pair kernel, mixed precision: one wave64 per home cell, with the
//     cutoff test reduced to three FMAs per candidate.
//
//   * coordinates are float, relative to the centre o of the chunk's bounding
//     box (the cell-local lpos minus the neighbour cell's image shift minus
//     o), so |x| stays below a few sigma;
//   * staged neighbours are stored slice-major as (r, |r|^2): slice s holds
//     staged entries s, s + S, s + 2S, ... in consecutive rows, so a lane
//     walks its rows at immediate LDS offsets;
//   * phase 1 evaluates |x_i - r|^2 < rc^2 as |r|^2 - 2 x_i.r < rc^2 - |x_i|^2
//     and shifts the outcome into a 64-bit mask; the entry of the particle
//     itself is recorded while staging and its bit is cleared, so phase 2
//     runs the pair law on every set bit without a test;
//   * phase 2 accumulates the force and, per stage, sum g, sum s^6 and
//     sum r^2 with g = 48 s^12 - 24 s^6, from which V = 4 s^12 - 4 s^6 +
//     (alpha - beta) + beta u = g / 12 - 2 s^6 + (alpha - beta) + beta u and
//     W = g - 2 beta u (u = r^2 / rc^2) follow in double.
// ---------------------------------------------------------------------------
constexpr int kStageM = 256;   // staged candidates per stage
constexpr int kRowsM = 324;    // >= max over slice counts S of S * rows(S) + 3

// Stage geometry of a chunk of ni particles (uniform per wave, looked up).
struct SliceGeom {
    int nsl;          // slices: 64 / ni
    int rows;         // rows per slice in the stage buffer (odd where it fits)
    unsigned magic;   // t / nsl == (t * magic) >> 20 for 0 <= t < 2048
    int cap;          // staged candidates per stage: min(kStageM, 64 nsl)
    unsigned magic_ni;   // lane / ni == (lane * magic_ni) >> 20 for lane < 64
};

constexpr SliceGeom slice_geom(int ni)
{
    const int nsl = 64 / ni;
    const int cap = kStageM < 64 * nsl ? kStageM : 64 * nsl;
    int rows = (cap + nsl - 1) / nsl;
    if (!(rows & 1) && nsl * (rows + 1) + 3 <= kRowsM) ++rows;
    return SliceGeom{nsl, rows, (1u << 20) / (unsigned)nsl + 1u, cap,
                     (1u << 20) / (unsigned)ni + 1u};
}

struct SliceGeomTable {
    SliceGeom g[65];
    constexpr SliceGeomTable() : g{}
    {
        for (int ni = 1; ni <= 64; ++ni) g[ni] = slice_geom(ni);
    }
};

__constant__ const SliceGeomTable kSliceGeom = SliceGeomTable();


// NWC > 1 (small systems, flag tiled 2): one home cell per workgroup of NWC
// waves, wave sub taking the neighbour cells q with q % NWC == sub; their
// partial forces are added in wave order before the kick.  The grid then runs
// cpb workgroups per cell-space block.
template <int LAW, int KICK, int NW, int NWC = 1, bool RES = false, bool UNIT = false>
__global__ __launch_bounds__(64 * NW * NWC, 6) void k_pair_cells_mixed(StepArgs p)
{
    static_assert(NW == 1 || NWC == 1, "several waves per cell: one cell per block");
    static_assert(!RES || KICK == 2, "resident cells: chained steps");
    // One record per wave, so that every per-lane LDS address is one base
    // register plus an immediate offset.
    struct alignas(16) WaveLds {
        float4 st[kRowsM];   // stage rows; slice sums; the block sums' scratch
        float4 tb_sh[27];    // image shift + chunk origin
        float4 tb_img[27];   // image shift of each neighbour cell
        int tb_delta[27];    // slot of a cell's first particle minus its offset
        int tb_off[32];      // entries 27..31 = INT_MAX (search padding)
        int selfe[64];       // staged entry of each chunk particle (-1: none)
    };
    __shared__ WaveLds wl[NW * NWC];
    __shared__ float4 comb[NWC > 1 ? NWC : 1][64];   // NWC > 1: each wave's partial F, count

    const int lb = xcd_slab_block();
    int b;
    if constexpr (NWC == 1)
        b = p.cblk_list ? p.cblk_list[lb] : lb;
    else
        b = p.lblk_list ? p.lblk_list[lb / p.cpb] : lb / p.cpb;
    const int k = p.cblk_sys[b];
    const SysDev& S = p.sys[k];
    if (S.tiled != (NWC > 1 ? 2 : 1)) return;   // this system's forces come from another kernel
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sub = wid % NWC;   // this wave's share of the neighbour cells
    WaveLds& W = wl[wid];
    const int lc = NWC == 1 ? (b - S.cblk_first) * NW + wid : (b - S.cblk_first) * p.cpb + lb % p.cpb;
    if (lc >= S.ncell) return;   // past the system's last cell (uniform over an NWC block)
    double wacc[4] = {0, 0, 0, 0};   // |v|^2, PE, W, |d|^2 of this lane's owned particles

    if (!KICK || S.active) {
        const int c = S.cell_off + lc;
        int cx, cy, cz;
        cell_xyz_of(p.cell_xyz, c, cx, cy, cz);
        const int h0 = __builtin_amdgcn_readfirstlane(p.start[c]);
        const int h1 = __builtin_amdgcn_readfirstlane(p.cend[c]);
        float4* sw = W.st;
        const float rc2f = S.rc2f;
        const float inv_s2 = S.inv_s2f;
        const float cw24 = S.cw24f;
        const float keepf = rc2f + 1e-3f;   // bounding-box pruning, wider than any rounding

        // ---- the 27 neighbour cells (reference lattice-image rule), one lane each
        int cntq = 0, j0 = 0;
        if (lane < 27) {
            const int ox = lane / 9 - 1, oy = (lane / 3) % 3 - 1, oz = lane % 3 - 1;
            int nbx, ix, nby, iy, nbz, iz;
            nb_axis(cx, ox, S.nc[0], nbx, ix);
            nb_axis(cy, oy, S.nc[1], nby, iy);
            nb_axis(cz, oz, S.nc[2], nbz, iz);
            const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
            j0 = p.start[nb];
            cntq = (NWC == 1 || lane % NWC == sub) ? p.cend[nb] - j0 : 0;
            // kept in LDS for the chunks of this cell (not in registers)
            W.tb_img[lane] = make_float4(S.shf[0][3 * (ox + 1) + ix + 1],
                                            S.shf[1][3 * (oy + 1) + iy + 1],
                                            S.shf[2][3 * (oz + 1) + iz + 1], 0.f);
        }
        const int incl = wave_incl_scan_dpp(cntq);
        if (lane < 27) {
            W.tb_off[lane] = incl - cntq;
            W.tb_delta[lane] = j0 - (incl - cntq);
        } else if (lane < 32) {
            W.tb_off[lane] = 0x7fffffff;
        }
        const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 26, 64));
        // offset of the home cell (offset 13) in the concatenation
        const int off_home = __builtin_amdgcn_readfirstlane(__shfl(incl - cntq, 13, 64));
        const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        bool res_own = false;   // resident step: this lane owns a particle (force in res_f)
        float3 res_f = make_float3(0.f, 0.f, 0.f);

        for (int i0 = h0; i0 < h1; i0 += 64) {
            const int ni = min(64, h1 - i0);
            const SliceGeom G = kSliceGeom.g[ni];
            const int nsl = G.nsl;
            const bool act = lane < nsl * ni;
            const int q_ni = (int)(((unsigned)lane * G.magic_ni) >> 20);   // lane / ni
            const int il = act ? lane - q_ni * ni : 0;
            const int sl = act ? q_ni : 0;
            const int si = i0 + il;
            const float4 li = p.lpos[si];

            // bounding box of the chunk, its centre o and half extents h
            // (cell-local coordinates are >= 0 up to rounding residues of
            // ~1e-14, so their bit patterns order as signed ints)
            float o3[3], h3[3];
            {
                const float c3[3] = {li.x, li.y, li.z};
#pragma unroll
                for (int a3 = 0; a3 < 3; ++a3) {
                    const int bits = __float_as_int(c3[a3]);
                    const float lo = __int_as_float(wave_extreme<false>(act ? bits : 0x7fffffff));
                    const float hi = __int_as_float(wave_extreme<true>(act ? bits : (int)0x80000000));
                    o3[a3] = 0.5f * (lo + hi);
                    h3[a3] = 0.5f * (hi - lo);
                }
            }
            const float xi = li.x - o3[0], yi = li.y - o3[1], zi = li.z - o3[2];
            const float m2x = -2.0f * xi, m2y = -2.0f * yi, m2z = -2.0f * zi;
            const float thr = rc2f - __builtin_fmaf(zi, zi, __builtin_fmaf(yi, yi, xi * xi));

            // slice-major stage geometry (uniform)
            const int cap = G.cap;
            const int SL = G.rows;
            const unsigned magic = G.magic;   // t / nsl = t * magic >> 20
            const float4* rowp = sw + sl * SL;

            wave_lds_sync();   // the previous chunk's reads of tb_sh / sw are done
            if (lane < 27) {
                const float4 im = W.tb_img[lane];
                W.tb_sh[lane] = make_float4(im.x + o3[0], im.y + o3[1], im.z + o3[2], 0.f);
            }
            wave_lds_sync();

            // candidate g of the concatenated neighbour cells (clamped to the
            // last one, so every lane loads): its cell q and cell-local position,
            // fetched one round ahead
            auto fetch = [&](int g, int& q, float4& v) {
                const int gc = min(g, total - 1);
                int pos = 0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1)
                    if (W.tb_off[pos + step] <= gc) pos += step;
                q = pos;
                const unsigned j = (unsigned)(W.tb_delta[pos] + gc);
                v = *(const float4*)((const char*)p.lpos + j * 16u);
            };
            int nq = 0;
            float4 nv = make_float4(0.f, 0.f, 0.f, 0.f);
            // (a wave of several per cell may have no candidate at all: then
            // nothing is fetched, as the staging loop below does not run)
            if (NWC == 1 || total > 0) fetch(lane, nq, nv);
            // concatenation index of the chunk's first particle (only the wave
            // that walks the home cell, table entry 13, meets the chunk itself)
            const int self0 = (NWC == 1 || 13 % NWC == sub) ? off_home + (i0 - h0) : 0x3fffffff;
            int g0 = 0;
            double pe = 0.0, w = 0.0;              // constant law
            double ga = 0.0, s6a = 0.0, r2a = 0.0;   // LJ: sums of g, s^6, r^2
            float ax = 0.f, ay = 0.f, az = 0.f;
            int nct = 0;
            while (g0 < total) {
                // ---- stage the surviving neighbours of the next candidates, 64 at a time
                if (lane < ni) W.selfe[lane] = -1;
                int m = 0;
                while (g0 < total && m <= cap - 64) {
                    const int cq = nq;
                    const float4 cv = nv;
                    const int g = g0 + lane;
                    g0 += 64;
                    fetch(g0 + lane, nq, nv);
                    const float4 sh = W.tb_sh[cq];
                    float4 rv;
                    rv.x = cv.x - sh.x;
                    rv.y = cv.y - sh.y;
                    rv.z = cv.z - sh.z;
                    const float ex = fmaxf(fabsf(rv.x) - h3[0], 0.f);
                    const float ey = fmaxf(fabsf(rv.y) - h3[1], 0.f);
                    const float ez = fmaxf(fabsf(rv.z) - h3[2], 0.f);
                    const bool keep =
                        g < total && __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)) < keepf;
                    rv.w = __builtin_fmaf(rv.z, rv.z, __builtin_fmaf(rv.y, rv.y, rv.x * rv.x));
                    const unsigned long long bal = __ballot(keep);
                    if (keep) {
                        const unsigned t = (unsigned)m + (unsigned)__popcll(bal & below);
                        const unsigned row = (t * magic) >> 20;
                        const unsigned e = (t - row * (unsigned)nsl) * (unsigned)SL + row;
                        sw[e] = rv;
                        // chunk particles occupy [self0, self0 + ni) of the home cell's range
                        const unsigned ic = (unsigned)(g - self0);
                        if (ic < (unsigned)ni) W.selfe[ic] = (int)e;
                    }
                    m += __popcll(bal);
                }
                const int kmax = __builtin_amdgcn_readfirstlane(
                    (int)(((unsigned)(m + nsl - 1) * magic) >> 20));
                // far-away fillers for the unused tail of the last row
                if (lane < kmax * nsl - m) {
                    const int t = m + lane;
                    const int row = (int)(((unsigned)t * magic) >> 20);
                    sw[(t - row * nsl) * SL + row] = make_float4(1e18f, 1e18f, 1e18f, 3e36f);
                }
                wave_lds_sync();
                if (kmax == 0) continue;

                // ---- phase 1: kmax rows per lane, four per step (rows past kmax are dropped)
                // Rows are visited from the last group of four down, so that
                // row r ends up in bit r (the first row is shifted in last).
                const int k4 = (kmax + 3) & ~3;
                const int nlo = min(k4, 32), nhi = k4 - nlo;
                unsigned lo = 0u, hi = 0u;
                for (int kb = 32 + nhi - 4; kb >= 32; kb -= 4) {
#pragma unroll
                    for (int u = 3; u >= 0; --u) {
                        const float4 d = rowp[kb + u];
                        const float v = __builtin_fmaf(m2x, d.x, __builtin_fmaf(m2y, d.y, __builtin_fmaf(m2z, d.z, d.w)));
                        // shift in the sign of v - thr (set iff v < thr)
                        hi = __builtin_amdgcn_alignbit(hi, (unsigned)__float_as_int(v - thr), 31);
                    }
                }
                for (int kb = nlo - 4; kb >= 0; kb -= 4) {
#pragma unroll
                    for (int u = 3; u >= 0; --u) {
                        const float4 d = rowp[kb + u];
                        const float v = __builtin_fmaf(m2x, d.x, __builtin_fmaf(m2y, d.y, __builtin_fmaf(m2z, d.z, d.w)));
                        lo = __builtin_amdgcn_alignbit(lo, (unsigned)__float_as_int(v - thr), 31);
                    }
                }
                // row r -> bit r; rows kmax .. k4 - 1 are dropped
                unsigned long long mask = ((unsigned long long)hi << 32) | lo;
                if (kmax < 64) mask &= (1ull << kmax) - 1ull;
                const int self_row = W.selfe[il] - sl * SL;
                if ((unsigned)self_row < (unsigned)SL) mask &= ~(1ull << self_row);
                if (!act) mask = 0ull;
                nct += __popcll(mask);

                // ---- phase 2: pair law on every set bit
                if (LAW == LAW_LJ) {
                    // g / 24 = 2 s^12 - s^6 = s^6 (2 s^6 - 1) (inline constants only);
                    // the force is accumulated as (F/r) / 24 = (g / 24) / r^2 + cw / 24
                    // and scaled by 24 in double once per particle
                    float gs = 0.f, s6s = 0.f, r2s = 0.f;
                    // the next set bit's row is read from LDS before the
                    // current pair's law runs
                    float4 dn = rowp[mask ? __builtin_ctzll(mask) : 0];
                    while (mask) {
                        mask &= mask - 1;
                        const float4 d = dn;
                        dn = rowp[mask ? __builtin_ctzll(mask) : 0];
                        const float dx = xi - d.x, dy = yi - d.y, dz = zi - d.z;
                        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float inv = __builtin_amdgcn_rcpf(r2);
                        const float t = UNIT ? inv : inv * inv_s2;   // UNIT: sigma = 1, a multiply by 1
                        const float s6 = t * t * t;
                        const float g = s6 * __builtin_fmaf(2.0f, s6, -1.0f);
                        gs += g;
                        s6s += s6;
                        r2s += r2;
                        const float fm = __builtin_fmaf(g, inv, cw24);
                        ax = __builtin_fmaf(fm, dx, ax);
                        ay = __builtin_fmaf(fm, dy, ay);
                        az = __builtin_fmaf(fm, dz, az);
                    }
                    ga += 24.0 * (double)gs;
                    s6a += (double)s6s;
                    r2a += (double)r2s;
                } else {
                    float spe = 0.f, sww = 0.f;
                    while (mask) {
                        const int bit = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const float4 d = rowp[bit];
                        const float dx = xi - d.x, dy = yi - d.y, dz = zi - d.z;
                        const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float nrm = sqrtf(r2);
                        const float ww = (float)S.strength * nrm;
                        spe += (float)S.strength * ((float)S.rc - nrm);
                        sww += ww;
                        const float fm = ww / r2;
                        ax += fm * dx; ay += fm * dy; az += fm * dz;
                    }
                    pe += (double)spe;
                    w += (double)sww;
                }
                wave_lds_sync();
            }
            if (LAW == LAW_LJ) {
                // V = 4 s^12 - 4 s^6 + (alpha - beta) + beta u = g / 12 - 2 s^6 + ...,
                // W = 48 s^12 - 24 s^6 - 2 beta u = g - 2 beta u
                const double u_sum = r2a * S.inv_rc2s;
                pe = ga * (1.0 / 12.0) - 2.0 * s6a + (S.alpha - S.beta) * (double)nct + S.beta * u_sum;
                w = ga - 2.0 * S.beta * u_sum;
                if (S.eps != 1.0) {
                    ax *= (float)S.eps; ay *= (float)S.eps; az *= (float)S.eps;
                    pe *= S.eps; w *= S.eps;
                }
            }
            // ---- sum the slices of each particle in a fixed order
            sw[lane] = make_float4(ax, ay, az, __int_as_float(nct));
            wave_lds_sync();
            if (act && sl == 0) {
                for (int t = 1; t < nsl; ++t) {
                    const float4 r = sw[il + t * ni];
                    ax += r.x; ay += r.y; az += r.z;
                    nct += __float_as_int(r.w);
                }
            }
            bool owner = act && sl == 0;
            if constexpr (NWC > 1) {
                // the waves' partial forces of each particle, added in wave order
                if (owner) comb[sub][il] = make_float4(ax, ay, az, __int_as_float(nct));
                __syncthreads();
                if (owner && sub == 0) {
                    float4 r = comb[0][il];
                    ax = r.x; ay = r.y; az = r.z;
                    nct = __float_as_int(r.w);
                    for (int w2 = 1; w2 < NWC; ++w2) {
                        r = comb[w2][il];
                        ax += r.x; ay += r.y; az += r.z;
                        nct += __float_as_int(r.w);
                    }
                }
                __syncthreads();   // comb is written again by the next chunk
                owner = owner && sub == 0;
            }
            if (owner) {
                // the LJ law accumulates F / 24 (exact rescaling in double)
                constexpr double fscale = (LAW == LAW_LJ) ? 24.0 : 1.0;
                const double fx = fscale * (double)ax, fy = fscale * (double)ay,
                             fz = fscale * (double)az;
                if (p.ncount) p.ncount[si] = nct;
                if constexpr (RES) {
                    // resident cells hold at most kResCap < 64 members: this is
                    // the only chunk, and the drift follows the loop (fewer
                    // registers live there)
                    res_own = true;
                    res_f = make_float3(ax, ay, az);
                } else {
                    if (KICK != 2) { p.force.x[si] = fx; p.force.y[si] = fy; p.force.z[si] = fz; }
                    double vx = p.vel.x[si], vy = p.vel.y[si], vz = p.vel.z[si];
                    if (KICK) kick_store(p, S, si, fx, fy, fz, vx, vy, vz, KICK == 2);
                    const double ddx = p.disp.x[si], ddy = p.disp.y[si], ddz = p.disp.z[si];
                    wacc[0] += (vx * vx + vy * vy) + vz * vz;
                    wacc[3] += (ddx * ddx + ddy * ddy) + ddz * ddz;
                }
            }
            if (act) {
                wacc[1] += (1. / 2.) * pe;
                wacc[2] += (1. / 2.) * w;
            }
        }
        if constexpr (RES) {
            if (res_own) {
                constexpr double fscale = (LAW == LAW_LJ) ? 24.0 : 1.0;
                const int si = h0 + lane;   // owners: lanes 0 .. n - 1 of the one chunk
                double vx = p.vel.x[si], vy = p.vel.y[si], vz = p.vel.z[si];
                const double ddx = p.disp.x[si], ddy = p.disp.y[si], ddz = p.disp.z[si];
                res_kick_drift<true>(p, S, c, si, fscale * (double)res_f.x,
                                     fscale * (double)res_f.y, fscale * (double)res_f.z, vx, vy,
                                     vz, ddx, ddy, ddz);
                wacc[0] += (vx * vx + vy * vy) + vz * vz;
                wacc[3] += (ddx * ddx + ddy * ddy) + ddz * ddz;
            }
        }
    }
    // this cell's sums (one record per cell, so the block shape does not
    // change any result)
    double out[4];
    block_sum4_lds<NWC>(wacc, reinterpret_cast<double*>(W.st), out);
    if (NWC == 1 ? lane == 0 : threadIdx.x == 0) {
        const int c = S.cell_off + lc;
#pragma unroll
        for (int q = 0; q < 4; ++q) p.partial_c[4 * (size_t)c + q] = out[q];
        if (KICK && lc == 0) p.pending[k] = 1;   // the step awaits its finalize
    }
}

// ---------------------------------------------------------------------------
// 6c. tiled pair kernel, double precision: the candidate search of 6b on
//     float copies, widened by a margin so that it keeps every pair, then the
//     reference's own double test and pair law for the survivors.
//
//   * staged entry: float4 (image position relative to the chunk centre,
//     |r|^2), the double position r_j and the index q of its cell in the
//     27-cell table (which holds the image shift i L);
//   * phase 1: |x - r|^2 < rc^2 + 1e-3 in float, a superset of the pairs
//     (float rounding stays below 1e-4 here);
//   * phase 2: sep = (r_i - r_j) - i L and r^2 = (x^2 + y^2) + z^2 exactly as
//     the reference forms them, kept if 0 < r^2 < rc^2, then the law; the
//     decisions are the reference's bit for bit.
// ---------------------------------------------------------------------------
constexpr int kRowBudgetD = 200;        // staged candidates per stage (at most)
constexpr int kRowsD = kRowBudgetD + 4;  // + the rows read past kmax in phase 1

constexpr SliceGeom slice_geom_d(int ni)
{
    const int nsl = 64 / ni;
    int rows = kRowBudgetD / nsl;
    if (rows > 64) rows = 64;
    if (!(rows & 1) && nsl > 1 && nsl * (rows - 1) >= 128) --rows;   // odd where it fits
    return SliceGeom{nsl, rows, (1u << 20) / (unsigned)nsl + 1u, nsl * rows,
                     (1u << 20) / (unsigned)ni + 1u};
}

struct SliceGeomTableD {
    SliceGeom g[65];
    constexpr SliceGeomTableD() : g{}
    {
        for (int ni = 1; ni <= 64; ++ni) g[ni] = slice_geom_d(ni);
    }
};

__constant__ const SliceGeomTableD kSliceGeomD = SliceGeomTableD();

// NWC > 1 (small systems): NWC waves share one home cell, wave sub taking the
// neighbour cells q with q % NWC == sub; their partial forces are added in
// wave order before the kick.  The system's flag names the kernel: tiled 1
// for NWC = 1, 2 for NWC > 1.
template <int LAW, int KICK, int NWC = 1, bool RES = false, bool UNIT = false>
__global__ __launch_bounds__(64 * kCellsPerBlockFp64 * NWC, 4) void k_pair_cells_fp64(StepArgs p)
{
    static_assert(!RES || KICK == 2, "resident cells: chained steps");
    static_assert(kCellsPerBlockFp64 == 1 || NWC == 1, "several waves per cell: one cell per block");
    constexpr int NW = kCellsPerBlockFp64 * NWC;
    __shared__ float4 st[NW][kRowsD];
    __shared__ double sd[NW][kRowsD * 3];      // r_j of each entry; slice and block sums
    __shared__ unsigned char sq[NW][kRowsD];   // table index of the entry's cell
    __shared__ int tb_delta[NW][27];
    __shared__ int tb_off[NW][32];             // entries 27..31 = INT_MAX (search padding)
    __shared__ double tb_shd[NW][27][3];       // image shift i L of each neighbour cell
    __shared__ double tb_rel[NW][27][3];       // i L - o (o: chunk centre)
    __shared__ double comb[NWC > 1 ? NWC : 1][64][4];   // NWC > 1: each wave's partial F, count

    const int* list = NWC > 1 ? p.lblk_list : p.cblk_list;
    const int b = list ? list[xcd_slab_block()] : xcd_slab_block();
    const int k = p.cblk_sys[b];
    const SysDev& S = p.sys[k];
    if (S.tiled != (NWC > 1 ? 2 : 1)) return;   // this system's forces come from another kernel
    const int lane = threadIdx.x & 63;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sub = wid % NWC;   // this wave's share of the neighbour cells
    const int lc = (b - S.cblk_first) * kCellsPerBlockFp64 + wid / NWC;
    if (lc >= S.ncell) return;   // past the system's last cell (uniform over the block)
    double wacc[4] = {0, 0, 0, 0};   // |v|^2, PE, W, |d|^2 of this lane's owned particles
    double* sdw = sd[wid];

    if (!KICK || S.active) {
        const int c = S.cell_off + lc;
        int cx, cy, cz;
        cell_xyz_of(p.cell_xyz, c, cx, cy, cz);
        const int h0 = __builtin_amdgcn_readfirstlane(p.start[c]);
        const int h1 = __builtin_amdgcn_readfirstlane(p.cend[c]);
        float4* sw = st[wid];
        const double rc2d = S.rc2;
        const double sig2 = S.sigma2;
        const double cwd24 = -2.0 * S.beta * S.inv_rc2s / 24.0;
        const float rc2f = S.rc2f;
        const float keepf = rc2f + 2e-3f;   // bounding-box pruning, wider than the phase-1 margin

        // ---- the 27 neighbour cells (reference lattice-image rule), one lane each
        int cntq = 0, j0 = 0;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        bool img = false;
        if (lane < 27) {
            const int ox = lane / 9 - 1, oy = (lane / 3) % 3 - 1, oz = lane % 3 - 1;
            int nbx, ix, nby, iy, nbz, iz;
            nb_axis(cx, ox, S.nc[0], nbx, ix);
            nb_axis(cy, oy, S.nc[1], nby, iy);
            nb_axis(cz, oz, S.nc[2], nbz, iz);
            const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
            j0 = p.start[nb];
            cntq = (NWC == 1 || lane % NWC == sub) ? p.cend[nb] - j0 : 0;
            sx = ix * S.L[0];
            sy = iy * S.L[1];
            sz = iz * S.L[2];
            tb_shd[wid][lane][0] = sx;
            tb_shd[wid][lane][1] = sy;
            tb_shd[wid][lane][2] = sz;
            img = (ix | iy | iz) != 0 && (NWC == 1 || cntq > 0);
        }
        // no neighbour cell is an image (cells away from the box faces): every
        // shift is +0.0, and (r_i - r_j) - 0.0 is r_i - r_j bit for bit
        const bool noimg = __builtin_amdgcn_ballot_w64(img) == 0ull;
        const int incl = wave_incl_scan_dpp(cntq);
        if (lane < 27) {
            tb_off[wid][lane] = incl - cntq;
            tb_delta[wid][lane] = j0 - (incl - cntq);
        } else if (lane < 32) {
            tb_off[wid][lane] = 0x7fffffff;
        }
        const int total = __builtin_amdgcn_readfirstlane(__shfl(incl, 26, 64));
        const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

        for (int i0 = h0; i0 < h1; i0 += 64) {
            const int ni = min(64, h1 - i0);
            const SliceGeom G = kSliceGeomD.g[ni];
            const int nsl = G.nsl;
            const bool act = lane < nsl * ni;
            const int q_ni = (int)(((unsigned)lane * G.magic_ni) >> 20);   // lane / ni
            const int il = act ? lane - q_ni * ni : 0;
            const int sl = act ? q_ni : 0;
            const int si = i0 + il;
            const double xi = p.pos.x[si], yi = p.pos.y[si], zi = p.pos.z[si];

            // bounding box of the chunk in float (positions are >= 0 after the
            // wrap, so their float bit patterns order as signed ints)
            float o3[3], h3[3];
            {
                const float c3[3] = {(float)xi, (float)yi, (float)zi};
#pragma unroll
                for (int a3 = 0; a3 < 3; ++a3) {
                    const int bits = __float_as_int(c3[a3]);
                    const float lo = __int_as_float(wave_extreme<false>(act ? bits : 0x7fffffff));
                    const float hi = __int_as_float(wave_extreme<true>(act ? bits : (int)0x80000000));
                    o3[a3] = 0.5f * (lo + hi);
                    h3[a3] = 0.5f * (hi - lo) + 1e-4f * (hi + 1.0f);   // + float rounding of r_i
                }
            }
            const float xf = (float)(xi - (double)o3[0]);
            const float yf = (float)(yi - (double)o3[1]);
            const float zf = (float)(zi - (double)o3[2]);
            const float m2x = -2.0f * xf, m2y = -2.0f * yf, m2z = -2.0f * zf;
            const float thr = (rc2f + 1e-3f) - __builtin_fmaf(zf, zf, __builtin_fmaf(yf, yf, xf * xf));

            const int cap = G.cap;
            const int SL = G.rows;
            const unsigned magic = G.magic;   // t / nsl = t * magic >> 20
            const float4* rowp = sw + sl * SL;

            wave_lds_sync();   // the previous chunk's reads of tb_rel / st / sd are done
            if (lane < 27) {
                tb_rel[wid][lane][0] = sx - (double)o3[0];
                tb_rel[wid][lane][1] = sy - (double)o3[1];
                tb_rel[wid][lane][2] = sz - (double)o3[2];
            }
            wave_lds_sync();

            // candidate g (clamped to the last one, so every lane loads): its
            // table index q and double position, fetched one round ahead
            auto fetch = [&](int g, int& q, double& x, double& y, double& z) {
                const int gc = min(g, total - 1);
                int pos = 0;
#pragma unroll
                for (int step = 16; step > 0; step >>= 1)
                    if (tb_off[wid][pos + step] <= gc) pos += step;
                q = pos;
                const unsigned j = (unsigned)(tb_delta[wid][pos] + gc);
                x = p.pos.x[j];
                y = p.pos.y[j];
                z = p.pos.z[j];
            };
            int nq = 0;
            double nx = 0.0, ny = 0.0, nz = 0.0;
            // (a wave of several per cell may have no candidate at all: then
            // nothing is fetched, as the staging loop below does not run)
            if (NWC == 1 || total > 0) fetch(lane, nq, nx, ny, nz);
            int g0 = 0;
            double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, w = 0.0;
            double ga = 0.0, s6a = 0.0, r2a = 0.0;   // LJ: sums of g / 24, s^6, r^2
            int nct = 0;
            while (g0 < total) {
                // ---- stage the surviving neighbours of the next candidates, 64 at a time
                int m = 0;
                // A stage is filled to exactly cap entries: of the batch that
                // overflows it, the survivors past the room are left for the
                // next stage, which restarts after the last one taken.
                while (g0 < total && m < cap) {
                    const int cq = nq;
                    const double cxd = nx, cyd = ny, czd = nz;
                    const int g = g0 + lane;
                    g0 += 64;
                    fetch(g0 + lane, nq, nx, ny, nz);
                    float4 rv;
                    rv.x = (float)(cxd + tb_rel[wid][cq][0]);
                    rv.y = (float)(cyd + tb_rel[wid][cq][1]);
                    rv.z = (float)(czd + tb_rel[wid][cq][2]);
                    const float ex = fmaxf(fabsf(rv.x) - h3[0], 0.f);
                    const float ey = fmaxf(fabsf(rv.y) - h3[1], 0.f);
                    const float ez = fmaxf(fabsf(rv.z) - h3[2], 0.f);
                    const bool keep =
                        g < total && __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)) < keepf;
                    rv.w = __builtin_fmaf(rv.z, rv.z, __builtin_fmaf(rv.y, rv.y, rv.x * rv.x));
                    const unsigned long long bal = __ballot(keep);
                    const int rank = __popcll(bal & below);
                    const int room = cap - m;   // >= 1
                    if (keep && rank < room) {
                        const unsigned t = (unsigned)(m + rank);
                        const unsigned row = (t * magic) >> 20;
                        const unsigned e = (t - row * (unsigned)nsl) * (unsigned)SL + row;
                        sw[e] = rv;
                        sdw[3 * e] = cxd;
                        sdw[3 * e + 1] = cyd;
                        sdw[3 * e + 2] = czd;
                        sq[wid][e] = (unsigned char)cq;
                    }
                    const int nk = __popcll(bal);
                    if (nk > room) {
                        const unsigned long long lastm = __ballot(keep && rank == room - 1);
                        g0 += __builtin_ctzll(lastm) + 1 - 64;   // after the last one taken
                        fetch(g0 + lane, nq, nx, ny, nz);
                        m = cap;
                    } else {
                        m += nk;
                    }
                }
                const int kmax = __builtin_amdgcn_readfirstlane(
                    (int)(((unsigned)(m + nsl - 1) * magic) >> 20));
                // far-away fillers for the unused tail of the last row
                if (lane < kmax * nsl - m) {
                    const int t = m + lane;
                    const int row = (int)(((unsigned)t * magic) >> 20);
                    sw[(t - row * nsl) * SL + row] = make_float4(1e18f, 1e18f, 1e18f, 3e36f);
                }
                wave_lds_sync();
                if (kmax == 0) continue;

                // ---- phase 1: kmax rows per lane, four per step (rows past kmax are dropped)
                const int k4 = (kmax + 3) & ~3;
                const int nlo = min(k4, 32), nhi = k4 - nlo;
                unsigned lo = 0u, hi = 0u;
                for (int kb = 0; kb < nlo; kb += 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float4 d = rowp[kb + u];
                        const float v = __builtin_fmaf(m2x, d.x, __builtin_fmaf(m2y, d.y, __builtin_fmaf(m2z, d.z, d.w)));
                        lo = __builtin_amdgcn_alignbit(lo, (unsigned)__float_as_int(v - thr), 31);
                    }
                }
                for (int kb = 32; kb < 32 + nhi; kb += 4) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float4 d = rowp[kb + u];
                        const float v = __builtin_fmaf(m2x, d.x, __builtin_fmaf(m2y, d.y, __builtin_fmaf(m2z, d.z, d.w)));
                        hi = __builtin_amdgcn_alignbit(hi, (unsigned)__float_as_int(v - thr), 31);
                    }
                }
                // row kb -> bit 63 - kb
                lo = (nlo == 32) ? lo : (lo << (32 - nlo));
                hi = (nhi == 0) ? 0u : ((nhi == 32) ? hi : (hi << (32 - nhi)));
                unsigned long long mask = (((unsigned long long)lo << 32) | hi) & (~0ull << (64 - kmax));
                if (!act) mask = 0ull;

                // ---- phase 2: the reference's test and law on every candidate that passed
                const int etop = sl * SL + 63;
                auto walk = [&](auto noimg_t) {
                  constexpr bool NOIMG = decltype(noimg_t)::value;
                  while (mask) {
                    const int bit = __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const int e = etop - bit;
                    // the entry's double position (24-bit multiply for the byte offset)
                    const double* pj = reinterpret_cast<const double*>(
                        reinterpret_cast<const char*>(sdw) + __umul24((unsigned)e, 24u));
                    double dx = xi - pj[0], dy = yi - pj[1], dz = zi - pj[2];
                    if (!NOIMG) {
                        const int q = sq[wid][e];
                        dx = dx - tb_shd[wid][q][0];
                        dy = dy - tb_shd[wid][q][1];
                        dz = dz - tb_shd[wid][q][2];
                    }
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (r2 < rc2d && r2 > 0.0) {
                        ++nct;
                        if (LAW == LAW_LJ) {
                            // s = (sigma / r)^6; W = 48 s^2 - 24 s - 2 beta u, F = W sep / r^2,
                            // accumulated as g / 24 = s (2 s - 1) and F / 24 = (g / 24) / r^2 +
                            // cw / 24 (the 24 is restored per particle)
                            const double inv = rcp_f64(r2);
                            const double t = UNIT ? inv : inv * sig2;   // UNIT: sigma = 1, a multiply by 1
                            const double s6 = t * t * t;
                            const double g = s6 * __builtin_fma(2.0, s6, -1.0);
                            ga += g;
                            s6a += s6;
                            r2a += r2;
                            const double fm = __builtin_fma(g, inv, cwd24);
                            fx = __builtin_fma(fm, dx, fx);
                            fy = __builtin_fma(fm, dy, fy);
                            fz = __builtin_fma(fm, dz, fz);
                        } else {
                            pair_fp64<LAW>(S, dx, dy, dz, r2, fx, fy, fz, pe, w);
                        }
                    }
                  }
                };
                if (noimg) walk(std::integral_constant<bool, true>{});
                else walk(std::integral_constant<bool, false>{});
                wave_lds_sync();
            }
            if (LAW == LAW_LJ) {
                // V = 4 s^2 - 4 s + (alpha - beta) + beta u = 2 g' - 2 s + ... and
                // W = 48 s^2 - 24 s - 2 beta u = 24 g' - 2 beta u (g' = g / 24),
                // times epsilon (physics/lennard_jones_force.cpp:107-121)
                const double u_sum = r2a * S.inv_rc2s;
                pe = S.eps * (2.0 * ga - 2.0 * s6a + (S.alpha - S.beta) * (double)nct + S.beta * u_sum);
                w = S.eps * (24.0 * ga - 2.0 * S.beta * u_sum);
                const double fs = 24.0 * S.eps;
                fx *= fs; fy *= fs; fz *= fs;
            }
            // ---- sum the slices of each particle in a fixed order
            sdw[4 * lane] = fx;
            sdw[4 * lane + 1] = fy;
            sdw[4 * lane + 2] = fz;
            sdw[4 * lane + 3] = (double)nct;
            wave_lds_sync();
            if (act && sl == 0) {
                for (int t = 1; t < nsl; ++t) {
                    const int o = 4 * (il + t * ni);
                    fx += sdw[o];
                    fy += sdw[o + 1];
                    fz += sdw[o + 2];
                    nct += (int)sdw[o + 3];
                }
            }
            bool owner = act && sl == 0;
            if constexpr (NWC > 1) {
                // the waves' partial forces of each particle, added in wave order
                if (owner) {
                    comb[sub][il][0] = fx;
                    comb[sub][il][1] = fy;
                    comb[sub][il][2] = fz;
                    comb[sub][il][3] = (double)nct;
                }
                __syncthreads();
                if (owner && sub == 0) {
                    fx = comb[0][il][0]; fy = comb[0][il][1]; fz = comb[0][il][2];
                    nct = (int)comb[0][il][3];
                    for (int w2 = 1; w2 < NWC; ++w2) {
                        fx += comb[w2][il][0];
                        fy += comb[w2][il][1];
                        fz += comb[w2][il][2];
                        nct += (int)comb[w2][il][3];
                    }
                }
                __syncthreads();   // comb is written again by the next chunk
                owner = owner && sub == 0;
            }
            if (owner) {
                if (p.ncount) p.ncount[si] = nct;
                if (KICK != 2) { p.force.x[si] = fx; p.force.y[si] = fy; p.force.z[si] = fz; }
                double vx = p.vel.x[si], vy = p.vel.y[si], vz = p.vel.z[si];
                const double ddx = p.disp.x[si], ddy = p.disp.y[si], ddz = p.disp.z[si];
                if constexpr (RES)
                    res_kick_drift<false>(p, S, c, si, fx, fy, fz, vx, vy, vz, ddx, ddy, ddz);
                else if (KICK)
                    kick_store(p, S, si, fx, fy, fz, vx, vy, vz, KICK == 2);
                wacc[0] += (vx * vx + vy * vy) + vz * vz;
                wacc[3] += (ddx * ddx + ddy * ddy) + ddz * ddz;
            }
            if (act) {
                wacc[1] += (1. / 2.) * pe;
                wacc[2] += (1. / 2.) * w;
            }
        }
    }
    double out[4];
    block_sum4_lds<NWC>(wacc, sdw, out);
    if (threadIdx.x == 0) {
        const int c = S.cell_off + lc;
#pragma unroll
        for (int q = 0; q < 4; ++q) p.partial_c[4 * (size_t)c + q] = out[q];
        if (KICK && lc == 0) p.pending[k] = 1;   // the step awaits its finalize
    }
}

// ---------------------------------------------------------------------------
// 7. per-system reduction of the block sums into a thermodynamics record
//    (physics/measurements.hpp:59-70; time += dt as engine/integrator.cpp:85)
//    mode 0: after a step; 1: after a force-only calculation; 2: measure only
// ---------------------------------------------------------------------------

// Reduces system k's block sums into its record (one NT-thread block).
// Mode 0 runs only while the system's step is pending and clears the flag.
template <int NT>
__device__ void finalize_range(const FinalArgs& f, int k, const double* part, int nb_all)
{
    const SysDev& S = f.sys[k];
    const int pend = f.mode == 0 ? f.pending[k] : 1;
    double acc[4] = {0, 0, 0, 0};
    const int nb = pend ? nb_all : 0;
#pragma unroll 16
    for (int b = threadIdx.x; b < nb; b += NT) {
        const double4 v = reinterpret_cast<const double4*>(part)[b];
        acc[0] += v.x; acc[1] += v.y; acc[2] += v.z; acc[3] += v.w;
    }
    double out[4];
    block_sum4<NT>(acc, out);
    if (threadIdx.x != 0 || !pend) return;
    if (f.mode == 0) {
        f.pending[k] = 0;
        if (!S.active) return;
    }
    double pe = f.pe[k], w = f.vir[k];
    if (f.mode != 2 && S.law != LAW_NONE) {
        pe = out[1];
        w = out[2];
        f.pe[k] = pe;
        f.vir[k] = w;
    }
    double t = f.time[k];
    if (f.mode == 0) {
        t = t + S.dt;
        f.time[k] = t;
    }
    const double ke = (1. / 2.) * out[0];
    double* r;
    if (f.mode == 0) {
        int st = f.step[k];
        f.step[k] = st + 1;
        if (st >= f.cap) return;
        r = f.records + 7 * ((size_t)st * f.nsys + k);
    } else if (f.mode == 2) {
        r = f.measure + 7 * k;
    } else {
        return;
    }
    r[0] = t;
    r[1] = ke;
    r[2] = pe;
    r[3] = pe + ke;
    r[4] = w;
    r[5] = (2. / 3.) * ke / (double)S.n;
    r[6] = out[3] / (double)S.n;
}

// Reduces system k's block sums (pair-kernel partials) into its record.
// System k's block sums: cell-space for tiled pair forces, else particle-space.
__device__ __forceinline__ const double* block_sums(const FinalArgs& f, const SysDev& S, int& nb)
{
    const bool cells = f.cell_blocks && S.tiled;
    nb = cells ? S.ncell : S.nblk;
    return cells ? f.partial_c + 4 * (size_t)S.cell_off : f.partial + 4 * (size_t)S.blk_first;
}

// First-level parts of system k's record sums: one per kPartSums partial
// sums (four per thread of a part's block), from system k's own count only,
// so that a system's records do not depend on the other systems of its engine.
__device__ int parts_of(const FinalArgs& f, int k)
{
    int nb;
    (void)block_sums(f, f.sys[k], nb);
    return max(1, min(kMaxParts, (nb + kPartSums - 1) / kPartSums));
}

// First level: part r of system k's parts_of(f, k) contiguous parts of its
// block sums, summed by one kBlock-thread block into f.partial2 (blocks with
// r beyond the system's parts do nothing).
__device__ void finalize_part(const FinalArgs& f, int k, int r)
{
    const int nfk = parts_of(f, k);
    if (r >= nfk) return;
    int nb;
    const double* part = block_sums(f, f.sys[k], nb);
    const int chunk = (nb + nfk - 1) / nfk;
    const int lo = min(nb, r * chunk), hi = min(nb, lo + chunk);
    double acc[4] = {0, 0, 0, 0};
#pragma unroll 8
    for (int b = lo + (int)threadIdx.x; b < hi; b += kBlock) {
        const double4 v = reinterpret_cast<const double4*>(part)[b];
        acc[0] += v.x; acc[1] += v.y; acc[2] += v.z; acc[3] += v.w;
    }
    double out[4];
    block_sum4<kBlock>(acc, out);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) f.partial2[4 * ((size_t)k * f.nf + r) + q] = out[q];
    }
}

// The two levels of the key pass's and the scan's record blocks as launches
// of their own, for records finished outside a cell build (the same sums in
// the same order, so a record does not depend on where a run was split).
__global__ __launch_bounds__(kBlock) void k_finalize_parts(FinalArgs f)
{
    finalize_part(f, blockIdx.x / f.nf, blockIdx.x % f.nf);
}

__global__ __launch_bounds__(kBlock) void k_finalize_reduced(FinalArgs f)
{
    const int k = blockIdx.x;
    finalize_range<kBlock>(f, k, f.partial2 + 4 * (size_t)k * f.nf, parts_of(f, k));
}

// Start of a step call: clears the n doubles of its record rows (rows of
// frozen systems are not written by the steps and read back as zeros) and
// resets every system's record counter.
__global__ __launch_bounds__(kBlock) void k_begin_call(double* v, long long n, int* step, int nsys)
{
    const long long t0 = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long k = t0; k < n; k += stride) v[k] = 0.0;
    for (long long k = t0; k < nsys; k += stride) step[k] = 0;
}

// The end of a step call: its n record doubles and, when given, the resident
// overflow flag (as the double after them) into `out`, pinned host memory
// written through the device's mapping; one launch in place of two copies to
// the host.
__global__ __launch_bounds__(kBlock) void k_records_out(const double* rec, long long n,
                                                        const int* flag, double* out)
{
    const long long t0 = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long k = t0; k < n; k += stride) out[k] = rec[k];
    if (t0 == 0) out[n] = flag ? (double)*flag : 0.0;
}

// Copies every system's time, PE and W (n each) to ck: the clocks of a
// resident snapshot, put back if a later call must be redone.
__global__ __launch_bounds__(kBlock) void k_res_checkpoint(FinalArgs f, double* ck)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= f.nsys) return;
    ck[k] = f.time[k];
    ck[f.nsys + k] = f.pe[k];
    ck[2 * f.nsys + k] = f.vir[k];
}

// v *= scale for the slots of one system
__global__ __launch_bounds__(kBlock) void k_scale_velocities(StepArgs p, int sysk, double scale)
{
    const SysDev& S = p.sys[sysk];
    const int s = S.atom_off + blockIdx.x * kBlock + threadIdx.x;
    if (s - S.atom_off >= S.n) return;
    p.vel.x[s] *= scale; p.vel.y[s] *= scale; p.vel.z[s] *= scale;
}

// ---------------------------------------------------------------------------
// parity hooks and observers
// ---------------------------------------------------------------------------

// Cell coordinates of every particle of one system, written by particle id.
__global__ __launch_bounds__(kBlock) void k_cell_keys(StepArgs p, int sysk, int* keys, int* clamped)
{
    const SysDev& S = p.sys[sysk];
    const int s = S.atom_off + blockIdx.x * kBlock + threadIdx.x;
    if (s - S.atom_off >= S.n) return;
    const double r[3] = {p.pos.x[s], p.pos.y[s], p.pos.z[s]};
    const int i = p.id[s];
    int nclamp = 0;
    for (int a = 0; a < 3; ++a) {
        int k = (int)floor(r[a] * S.q[a]);
        if (k < 0 || k >= S.nc[a]) ++nclamp;
        keys[3 * i + a] = cell_axis(r[a], S.q[a], S.nc[a]);
    }
    if (nclamp) atomicAdd(clamped, nclamp);
}

// In-cutoff pairs (first id < second id) of one system.
__global__ __launch_bounds__(kBlock) void k_pairs_emit(StepArgs p, int sysk, long long cap,
                                                       int* idx, double* sep,
                                                       unsigned long long* counter)
{
    const SysDev& S = p.sys[sysk];
    const int s = S.atom_off + blockIdx.x * kBlock + threadIdx.x;
    if (s - S.atom_off >= S.n) return;
    const int me = p.id[s];
    for_each_neighbour_fp64(S, p, s, [&](int j, double dx, double dy, double dz, double,
                                         int ix, int iy, int iz) {
        const int other = p.id[j];
        if (me < other) {
            unsigned long long k = atomicAdd(counter, 1ull);
            if ((long long)k < cap) {
                idx[5 * k] = me; idx[5 * k + 1] = other;
                idx[5 * k + 2] = ix; idx[5 * k + 3] = iy; idx[5 * k + 4] = iz;
                sep[3 * k] = dx; sep[3 * k + 1] = dy; sep[3 * k + 2] = dz;
            }
        }
    });
}

// g(r) histogram: ordered pairs i != j (all images) with r < rmax, per system.
// Stencil of ceil(rmax / a) cells per axis around the home cell; each
// (cell, image) combination is visited once even when the stencil wraps.
__global__ __launch_bounds__(kBlock) void k_rdf(StepArgs p, double rmax, int bins,
                                                unsigned long long* hist)
{
    extern __shared__ unsigned int lh[];
    const int b = blockIdx.x;
    const int k = p.blk_sys[b];
    const SysDev& S = p.sys[k];
    for (int t = threadIdx.x; t < bins; t += kBlock) lh[t] = 0;
    __syncthreads();
    const int s = b * kBlock + threadIdx.x;
    if (s - S.atom_off < S.n && S.law != LAW_NONE) {
        int c[3];
        cell_xyz_of(p.cell_xyz, p.scell[s], c[0], c[1], c[2]);
        int R[3];
        for (int a = 0; a < 3; ++a) R[a] = (int)ceil(rmax / S.a[a]);
        const double xi = p.pos.x[s], yi = p.pos.y[s], zi = p.pos.z[s];
        const double rmax2 = rmax * rmax, inv_dr = (double)bins / rmax;
        // position inside the home cell, for the cell-box distance bound below
        const double l3[3] = {xi - c[0] * S.a[0], yi - c[1] * S.a[1], zi - c[2] * S.a[2]};
        // distance from the particle to the box of the cell at offset o along
        // axis q, less a margin far above the rounding of cell boundaries
        auto gap = [&](int q, int o) {
            const double g = o > 0 ? o * S.a[q] - l3[q] : (o < 0 ? l3[q] - (o + 1) * S.a[q] : 0.0);
            return fmax(g - 1e-9, 0.0);
        };
        // Half shell: every stencil offset is a distinct (cell, image), and the
        // pair (j, i) at the mirrored offset has exactly the negated separation
        // (IEEE subtraction is sign-symmetric), so it falls in the same bin.
        // Each unordered pair is visited once -- positive offsets, and j > i in
        // the home cell -- and counted twice.
        for (int ox = -R[0]; ox <= R[0]; ++ox) {
            int tx = c[0] + ox;
            int ix = (tx >= 0) ? tx / S.nc[0] : -((-tx + S.nc[0] - 1) / S.nc[0]);
            int nbx = tx - ix * S.nc[0];
            const double gx = gap(0, ox);
            for (int oy = -R[1]; oy <= R[1]; ++oy) {
                if (ox < 0 || (ox == 0 && oy < 0)) continue;
                const double gy = gap(1, oy);
                int ty = c[1] + oy;
                int iy = (ty >= 0) ? ty / S.nc[1] : -((-ty + S.nc[1] - 1) / S.nc[1]);
                int nby = ty - iy * S.nc[1];
                for (int oz = -R[2]; oz <= R[2]; ++oz) {
                    const bool home = ox == 0 && oy == 0 && oz == 0;
                    if (ox == 0 && oy == 0 && oz < 0) continue;
                    const double gz = gap(2, oz);
                    // no particle of that cell can be within rmax
                    if ((gx * gx + gy * gy) + gz * gz >= rmax2) continue;
                    int tz = c[2] + oz;
                    int iz = (tz >= 0) ? tz / S.nc[2] : -((-tz + S.nc[2] - 1) / S.nc[2]);
                    int nbz = tz - iz * S.nc[2];
                    const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
                    const double shx = ix * S.L[0], shy = iy * S.L[1], shz = iz * S.L[2];
                    for (int j = home ? s + 1 : p.start[nb]; j < p.start[nb + 1]; ++j) {
                        const double dx = (xi - p.pos.x[j]) - shx;
                        const double dy = (yi - p.pos.y[j]) - shy;
                        const double dz = (zi - p.pos.z[j]) - shz;
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        if (r2 < rmax2) {
                            int bin = (int)(sqrt(r2) * inv_dr);
                            if (bin < bins) atomicAdd(&lh[bin], 2u);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < bins; t += kBlock)
        if (lh[t]) atomicAdd(&hist[(size_t)k * bins + t], (unsigned long long)lh[t]);
}


// ===========================================================================
// Adjoint (reverse-mode) kernels for the FP64 LJ trajectory of one system.
//
// With s = r^2, x^2 = s / sigma^2, S6 = x^-6, u = s / (sigma^2 rc^2):
//   pair energy  e(s) = eps [4 S6 (S6 - 1) + alpha + beta (u - 1)]
//   pair virial  w(s) = eps [24 S6 (2 S6 - 1) - 2 beta u]
//   force        f = g(s) sep,  g = w / s
//   w'(s) = eps [-72 S6 (4 S6 - 1) / s - 2 beta / (sigma^2 rc^2)],  g' = (w' - g) / s
// so the force Jacobian acting on a per-particle vector phi is
//   (J phi)_i = sum_j g (phi_i - phi_j) + 2 g' (sep . (phi_i - phi_j)) sep,
// grad_i W = sum_j 2 w' sep, and the parameter derivatives are
//   d/d eps: e / eps, w / eps, f / eps;
//   d/d sigma: de = w / sigma, dw = eps [144 S6 (4 S6 - 1) + 4 beta u] / sigma,
//              df = (dw / s) sep.
// Vectors indexed by particle id live in three planes of length n.
// ===========================================================================
struct AdjArgs {
    StepArgs p;
    int sys;
    int n;
    const double* phi;     // [3][n] id order: force adjoint of this time
    double* jphi;          // [3][n] out: J phi
    double* gradw;         // [3][n] out: grad W
    double* force;         // [3][n] out: F
    double* part;          // per block: sum df/deps.phi, sum df/dsig.phi, U, W, sum dw/dsig / 2
};

// One particle per kAdjLanes lanes, lane q walking neighbour cell q of the
// 27 (reference lattice-image rule, the pair set of for_each_neighbour_fp64),
// so that a system of a few thousand particles still fills the GPU; the
// lanes' sums are combined by a fixed butterfly.
constexpr int kAdjLanes = 32;
constexpr int kAdjPerBlock = kBlock / kAdjLanes;   // particles per block

__global__ __launch_bounds__(kBlock) void k_adjoint_pair(AdjArgs a)
{
    const StepArgs& p = a.p;
    const SysDev& S = p.sys[a.sys];
    const int q = threadIdx.x & (kAdjLanes - 1);
    const int g = threadIdx.x / kAdjLanes;
    const int pi = blockIdx.x * kAdjPerBlock + g;
    const bool active = pi < S.n;
    const int s = S.atom_off + (active ? pi : 0);
    const int n = a.n;
    double v[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // J phi, grad W, F, 5 sums
    if (active && q < 27) {
        const int i = p.id[s];
        const double pxi = a.phi[i], pyi = a.phi[n + i], pzi = a.phi[2 * n + i];
        const double eps = S.eps, sig = sqrt(S.sigma2);
        int cx, cy, cz;
        cell_xyz_of(p.cell_xyz, p.scell[s], cx, cy, cz);
        const int ox = q / 9 - 1, oy = (q / 3) % 3 - 1, oz = q % 3 - 1;
        int nbx, ix, nby, iy, nbz, iz;
        nb_axis(cx, ox, S.nc[0], nbx, ix);
        nb_axis(cy, oy, S.nc[1], nby, iy);
        nb_axis(cz, oz, S.nc[2], nbz, iz);
        const double shx = ix * S.L[0], shy = iy * S.L[1], shz = iz * S.L[2];
        const int nb = S.cell_off + (nbx * S.nc[1] + nby) * S.nc[2] + nbz;
        const double xi = p.pos.x[s], yi = p.pos.y[s], zi = p.pos.z[s];
        const int j1 = p.start[nb + 1];
        for (int js = p.start[nb]; js < j1; ++js) {
            if (js == s) continue;
            const double dx = (xi - p.pos.x[js]) - shx;
            const double dy = (yi - p.pos.y[js]) - shy;
            const double dz = (zi - p.pos.z[js]) - shz;
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            if (!(r2 < S.rc2)) continue;
            const int j = p.id[js];
            const double inv = 1.0 / r2;
            const double t = S.sigma2 * inv;
            const double s6 = t * t * t;
            const double u = r2 * S.inv_rc2s;
            const double e = eps * (4.0 * s6 * (s6 - 1.0) + S.alpha + S.beta * (u - 1.0));
            const double w = eps * (24.0 * s6 * (2.0 * s6 - 1.0) - 2.0 * S.beta * u);
            const double gg = w * inv;
            const double wp = eps * (-72.0 * s6 * (4.0 * s6 - 1.0) * inv - 2.0 * S.beta * S.inv_rc2s);
            const double gp = (wp - gg) * inv;
            const double ddx = pxi - a.phi[j], ddy = pyi - a.phi[n + j], ddz = pzi - a.phi[2 * n + j];
            const double sd = dx * ddx + dy * ddy + dz * ddz;
            v[0] += gg * ddx + 2.0 * gp * sd * dx;
            v[1] += gg * ddy + 2.0 * gp * sd * dy;
            v[2] += gg * ddz + 2.0 * gp * sd * dz;
            v[3] += 2.0 * wp * dx; v[4] += 2.0 * wp * dy; v[5] += 2.0 * wp * dz;
            v[6] += gg * dx; v[7] += gg * dy; v[8] += gg * dz;
            const double sphi = dx * pxi + dy * pyi + dz * pzi;
            const double dw_dsig = eps * (144.0 * s6 * (4.0 * s6 - 1.0) + 4.0 * S.beta * u) / sig;
            v[9] += (gg / eps) * sphi;
            v[10] += dw_dsig * inv * sphi;
            v[11] += 0.5 * e;
            v[12] += 0.5 * w;
            v[13] += 0.5 * dw_dsig;
        }
    }
    // every lane of the particle ends with the same sums (each butterfly
    // stage adds the same two values on both sides)
#pragma unroll
    for (int k = 0; k < 14; ++k)
#pragma unroll
        for (int m = kAdjLanes / 2; m > 0; m >>= 1) v[k] += __shfl_xor(v[k], m, kAdjLanes);
    if (active && q == 0) {
        const int i = p.id[s];
        a.jphi[i] = v[0]; a.jphi[n + i] = v[1]; a.jphi[2 * n + i] = v[2];
        a.gradw[i] = v[3]; a.gradw[n + i] = v[4]; a.gradw[2 * n + i] = v[5];
        a.force[i] = v[6]; a.force[n + i] = v[7]; a.force[2 * n + i] = v[8];
    }
    __shared__ double red[5][kAdjPerBlock];
    if (q == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) red[k][g] = v[9 + k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double t = 0;
        for (int w = 0; w < kAdjPerBlock; ++w) t += red[threadIdx.x][w];
        a.part[5 * blockIdx.x + threadIdx.x] = t;
    }
}

// Copies the state of one system from its slots to id order (pos, vel, and
// the displacements when disp is not null).
__global__ __launch_bounds__(kBlock) void k_to_id(StepArgs p, int sysk, int n, double* pos,
                                                  double* vel, double* disp)
{
    const SysDev& S = p.sys[sysk];
    const int s = S.atom_off + blockIdx.x * kBlock + threadIdx.x;
    if (s - S.atom_off >= S.n) return;
    const int i = p.id[s];
    pos[i] = p.pos.x[s]; pos[n + i] = p.pos.y[s]; pos[2 * n + i] = p.pos.z[s];
    vel[i] = p.vel.x[s]; vel[n + i] = p.vel.y[s]; vel[2 * n + i] = p.vel.z[s];
    if (disp) { disp[i] = p.disp.x[s]; disp[n + i] = p.disp.y[s]; disp[2 * n + i] = p.disp.z[s]; }
}

// Loads one system from id-ordered arrays into its slots (slot atom_off + i
// holds particle i); displacements come from disp (cleared when it is null),
// forces are cleared.
__global__ __launch_bounds__(kBlock) void k_from_id(StepArgs p, int sysk, int n, const double* pos,
                                                    const double* vel, const double* disp)
{
    const SysDev& S = p.sys[sysk];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= S.n) return;
    const int s = S.atom_off + i;
    p.pos.x[s] = pos[i]; p.pos.y[s] = pos[n + i]; p.pos.z[s] = pos[2 * n + i];
    p.vel.x[s] = vel[i]; p.vel.y[s] = vel[n + i]; p.vel.z[s] = vel[2 * n + i];
    if (disp) {
        p.disp.x[s] = disp[i]; p.disp.y[s] = disp[n + i]; p.disp.z[s] = disp[2 * n + i];
    } else {
        p.disp.x[s] = 0; p.disp.y[s] = 0; p.disp.z[s] = 0;
    }
    p.force.x[s] = 0; p.force.y[s] = 0; p.force.z[s] = 0;
    p.id[s] = i;
}

// Reverse velocity Verlet, before the pair kernel at time t:
//   phi = (dt/2) (lam_v + cT v_t) + pend
__global__ __launch_bounds__(kBlock) void k_adj_pre(int n3, double hdt, double cT, const double* lam_v,
                                                    const double* v, const double* pend, double* phi)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n3) return;
    phi[k] = hdt * (lam_v[k] + cT * v[k]) + pend[k];
}

// ... and after it:
//   g = lam_r + c1 (-F + cW gradW) + J phi        (dL/dr_t)
//   h = lam_v + cT v_t + dt g                      (dL/d v_half)
//   lam_v <- h, pend <- (dt/2) h, lam_r <- g
__global__ __launch_bounds__(kBlock) void k_adj_post(int n3, double dt, double c1, double cW,
                                                     double cT, const double* F, const double* gw,
                                                     const double* jphi, const double* v,
                                                     double* lam_r, double* lam_v, double* pend)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n3) return;
    const double g = lam_r[k] + c1 * (-F[k] + cW * gw[k]) + jphi[k];
    const double h = lam_v[k] + cT * v[k] + dt * g;
    lam_r[k] = g;
    lam_v[k] = h;
    pend[k] = (0.5 * dt) * h;
}


// Accumulates one adjoint step's parameter terms (fixed-order block sums):
//   acc[0] += sum df/deps.phi + c1 (U / eps + cW W / eps)
//   acc[1] += sum df/dsig.phi + c1 (W / sig + cW dW/dsig)
__global__ __launch_bounds__(kBlock) void k_adj_acc(const double* part, int nb, double c1, double cW,
                                                    double eps, double sig, double* acc)
{
    double v[4] = {0, 0, 0, 0};
    double v4 = 0;
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        v[0] += part[5 * b]; v[1] += part[5 * b + 1]; v[2] += part[5 * b + 2];
        v[3] += part[5 * b + 3]; v4 += part[5 * b + 4];
    }
    double out[4];
    block_sum4(v, out);
    __shared__ double s4[kBlock / 64];
    const double t4 = wave_sum(v4);
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = t4;
    __syncthreads();
    if (threadIdx.x == 0) {
        double dws = 0;
        for (int w = 0; w < kBlock / 64; ++w) dws += s4[w];
        const double U = out[2], W = out[3];
        acc[0] += out[0] + c1 * (U / eps + cW * W / eps);
        acc[1] += out[1] + c1 * (W / sig + cW * dws);
    }
}

}  // namespace ljk
