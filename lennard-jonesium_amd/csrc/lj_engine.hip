// lj_engine.hip -- host side of the C ABI declared in include/lj_engine.h.
//
// One engine = device buffers + one HIP stream for one or more independent
// periodic systems.  A step is the kernel chain
//   integrate+key -> scan tiles -> scan apply -> scatter -> sort/reorder
//   -> pair forces + second kick + sums -> finalize
// (see lj_kernels.hpp); the chain for two steps (the sort alternates between
// two buffer sets) is captured once into a HIP graph and replayed.

#include "lj_engine.h"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lj_kernels.hpp"

using ljk::kBlock;
using ljk::kScanTile;
using ljk::SysDev;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(LJ_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

int round_up(int x, int m) { return (x + m - 1) / m * m; }

template <class T>
int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc((void**)p, count * sizeof(T)));
    return LJ_OK;
}

struct DevBuf {
    double* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

constexpr int kMaxSlots = 1 << 30;   // particle slots, and cells, of one engine

struct KernelTimer {
    const char* name;
    double total_ms = 0;
    long long calls = 0;
};

// Record rows (one lj_thermo per step and system) of lj_engine_step_noalloc,
// which allocates nothing: a fixed array on every device the library runs on,
// shared by all engines of that device and kept until the process ends.
constexpr int kScratchRows = LJ_NOALLOC_MAX_ROWS;

// Record doubles a step call hands back through its engine's pinned buffer
// (about 2,300 rows: a 20-step call of up to 117 systems); longer record
// sets are copied to the host directly.
constexpr long long kHostOut = 16384;

}  // namespace

static __device__ __align__(256) double scratch[kScratchRows * 7];

namespace {

// Device address of `scratch` on each device, taken once per device (with
// that device current) by hipGetSymbolAddress: the array's name in host code
// is a host-side stand-in, not a device address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_address(int dev, double** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxDevices) return fail(LJ_E_HIP, "device index beyond 64");
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t he = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !ptr)
            return fail(LJ_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he));
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch[dev];
    return LJ_OK;
}

}  // namespace

struct lj_engine {
    int nsys = 0;
    int precision = 0;
    bool has_cells = true;      // false when every system uses LJ_LAW_NONE
    std::vector<SysDev> sys;
    int nslot = 0, nblk = 0;
    int ncell_total = 0, ntile = 0;
    int ncblk = 0;                  // cell-space blocks (tiled pair kernel)
    int nsblk = 0;                  // sort blocks (groups of sort_k cells per wave)
    int cap_cblk = 0, cap_sblk = 0;  // allocated sizes of the block maps (every shape fits)
    int cells_per_block = ljk::kCellsPerBlockMixed;   // mixed tiled kernel: waves per workgroup
    int sort_wave = 44;             // sort: target particles per wave
    int pair_variant = LJ_PAIR_CELL_TILED;
    bool unit_sigma = false;        // every system has sigma = 1 (the tiled kernels skip a multiply)
    bool last_pair_force = false;   // the last pair launch computed forces (not sums only)
    int n_tiled = 0;                // systems whose pair forces come from the tiled kernel
    int n_multi = 0;                // ... from the tiled kernel with several waves per cell
    double split_occupancy = 6.0;   // LJ_PAIR_BY_OCCUPANCY: tiled at >= this many particles per cell
    int* d_cblk_list = nullptr;     // cell-space blocks of the tiled systems (when split)
    int* d_blk_list = nullptr;      // particle-space blocks of the other systems (when split)
    int* d_lblk_list = nullptr;     // cell-space blocks of the several-wave systems (when split)
    int n_cblk_list = 0, n_blk_list = 0, n_lblk_list = 0;

    SysDev* d_sys = nullptr;
    int* d_blk_sys = nullptr;
    int* d_cblk_sys = nullptr;
    int* d_sblk_sys = nullptr;
    double* d_pos[2] = {nullptr, nullptr};
    double* d_vel[2] = {nullptr, nullptr};
    double* d_disp[2] = {nullptr, nullptr};
    int* d_id[2] = {nullptr, nullptr};
    double* d_force = nullptr;
    int* d_scell = nullptr;
    int* d_members = nullptr;    // [cell][cell_cap] first arrivals of each cell
    int2* d_ovf = nullptr;       // (slot, cell) of later arrivals
    int* d_novf = nullptr;       // overflow entries: [0] this key pass, [1] the last scanned one
    int cell_cap = 0;            // member records per cell
    float4* d_lpos = nullptr;
    int *d_count = nullptr, *d_count_init = nullptr, *d_start = nullptr, *d_tile = nullptr;
    int* d_cell_xyz = nullptr;   // per cell: x | y << 10 | z << 20
    double* d_partial = nullptr;     // particle-space block sums
    double* d_partial_c = nullptr;   // per-cell sums (tiled pair kernels)
    double *d_time = nullptr, *d_pe = nullptr, *d_vir = nullptr;
    int* d_step = nullptr;
    int* d_pending = nullptr;   // per system: a step awaits its finalize
    double* d_partial2 = nullptr;   // [nsys][kMaxParts][4] first-level record sums
    double* d_records = nullptr;
    int rec_cap = 0;
    // record rows the step kernels write: d_records / rec_cap unless an
    // entry point points them elsewhere for the duration of its call
    double* rec_buf = nullptr;
    int rec_rows = 0;
    int device = 0;               // the HIP device current at create
    double* scratch = nullptr;    // that device's address of the scratch record rows
    double* d_measure = nullptr;
    unsigned long long* d_rdf = nullptr;
    int rdf_bins = 0, rdf_samples = 0;
    double rdf_rmax = 0;

    // Pinned host memory for a call's records and flag (kHostOut doubles and
    // one more), written by k_records_out through its device address h_out_d
    double* h_out = nullptr;
    double* h_out_d = nullptr;
    int* d_ncount = nullptr;    // set only while a neighbour-count hook runs
    int cur = 0;                // which buffer set holds the current state
    bool sorted = false;        // slots are in cell order (after a cell build)
    // Inside a chained step sequence: the stored velocities already carry the
    // next step's first half kick and the stored forces are stale.  False
    // between API calls.
    bool prekicked = false;
    hipStream_t stream = nullptr;
    bool use_graphs = true;
    // graphs[s][l]: 2^(l+1) steps starting from buffer set s (2, 4, ..., 32)
    hipGraphExec_t graphs[2][5] = {};
    int graph_cap = -1;               // record capacity baked into the graph
    const double* graph_rec = nullptr;   // record buffer baked into the graph

    // Resident cells (chained steps without a cell build; lj_kernels.hpp
    // section 4b): buffers allocated on first use, graphs of resident steps.
    int res_mode = LJ_RESIDENT_AUTO;
    bool res_ready = false;
    int res_nslot = 0;                 // kResCap slots per real cell
    int res_nfix = 0;                  // fix-up blocks (the listed cells, one wave each)
    double* r_pos[2] = {nullptr, nullptr};
    double* r_vel[2] = {nullptr, nullptr};
    double* r_disp[2] = {nullptr, nullptr};
    int* r_id[2] = {nullptr, nullptr};
    float4* r_lpos[2] = {nullptr, nullptr};
    int *r_base = nullptr, *r_count = nullptr, *r_end = nullptr, *r_dirty = nullptr;
    int *r_narr = nullptr, *r_flag = nullptr, *d_cell_sys = nullptr;
    ljk::ResMover* r_arr = nullptr;
    double* r_force = nullptr;         // forces of the last step of a call that ended resident
    double* r_ck = nullptr;            // time, PE, W of the snapshot (for a fallback)
    hipGraphExec_t rgraphs[2][5] = {}; // resident steps: 2^(l+1) steps from parity s
    long long res_calls = 0, res_fallbacks = 0;
    // Live resident cells: the resident buffers of parity res_par hold the
    // current state (the last call ended there, its forces in r_force).  The
    // packed set `cur` and d_force then still hold the snapshot the resident
    // calls started from, res_since steps earlier, with its clocks in r_ck.
    bool res_live = false;
    int res_par = 0;
    long long res_since = 0;

    bool profiling = false;
    std::vector<KernelTimer> timers;
    // Profiling events: pairs (2k, 2k+1) bracket the k-th pending launch;
    // pending[k] is its timer slot.  Read back only when flushed.
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> pending;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;   // events of the launch being timed
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    ljk::StepArgs args(int from) const
    {
        ljk::StepArgs p;
        p.sys = d_sys;
        p.blk_sys = d_blk_sys;
        p.cblk_sys = d_cblk_sys;
        p.sblk_sys = d_sblk_sys;
        auto planes = [&](double* base) {
            return ljk::Planes{base, base + nslot, base + 2 * (size_t)nslot};
        };
        p.pos = planes(d_pos[from]);
        p.vel = planes(d_vel[from]);
        p.disp = planes(d_disp[from]);
        p.force = planes(d_force);
        p.pos_n = planes(d_pos[from ^ 1]);
        p.vel_n = planes(d_vel[from ^ 1]);
        p.disp_n = planes(d_disp[from ^ 1]);
        p.id = d_id[from];
        p.id_n = d_id[from ^ 1];
        p.count = d_count;
        p.count_init = d_count_init;
        p.start = d_start;
        p.tile_sum = d_tile;
        p.members = d_members;
        p.ovf = d_ovf;
        p.novf = d_novf;
        p.cell_cap = cell_cap;
        p.scell = d_scell;
        p.lpos = d_lpos;
        p.partial = d_partial;
        p.partial_c = d_partial_c;
        p.skip_tiled = 0;
        p.cblk_list = nullptr;
        p.blk_list = nullptr;
        p.lblk_list = nullptr;
        p.cpb = precision == LJ_PRECISION_MIXED ? cells_per_block : ljk::kCellsPerBlockFp64;
        p.pending = d_pending;
        p.cell_xyz = d_cell_xyz;
        p.ncount = d_ncount;
        p.ncell_total = ncell_total;
        p.nslot = nslot;
        p.cend = d_start ? d_start + 1 : nullptr;
        p.lpos_n = d_lpos;
        p.rc_base = r_base;
        p.rc_count = r_count;
        p.rc_end = r_end;
        p.rc_dirty = r_dirty;
        p.rc_narr = r_narr;
        p.rc_arr = r_arr;
        p.rc_flag = r_flag;
        p.cell_sys = d_cell_sys;
        p.rc_par = 0;
        return p;
    }

    // A resident step of parity r (record parts and pending flags): positions
    // and cell-local coordinates from set r, the next ones into set r ^ 1;
    // velocities, displacements and ids in place.  next: the fix-up's view,
    // positions and coordinates of set r ^ 1.
    ljk::StepArgs args_res(int r, bool next = false) const
    {
        ljk::StepArgs p = args(cur);
        auto planes = [&](double* base) {
            return ljk::Planes{base, base + res_nslot, base + 2 * (size_t)res_nslot};
        };
        const int a = next ? r ^ 1 : r;
        p.pos = planes(r_pos[a]);
        p.pos_n = planes(r_pos[r ^ 1]);
        p.vel = p.vel_n = planes(r_vel[0]);
        p.disp = p.disp_n = planes(r_disp[0]);
        p.id = p.id_n = r_id[0];
        p.lpos = r_lpos[a];
        p.lpos_n = r_lpos[r ^ 1];
        p.force = planes(r_force);
        p.start = const_cast<int*>(r_base);
        p.cend = r_end;
        p.pending = d_pending + (size_t)r * nsys;
        p.nslot = res_nslot;
        p.rc_par = r;
        return p;
    }

    // fargs(0) with the record parts and pending flags of parity par (the
    // packed path uses parity 0).
    ljk::FinalArgs fargs_par(int par) const
    {
        ljk::FinalArgs f = fargs(0);
        f.partial2 = d_partial2 + (size_t)par * nsys * ljk::kMaxParts * 4;
        f.pending = d_pending + (size_t)par * nsys;
        return f;
    }

    ljk::FinalArgs fargs(int mode) const
    {
        ljk::FinalArgs f;
        f.sys = d_sys;
        f.partial = d_partial;
        f.partial_c = d_partial_c;
        f.time = d_time;
        f.pe = d_pe;
        f.vir = d_vir;
        f.step = d_step;
        f.records = rec_buf ? rec_buf : d_records;
        f.measure = d_measure;
        f.nsys = nsys;
        f.cap = rec_buf ? rec_rows : rec_cap;
        f.mode = mode;
        // a step's sums come from the force kernels (per-system layout); a
        // force-only or measure call's from its own last pair launch
        f.cell_blocks = (mode == 0 ? sys[0].law != ljk::LAW_NONE : last_pair_force) ? 1 : 0;
        f.pending = d_pending;
        f.partial2 = d_partial2;
        // first-level part slots: the most parts any system uses (parts_of)
        int nbmax = 1;
        for (const SysDev& s : sys)
            nbmax = std::max(nbmax, (f.cell_blocks && s.tiled) ? s.ncell : s.nblk);
        f.nf = std::min(ljk::kMaxParts, (nbmax + ljk::kPartSums - 1) / ljk::kPartSums);
        return f;
    }

    // Grids above the single-block scan's reach scan in two launches.
    bool two_level_scan() const
    {
        return has_cells && ncell_total > ljk::kScanOneMax * 4 * ljk::kScanOneThreads;
    }

};

namespace {

// Per-kernel timing: while profiling, every launch carries a start and a stop
// event of its own (hipExtLaunchKernel), which HIP stamps at the kernel's own
// start and end, so a bracket is the kernel alone, without the dispatch gap
// in front of it.  No host wait sits between launches; flush_timers() waits
// once and accumulates the brackets.
constexpr int kMaxPendingTimers = 2048;

int flush_timers(lj_engine* e)
{
    if (e->pending.empty()) return LJ_OK;
    HIP_TRY(hipEventSynchronize(e->prof_ev[2 * (e->pending.size() - 1) + 1]));
    for (size_t k = 0; k < e->pending.size(); ++k) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e->prof_ev[2 * k], e->prof_ev[2 * k + 1]));
        KernelTimer& t = e->timers[e->pending[k]];
        t.total_ms += ms;
        t.calls += 1;
    }
    e->pending.clear();
    return LJ_OK;
}

template <class F>
int timed(lj_engine* e, int slot, const char* name, F&& launch)
{
    if (!e->profiling) {
        launch();
        return LJ_OK;
    }
    if ((int)e->timers.size() <= slot) e->timers.resize(slot + 1);
    e->timers[slot].name = name;
    if ((int)e->pending.size() == kMaxPendingTimers) {
        int rc = flush_timers(e);
        if (rc) return rc;
    }
    const size_t k = e->pending.size();
    while (e->prof_ev.size() < 2 * k + 2) {
        hipEvent_t ev = nullptr;
        HIP_TRY(hipEventCreate(&ev));
        e->prof_ev.push_back(ev);
    }
    e->ev_start = e->prof_ev[2 * k];
    e->ev_stop = e->prof_ev[2 * k + 1];
    launch();
    e->ev_start = e->ev_stop = nullptr;
    HIP_TRY(hipGetLastError());
    e->pending.push_back(slot);
    return LJ_OK;
}

// Launches kernel k on the engine stream, with the timing events when a
// timed() bracket is open.
template <typename... Params, typename... Args>
void launch(lj_engine* e, void (*k)(Params...), dim3 grid, dim3 block, Args... args)
{
    if (e->ev_start)
        hipExtLaunchKernelGGL(k, grid, block, 0, e->stream, e->ev_start, e->ev_stop, 0, args...);
    else
        hipLaunchKernelGGL(k, grid, block, 0, e->stream, args...);
}

int enqueue_finalize(lj_engine* e, int mode);

// Cell build, optionally for the integrated positions (integrate 1: first
// half kick, drift, wrap; 2: drift and wrap of velocities that already carry
// the half kick): the key pass evaluates the integrated position for the keys
// only (and records each cell's arrivals), and the sort's move into the other
// buffer set applies the same integration.  Between them: the scan.
// Leaves e->cur pointing at the sorted set.
int enqueue_cells(lj_engine* e, int integrate)
{
    const ljk::StepArgs p = e->args(e->cur);
    const dim3 gb(e->nblk), tb(kBlock);
    // a step also finishes the previous step's records (pending flags
    // decide): first-level sums in blocks of the key pass, the second level
    // in blocks of the scan
    ljk::FinalArgs f0 = e->fargs(0);
    const int nfin = integrate ? e->nsys * f0.nf : 0;
    int rc = LJ_OK;
    if (integrate == 2)
        rc = timed(e, 0, "integrate_key", [&] {
            launch(e, ljk::k_integrate_key<3, true>, dim3(e->nblk + nfin), tb, p, f0, nfin);
        });
    else if (integrate)
        rc = timed(e, 0, "integrate_key", [&] {
            launch(e, ljk::k_integrate_key<2, true>, dim3(e->nblk + nfin), tb, p, f0, nfin);
        });
    else
        rc = timed(e, 0, "integrate_key", [&] {
            launch(e, ljk::k_integrate_key<0, true>, gb, tb, p, f0, 0);
        });
    if (rc) return rc;
    const int nred = integrate ? e->nsys : 0;
    if (!e->two_level_scan()) {
        if ((rc = timed(e, 1, "scan", [&] {
                 launch(e, ljk::k_scan_one, dim3(1 + nred), dim3(ljk::kScanOneThreads), p, f0);
             })))
            return rc;
    } else {
        if ((rc = timed(e, 1, "scan_tiles", [&] {
                 launch(e, ljk::k_scan_tiles, dim3(e->ntile), tb, p);
             })))
            return rc;
        if ((rc = timed(e, 2, "scan_apply", [&] {
                 launch(e, ljk::k_scan_apply, dim3(e->ntile + nred), tb, p, f0);
             })))
            return rc;
    }
    // sort blocks: waves of sort_k cells
    const bool mixed = e->precision == LJ_PRECISION_MIXED;
    const dim3 gc(e->nsblk);
    rc = timed(e, 4, "sort", [&] {
        if (mixed) {
            if (integrate == 2) launch(e, ljk::k_sort_cells<true, 2>, gc, tb, p);
            else if (integrate) launch(e, ljk::k_sort_cells<true, 1>, gc, tb, p);
            else launch(e, ljk::k_sort_cells<true, 0>, gc, tb, p);
        } else {
            if (integrate == 2) launch(e, ljk::k_sort_cells<false, 2>, gc, tb, p);
            else if (integrate) launch(e, ljk::k_sort_cells<false, 1>, gc, tb, p);
            else launch(e, ljk::k_sort_cells<false, 0>, gc, tb, p);
        }
    });
    if (rc) return rc;
    e->cur ^= 1;
    e->sorted = true;
    return LJ_OK;
}

template <int PREC, int LAW, int KICK>
void launch_pair(lj_engine* e, const ljk::StepArgs& p)
{
    const int grid = p.blk_list ? e->n_blk_list : e->nblk;
    launch(e, ljk::k_pair<PREC, LAW, KICK>, dim3(grid), dim3(kBlock), p);
}

// The one-wave tiled kernel of a precision over `grid` blocks, for the launch
// shape and, when every system has sigma = 1, the law without the sigma^2
// multiply (bit-identical: the product with 1 is exact).
template <int PREC, int LAW, int KICK, bool RES>
void launch_tiled_shape(lj_engine* e, const ljk::StepArgs& p, dim3 grid)
{
    auto go = [&](auto unit_t) {
        constexpr bool U = decltype(unit_t)::value && LAW == ljk::LAW_LJ;
        if constexpr (PREC == ljk::MIXED) {
            if (e->cells_per_block == 2)
                launch(e, ljk::k_pair_cells_mixed<LAW, KICK, 2, 1, RES, U>, grid, dim3(128), p);
            else if (e->cells_per_block == 8)
                launch(e, ljk::k_pair_cells_mixed<LAW, KICK, 8, 1, RES, U>, grid, dim3(512), p);
            else
                launch(e, ljk::k_pair_cells_mixed<LAW, KICK, 4, 1, RES, U>, grid, dim3(256), p);
        } else {
            launch(e, ljk::k_pair_cells_fp64<LAW, KICK, 1, RES, U>, grid,
                   dim3(64 * ljk::kCellsPerBlockFp64), p);
        }
    };
    if (e->unit_sigma) go(std::true_type{});
    else go(std::false_type{});
}

template <int PREC, int LAW, int KICK>
void launch_pair_cells(lj_engine* e, const ljk::StepArgs& p)
{
    const int grid = p.cblk_list ? e->n_cblk_list : e->ncblk;
    launch_tiled_shape<PREC, LAW, KICK, false>(e, p, dim3(grid));
}

// Systems flagged 2 (small systems): the tiled kernel of their precision with
// several waves per home cell, each over a share of the neighbour cells.
template <int PREC, int LAW, int KICK>
void launch_pair_multiwave(lj_engine* e, const ljk::StepArgs& p)
{
    const int grid = (p.lblk_list ? e->n_lblk_list : e->ncblk) * p.cpb;
    if constexpr (PREC == ljk::MIXED)
        launch(e, ljk::k_pair_cells_mixed<LAW, KICK, 1, ljk::kWavesPerCellSmall>, dim3(grid),
               dim3(64 * ljk::kWavesPerCellSmall), p);
    else
        launch(e, ljk::k_pair_cells_fp64<LAW, KICK, ljk::kWavesPerCellSmall>, dim3(grid),
               dim3(64 * ljk::kWavesPerCellSmall), p);
}

template <int LAW, int KICK>
void launch_multiwave(lj_engine* e, const ljk::StepArgs& p, bool mixed)
{
    if (mixed) launch_pair_multiwave<ljk::MIXED, LAW, KICK>(e, p);
    else launch_pair_multiwave<ljk::FP64, LAW, KICK>(e, p);
}

template <int LAW, int KICK>
void launch_tiled(lj_engine* e, const ljk::StepArgs& p, bool mixed)
{
    if (mixed) launch_pair_cells<ljk::MIXED, LAW, KICK>(e, p);
    else launch_pair_cells<ljk::FP64, LAW, KICK>(e, p);
}

template <int LAW, int KICK>
void launch_particle(lj_engine* e, const ljk::StepArgs& p, bool mixed)
{
    if (mixed) launch_pair<ljk::MIXED, LAW, KICK>(e, p);
    else launch_pair<ljk::FP64, LAW, KICK>(e, p);
}

// Pair forces (or, for LJ_LAW_NONE engines / measure, just the sums): the
// tiled kernel for systems flagged tiled, the per-particle kernel for the rest.
// kick 1: second half kick; 2: that kick and the next step's first half kick
// (a chained step; forces are not stored).
int enqueue_pair(lj_engine* e, int kick, bool measure_only)
{
    ljk::StepArgs p = e->args(e->cur);
    const int law = measure_only ? ljk::LAW_NONE : e->sys[0].law;
    const bool mixed = e->precision == LJ_PRECISION_MIXED;
    e->last_pair_force = law != ljk::LAW_NONE;
    if (law == ljk::LAW_NONE)
        return timed(e, 5, "pair", [&] {
            if (kick) launch_pair<ljk::FP64, ljk::LAW_NONE, 1>(e, p);
            else launch_pair<ljk::FP64, ljk::LAW_NONE, 0>(e, p);
        });
    auto tiled = [&] {
        if (law == ljk::LAW_LJ) {
            if (kick == 2) launch_tiled<ljk::LAW_LJ, 2>(e, p, mixed);
            else if (kick) launch_tiled<ljk::LAW_LJ, 1>(e, p, mixed);
            else launch_tiled<ljk::LAW_LJ, 0>(e, p, mixed);
        } else {
            if (kick == 2) launch_tiled<ljk::LAW_CONSTANT, 2>(e, p, mixed);
            else if (kick) launch_tiled<ljk::LAW_CONSTANT, 1>(e, p, mixed);
            else launch_tiled<ljk::LAW_CONSTANT, 0>(e, p, mixed);
        }
    };
    auto particle = [&] {
        if (law == ljk::LAW_LJ) {
            if (kick == 2) launch_particle<ljk::LAW_LJ, 2>(e, p, mixed);
            else if (kick) launch_particle<ljk::LAW_LJ, 1>(e, p, mixed);
            else launch_particle<ljk::LAW_LJ, 0>(e, p, mixed);
        } else {
            if (kick == 2) launch_particle<ljk::LAW_CONSTANT, 2>(e, p, mixed);
            else if (kick) launch_particle<ljk::LAW_CONSTANT, 1>(e, p, mixed);
            else launch_particle<ljk::LAW_CONSTANT, 0>(e, p, mixed);
        }
    };
    auto multiwave = [&] {
        if (law == ljk::LAW_LJ) {
            if (kick == 2) launch_multiwave<ljk::LAW_LJ, 2>(e, p, mixed);
            else if (kick) launch_multiwave<ljk::LAW_LJ, 1>(e, p, mixed);
            else launch_multiwave<ljk::LAW_LJ, 0>(e, p, mixed);
        } else {
            if (kick == 2) launch_multiwave<ljk::LAW_CONSTANT, 2>(e, p, mixed);
            else if (kick) launch_multiwave<ljk::LAW_CONSTANT, 1>(e, p, mixed);
            else launch_multiwave<ljk::LAW_CONSTANT, 0>(e, p, mixed);
        }
    };
    // with several kernels in use, each launches only over its own systems'
    // blocks; the first one launched is timed as "pair"
    const int n_particle = e->nsys - e->n_tiled - e->n_multi;
    int rc = LJ_OK;
    bool first = true;
    if (e->n_tiled > 0) {
        if (e->n_tiled < e->nsys) p.cblk_list = e->d_cblk_list;
        rc = timed(e, 5, "pair", tiled);
        p.cblk_list = nullptr;
        first = false;
    }
    if (!rc && e->n_multi > 0) {
        if (e->n_multi < e->nsys) p.lblk_list = e->d_lblk_list;
        rc = first ? timed(e, 5, "pair", multiwave) : timed(e, 8, "pair_multiwave", multiwave);
        p.lblk_list = nullptr;
        first = false;
    }
    if (!rc && n_particle > 0) {
        if (n_particle < e->nsys) {
            p.skip_tiled = 1;
            p.blk_list = e->d_blk_list;
        }
        rc = first ? timed(e, 5, "pair", particle) : timed(e, 7, "pair_particle", particle);
    }
    return rc;
}

int enqueue_finalize(lj_engine* e, int mode)
{
    const ljk::FinalArgs f = e->fargs(mode);
    return timed(e, 6, "finalize", [&] {
        launch(e, ljk::k_finalize_parts, dim3(e->nsys * f.nf), dim3(kBlock), f);
        launch(e, ljk::k_finalize_reduced, dim3(e->nsys), dim3(kBlock), f);
    });
}

// The end of a step call: copies its nsteps record rows to `records`
// (optional) and the resident overflow flag to *flag (optional), then waits
// for the stream.  Up to kHostOut doubles go through one k_records_out launch
// into the pinned buffer; more through copies to the host.
int finish_call(lj_engine* e, int nsteps, lj_thermo* records, int* flag)
{
    const double* rec = e->fargs(0).records;
    const long long n = records ? (long long)nsteps * e->nsys * 7 : 0;
    if (n <= kHostOut) {
        if (records || flag) {
            const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(64, (n + kBlock - 1) / kBlock));
            hipLaunchKernelGGL(ljk::k_records_out, dim3(grid), dim3(kBlock), 0, e->stream, rec, n,
                               flag ? (const int*)e->r_flag : nullptr, e->h_out_d);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (records) std::memcpy(records, e->h_out, (size_t)n * sizeof(double));
        if (flag) *flag = e->h_out[n] != 0.0 ? 1 : 0;
        return LJ_OK;
    }
    if (flag) HIP_TRY(hipMemcpyAsync(flag, e->r_flag, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(records, rec, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return LJ_OK;
}

// One velocity-Verlet step.  Its record is written during the next step's
// cell build (or before the next step's integration without cells) or by
// enqueue_finalize(e, 0), which must follow the last step of a sequence
// before anything else reads records, time or the block sums.
// chain: another step follows in the same launch sequence with nothing in
// between, so this step's pair kernel may apply that step's first half kick
// (forces are then not stored).  The last step of a sequence leaves the
// state as the reference has it: full-step velocities and forces.
int enqueue_step(lj_engine* e, bool chain = false)
{
    int rc;
    if (e->has_cells) {
        if ((rc = enqueue_cells(e, e->prekicked ? 2 : 1))) return rc;
        e->prekicked = chain;
        return enqueue_pair(e, chain ? 2 : 1, false);
    } else {
        const ljk::StepArgs p = e->args(e->cur);
        if ((rc = enqueue_finalize(e, 0))) return rc;
        if ((rc = timed(e, 0, "integrate_key", [&] {
                 launch(e, ljk::k_integrate_key<1, false>, dim3(e->nblk), dim3(kBlock), p,
                        e->fargs(0), 0);
             })))
            return rc;
    }
    return enqueue_pair(e, true, false);
}

void drop_graphs(lj_engine* e);

// Grows the record buffer to hold nsteps steps.  The step graphs carry the
// buffer's address in their kernel arguments, so they are destroyed before
// the old buffer is freed (and captured again on their next use).
int ensure_records(lj_engine* e, int nsteps)
{
    if (nsteps <= e->rec_cap) return LJ_OK;
    if ((long long)nsteps * e->nsys * 7 > (1LL << 40))
        return fail(LJ_E_INVALID, "records of nsteps steps x systems exceed 8 TiB");
    int cap = std::max(nsteps, 64);
    drop_graphs(e);
    if (e->d_records) {
        HIP_TRY(hipStreamSynchronize(e->stream));
        HIP_TRY(hipFree(e->d_records));
    }
    e->d_records = nullptr;
    e->rec_cap = 0;
    int rc = dev_alloc(&e->d_records, (size_t)cap * e->nsys * 7);
    if (rc) return rc;
    e->rec_cap = cap;
    return LJ_OK;
}

// Step graphs of 2, 4, 8, 16 and 32 steps from either buffer set.  Each graph
// launch ends with a release of the caches that costs several microseconds
// once the pair kernel has left the L2 full of dirty lines, so a call replays
// the fewest launches that add up to its step count (20 steps: 16 + 4).  An
// even count returns the state to the buffer set it started from.
constexpr int kGraphLevels = 5;

// Captures nsteps steps starting from buffer set `from` into *out.
int capture_steps(lj_engine* e, int from, int nsteps, hipGraphExec_t* out)
{
    const int saved = e->cur;
    const bool saved_sorted = e->sorted;
    e->cur = from;
    e->prekicked = false;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = LJ_OK;
    for (int s = 0; s < nsteps && !rc; ++s) rc = enqueue_step(e, s + 1 < nsteps);
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    e->cur = saved;
    e->sorted = saved_sorted;
    e->prekicked = false;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return fail(LJ_E_HIP, std::string("hipStreamEndCapture: ") +
                                                    hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(LJ_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    return LJ_OK;
}

void drop_graphs(lj_engine* e)
{
    for (auto* sets : {&e->graphs, &e->rgraphs})
        for (auto& set : *sets)
            for (hipGraphExec_t& g : set) {
                if (g) (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

int build_graph(lj_engine* e)
{
    if (e->graphs[1][kGraphLevels - 1] && e->graph_cap == e->rec_cap &&
        e->graph_rec == e->d_records && !e->rec_buf)
        return LJ_OK;
    if (e->rec_buf) return fail(LJ_E_STATE, "step graphs are captured on the engine's own records only");
    drop_graphs(e);
    int rc = LJ_OK;
    for (int from = 0; from < 2 && !rc; ++from)
        for (int l = 0; l < kGraphLevels && !rc; ++l)
            rc = capture_steps(e, from, 2 << l, &e->graphs[from][l]);
    if (rc) {
        drop_graphs(e);
        return rc;
    }
    e->graph_cap = e->rec_cap;
    e->graph_rec = e->d_records;
    return LJ_OK;
}

// Cell-space blocks of the tiled pair kernel (cells_per_block cells per block
// in mixed precision, one in FP64) and sort blocks (kBlock / 64 waves of
// sort_k cells, about sort_wave particles per wave) of every system, into the
// host copies; block maps of at most cap_cblk / cap_sblk entries.
void layout_blocks(lj_engine* e, std::vector<int>& cblk_sys, std::vector<int>& sblk_sys)
{
    const int cpb = e->precision == LJ_PRECISION_MIXED ? e->cells_per_block
                                                       : ljk::kCellsPerBlockFp64;
    int cblk = 0, sblk = 0;
    cblk_sys.clear();
    sblk_sys.clear();
    for (int k = 0; k < e->nsys; ++k) {
        SysDev& s = e->sys[k];
        s.cblk_first = cblk;
        s.ncblk = std::max(1, (s.ncell + cpb - 1) / cpb);
        cblk += s.ncblk;
        cblk_sys.resize(cblk, k);
        // a wave's cells rarely hold more than its 64 lanes
        const double occ = s.ncell > 0 ? (double)s.n / s.ncell : 1.0;
        s.sort_k = (int)std::min(63.0, std::max(1.0, std::floor((double)e->sort_wave / occ)));
        s.sblk_first = sblk;
        const int waves = std::max(1, (s.ncell + s.sort_k - 1) / s.sort_k);
        sblk += (waves + kBlock / 64 - 1) / (kBlock / 64);
        sblk_sys.resize(sblk, k);
    }
    e->ncblk = cblk;
    e->nsblk = sblk;
}

// Uploads the block layout of the current launch shape (and the systems'
// parameters that carry it); the step graphs are rebuilt on their next use.
int upload_layout(lj_engine* e)
{
    std::vector<int> cblk_sys, sblk_sys;
    layout_blocks(e, cblk_sys, sblk_sys);
    if (e->ncblk > e->cap_cblk || e->nsblk > e->cap_sblk)
        return fail(LJ_E_INVALID, "launch shape needs more blocks than were allocated");
    drop_graphs(e);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(e->d_sys, e->sys.data(), e->nsys * sizeof(SysDev), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_cblk_sys, cblk_sys.data(), cblk_sys.size() * sizeof(int),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_sblk_sys, sblk_sys.data(), sblk_sys.size() * sizeof(int),
                      hipMemcpyHostToDevice));
    return LJ_OK;
}

int assign_pair_kernels(lj_engine* e, bool force_lists = false);

// Per-system pair kernel from the engine's policy: all tiled, all per-particle,
// or tiled where the mean cell occupancy N / cells reaches split_occupancy.
// Uploads the flags and block lists (also when force_lists is set, after the
// block layout changed); the step graphs are rebuilt on their next use.
int assign_pair_kernels(lj_engine* e, bool force_lists)
{
    int nt = 0, nl = 0;
    bool changed = false;
    for (SysDev& s : e->sys) {
        int t = e->pair_variant == LJ_PAIR_CELL_TILED ? 1 : 0;
        if (e->pair_variant == LJ_PAIR_CELL_MULTIWAVE) t = 2;
        if (e->pair_variant == LJ_PAIR_BY_OCCUPANCY)
            t = (s.ncell > 0 && (double)s.n >= e->split_occupancy * (double)s.ncell) ? 1 : 0;
        if (!e->has_cells || s.law == ljk::LAW_NONE) t = 0;
        changed |= t != s.tiled;
        s.tiled = t;
        nt += t == 1;
        nl += t == 2;
    }
    e->n_tiled = nt;
    e->n_multi = nl;
    if (changed || force_lists) {
        drop_graphs(e);
        std::vector<int> cl, ll, bl;
        for (const SysDev& s : e->sys) {
            if (s.tiled)
                for (int b = 0; b < s.ncblk; ++b) (s.tiled == 1 ? cl : ll).push_back(s.cblk_first + b);
            else
                for (int b = 0; b < s.nblk; ++b) bl.push_back(s.blk_first + b);
        }
        e->n_cblk_list = (int)cl.size();
        e->n_lblk_list = (int)ll.size();
        e->n_blk_list = (int)bl.size();
        HIP_TRY(hipStreamSynchronize(e->stream));
        HIP_TRY(hipMemcpy(e->d_sys, e->sys.data(), e->nsys * sizeof(SysDev), hipMemcpyHostToDevice));
        if (!cl.empty())
            HIP_TRY(hipMemcpy(e->d_cblk_list, cl.data(), cl.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!ll.empty())
            HIP_TRY(hipMemcpy(e->d_lblk_list, ll.data(), ll.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!bl.empty())
            HIP_TRY(hipMemcpy(e->d_blk_list, bl.data(), bl.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return LJ_OK;
}

int check_engine(const lj_engine* e) { return e ? LJ_OK : fail(LJ_E_INVALID, "null engine"); }

int check_sys(const lj_engine* e, int sys)
{
    if (!e) return fail(LJ_E_INVALID, "null engine");
    if (sys < 0 || sys >= e->nsys) return fail(LJ_E_INVALID, "system index out of range");
    return LJ_OK;
}

int run_steps(lj_engine* e, int nsteps, bool graphs, lj_thermo* records);
bool res_eligible(const lj_engine* e, int nsteps);
int res_ensure(lj_engine* e);
int res_settle(lj_engine* e);
int build_res_graphs(lj_engine* e);

// Ensures slots are in cell order without moving particles (needed before
// any traversal when the state was just loaded).
int ensure_sorted(lj_engine* e)
{
    if (!e->has_cells || e->sorted) return LJ_OK;
    return enqueue_cells(e, false);
}

}  // namespace

extern "C" {

const char* lj_last_error(void) { return g_err.c_str(); }

int lj_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lj_engine_create(const lj_system_desc* systems, int nsys, int precision, lj_engine** out)
{
    if (!out) return fail(LJ_E_INVALID, "out is null");
    *out = nullptr;
    if (!systems || nsys < 1) return fail(LJ_E_INVALID, "need at least one system");
    if (precision != LJ_PRECISION_FP64 && precision != LJ_PRECISION_MIXED)
        return fail(LJ_E_INVALID, "unknown precision");

    // the descriptors are checked first (host only), then the device is looked for
    auto e = new lj_engine();
    e->nsys = nsys;
    e->precision = precision;
    int nnone = 0;
    for (int k = 0; k < nsys; ++k) nnone += systems[k].force_law == LJ_LAW_NONE;
    if (nnone != 0 && nnone != nsys) {
        delete e;
        return fail(LJ_E_INVALID, "LJ_LAW_NONE systems cannot share an engine with others");
    }
    e->has_cells = nnone == 0;
    for (int k = 1; k < nsys; ++k)
        if (systems[k].force_law != systems[0].force_law) {
            delete e;
            return fail(LJ_E_INVALID, "all systems of an engine must use the same force law");
        }

    int atom_off = 0, cell_off = 0, blk = 0, cblk = 0, sblk = 0;
    std::vector<int> count_init;
    for (int k = 0; k < nsys; ++k) {
        const lj_system_desc& d = systems[k];
        SysDev s;
        std::memset(&s, 0, sizeof(s));
        if (d.particle_count < 1) { delete e; return fail(LJ_E_INVALID, "particle_count < 1"); }
        if (d.particle_count > kMaxSlots) {
            delete e;
            return fail(LJ_E_INVALID, "particle_count exceeds 2^30");
        }
        if (!(d.time_delta > 0)) { delete e; return fail(LJ_E_INVALID, "time_delta must be > 0"); }
        if (d.force_law < 0 || d.force_law > 2) { delete e; return fail(LJ_E_INVALID, "bad force_law"); }
        if (!d.periodic && d.force_law != LJ_LAW_NONE) {
            delete e;
            return fail(LJ_E_INVALID, "a force calculation needs the periodic box");
        }
        s.n = d.particle_count;
        s.active = 1;
        s.tiled = (d.force_law != LJ_LAW_NONE && e->has_cells) ? 1 : 0;   // default: all tiled
        s.law = d.force_law;
        s.periodic = d.periodic ? 1 : 0;
        s.dt = d.time_delta;
        s.eps = d.epsilon;
        s.sigma2 = d.sigma * d.sigma;
        s.strength = d.strength;
        for (int a = 0; a < 3; ++a) s.L[a] = d.box[a];
        double rc_abs = 0;
        if (d.force_law == LJ_LAW_LJ) {
            if (!(d.sigma > 0) || !(d.epsilon > 0)) {
                delete e;
                return fail(LJ_E_INVALID, "epsilon and sigma must be > 0");
            }
            if (!(d.cutoff * d.cutoff > std::cbrt(2.0))) {
                delete e;
                return fail(LJ_E_INVALID, "Cutoff distance is too short");
            }
            // physics/lennard_jones_force.cpp:68-73 on the reduced cutoff
            s.rc2_red = d.cutoff * d.cutoff;
            s.inv_rc2s = 1.0 / (s.sigma2 * s.rc2_red);
            double rcm6 = 1 / (s.rc2_red * s.rc2_red * s.rc2_red);
            s.alpha = -4.0 * rcm6 * (rcm6 - 1.0);
            s.beta = 12.0 * rcm6 * (2.0 * rcm6 - 1.0);
            rc_abs = d.cutoff * d.sigma;
        } else if (d.force_law == LJ_LAW_CONSTANT) {
            rc_abs = d.cutoff;
            s.rc = d.cutoff;
        }
        s.rc2 = rc_abs * rc_abs;
        s.rc2f = (float)s.rc2;
        s.inv_s2f = (float)(1.0 / s.sigma2);
        s.cwf = (float)(-2.0 * s.beta * s.inv_rc2s);
        s.cw24f = (float)(-2.0 * s.beta * s.inv_rc2s / 24.0);
        if (d.force_law != LJ_LAW_NONE) {
            for (int a = 0; a < 3; ++a) {
                if (!(d.box[a] >= rc_abs) || !(rc_abs > 0)) {
                    delete e;
                    return fail(LJ_E_INVALID, "Simulation box size is less than the cutoff distance");
                }
                s.nc[a] = (int)std::floor(d.box[a] / rc_abs);   // tools/cell_list_array.cpp:48
                s.q[a] = (double)s.nc[a] / d.box[a];
                s.a[a] = d.box[a] / (double)s.nc[a];
            }
            s.ncell = s.nc[0] * s.nc[1] * s.nc[2];
            for (int a = 0; a < 3; ++a) {
                if (s.nc[a] > 1023) {
                    delete e;
                    return fail(LJ_E_INVALID, "more than 1023 cells along a box side");
                }
                for (int o = -1; o <= 1; ++o)
                    for (int i = -1; i <= 1; ++i)
                        s.shf[a][3 * (o + 1) + (i + 1)] =
                            (float)((i * s.nc[a] - o) * s.a[a] - i * s.L[a]);
            }
        }
        const int slots = round_up(s.n, kBlock);
        // slot, cell and block indices are 32-bit on the device
        if ((long long)atom_off + slots > kMaxSlots ||
            (long long)cell_off + s.ncell + 1 > kMaxSlots) {
            delete e;
            return fail(LJ_E_INVALID, "the systems exceed 2^30 particle slots or cells in one engine");
        }
        s.atom_off = atom_off;
        s.blk_first = blk;
        s.nblk = slots / kBlock;
        // block maps sized for every launch shape: one cell per tiled block,
        // one cell per sort wave
        cblk += std::max(1, s.ncell);
        sblk += (std::max(1, s.ncell) + kBlock / 64 - 1) / (kBlock / 64);
        if (e->has_cells) {
            s.cell_off = cell_off;
            count_init.resize(cell_off + s.ncell + 1, 0);
            count_init[cell_off + s.ncell] = slots - s.n;   // tail cell: padding slots
            cell_off += s.ncell + 1;
        }
        atom_off += slots;
        blk += s.nblk;
        e->sys.push_back(s);
    }
    e->nslot = atom_off;
    e->unit_sigma = true;
    for (const SysDev& s : e->sys) e->unit_sigma = e->unit_sigma && s.sigma2 == 1.0;
    for (const SysDev& s : e->sys) {
        e->n_tiled += s.tiled == 1;
        e->n_multi += s.tiled == 2;
    }
    e->nblk = blk;
    e->cap_cblk = cblk;
    e->cap_sblk = sblk;
    e->ncell_total = cell_off;
    e->ntile = std::max(1, (cell_off + kScanTile - 1) / kScanTile);

    std::vector<int> blk_sys(blk), cblk_sys, sblk_sys;
    for (int k = 0; k < nsys; ++k)
        for (int b = 0; b < e->sys[k].nblk; ++b) blk_sys[e->sys[k].blk_first + b] = k;
    layout_blocks(e, cblk_sys, sblk_sys);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete e;
        return fail(LJ_E_HIP, "no HIP device visible");
    }
    auto cleanup_fail = [&](int rc) {
        lj_engine_destroy(e);
        return rc;
    };
    int rc;
    if (hipGetDevice(&e->device) != hipSuccess)
        return cleanup_fail(fail(LJ_E_HIP, "hipGetDevice failed"));
    if ((rc = scratch_address(e->device, &e->scratch))) return cleanup_fail(rc);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(LJ_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
        return cleanup_fail(fail(LJ_E_HIP, "hipEventCreate failed"));
    if (hipHostMalloc((void**)&e->h_out, (kHostOut + 1) * sizeof(double), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&e->h_out_d, e->h_out, 0) != hipSuccess)
        return cleanup_fail(fail(LJ_E_HIP, "pinned record output: hipHostMalloc failed"));
    const size_t ns = e->nslot;
    if ((rc = dev_alloc(&e->d_sys, nsys)) || (rc = dev_alloc(&e->d_blk_sys, blk)) ||
        (rc = dev_alloc(&e->d_cblk_sys, cblk)) || (rc = dev_alloc(&e->d_sblk_sys, sblk)) ||
        (rc = dev_alloc(&e->d_force, 3 * ns)) ||
        (rc = dev_alloc(&e->d_partial, 4 * (size_t)blk)) ||
        (rc = dev_alloc(&e->d_partial_c, 4 * (size_t)std::max(1, cell_off))) ||
        (rc = dev_alloc(&e->d_cblk_list, (size_t)cblk)) || (rc = dev_alloc(&e->d_blk_list, (size_t)blk)) ||
        (rc = dev_alloc(&e->d_lblk_list, (size_t)cblk)) ||
        (rc = dev_alloc(&e->d_time, nsys)) || (rc = dev_alloc(&e->d_pe, nsys)) ||
        (rc = dev_alloc(&e->d_vir, nsys)) || (rc = dev_alloc(&e->d_step, nsys)) ||
        (rc = dev_alloc(&e->d_pending, 2 * (size_t)nsys)) ||
        (rc = dev_alloc(&e->d_partial2, 2 * (size_t)nsys * ljk::kMaxParts * 4)) ||
        (rc = dev_alloc(&e->d_measure, 7 * (size_t)nsys)))
        return cleanup_fail(rc);
    for (int b = 0; b < 2; ++b) {
        if ((rc = dev_alloc(&e->d_pos[b], 3 * ns)) || (rc = dev_alloc(&e->d_vel[b], 3 * ns)) ||
            (rc = dev_alloc(&e->d_disp[b], 3 * ns)) || (rc = dev_alloc(&e->d_id[b], ns)))
            return cleanup_fail(rc);
    }
    if (e->has_cells) {
        // Member records: 64 per cell unless the grid is much larger than the
        // particle count (dilute systems); later arrivals of a cell go to the
        // overflow list, which holds one entry per slot at most.
        const size_t budget = std::max<size_t>((size_t)64 * ns, (size_t)1 << 24);
        e->cell_cap = (int)std::max<size_t>(
            1, std::min<size_t>(64, budget / std::max(1, e->ncell_total)));
        if ((rc = dev_alloc(&e->d_members, (size_t)e->ncell_total * e->cell_cap)) ||
            (rc = dev_alloc(&e->d_ovf, ns)) || (rc = dev_alloc(&e->d_novf, 2)) ||
            (rc = dev_alloc(&e->d_scell, ns)) ||
            (rc = dev_alloc(&e->d_count, e->ncell_total)) ||
            (rc = dev_alloc(&e->d_count_init, e->ncell_total)) ||
            (rc = dev_alloc(&e->d_cell_xyz, e->ncell_total)) ||
            (rc = dev_alloc(&e->d_start, e->ncell_total + 1)) ||
            (rc = dev_alloc(&e->d_tile, e->ntile)))
            return cleanup_fail(rc);
        if (precision == LJ_PRECISION_MIXED && (rc = dev_alloc(&e->d_lpos, ns)))
            return cleanup_fail(rc);
    }
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    std::vector<double> zeros(std::max<size_t>(3 * ns, (size_t)nsys), 0.0);
    // particle ids are local to their system: slot atom_off + i holds id i
    std::vector<int> ident(ns, 0);
    for (int k = 0; k < nsys; ++k)
        for (int i = 0; i < e->sys[k].nblk * kBlock; ++i)
            ident[e->sys[k].atom_off + i] = i < e->sys[k].n ? i : 0;
    bool ok = h2d(e->d_sys, e->sys.data(), nsys * sizeof(SysDev)) &&
              h2d(e->d_blk_sys, blk_sys.data(), blk * sizeof(int)) &&
              h2d(e->d_cblk_sys, cblk_sys.data(), cblk_sys.size() * sizeof(int)) &&
              h2d(e->d_sblk_sys, sblk_sys.data(), sblk_sys.size() * sizeof(int)) &&
              h2d(e->d_force, zeros.data(), 3 * ns * sizeof(double)) &&
              h2d(e->d_time, zeros.data(), nsys * sizeof(double)) &&
              h2d(e->d_pe, zeros.data(), nsys * sizeof(double)) &&
              h2d(e->d_vir, zeros.data(), nsys * sizeof(double)) &&
              h2d(e->d_pending, std::vector<int>(2 * nsys, 0).data(), 2 * nsys * sizeof(int));
    for (int b = 0; b < 2 && ok; ++b)
        ok = h2d(e->d_pos[b], zeros.data(), 3 * ns * sizeof(double)) &&
             h2d(e->d_vel[b], zeros.data(), 3 * ns * sizeof(double)) &&
             h2d(e->d_disp[b], zeros.data(), 3 * ns * sizeof(double)) &&
             h2d(e->d_id[b], ident.data(), ns * sizeof(int));
    if (ok && e->has_cells) {
        std::vector<int> xyz(e->ncell_total, 0);
        for (const SysDev& sd : e->sys)
            for (int lc = 0; lc < sd.ncell; ++lc) {
                const int cz = lc % sd.nc[2], cy = (lc / sd.nc[2]) % sd.nc[1];
                const int cx = lc / (sd.nc[2] * sd.nc[1]);
                xyz[sd.cell_off + lc] = cx | (cy << 10) | (cz << 20);
            }
        const int novf0[2] = {0, 0};
        // member records start as slot 0 (only records of the current pass are read)
        ok = hipMemset(e->d_members, 0, (size_t)e->ncell_total * e->cell_cap * sizeof(int)) ==
                 hipSuccess &&
             h2d(e->d_novf, novf0, sizeof(novf0)) &&
             h2d(e->d_cell_xyz, xyz.data(), e->ncell_total * sizeof(int)) &&
             h2d(e->d_count, count_init.data(), e->ncell_total * sizeof(int)) &&
             h2d(e->d_count_init, count_init.data(), e->ncell_total * sizeof(int));
    }
    if (!ok) return cleanup_fail(fail(LJ_E_HIP, "initial upload failed"));
    if ((rc = ensure_records(e, 64))) return cleanup_fail(rc);
    e->sorted = false;
    *out = e;
    return LJ_OK;
}

int lj_engine_destroy(lj_engine* e)
{
    if (!e) return LJ_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    void* bufs[] = {e->d_sys, e->d_blk_sys, e->d_cblk_sys, e->d_sblk_sys, e->d_force, e->d_members, e->d_ovf, e->d_novf,
                    e->d_scell, e->d_lpos, e->d_count, e->d_count_init, e->d_cell_xyz, e->d_start, e->d_tile,
                    e->d_partial, e->d_partial_c, e->d_cblk_list, e->d_blk_list, e->d_lblk_list, e->d_time, e->d_pe, e->d_vir, e->d_step, e->d_pending, e->d_partial2, e->d_records,
                    e->d_measure, e->d_rdf};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int b = 0; b < 2; ++b) {
        void* more[] = {e->d_pos[b], e->d_vel[b], e->d_disp[b], e->d_id[b],
                        e->r_pos[b], e->r_vel[b], e->r_disp[b], e->r_id[b], e->r_lpos[b]};
        for (void* m : more)
            if (m) (void)hipFree(m);
    }
    void* res[] = {e->r_base, e->r_count, e->r_end, e->r_dirty, e->r_narr,
                   e->r_arr, e->r_flag, e->d_cell_sys, e->r_ck, e->r_force};
    for (void* b : res)
        if (b) (void)hipFree(b);
    for (hipEvent_t ev : e->prof_ev) (void)hipEventDestroy(ev);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->h_out) (void)hipHostFree(e->h_out);
    delete e;
    return LJ_OK;
}

int lj_engine_system_count(const lj_engine* e) { return e ? e->nsys : 0; }

int lj_engine_set_state(lj_engine* e, int sys, const double* pos, const double* vel,
                        const double* disp, const double* force, double time)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (!pos || !vel) return fail(LJ_E_INVALID, "pos and vel are required");
    const SysDev& S = e->sys[sys];
    const int n = S.n;
    HIP_TRY(hipStreamSynchronize(e->stream));
    // Gather the current slot -> id map so other systems stay untouched; the
    // loaded system goes into its slots in id order (unsorted).
    std::vector<double> planes(3 * (size_t)n);
    auto put = [&](double* base, const double* src) -> int {
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < n; ++i) planes[(size_t)a * n + i] = src ? src[3 * (size_t)i + a] : 0.0;
        for (int a = 0; a < 3; ++a)
            HIP_TRY(hipMemcpy(base + (size_t)a * e->nslot + S.atom_off, planes.data() + (size_t)a * n,
                              n * sizeof(double), hipMemcpyHostToDevice));
        return LJ_OK;
    };
    if ((rc = put(e->d_pos[e->cur], pos)) || (rc = put(e->d_vel[e->cur], vel)) ||
        (rc = put(e->d_disp[e->cur], disp)) || (rc = put(e->d_force, force)))
        return rc;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = i;
    HIP_TRY(hipMemcpy(e->d_id[e->cur] + S.atom_off, ids.data(), n * sizeof(int),
                      hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_time + sys, &time, sizeof(double), hipMemcpyHostToDevice));
    double zero = 0.0;
    HIP_TRY(hipMemcpy(e->d_pe + sys, &zero, sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_vir + sys, &zero, sizeof(double), hipMemcpyHostToDevice));
    e->sorted = false;
    // full-step velocities and forces (only a failed step call can leave the
    // engine inside a chained sequence)
    e->prekicked = false;
    return LJ_OK;
}

int lj_engine_get_state(lj_engine* e, int sys, double* pos, double* vel, double* disp,
                        double* force, double* time, double* potential_energy, double* virial)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    const SysDev& S = e->sys[sys];
    const int n = S.n;
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<int> ids(n);
    HIP_TRY(hipMemcpy(ids.data(), e->d_id[e->cur] + S.atom_off, n * sizeof(int),
                      hipMemcpyDeviceToHost));
    std::vector<double> planes(3 * (size_t)n);
    auto get = [&](const double* base, double* dst) -> int {
        if (!dst) return LJ_OK;
        for (int a = 0; a < 3; ++a)
            HIP_TRY(hipMemcpy(planes.data() + (size_t)a * n, base + (size_t)a * e->nslot + S.atom_off,
                              n * sizeof(double), hipMemcpyDeviceToHost));
        for (int s = 0; s < n; ++s)
            for (int a = 0; a < 3; ++a) dst[3 * (size_t)ids[s] + a] = planes[(size_t)a * n + s];
        return LJ_OK;
    };
    if ((rc = get(e->d_pos[e->cur], pos)) || (rc = get(e->d_vel[e->cur], vel)) ||
        (rc = get(e->d_disp[e->cur], disp)) || (rc = get(e->d_force, force)))
        return rc;
    if (time) HIP_TRY(hipMemcpy(time, e->d_time + sys, sizeof(double), hipMemcpyDeviceToHost));
    if (potential_energy)
        HIP_TRY(hipMemcpy(potential_energy, e->d_pe + sys, sizeof(double), hipMemcpyDeviceToHost));
    if (virial) HIP_TRY(hipMemcpy(virial, e->d_vir + sys, sizeof(double), hipMemcpyDeviceToHost));
    return LJ_OK;
}

int lj_engine_step(lj_engine* e, int nsteps, lj_thermo* records)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (nsteps < 0) return fail(LJ_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return LJ_OK;
    if ((rc = ensure_records(e, nsteps))) return rc;
    const bool graphs = e->use_graphs && !e->profiling && e->has_cells;
    return run_steps(e, nsteps, graphs, records);
}

int lj_engine_step_noalloc(lj_engine* e, int nsteps, lj_thermo* records)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (nsteps < 0) return fail(LJ_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return LJ_OK;
    if ((long long)nsteps * e->nsys > kScratchRows)
        return fail(LJ_E_INVALID, "lj_engine_step_noalloc: nsteps * systems = " +
                                      std::to_string((long long)nsteps * e->nsys) +
                                      " exceeds the " + std::to_string(kScratchRows) +
                                      " record rows of its scratch array");
    if (e->profiling)
        return fail(LJ_E_STATE, "lj_engine_step_noalloc: switch profiling off first "
                                "(profiling creates timing events)");
    // the step kernels write their records into this device's scratch rows
    // for the duration of the call; eager launches (no graph capture)
    e->rec_buf = e->scratch;
    e->rec_rows = kScratchRows / e->nsys;
    rc = run_steps(e, nsteps, false, records);
    e->rec_buf = nullptr;
    e->rec_rows = 0;
    return rc;
}

int lj_engine_reserve(lj_engine* e, int max_steps)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (max_steps < 0) return fail(LJ_E_INVALID, "max_steps < 0");
    if ((rc = ensure_records(e, max_steps))) return rc;
    if (e->use_graphs && !e->profiling && e->has_cells && (rc = build_graph(e))) return rc;
    if (res_eligible(e, std::max(max_steps, 4)) && (rc = res_ensure(e))) return rc;
    if (e->res_ready && e->use_graphs && !e->profiling && res_eligible(e, 4) &&
        (rc = build_res_graphs(e)))
        return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return LJ_OK;
}
}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Resident cells (lj_kernels.hpp section 4b).  On an engine whose systems all
// use the tiled kernel (all with one wave per cell, or all with several),
// step calls run on resident cells and stay there between calls:
//   entering (a call of n >= kResMinSteps steps from the packed set):
//                key pass -> k_res_enter (the first cell build, into the resident slots)
//   entering a call on live cells: k_res_kick_drift (stored forces) -> fix-up
//   steps 1..n-1: pair with the drift -> fix-up (replayed from graphs)
//   step n:       the pair kernel of a call's last step on the resident
//                 layout (forces into r_force), then the record
// Any other entry point first packs the live state back (res_settle: a scan of
// the resident counts, k_res_leave with the forces).  The pair kernels see the
// same cells in the same order as on the packed path, so records and state are
// the same bit for bit, however the steps are split into calls.  The packed
// set is not written while cells are live, so when a capacity was exceeded the
// steps since it are redone on the packed path (its clocks are checkpointed).
// ---------------------------------------------------------------------------
constexpr int kResMinSteps = 4;
// LJ_RESIDENT_AUTO: particles per cell at which resident cells are used.
// Measured on the liquid (14.6 per cell: step 74.4 -> 64.7 us) and on the
// coexistence workload (8.2 per cell, half of the cells nearly empty: 521 ->
// 599 us, the fix-up and the per-cell drift outweighing the saved sort).
constexpr double kResMinOccupancy = 10.0;
// Steps on live resident cells after which the next call packs them back and
// enters again (the same results), so that a capacity overflow never redoes
// more than about this many steps.
constexpr long long kResMaxSince = 16384;

bool res_eligible(const lj_engine* e, int nsteps)
{
    if (e->res_mode == LJ_RESIDENT_OFF || !e->has_cells || nsteps < kResMinSteps) return false;
    // the tiled kernel for every system, with one or with several waves per cell
    if (e->n_tiled != e->nsys && e->n_multi != e->nsys) return false;
    long long slots = 0, n = 0, cells = 0;
    for (const SysDev& s : e->sys) {
        if (!s.active || s.law == ljk::LAW_NONE) return false;
        slots += (long long)s.ncell * ljk::kResCap;
        n += s.n;
        cells += s.ncell;
    }
    // the drift's and the fix-up's work per cell pays off at liquid occupancies
    if (e->res_mode == LJ_RESIDENT_AUTO && (double)n < kResMinOccupancy * (double)cells)
        return false;
    // resident slots: within the slot limit and at most 16 per particle
    return slots <= kMaxSlots && slots <= 16LL * e->nslot + (1 << 20);
}

// Allocates and initialises the resident buffers (once per engine).
int res_ensure(lj_engine* e)
{
    if (e->res_ready) return LJ_OK;
    const int nc = e->ncell_total;
    std::vector<int> base(nc, 0), count(nc, 0), csys(nc, 0);
    std::vector<int> count_init(nc, 0);
    HIP_TRY(hipMemcpy(count_init.data(), e->d_count_init, nc * sizeof(int), hipMemcpyDeviceToHost));
    long long off = 0;
    for (int k = 0; k < e->nsys; ++k) {
        const SysDev& s = e->sys[k];
        for (int lc = 0; lc <= s.ncell; ++lc) {
            const int c = s.cell_off + lc;
            csys[c] = k;
            base[c] = (int)(off + (long long)std::min(lc, s.ncell) * ljk::kResCap);
            if (lc == s.ncell) count[c] = count_init[c];   // tail cell: padding slots
        }
        off += (long long)s.ncell * ljk::kResCap;
    }
    const size_t ns = (size_t)off;
    int rc;
    for (int b = 0; b < 2; ++b)
        if ((rc = dev_alloc(&e->r_pos[b], 3 * ns)) ||
            (e->precision == LJ_PRECISION_MIXED && (rc = dev_alloc(&e->r_lpos[b], ns))))
            return rc;
    if ((rc = dev_alloc(&e->r_vel[0], 3 * ns)) || (rc = dev_alloc(&e->r_disp[0], 3 * ns)) ||
        (rc = dev_alloc(&e->r_force, 3 * ns)) || (rc = dev_alloc(&e->r_id[0], ns)))
        return rc;
    if ((rc = dev_alloc(&e->r_base, nc)) || (rc = dev_alloc(&e->r_count, nc)) ||
        (rc = dev_alloc(&e->r_end, nc)) || (rc = dev_alloc(&e->r_dirty, nc)) ||
        (rc = dev_alloc(&e->r_narr, nc)) || (rc = dev_alloc(&e->r_arr, (size_t)nc * ljk::kResArr)) ||
        (rc = dev_alloc(&e->r_flag, 1)) || (rc = dev_alloc(&e->d_cell_sys, nc)) ||
        (rc = dev_alloc(&e->r_ck, 3 * (size_t)e->nsys)))
        return rc;
    // every slot starts as zeros (finite values wherever a kernel may read)
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMemset(e->r_pos[b], 0, 3 * ns * sizeof(double)));
        if (e->r_lpos[b]) HIP_TRY(hipMemset(e->r_lpos[b], 0, ns * sizeof(float4)));
    }
    HIP_TRY(hipMemset(e->r_vel[0], 0, 3 * ns * sizeof(double)));
    HIP_TRY(hipMemset(e->r_disp[0], 0, 3 * ns * sizeof(double)));
    HIP_TRY(hipMemset(e->r_force, 0, 3 * ns * sizeof(double)));
    HIP_TRY(hipMemset(e->r_id[0], 0, ns * sizeof(int)));
    HIP_TRY(hipMemcpy(e->r_base, base.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->r_end, base.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->r_count, count.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->d_cell_sys, csys.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(e->r_dirty, 0, nc * sizeof(int)));
    HIP_TRY(hipMemset(e->r_narr, 0, nc * sizeof(int)));
    HIP_TRY(hipMemset(e->r_arr, 0, (size_t)nc * ljk::kResArr * sizeof(ljk::ResMover)));
    HIP_TRY(hipMemset(e->r_flag, 0, sizeof(int)));
    e->res_nslot = (int)ns;
    // fix-up waves of kResFixCells cells
    const int waves = (nc + ljk::kResFixCells - 1) / ljk::kResFixCells;
    e->res_nfix = (waves + kBlock / 64 - 1) / (kBlock / 64);
    e->res_ready = true;
    return LJ_OK;
}

template <int LAW>
void launch_res_pair(lj_engine* e, const ljk::StepArgs& p)
{
    const dim3 grid(e->ncblk);
    if (e->n_multi == e->nsys) {   // several waves per home cell (small systems)
        const dim3 gm(e->ncblk * p.cpb), bm(64 * ljk::kWavesPerCellSmall);
        if (e->precision == LJ_PRECISION_MIXED)
            launch(e, ljk::k_pair_cells_mixed<LAW, 2, 1, ljk::kWavesPerCellSmall, true>, gm, bm, p);
        else
            launch(e, ljk::k_pair_cells_fp64<LAW, 2, ljk::kWavesPerCellSmall, true>, gm, bm, p);
        return;
    }
    if (e->precision == LJ_PRECISION_MIXED) launch_tiled_shape<ljk::MIXED, LAW, 2, true>(e, p, grid);
    else launch_tiled_shape<ljk::FP64, LAW, 2, true>(e, p, grid);
}

// The pair kernel of a call's last step on the resident layout: the tiled
// kernels of the packed path's last step (KICK 1: second half kick, forces
// stored), with the resident cells' ranges and p.force = r_force.
template <int LAW>
void launch_res_last(lj_engine* e, const ljk::StepArgs& p)
{
    const dim3 grid(e->ncblk);
    if (e->n_multi == e->nsys) {
        const dim3 gm(e->ncblk * p.cpb), bm(64 * ljk::kWavesPerCellSmall);
        if (e->precision == LJ_PRECISION_MIXED)
            launch(e, ljk::k_pair_cells_mixed<LAW, 1, 1, ljk::kWavesPerCellSmall, false>, gm, bm, p);
        else
            launch(e, ljk::k_pair_cells_fp64<LAW, 1, ljk::kWavesPerCellSmall, false>, gm, bm, p);
        return;
    }
    if (e->precision == LJ_PRECISION_MIXED) launch_tiled_shape<ljk::MIXED, LAW, 1, false>(e, p, grid);
    else launch_tiled_shape<ljk::FP64, LAW, 1, false>(e, p, grid);
}

// One resident step of parity r: the pair kernel with the drift and the
// fix-up (first: no previous resident step, whose record the fix-up would
// finish).
int enqueue_res_step(lj_engine* e, int r, bool first)
{
    const ljk::StepArgs p = e->args_res(r);
    const bool mixed = e->precision == LJ_PRECISION_MIXED;
    e->last_pair_force = true;
    int rc = timed(e, 5, "pair", [&] {
        if (e->sys[0].law == ljk::LAW_LJ) launch_res_pair<ljk::LAW_LJ>(e, p);
        else launch_res_pair<ljk::LAW_CONSTANT>(e, p);
    });
    if (rc) return rc;
    const ljk::StepArgs q = e->args_res(r, true);
    const ljk::FinalArgs f1 = e->fargs_par(r), f2 = e->fargs_par(r ^ 1);
    const int nred = first ? 0 : e->nsys;
    const int nfin = e->nsys * f1.nf;
    return timed(e, 9, "fixup", [&] {
        const dim3 grid(nred + nfin + e->res_nfix);
        if (mixed) launch(e, ljk::k_res_fixup<true>, grid, dim3(kBlock), q, f1, f2, nred, nfin);
        else launch(e, ljk::k_res_fixup<false>, grid, dim3(kBlock), q, f1, f2, nred, nfin);
    });
}

// Captures nsteps resident steps starting at parity `from` (none the first).
int capture_res_steps(lj_engine* e, int from, int nsteps, hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = LJ_OK;
    int r = from;
    for (int s = 0; s < nsteps && !rc; ++s, r ^= 1) rc = enqueue_res_step(e, r, false);
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return fail(LJ_E_HIP, std::string("hipStreamEndCapture: ") +
                                                    hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(LJ_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    return LJ_OK;
}

int build_res_graphs(lj_engine* e)
{
    // the packed graphs' checks decide when both kinds are stale (records)
    int rc = build_graph(e);
    if (rc) return rc;
    if (e->rgraphs[1][kGraphLevels - 1]) return LJ_OK;
    for (int from = 0; from < 2 && !rc; ++from)
        for (int l = 0; l < kGraphLevels && !rc; ++l)
            rc = capture_res_steps(e, from, 2 << l, &e->rgraphs[from][l]);
    if (rc) drop_graphs(e);
    return rc;
}

// Packs live resident cells back into the other packed set (with the last
// step's forces) and makes it current; nothing to do when none are live.
int res_settle(lj_engine* e)
{
    if (!e->res_live) return LJ_OK;
    int rc;
    ljk::StepArgs p = e->args(e->cur ^ 1);
    p.count = e->r_count;
    p.count_init = e->r_count;   // the two-level scan's count reset leaves them as they are
    const ljk::FinalArgs f = e->fargs(0);   // (no record blocks in these launches)
    if (!e->two_level_scan()) {
        if ((rc = timed(e, 1, "scan", [&] {
                 launch(e, ljk::k_scan_one, dim3(1), dim3(ljk::kScanOneThreads), p, f);
             })))
            return rc;
    } else {
        if ((rc = timed(e, 1, "scan_tiles", [&] {
                 launch(e, ljk::k_scan_tiles, dim3(e->ntile), dim3(kBlock), p);
             })))
            return rc;
        if ((rc = timed(e, 2, "scan_apply", [&] {
                 launch(e, ljk::k_scan_apply, dim3(e->ntile), dim3(kBlock), p, f);
             })))
            return rc;
    }
    ljk::StepArgs q = e->args_res(e->res_par);   // the live state
    const ljk::StepArgs pk = e->args(e->cur ^ 1);
    q.pos_n = pk.pos;
    q.vel_n = pk.vel;
    q.disp_n = pk.disp;
    q.id_n = pk.id;
    q.lpos_n = e->d_lpos;
    q.start = e->d_start;
    const ljk::Planes fdst = pk.force;
    const dim3 grid((e->ncell_total + kBlock / 64 - 1) / (kBlock / 64));
    if ((rc = timed(e, 11, "res_leave", [&] {
             if (e->precision == LJ_PRECISION_MIXED)
                 launch(e, ljk::k_res_leave<true, true>, grid, dim3(kBlock), q, fdst);
             else
                 launch(e, ljk::k_res_leave<false, true>, grid, dim3(kBlock), q, fdst);
         })))
        return rc;
    HIP_TRY(hipGetLastError());
    e->cur ^= 1;
    e->sorted = true;
    e->prekicked = false;
    e->res_live = false;
    e->res_since = 0;
    return LJ_OK;
}

// The step loop on resident cells (see above).  Returns 1 (with nothing to
// report) when a capacity was exceeded: the engine is then back at the
// snapshot, and the caller redoes on the packed path the `res_since` steps
// taken since it and this call's steps.
int run_steps_res(lj_engine* e, int nsteps, bool graphs, lj_thermo* records)
{
    int rc;
    const bool mixed = e->precision == LJ_PRECISION_MIXED;
    const int nsys = e->nsys;
    e->res_calls++;
    if (graphs && (rc = build_res_graphs(e))) return rc;
    const ljk::FinalArgs f = e->fargs(0);
    const long long nrec = (long long)nsteps * nsys * 7;
    const long long nthr = std::max<long long>(nrec, nsys);
    const dim3 cell_grid((e->ncell_total + kBlock / 64 - 1) / (kBlock / 64));
    int r;
    if (!e->res_live) {
        hipLaunchKernelGGL(ljk::k_begin_call,
                           dim3((unsigned)std::min<long long>(4096, (nthr + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, e->stream, f.records, nrec, e->d_step, nsys);
        // the snapshot: the packed set e->cur and d_force stay as they are;
        // the clocks are checkpointed
        hipLaunchKernelGGL(ljk::k_res_checkpoint, dim3((nsys + kBlock - 1) / kBlock), dim3(kBlock),
                           0, e->stream, f, e->r_ck);
        e->res_since = 0;
        // the first half kick and drift: key pass, then into the resident slots
        const ljk::StepArgs p = e->args(e->cur);
        if ((rc = timed(e, 0, "integrate_key", [&] {
                 launch(e, ljk::k_integrate_key<2, true>, dim3(e->nblk), dim3(kBlock), p, f, 0);
             })))
            return rc;
        ljk::StepArgs q = p;
        const ljk::StepArgs r0 = e->args_res(0);
        q.pos_n = r0.pos;
        q.vel_n = r0.vel;
        q.disp_n = r0.disp;
        q.id_n = r0.id;
        q.lpos_n = r0.lpos;
        q.rc_par = 0;
        if ((rc = timed(e, 10, "res_enter", [&] {
                 if (mixed) launch(e, ljk::k_res_enter<true, 1>, cell_grid, dim3(kBlock), q);
                 else launch(e, ljk::k_res_enter<false, 1>, cell_grid, dim3(kBlock), q);
             })))
            return rc;
        r = 0;
    } else {
        // the first half kick and drift from the stored forces, then the fix-up
        r = e->res_par;
        const ljk::StepArgs p = e->args_res(r);
        constexpr int cpb = ljk::kResDriftCells * (kBlock / 64);   // cells per block
        const dim3 drift_grid((e->ncell_total + cpb - 1) / cpb);
        // (this launch also clears the call's records: k_begin_call's work)
        if ((rc = timed(e, 12, "res_drift", [&] {
                 if (mixed)
                     launch(e, ljk::k_res_kick_drift<true>, drift_grid, dim3(kBlock), p, f.records,
                            nrec, e->d_step, nsys);
                 else
                     launch(e, ljk::k_res_kick_drift<false>, drift_grid, dim3(kBlock), p, f.records,
                            nrec, e->d_step, nsys);
             })))
            return rc;
        const ljk::StepArgs q = e->args_res(r, true);
        const ljk::FinalArgs f1 = e->fargs_par(r), f2 = e->fargs_par(r ^ 1);
        if ((rc = timed(e, 9, "fixup", [&] {
                 if (mixed) launch(e, ljk::k_res_fixup<true>, dim3(e->res_nfix), dim3(kBlock), q, f1, f2, 0, 0);
                 else launch(e, ljk::k_res_fixup<false>, dim3(e->res_nfix), dim3(kBlock), q, f1, f2, 0, 0);
             })))
            return rc;
        r ^= 1;
    }
    e->res_live = true;
    // steps 1 .. n - 1: pair kernel with the drift, fix-up
    const int middle_end = nsteps - 1;
    int done = 0;
    if (middle_end > 0) {
        if ((rc = enqueue_res_step(e, r, true))) return rc;
        r ^= 1;
        done = 1;
    }
    if (graphs) {
        for (int l = kGraphLevels - 1; l >= 0; --l) {
            const int len = 2 << l;
            while (middle_end - done >= len) {
                HIP_TRY(hipGraphLaunch(e->rgraphs[r][l], e->stream));
                done += len;   // even: r unchanged
            }
        }
    }
    for (; done < middle_end; ++done, r ^= 1)
        if ((rc = enqueue_res_step(e, r, false))) return rc;
    // step n: forces into r_force and the second half kick, then the records
    // (the second level of step n - 1's, both levels of step n's)
    {
        const ljk::StepArgs p = e->args_res(r);
        e->last_pair_force = true;
        if ((rc = timed(e, 5, "pair", [&] {
                 if (e->sys[0].law == ljk::LAW_LJ) launch_res_last<ljk::LAW_LJ>(e, p);
                 else launch_res_last<ljk::LAW_CONSTANT>(e, p);
             })))
            return rc;
        const ljk::FinalArgs f1 = e->fargs_par(r), f2 = e->fargs_par(r ^ 1);
        const int nred = middle_end > 0 ? nsys : 0;
        const int nfin = nsys * f1.nf;
        if ((rc = timed(e, 6, "finalize", [&] {
                 if (mixed) launch(e, ljk::k_res_fixup<true>, dim3(nred + nfin), dim3(kBlock), p, f1, f2, nred, nfin);
                 else launch(e, ljk::k_res_fixup<false>, dim3(nred + nfin), dim3(kBlock), p, f1, f2, nred, nfin);
                 launch(e, ljk::k_finalize_reduced, dim3(nsys), dim3(kBlock), f1);
             })))
            return rc;
    }
    e->res_par = r;
    HIP_TRY(hipGetLastError());
    int flag = 0;
    if ((rc = finish_call(e, nsteps, records, &flag))) return rc;
    if (!flag) {
        e->res_since += nsteps;
        return LJ_OK;
    }
    // a capacity was exceeded: back to the snapshot (the packed set e->cur,
    // d_force and the checkpointed clocks)
    e->res_fallbacks++;
    e->res_live = false;
    e->prekicked = false;
    HIP_TRY(hipMemcpyAsync(e->d_time, e->r_ck, nsys * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->d_pe, e->r_ck + nsys, nsys * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->d_vir, e->r_ck + 2 * nsys, nsys * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(e->d_pending, 0, 2 * nsys * sizeof(int), e->stream));
    // the entering key pass's overflow entries (cells beyond the member
    // records), which no scan consumed
    HIP_TRY(hipMemsetAsync(e->d_novf, 0, 2 * sizeof(int), e->stream));
    HIP_TRY(hipMemsetAsync(e->r_dirty, 0, e->ncell_total * sizeof(int), e->stream));
    HIP_TRY(hipMemsetAsync(e->r_narr, 0, e->ncell_total * sizeof(int), e->stream));
    HIP_TRY(hipMemsetAsync(e->r_flag, 0, sizeof(int), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return 1;
}

int run_steps_packed(lj_engine* e, int nsteps, bool graphs, lj_thermo* records);

// The step loop shared by the step entry points: nsteps steps of every
// system, records into the engine's current record buffer (fargs), replayed
// from the captured step graphs when `graphs` is set, then copied to
// `records` (optional).  On live resident cells, or entering them (allocating
// callers always, the others once allocated); after a capacity was exceeded
// there, the steps since the snapshot are redone on the packed path (in calls
// of at most the record capacity, whose records are dropped: a run's results
// do not depend on how it is split into calls), then this call's.
int run_steps(lj_engine* e, int nsteps, bool graphs, lj_thermo* records)
{
    int rc;
    // a fresh snapshot now and then bounds what an overflow has to redo
    if (e->res_live && e->res_since >= kResMaxSince && (rc = res_settle(e))) return rc;
    if (e->res_live || (res_eligible(e, nsteps) && (e->res_ready || !e->rec_buf))) {
        if ((rc = res_ensure(e))) return rc;
        const long long since = e->res_live ? e->res_since : 0;
        rc = run_steps_res(e, nsteps, graphs, records);
        if (rc != 1) return rc;
        const int cap = std::max(1, e->fargs(0).cap);
        for (long long left = since; left > 0;) {
            const int n = (int)std::min<long long>(left, cap);
            if ((rc = run_steps_packed(e, n, graphs, nullptr))) return rc;
            left -= n;
        }
    }
    return run_steps_packed(e, nsteps, graphs, records);
}

// The packed path's step loop (a cell build every step).
int run_steps_packed(lj_engine* e, int nsteps, bool graphs, lj_thermo* records)
{
    int rc;
    const ljk::FinalArgs f = e->fargs(0);
    const long long nrec = (long long)nsteps * e->nsys * 7;
    const long long nthr = std::max<long long>(nrec, e->nsys);
    hipLaunchKernelGGL(ljk::k_begin_call,
                       dim3((unsigned)std::min<long long>(4096, (nthr + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, e->stream, f.records, nrec, e->d_step, e->nsys);
    int done = 0;
    if (graphs && nsteps >= 2) {
        if ((rc = build_graph(e))) return rc;
        // the largest graph that fits, repeatedly (an even count keeps e->cur)
        for (int l = kGraphLevels - 1; l >= 0; --l) {
            const int len = 2 << l;
            while (nsteps - done >= len) {
                HIP_TRY(hipGraphLaunch(e->graphs[e->cur][l], e->stream));
                done += len;
                e->sorted = true;
            }
        }
    }
    while (done < nsteps) {
        if ((rc = enqueue_step(e, done + 1 < nsteps))) return rc;
        ++done;
    }
    if ((rc = enqueue_finalize(e, 0))) return rc;   // the last step's record
    HIP_TRY(hipGetLastError());
    return finish_call(e, nsteps, records, nullptr);
}

}  // namespace

extern "C" {

int lj_engine_compute_forces(lj_engine* e, double* pe, double* vir)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (e->has_cells) {
        if ((rc = enqueue_cells(e, false))) return rc;
        if ((rc = enqueue_pair(e, false, false))) return rc;
        if ((rc = enqueue_finalize(e, 1))) return rc;
    }
    HIP_TRY(hipGetLastError());
    if (pe) HIP_TRY(hipMemcpyAsync(pe, e->d_pe, e->nsys * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (vir) HIP_TRY(hipMemcpyAsync(vir, e->d_vir, e->nsys * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return LJ_OK;
}

int lj_engine_measure(lj_engine* e, lj_thermo* out)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if ((rc = enqueue_pair(e, false, true))) return rc;
    if ((rc = enqueue_finalize(e, 2))) return rc;
    HIP_TRY(hipGetLastError());
    if (out)
        HIP_TRY(hipMemcpyAsync(out, e->d_measure, e->nsys * sizeof(lj_thermo), hipMemcpyDeviceToHost,
                               e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return LJ_OK;
}

int lj_engine_set_temperature(lj_engine* e, int sys, double target, double* measured)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    std::vector<lj_thermo> m(e->nsys);
    if ((rc = lj_engine_measure(e, m.data()))) return rc;
    const double t = m[sys].temperature;
    if (measured) *measured = t;
    if (!(t > 0)) return fail(LJ_E_STATE, "Cannot scale temperature of zero-temperature state");
    const double scale = std::sqrt(target / t);
    const ljk::StepArgs p = e->args(e->cur);
    hipLaunchKernelGGL(ljk::k_scale_velocities, dim3(e->sys[sys].nblk), dim3(kBlock), 0, e->stream,
                       p, sys, scale);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    return LJ_OK;
}

int lj_engine_cell_keys(lj_engine* e, int sys, int* keys, int* nc, int* clamped)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    const SysDev& S = e->sys[sys];
    if (S.law == LJ_LAW_NONE) return fail(LJ_E_STATE, "system has no cell grid");
    if (nc)
        for (int a = 0; a < 3; ++a) nc[a] = S.nc[a];
    int* d_keys = nullptr;
    int* d_cl = nullptr;
    if ((rc = dev_alloc(&d_keys, 3 * (size_t)S.n)) || (rc = dev_alloc(&d_cl, 1))) {
        if (d_keys) (void)hipFree(d_keys);
        return rc;
    }
    int zero = 0;
    hipError_t he = hipMemcpy(d_cl, &zero, sizeof(int), hipMemcpyHostToDevice);
    if (he == hipSuccess) {
        const ljk::StepArgs p = e->args(e->cur);
        hipLaunchKernelGGL(ljk::k_cell_keys, dim3(S.nblk), dim3(kBlock), 0, e->stream, p, sys,
                           d_keys, d_cl);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he == hipSuccess && keys)
        he = hipMemcpy(keys, d_keys, 3 * (size_t)S.n * sizeof(int), hipMemcpyDeviceToHost);
    int cl = 0;
    if (he == hipSuccess) he = hipMemcpy(&cl, d_cl, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d_keys);
    (void)hipFree(d_cl);
    if (he != hipSuccess) return fail(LJ_E_HIP, std::string("cell_keys: ") + hipGetErrorString(he));
    if (clamped) *clamped = cl;
    return LJ_OK;
}

int lj_engine_cell_list(lj_engine* e, int sys, int* start, int* atoms)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    const SysDev& S = e->sys[sys];
    if (S.law == LJ_LAW_NONE) return fail(LJ_E_STATE, "system has no cell grid");
    if ((rc = enqueue_cells(e, false))) return rc;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<int> st(S.ncell + 1);
    HIP_TRY(hipMemcpy(st.data(), e->d_start + S.cell_off, (S.ncell + 1) * sizeof(int),
                      hipMemcpyDeviceToHost));
    if (start)
        for (int c = 0; c <= S.ncell; ++c) start[c] = st[c] - S.atom_off;
    if (atoms)
        HIP_TRY(hipMemcpy(atoms, e->d_id[e->cur] + S.atom_off, S.n * sizeof(int),
                          hipMemcpyDeviceToHost));
    return LJ_OK;
}

int lj_engine_pairs(lj_engine* e, int sys, long long cap, int* idx, double* sep, long long* count)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    const SysDev& S = e->sys[sys];
    if (S.law == LJ_LAW_NONE) return fail(LJ_E_STATE, "system has no cell grid");
    if (cap < 0) return fail(LJ_E_INVALID, "cap < 0");
    if ((rc = ensure_sorted(e))) return rc;
    int* d_idx = nullptr;
    double* d_sep = nullptr;
    unsigned long long* d_cnt = nullptr;
    if ((rc = dev_alloc(&d_idx, 5 * (size_t)std::max(cap, 1LL))) ||
        (rc = dev_alloc(&d_sep, 3 * (size_t)std::max(cap, 1LL))) || (rc = dev_alloc(&d_cnt, 1))) {
        if (d_idx) (void)hipFree(d_idx);
        if (d_sep) (void)hipFree(d_sep);
        return rc;
    }
    unsigned long long zero = 0, total = 0;
    hipError_t he = hipMemcpy(d_cnt, &zero, sizeof(zero), hipMemcpyHostToDevice);
    if (he == hipSuccess) {
        const ljk::StepArgs p = e->args(e->cur);
        hipLaunchKernelGGL(ljk::k_pairs_emit, dim3(S.nblk), dim3(kBlock), 0, e->stream, p, sys, cap,
                           d_idx, d_sep, d_cnt);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he == hipSuccess) he = hipMemcpy(&total, d_cnt, sizeof(total), hipMemcpyDeviceToHost);
    const long long w = std::min<long long>((long long)total, cap);
    if (he == hipSuccess && idx && w > 0)
        he = hipMemcpy(idx, d_idx, 5 * (size_t)w * sizeof(int), hipMemcpyDeviceToHost);
    if (he == hipSuccess && sep && w > 0)
        he = hipMemcpy(sep, d_sep, 3 * (size_t)w * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d_idx);
    (void)hipFree(d_sep);
    (void)hipFree(d_cnt);
    if (he != hipSuccess) return fail(LJ_E_HIP, std::string("pairs: ") + hipGetErrorString(he));
    if (count) *count = (long long)total;
    return LJ_OK;
}

int lj_engine_neighbor_counts(lj_engine* e, int sys, int* counts)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    const SysDev& S = e->sys[sys];
    if (S.law == LJ_LAW_NONE) return fail(LJ_E_STATE, "system has no cell grid");
    if ((rc = dev_alloc(&e->d_ncount, e->nslot))) return rc;
    rc = enqueue_cells(e, false);
    if (!rc) rc = enqueue_pair(e, false, false);
    if (!rc) rc = enqueue_finalize(e, 1);
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    std::vector<int> cnt(S.n), ids(S.n);
    if (he == hipSuccess)
        he = hipMemcpy(cnt.data(), e->d_ncount + S.atom_off, S.n * sizeof(int), hipMemcpyDeviceToHost);
    if (he == hipSuccess)
        he = hipMemcpy(ids.data(), e->d_id[e->cur] + S.atom_off, S.n * sizeof(int),
                       hipMemcpyDeviceToHost);
    (void)hipFree(e->d_ncount);
    e->d_ncount = nullptr;
    if (rc) return rc;
    if (he != hipSuccess) return fail(LJ_E_HIP, std::string("neighbor_counts: ") + hipGetErrorString(he));
    if (counts)
        for (int s = 0; s < S.n; ++s) counts[ids[s]] = cnt[s];
    return LJ_OK;
}

int lj_engine_rdf_sample(lj_engine* e, double rmax, int bins)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (!e->has_cells) return fail(LJ_E_STATE, "no cell grid");
    if (!(rmax > 0) || bins < 1 || bins > 4096) return fail(LJ_E_INVALID, "bad rmax / bins");
    if (e->d_rdf && (bins != e->rdf_bins || rmax != e->rdf_rmax))
        return fail(LJ_E_STATE, "read the accumulated g(r) before changing rmax / bins");
    if (!e->d_rdf) {
        if ((rc = dev_alloc(&e->d_rdf, (size_t)e->nsys * bins))) return rc;
        HIP_TRY(hipMemset(e->d_rdf, 0, (size_t)e->nsys * bins * sizeof(unsigned long long)));
        e->rdf_bins = bins;
        e->rdf_rmax = rmax;
        e->rdf_samples = 0;
    }
    if ((rc = ensure_sorted(e))) return rc;
    const ljk::StepArgs p = e->args(e->cur);
    hipLaunchKernelGGL(ljk::k_rdf, dim3(e->nblk), dim3(kBlock), bins * sizeof(unsigned int),
                       e->stream, p, rmax, bins, e->d_rdf);
    HIP_TRY(hipGetLastError());
    e->rdf_samples += 1;
    return LJ_OK;
}

int lj_engine_rdf_read(lj_engine* e, double* counts, int* samples)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (!e->d_rdf) {
        if (samples) *samples = 0;
        return LJ_OK;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<unsigned long long> h((size_t)e->nsys * e->rdf_bins);
    HIP_TRY(hipMemcpy(h.data(), e->d_rdf, h.size() * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
    if (counts)
        for (size_t k = 0; k < h.size(); ++k) counts[k] = (double)h[k];
    if (samples) *samples = e->rdf_samples;
    HIP_TRY(hipFree(e->d_rdf));
    e->d_rdf = nullptr;
    e->rdf_samples = 0;
    return LJ_OK;
}

int lj_engine_set_active(lj_engine* e, int sys, int active)
{
    int rc = check_sys(e, sys);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->sys[sys].active = active ? 1 : 0;
    HIP_TRY(hipMemcpy(e->d_sys + sys, &e->sys[sys], sizeof(SysDev), hipMemcpyHostToDevice));
    return LJ_OK;
}

int lj_engine_set_graphs(lj_engine* e, int enabled)
{
    int rc = check_engine(e);
    if (rc) return rc;
    e->use_graphs = enabled != 0;
    return LJ_OK;
}

int lj_engine_set_resident_cells(lj_engine* e, int mode)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (mode != LJ_RESIDENT_OFF && mode != LJ_RESIDENT_AUTO && mode != LJ_RESIDENT_ON)
        return fail(LJ_E_INVALID, "resident cells: mode must be LJ_RESIDENT_OFF, _AUTO or _ON");
    e->res_mode = mode;
    return LJ_OK;
}

int lj_engine_resident_cells(const lj_engine* e)
{
    return e ? e->res_mode : LJ_RESIDENT_OFF;
}

int lj_engine_resident_stats(const lj_engine* e, long long* calls, long long* fallbacks)
{
    if (!e) return fail(LJ_E_INVALID, "engine is null");
    if (calls) *calls = e->res_calls;
    if (fallbacks) *fallbacks = e->res_fallbacks;
    return LJ_OK;
}

int lj_engine_set_pair_kernel(lj_engine* e, int variant)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (variant != LJ_PAIR_PER_PARTICLE && variant != LJ_PAIR_CELL_TILED &&
        variant != LJ_PAIR_BY_OCCUPANCY && variant != LJ_PAIR_CELL_MULTIWAVE)
        return fail(LJ_E_INVALID, "unknown pair kernel variant");
    e->pair_variant = variant;
    return assign_pair_kernels(e);
}

int lj_engine_set_pair_split(lj_engine* e, double min_occupancy)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (!(min_occupancy > 0)) return fail(LJ_E_INVALID, "min_occupancy must be > 0");
    e->split_occupancy = min_occupancy;
    e->pair_variant = LJ_PAIR_BY_OCCUPANCY;
    return assign_pair_kernels(e);
}

int lj_engine_set_launch_shape(lj_engine* e, int cells_per_block, int sort_wave_particles)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (cells_per_block != 2 && cells_per_block != 4 && cells_per_block != 8)
        return fail(LJ_E_INVALID, "cells_per_block must be 2, 4 or 8");
    if (sort_wave_particles < 8 || sort_wave_particles > 64)
        return fail(LJ_E_INVALID, "sort_wave_particles must be in [8, 64]");
    if (cells_per_block == e->cells_per_block && sort_wave_particles == e->sort_wave) return LJ_OK;
    e->cells_per_block = cells_per_block;
    e->sort_wave = sort_wave_particles;
    if ((rc = upload_layout(e))) return rc;
    return assign_pair_kernels(e, true);
}

int lj_engine_launch_shape(lj_engine* e, int* cells_per_block, int* sort_wave_particles)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (cells_per_block) *cells_per_block = e->cells_per_block;
    if (sort_wave_particles) *sort_wave_particles = e->sort_wave;
    return LJ_OK;
}

int lj_engine_pair_kernels(lj_engine* e, int* per_system)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if (!per_system) return fail(LJ_E_INVALID, "null output");
    for (int k = 0; k < e->nsys; ++k)
        per_system[k] = e->sys[k].tiled == 1 ? LJ_PAIR_CELL_TILED
                        : e->sys[k].tiled == 2 ? LJ_PAIR_CELL_MULTIWAVE : LJ_PAIR_PER_PARTICLE;
    return LJ_OK;
}

int lj_engine_autotune(lj_engine* e, int steps, int* chosen, double* ms_per_step)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (!e->has_cells) {
        if (chosen) *chosen = e->pair_variant;
        return LJ_OK;
    }
    if (steps < kResMinSteps) steps = kResMinSteps;   // both trial calls alike (resident or not)
    // save the state (current buffer set, forces, clocks) so trials leave no trace
    const size_t ns = e->nslot;
    DevBuf sp, sv, sd, sf, st;
    int* sid = nullptr;
    if ((rc = dev_alloc(&sp.p, 3 * ns)) || (rc = dev_alloc(&sv.p, 3 * ns)) ||
        (rc = dev_alloc(&sd.p, 3 * ns)) || (rc = dev_alloc(&sf.p, 3 * ns)) ||
        (rc = dev_alloc(&st.p, 3 * (size_t)e->nsys)) || (rc = dev_alloc(&sid, ns))) {
        if (sid) (void)hipFree(sid);
        return rc;
    }
    const int cur0 = e->cur;
    const int variant0 = e->pair_variant;
    const double occ0 = e->split_occupancy;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, e->stream);
    };
    auto save = [&]() {
        hipError_t h = cp(sp.p, e->d_pos[e->cur], 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(sv.p, e->d_vel[e->cur], 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(sd.p, e->d_disp[e->cur], 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(sf.p, e->d_force, 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(sid, e->d_id[e->cur], ns * sizeof(int));
        if (h == hipSuccess) h = cp(st.p, e->d_time, e->nsys * sizeof(double));
        if (h == hipSuccess) h = cp(st.p + e->nsys, e->d_pe, e->nsys * sizeof(double));
        if (h == hipSuccess) h = cp(st.p + 2 * e->nsys, e->d_vir, e->nsys * sizeof(double));
        return h;
    };
    // Puts the saved state back into buffer set cur0.  The cell ranges, cell
    // of each slot and mixed-precision coordinates still describe a trial's
    // last sort, so the state counts as unsorted: the next traversal
    // (pairs, g(r), neighbour counts) rebuilds the cells first.
    auto restore = [&]() {
        e->cur = cur0;
        e->sorted = false;
        e->prekicked = false;
        e->res_live = false;   // a trial's resident cells are dropped
        e->res_since = 0;
        hipError_t h = cp(e->d_pos[cur0], sp.p, 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_vel[cur0], sv.p, 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_disp[cur0], sd.p, 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_force, sf.p, 3 * ns * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_id[cur0], sid, ns * sizeof(int));
        if (h == hipSuccess) h = cp(e->d_time, st.p, e->nsys * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_pe, st.p + e->nsys, e->nsys * sizeof(double));
        if (h == hipSuccess) h = cp(e->d_vir, st.p + 2 * e->nsys, e->nsys * sizeof(double));
        if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
        return h;
    };
    // every exit after the save goes through here: the state and the kernel
    // policy the engine had on entry are put back unless a winner was chosen
    const int cpb_entry = e->cells_per_block, sw_entry = e->sort_wave;
    const int res_entry = e->res_mode;
    auto finish = [&](int code, bool restored) {
        if (!restored) (void)restore();
        (void)hipFree(sid);
        if (code) {
            const std::string msg = g_err;
            e->res_mode = res_entry;
            e->split_occupancy = occ0;
            e->pair_variant = variant0;
            (void)lj_engine_set_launch_shape(e, cpb_entry, sw_entry);
            (void)assign_pair_kernels(e);
            g_err = msg;
        }
        return code;
    };
    if (save() != hipSuccess) return finish(fail(LJ_E_HIP, "autotune: save failed"), true);
    // candidates: all tiled, all per-particle, all several-wave tiled, and
    // splits by cell occupancy
    // whose tiled sets differ from those and from each other
    struct Candidate {
        int variant;
        double occupancy;
    };
    std::vector<Candidate> cands = {{LJ_PAIR_CELL_TILED, 0.0}, {LJ_PAIR_PER_PARTICLE, 0.0},
                                    {LJ_PAIR_CELL_MULTIWAVE, 0.0}};
    {
        std::vector<int> seen = {0, e->nsys};
        for (double thr : {1.5, 3.0, 6.0, 12.0}) {
            int nt = 0;
            for (const SysDev& s : e->sys)
                nt += (s.ncell > 0 && (double)s.n >= thr * (double)s.ncell) ? 1 : 0;
            if (std::find(seen.begin(), seen.end(), nt) != seen.end()) continue;
            seen.push_back(nt);
            cands.push_back({LJ_PAIR_BY_OCCUPANCY, thr});
        }
    }
    // One candidate: calls of `steps` and of 2 `steps` steps, timed with
    // events; their difference over `steps` is the time per step of long
    // calls (what a call costs once cancels out).
    auto trial = [&](double* per_step) -> int {
        int trc;
        if ((trc = lj_engine_reserve(e, 2 * steps))) return trc;   // graphs built untimed
        // warm-up (resident trials: the timed calls then both start on live cells)
        if ((trc = lj_engine_step(e, kResMinSteps, nullptr))) return trc;
        float ms[2] = {0, 0};
        for (int r = 0; r < 2; ++r) {
            hipError_t h = hipEventRecord(e->ev0, e->stream);
            if (h != hipSuccess) return fail(LJ_E_HIP, "autotune: event record failed");
            if ((trc = lj_engine_step(e, (r + 1) * steps, nullptr))) return trc;
            h = hipEventRecord(e->ev1, e->stream);
            if (h == hipSuccess) h = hipEventSynchronize(e->ev1);
            if (h == hipSuccess) h = hipEventElapsedTime(&ms[r], e->ev0, e->ev1);
            if (h != hipSuccess) return fail(LJ_E_HIP, "autotune: event timing failed");
        }
        *per_step = std::max(0.0, (double)(ms[1] - ms[0])) / steps;
        return LJ_OK;
    };
    double best = 1e300;
    Candidate best_c = {variant0, occ0};
    int best_res = e->res_mode;
    for (const Candidate& c : cands) {
        e->split_occupancy = c.variant == LJ_PAIR_BY_OCCUPANCY ? c.occupancy : occ0;
        if ((rc = lj_engine_set_pair_kernel(e, c.variant))) return finish(rc, false);
        // tiled assignments with and without resident cells
        for (int res : {LJ_RESIDENT_ON, LJ_RESIDENT_OFF}) {
            e->res_mode = res;
            if (res == LJ_RESIDENT_ON && !res_eligible(e, 2 * steps)) continue;
            double ms = 0;
            if ((rc = trial(&ms))) { e->res_mode = res_entry; return finish(rc, false); }
            if (ms < best) { best = ms; best_c = c; best_res = res; }
            if (restore() != hipSuccess) {
                e->res_mode = res_entry;
                return finish(fail(LJ_E_HIP, "autotune: restore failed"), true);
            }
        }
    }
    e->res_mode = best_res;
    e->split_occupancy = best_c.variant == LJ_PAIR_BY_OCCUPANCY ? best_c.occupancy : occ0;
    if ((rc = lj_engine_set_pair_kernel(e, best_c.variant))) return finish(rc, true);
    // then the launch shape with that assignment: waves per workgroup of the
    // mixed tiled kernel and particles per sort wave (neither changes any
    // result: the tiled kernels keep one record of sums per cell, and the
    // sort's order does not depend on how cells are grouped into waves)
    const int cpb0 = e->cells_per_block, sw0 = e->sort_wave;
    int best_cpb = cpb0, best_sw = sw0;
    std::vector<std::pair<int, int>> shapes;
    const bool tiled_mixed = e->precision == LJ_PRECISION_MIXED && e->n_tiled > 0;
    for (int cpb : {2, 4, 8})
        if (tiled_mixed || cpb == cpb0)
            for (int sw : {32, 44, 56}) shapes.push_back({cpb, sw});
    for (const auto& sh : shapes) {
        if (sh.first == cpb0 && sh.second == sw0) continue;   // timed above
        if ((rc = lj_engine_set_launch_shape(e, sh.first, sh.second))) return finish(rc, true);
        double ms = 0;
        if ((rc = trial(&ms))) return finish(rc, false);
        if (ms < best) { best = ms; best_cpb = sh.first; best_sw = sh.second; }
        if (restore() != hipSuccess) return finish(fail(LJ_E_HIP, "autotune: restore failed"), true);
    }
    if ((rc = lj_engine_set_launch_shape(e, best_cpb, best_sw))) return finish(rc, true);
    (void)finish(LJ_OK, true);
    if (chosen) *chosen = best_c.variant;
    if (ms_per_step) *ms_per_step = best;
    return LJ_OK;
}

int lj_engine_set_profiling(lj_engine* e, int enabled)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = flush_timers(e))) return rc;
    e->profiling = enabled != 0;
    e->timers.clear();
    return LJ_OK;
}

int lj_engine_kernel_times(lj_engine* e, int cap, const char** names, double* ms, int* count)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = flush_timers(e))) return rc;
    int k = 0;
    for (const KernelTimer& t : e->timers) {
        if (!t.calls) continue;
        if (k < cap) {
            if (names) names[k] = t.name;
            if (ms) ms[k] = t.total_ms / (double)t.calls;
        }
        ++k;
    }
    if (count) *count = k;
    return LJ_OK;
}

int lj_engine_kernel_launches(lj_engine* e, int cap, long long* launches, int* count)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = flush_timers(e))) return rc;
    int k = 0;
    for (const KernelTimer& t : e->timers) {
        if (!t.calls) continue;
        if (k < cap && launches) launches[k] = t.calls;
        ++k;
    }
    if (count) *count = k;
    return LJ_OK;
}

}  // extern "C"


extern "C" int lj_engine_adjoint(lj_engine* e, int nsteps, int every, double density, double* loss,
                                 double* grad_eps, double* grad_sigma, double* grad_pos,
                                 double* grad_vel)
{
    int rc = check_engine(e);
    if (rc) return rc;
    if ((rc = res_settle(e))) return rc;
    if (e->nsys != 1 || e->precision != LJ_PRECISION_FP64 || e->sys[0].law != LJ_LAW_LJ)
        return fail(LJ_E_INVALID, "the adjoint needs a single-system FP64 LJ engine");
    if (nsteps < 1 || every < 1) return fail(LJ_E_INVALID, "nsteps and checkpoint interval must be >= 1");
    const SysDev S = e->sys[0];
    const int n = S.n, n3 = 3 * n;
    const int nck = nsteps / every + 1;
    const int nadj = (n + ljk::kAdjPerBlock - 1) / ljk::kAdjPerBlock;   // adjoint pair blocks
    const double dt = S.dt, c1 = 1.0 / nsteps;
    const double cT = c1 * 2.0 * density / (3.0 * n);   // d/dv of c1 * rho * T
    const double cTke = 2.0 * density / (3.0 * n);       // rho * T = cTke * KE
    const double cW = density / (3.0 * n);
    const double eps = S.eps, sig = std::sqrt(S.sigma2);
    DevBuf ck_pos, ck_vel, seg_pos, seg_vel, lam_r, lam_v, pend, phi, jphi, gradw, frc, part, acc;
    DevBuf disp0, time0;
    if ((rc = dev_alloc(&disp0.p, n3)) || (rc = dev_alloc(&time0.p, 1)) ||
        (rc = dev_alloc(&ck_pos.p, (size_t)nck * n3)) || (rc = dev_alloc(&ck_vel.p, (size_t)nck * n3)) ||
        (rc = dev_alloc(&seg_pos.p, (size_t)(every + 1) * n3)) ||
        (rc = dev_alloc(&seg_vel.p, (size_t)(every + 1) * n3)) || (rc = dev_alloc(&lam_r.p, n3)) ||
        (rc = dev_alloc(&lam_v.p, n3)) || (rc = dev_alloc(&pend.p, n3)) || (rc = dev_alloc(&phi.p, n3)) ||
        (rc = dev_alloc(&jphi.p, n3)) || (rc = dev_alloc(&gradw.p, n3)) || (rc = dev_alloc(&frc.p, n3)) ||
        (rc = dev_alloc(&part.p, 5 * (size_t)nadj)) || (rc = dev_alloc(&acc.p, 2)))
        return rc;
    const dim3 gb(S.nblk), tb(kBlock), gv((n3 + kBlock - 1) / kBlock), ga(nadj);
    auto to_id = [&](double* pos, double* vel, double* disp = nullptr) {
        const ljk::StepArgs p = e->args(e->cur);
        hipLaunchKernelGGL(ljk::k_to_id, gb, tb, 0, e->stream, p, 0, n, pos, vel, disp);
    };
    auto from_id = [&](const double* pos, const double* vel, const double* disp = nullptr) {
        const ljk::StepArgs p = e->args(e->cur);
        hipLaunchKernelGGL(ljk::k_from_id, dim3((n + kBlock - 1) / kBlock), tb, 0, e->stream, p, 0,
                           n, pos, vel, disp);
        e->sorted = false;
    };
    auto forces_now = [&]() -> int {
        int r = enqueue_cells(e, false);
        if (!r) r = enqueue_pair(e, false, false);
        if (!r) r = enqueue_finalize(e, 1);
        return r;
    };
    auto adj_pair = [&](const double* ph) {
        ljk::AdjArgs a;
        a.p = e->args(e->cur);
        a.sys = 0;
        a.n = n;
        a.phi = ph;
        a.jphi = jphi.p;
        a.gradw = gradw.p;
        a.force = frc.p;
        a.part = part.p;
        hipLaunchKernelGGL(ljk::k_adjoint_pair, ga, tb, 0, e->stream, a);
    };
    const bool saved_graphs = e->use_graphs;
    // the displacements and the clock on entry, given back at the end (the
    // segments below integrate from checkpoints with cleared displacements)
    HIP_TRY(hipMemcpyAsync(time0.p, e->d_time, sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    // ---- forward: F(r0), then nsteps with checkpoints every `every` steps
    if ((rc = forces_now())) return rc;
    to_id(ck_pos.p, ck_vel.p, disp0.p);
    std::vector<lj_thermo> rec(nsteps);
    for (int t = 0; t < nsteps;) {
        const int k = std::min(every, nsteps - t);
        if ((rc = lj_engine_step(e, k, rec.data() + t))) return rc;
        if ((rc = res_settle(e))) return rc;   // the checkpoint reads the packed set
        t += k;
        if (t % every == 0) to_id(ck_pos.p + (size_t)(t / every) * n3, ck_vel.p + (size_t)(t / every) * n3);
    }
    double L = 0;
    for (int t = 0; t < nsteps; ++t)
        L += rec[t].potential_energy + cTke * rec[t].kinetic_energy + cW * rec[t].virial;
    L *= c1;
    // ---- reverse sweep over segments, recomputing each from its checkpoint
    HIP_TRY(hipMemsetAsync(lam_r.p, 0, n3 * sizeof(double), e->stream));
    HIP_TRY(hipMemsetAsync(lam_v.p, 0, n3 * sizeof(double), e->stream));
    HIP_TRY(hipMemsetAsync(pend.p, 0, n3 * sizeof(double), e->stream));
    HIP_TRY(hipMemsetAsync(acc.p, 0, 2 * sizeof(double), e->stream));
    // one segment of len steps from the checkpoint just loaded: the forces,
    // the recomputed trajectory, then the reverse sweep over it
    auto segment = [&](int len) -> int {
        int r;
        if ((r = forces_now())) return r;
        to_id(seg_pos.p, seg_vel.p);
        for (int k = 1; k <= len; ++k) {
            if ((r = enqueue_step(e))) return r;
            to_id(seg_pos.p + (size_t)k * n3, seg_vel.p + (size_t)k * n3);
        }
        if ((r = enqueue_finalize(e, 0))) return r;
        for (int k = len; k >= 1; --k) {
            const double* vt = seg_vel.p + (size_t)k * n3;
            hipLaunchKernelGGL(ljk::k_adj_pre, gv, tb, 0, e->stream, n3, 0.5 * dt, cT, lam_v.p, vt,
                               pend.p, phi.p);
            from_id(seg_pos.p + (size_t)k * n3, vt);
            if ((r = enqueue_cells(e, false))) return r;
            adj_pair(phi.p);
            hipLaunchKernelGGL(ljk::k_adj_acc, dim3(1), tb, 0, e->stream, part.p, nadj, c1, cW, eps,
                               sig, acc.p);
            hipLaunchKernelGGL(ljk::k_adj_post, gv, tb, 0, e->stream, n3, dt, c1, cW, cT, frc.p,
                               gradw.p, jphi.p, vt, lam_r.p, lam_v.p, pend.p);
        }
        return LJ_OK;
    };
    const int nseg = (nsteps + every - 1) / every;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int t0 = sg * every, t1 = std::min(t0 + every, nsteps);
        from_id(ck_pos.p + (size_t)sg * n3, ck_vel.p + (size_t)sg * n3);
        if ((rc = segment(t1 - t0))) return rc;
    }
    // ---- t = 0: the first half kick used F(r0)
    from_id(ck_pos.p, ck_vel.p);
    if ((rc = enqueue_cells(e, false))) return rc;
    adj_pair(pend.p);
    hipLaunchKernelGGL(ljk::k_adj_acc, dim3(1), tb, 0, e->stream, part.p, nadj, 0.0, 0.0, eps, sig,
                       acc.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(e->stream));
    std::vector<double> lr(n3), jp(n3), lv(n3);
    double hacc[2];
    HIP_TRY(hipMemcpy(lr.data(), lam_r.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(jp.data(), jphi.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lv.data(), lam_v.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hacc, acc.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
    // leave the engine holding the initial state (positions, velocities,
    // displacements, clock) with F(r0)
    from_id(ck_pos.p, ck_vel.p, disp0.p);
    if ((rc = forces_now())) return rc;
    HIP_TRY(hipMemcpyAsync(e->d_time, time0.p, sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->use_graphs = saved_graphs;
    if (loss) *loss = L;
    if (grad_eps) *grad_eps = hacc[0];
    if (grad_sigma) *grad_sigma = hacc[1];
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            if (grad_pos) grad_pos[3 * (size_t)i + a] = lr[(size_t)a * n + i] + jp[(size_t)a * n + i];
            if (grad_vel) grad_vel[3 * (size_t)i + a] = lv[(size_t)a * n + i];
        }
    return LJ_OK;
}
