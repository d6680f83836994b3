// simulation.cpp -- api::Simulation / SimulationController / phases / sinks
// over the GPU engine.
//
// Follows (paths relative to the reference repository root):
//   api/simulation.cpp:48-155           construction and run()
//   control/simulation_controller.cpp:39-146   command loop and log ordering
//   control/simulation_phase.cpp:40-135        equilibration / observation decisions
//   control/simulation_phase.hpp:120-244       phase clocks
//   physics/analyzers.cpp:44-159, physics/analyzers.hpp:70-110   statistics
//   tools/moving_sample.hpp:81-189      fixed window, Bessel-corrected moments
//   output/sinks.cpp:40-204             file formats
//
// The engine advances straight to the next time step at which a phase can
// act (an adjustment check, an observation, steady state, timeout); between
// those steps a phase only collects samples and asks to advance, so the
// result equals evaluating every step.  The per-step records come back in
// one copy and feed the analyzers in order.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "lj_engine.h"
#include "lj_host.h"

namespace lj_host {
std::string format_double(double v);
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m)
{
    g_err = m;
    return code;
}

std::string fmt_g4(double v)
{
    char b[64];
    std::snprintf(b, sizeof(b), "%.4g", v);
    return b;
}

double relative_error(double measured, double desired) { return std::abs(measured / desired - 1.0); }

// tools::MovingSample<double>: keeps the last `size` values, oldest first.
struct Sample {
    explicit Sample(int size) : cap(size) {}
    void push(double v)
    {
        if ((int)buf.size() == cap) buf.pop_front();
        buf.push_back(v);
    }
    double mean() const
    {
        return std::reduce(buf.begin(), buf.end(), 0.0) / static_cast<double>(buf.size());
    }
    double variance(double m) const
    {
        return std::transform_reduce(buf.begin(), buf.end(), 0.0, std::plus<double>(),
                                     [m](double v) { return (v - m) * (v - m); }) /
               static_cast<double>(buf.size() - 1);
    }
    int cap;
    std::deque<double> buf;
};

// tools::MovingSample<Eigen::Vector2d> for (time, msd): mean vector and
// Bessel-corrected covariance matrix.
struct V2 {
    double a = 0, b = 0;
};
V2 operator+(V2 x, V2 y) { return {x.a + y.a, x.b + y.b}; }
struct M2 {
    double aa = 0, ab = 0, bb = 0;
};
M2 operator+(M2 x, M2 y) { return {x.aa + y.aa, x.ab + y.ab, x.bb + y.bb}; }

struct Sample2 {
    explicit Sample2(int size) : cap(size) {}
    void push(V2 v)
    {
        if ((int)buf.size() == cap) buf.pop_front();
        buf.push_back(v);
    }
    // returns covariance(0,1) / covariance(0,0)
    double slope() const
    {
        const double n = static_cast<double>(buf.size());
        V2 s = std::reduce(buf.begin(), buf.end(), V2{});
        V2 m{s.a / n, s.b / n};
        M2 c = std::transform_reduce(buf.begin(), buf.end(), M2{}, std::plus<M2>(), [m](V2 v) {
            const double da = v.a - m.a, db = v.b - m.b;
            return M2{da * da, da * db, db * db};
        });
        const double d = n - 1;
        return (c.ab / d) / (c.aa / d);
    }
    int cap;
    std::deque<V2> buf;
};

struct Observation {
    double temperature, density, total_energy, pressure, specific_heat, diffusion_coefficient;
};

// physics/analyzers.cpp:44-159: the observation of one window of samples --
// means of T and E, P = rho (<T> + <W> / 3N) (:73-75), the heat capacity from
// the temperature variance (Lebowitz-Percus-Verlet, :113-122) and D = slope of
// MSD against time / 6 (:146-148).
Observation analyze(const Sample& temps, const Sample& energies, const Sample& virials,
                    const Sample2& msd, double density, int n)
{
    const double T = temps.mean();
    const double E = energies.mean();
    const double Wm = virials.mean();
    const double P = density * (T + Wm / (3.0 * n));
    const double rel = temps.variance(T) / (T * T);
    const double cv = (3. / 2.) / (1 - (3. / 2.) * n * rel);
    const double D = (1. / 6.) * msd.slope();
    return {T, density, E, P, cv, D};
}

enum class Cmd { Advance, Observe, Adjust, Complete, Abort };

struct Command {
    Cmd kind;
    double measured = 0, target = 0;
    Observation obs{};
    std::string reason;
};

struct Phase {
    virtual ~Phase() = default;
    virtual void set_start_time(int t) = 0;
    // next time step (> t) at which evaluate can issue anything but AdvanceTime
    virtual int next_event(int t) const = 0;
    virtual void collect(const lj_thermo& r) = 0;
    virtual void evaluate(std::deque<Command>& q, int t) = 0;
    std::string name;
};

// control/simulation_phase.cpp:40-84
struct Equilibration : Phase {
    double temperature, tolerance;
    int interval, steady, timeout;
    Sample temps;
    int start = 0, last_check = 0, last_adjust = 0;
    Equilibration(std::string nm, double T, double tol, int sample, int adj, int st, int to)
        : temperature(T), tolerance(tol), interval(adj), steady(st), timeout(to), temps(sample)
    {
        name = std::move(nm);
    }
    void set_start_time(int t) override { start = last_check = last_adjust = t; }
    int next_event(int t) const override
    {
        long e = std::min<long>({(long)last_check + interval, (long)last_adjust + steady,
                                 (long)start + timeout});
        return (int)std::max<long>(e, (long)t + 1);
    }
    void collect(const lj_thermo& r) override { temps.push(r.temperature); }
    void evaluate(std::deque<Command>& q, int t) override
    {
        if (t - last_check >= interval) {
            last_check = t;
            const double measured = temps.mean();
            if (relative_error(measured, temperature) >= tolerance) {
                last_adjust = t;
                Command c{Cmd::Adjust};
                c.measured = measured;
                c.target = temperature;
                q.push_back(c);
            }
        }
        if (t - last_adjust >= steady) {
            q.push_back({Cmd::Complete});
            return;
        }
        if (t - start >= timeout) {
            Command c{Cmd::Abort};
            c.reason = "Could not equilibrate";
            q.push_back(c);
            return;
        }
        q.push_back({Cmd::Advance});
    }
};

// control/simulation_phase.cpp:94-135 with physics/analyzers.cpp:44-159
struct ObservationP : Phase {
    double temperature, density, tolerance;
    int n, interval, count_target;
    Sample temps, energies, virials;
    Sample2 msd;
    int start = 0, last_obs = 0, count = 0;
    ObservationP(std::string nm, double T, double rho, int N, double tol, int sample, int iv,
                 int cnt)
        : temperature(T), density(rho), tolerance(tol), n(N), interval(iv), count_target(cnt),
          temps(sample), energies(sample), virials(sample), msd(sample)
    {
        name = std::move(nm);
    }
    void set_start_time(int t) override { start = last_obs = t; }
    int next_event(int t) const override
    {
        if (count >= count_target) return t + 1;
        return std::max(last_obs + interval, t + 1);
    }
    void collect(const lj_thermo& r) override
    {
        temps.push(r.temperature);
        energies.push(r.total_energy);
        virials.push(r.virial);
        msd.push({r.time, r.mean_square_displacement});
    }
    Observation result() const { return analyze(temps, energies, virials, msd, density, n); }
    void evaluate(std::deque<Command>& q, int t) override
    {
        if (t - last_obs >= interval) {
            last_obs = t;
            Observation o = result();
            if (relative_error(o.temperature, temperature) >= tolerance) {
                Command c{Cmd::Abort};
                c.reason = "Temperature drifted outside acceptable range";
                q.push_back(c);
                return;
            }
            ++count;
            Command c{Cmd::Observe};
            c.obs = o;
            q.push_back(c);
        }
        if (count >= count_target) {
            q.push_back({Cmd::Complete});
            return;
        }
        q.push_back({Cmd::Advance});
    }
};

std::string F(double v) { return lj_host::format_double(v); }

}  // namespace

struct lj_simulation {
    // configuration (api/configuration.hpp:35-102)
    double temperature, density;
    int n;
    unsigned seed;
    double cutoff, dt;
    std::string eq_name, ob_name;
    double eq_tol, ob_tol;
    int eq_sample, eq_interval, eq_steady, eq_timeout;
    int ob_sample, ob_interval, ob_count;
    std::string f_events, f_thermo, f_obs, f_snap;
    int precision;
    // initial condition (engine/initial_condition.cpp:49-104)
    std::vector<double> pos0, vel0;
    double box = 0;
    double alpha = 0, beta = 0;
};

namespace {

// One simulation's controller state while a (batched) run is in progress.
struct Run {
    lj_simulation* sim;
    int sys;
    std::ofstream ev, th, ob, sn;
    bool echo;
    std::vector<std::unique_ptr<Phase>> phases;   // front = current
    std::deque<Command> queue;
    int time_step = 0;
    bool done = false;
    int status = 0;   // 0 completed, 1 equilibration aborted, 2 observation aborted
    bool in_observation = false;

    void event(const std::string& line)
    {
        ev << line << '\n';
        ev.flush();
        if (echo) std::cout << line << std::endl;
    }
};

int open_run(Run& r)
{
    lj_simulation* s = r.sim;
    r.ev.open(s->f_events);
    r.th.open(s->f_thermo);
    r.ob.open(s->f_obs);
    r.sn.open(s->f_snap);
    if (!r.ev || !r.th || !r.ob || !r.sn) return fail(LJ_E_INVALID, "cannot open output files");
    // output/sinks.cpp:100-112, 132-144, 162-189 (headers)
    r.th << "TimeStep,Time,KineticEnergy,PotentialEnergy,TotalEnergy,Virial,Temperature,"
            "MeanSquareDisplacement\n";
    r.ob << "TimeStep,Temperature,Density,TotalEnergy,Pressure,SpecificHeat,DiffusionCoefficient\n";
    r.sn << "TimeStep,ParticleID,Position,Position,Position,Velocity,Velocity,Velocity,Force,"
            "Force,Force\n";
    r.sn << "TimeStep,ParticleID,X,Y,Z,X,Y,Z,X,Y,Z\n";
    r.phases.clear();
    r.phases.emplace_back(new Equilibration(s->eq_name, s->temperature, s->eq_tol, s->eq_sample,
                                            s->eq_interval, s->eq_steady, s->eq_timeout));
    r.phases.emplace_back(new ObservationP(s->ob_name, s->temperature, s->density, s->n,
                                           s->ob_tol, s->ob_sample, s->ob_interval, s->ob_count));
    r.time_step = 0;
    r.phases.front()->set_start_time(0);
    r.event("0: Phase started: " + r.phases.front()->name);
    r.queue.clear();
    r.queue.push_back({Cmd::Advance});
    return LJ_OK;
}

int snapshot(lj_engine* e, Run& r)
{
    const int n = r.sim->n;
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), frc(3 * (size_t)n);
    int rc = lj_engine_get_state(e, r.sys, pos.data(), vel.data(), nullptr, frc.data(), nullptr,
                                 nullptr, nullptr);
    if (rc) return fail(rc, lj_last_error());
    std::string line;
    for (int i = 0; i < n; ++i) {
        line = std::to_string(r.time_step) + "," + std::to_string(i);
        for (const std::vector<double>* a : {&pos, &vel, &frc})
            for (int k = 0; k < 3; ++k) line += "," + F((*a)[3 * (size_t)i + k]);
        r.sn << line << '\n';
    }
    return LJ_OK;
}

// Runs queued non-advance commands of one simulation until it needs to
// advance (returns the number of steps it can take before its next event)
// or it has finished (returns 0).
int drain(lj_engine* e, Run& r, int* rc)
{
    *rc = LJ_OK;
    while (!r.queue.empty()) {
        Command c = r.queue.front();
        if (c.kind == Cmd::Advance) {
            return r.phases.front()->next_event(r.time_step) - r.time_step;
        }
        r.queue.pop_front();
        switch (c.kind) {
        case Cmd::Observe: {
            const Observation& o = c.obs;
            r.ob << r.time_step << "," << F(o.temperature) << "," << F(o.density) << ","
                 << F(o.total_energy) << "," << F(o.pressure) << "," << F(o.specific_heat) << ","
                 << F(o.diffusion_coefficient) << '\n';
            r.event(std::to_string(r.time_step) + ": Observation recorded");
            break;
        }
        case Cmd::Adjust: {
            int st = lj_engine_set_temperature(e, r.sys, c.target, nullptr);
            if (st) { *rc = fail(st, lj_last_error()); return 0; }
            r.event(std::to_string(r.time_step) + ": Temperature measured at: " +
                    fmt_g4(c.measured) + ", adjusted to: " + fmt_g4(c.target));
            break;
        }
        case Cmd::Complete: {
            r.event(std::to_string(r.time_step) + ": Phase complete: " + r.phases.front()->name);
            r.phases.erase(r.phases.begin());
            if (r.phases.empty()) {
                if ((*rc = snapshot(e, r))) return 0;
                r.done = true;
                r.status = 0;
            } else {
                r.in_observation = true;
                r.phases.front()->set_start_time(r.time_step);
                r.queue.push_back({Cmd::Advance});
                r.event(std::to_string(r.time_step) + ": Phase started: " + r.phases.front()->name);
            }
            break;
        }
        case Cmd::Abort: {
            r.event(std::to_string(r.time_step) + ": Simulation aborted: " + c.reason);
            if ((*rc = snapshot(e, r))) return 0;
            r.done = true;
            r.status = r.in_observation ? 2 : 1;
            break;
        }
        default:
            break;
        }
    }
    r.done = true;
    return 0;
}

void thermo_line(Run& r, int ts, const lj_thermo& t)
{
    r.th << ts << "," << F(t.time) << "," << F(t.kinetic_energy) << "," << F(t.potential_energy)
         << "," << F(t.total_energy) << "," << F(t.virial) << "," << F(t.temperature) << ","
         << F(t.mean_square_displacement) << '\n';
}

int run_batch(lj_simulation** sims, int nsim, int echo, int* status)
{
    if (nsim < 1) return LJ_OK;
    const int prec = sims[0]->precision;
    std::vector<lj_system_desc> desc(nsim);
    for (int k = 0; k < nsim; ++k) {
        lj_simulation* s = sims[k];
        if (s->precision != prec) return fail(LJ_E_INVALID, "mixed precisions in one batch");
        lj_system_desc& d = desc[k];
        std::memset(&d, 0, sizeof(d));
        d.particle_count = s->n;
        d.box[0] = d.box[1] = d.box[2] = s->box;
        d.cutoff = s->cutoff;
        d.time_delta = s->dt;
        d.epsilon = 1.0;
        d.sigma = 1.0;
        d.force_law = LJ_LAW_LJ;
        d.periodic = 1;
    }
    lj_engine* e = nullptr;
    int rc = lj_engine_create(desc.data(), nsim, prec, &e);
    if (rc) return fail(rc, lj_last_error());
    // tiled pair kernel for dense systems, per-particle for dilute ones: a
    // rule of each system alone, so a batch and a single run of the same
    // system give identical output
    if ((rc = lj_engine_set_pair_kernel(e, LJ_PAIR_BY_OCCUPANCY))) {
        lj_engine_destroy(e);
        return fail(rc, lj_last_error());
    }
    std::vector<Run> runs(nsim);
    for (int k = 0; k < nsim; ++k) {
        runs[k].sim = sims[k];
        runs[k].sys = k;
        runs[k].echo = echo != 0;
        if ((rc = lj_engine_set_state(e, k, sims[k]->pos0.data(), sims[k]->vel0.data(), nullptr,
                                      nullptr, 0.0))) {
            lj_engine_destroy(e);
            return fail(rc, lj_last_error());
        }
        if ((rc = open_run(runs[k]))) { lj_engine_destroy(e); return rc; }
        if (status) status[k] = 3;
    }
    std::vector<lj_thermo> rec;
    std::vector<int> want(nsim);
    while (true) {
        int chunk = std::numeric_limits<int>::max();
        int live = 0;
        for (int k = 0; k < nsim; ++k) {
            Run& r = runs[k];
            if (r.done) { want[k] = 0; continue; }
            int drc;
            want[k] = drain(e, r, &drc);
            if (drc) { lj_engine_destroy(e); return drc; }
            if (r.done) {
                if (status) status[k] = r.status;
                if ((rc = lj_engine_set_active(e, k, 0))) { lj_engine_destroy(e); return fail(rc, lj_last_error()); }
                continue;
            }
            ++live;
            chunk = std::min(chunk, want[k]);
        }
        if (!live) break;
        chunk = std::max(1, std::min(chunk, 1 << 16));
        rec.resize((size_t)chunk * nsim);
        if ((rc = lj_engine_step(e, chunk, rec.data()))) {
            lj_engine_destroy(e);
            return fail(rc, lj_last_error());
        }
        // feed the records in time order: AdvanceTime logs with the old
        // time step, then evaluates at time_step + 1 (simulation_controller.cpp:60-69)
        for (int k = 0; k < nsim; ++k) {
            Run& r = runs[k];
            if (r.done) continue;
            for (int s = 0; s < chunk; ++s) {
                const lj_thermo& t = rec[(size_t)s * nsim + k];
                thermo_line(r, r.time_step, t);
                r.time_step += 1;
                r.phases.front()->collect(t);
                if (s == chunk - 1) {
                    r.queue.pop_front();   // the pending AdvanceTime
                    r.phases.front()->evaluate(r.queue, r.time_step);
                }
            }
        }
    }
    for (int k = 0; k < nsim; ++k) {
        runs[k].ev.close();
        runs[k].th.close();
        runs[k].ob.close();
        runs[k].sn.close();
        if (status) status[k] = runs[k].status;
    }
    lj_engine_destroy(e);
    return LJ_OK;
}

}  // namespace

extern "C" {

const char* lj_host_last_error(void) { return g_err.c_str(); }

int lj_simulation_create(const lj_sim_config* c, lj_simulation** out)
{
    if (!out || !c) return fail(LJ_E_INVALID, "null argument");
    *out = nullptr;
    if (c->particle_count < 1 || !(c->density > 0) || !(c->temperature > 0) || !(c->time_delta > 0))
        return fail(LJ_E_INVALID, "invalid system parameters");
    if (!(c->cutoff_distance * c->cutoff_distance > std::cbrt(2.0)))
        return fail(LJ_E_INVALID, "Cutoff distance is too short");
    auto s = new lj_simulation();
    s->temperature = c->temperature;
    s->density = c->density;
    s->n = c->particle_count;
    s->seed = c->random_seed;
    s->cutoff = c->cutoff_distance;
    s->dt = c->time_delta;
    s->eq_name = c->equilibration_name ? c->equilibration_name : "Equilibration Phase";
    s->ob_name = c->observation_name ? c->observation_name : "Observation Phase";
    s->eq_tol = c->equilibration_tolerance;
    s->eq_sample = c->equilibration_sample_size;
    s->eq_interval = c->adjustment_interval;
    s->eq_steady = c->steady_state_time;
    s->eq_timeout = c->equilibration_timeout;
    s->ob_tol = c->observation_tolerance;
    s->ob_sample = c->observation_sample_size;
    s->ob_interval = c->observation_interval;
    s->ob_count = c->observation_count;
    s->f_events = c->event_log ? c->event_log : "events.log";
    s->f_thermo = c->thermodynamic_log ? c->thermodynamic_log : "thermodynamics.csv";
    s->f_obs = c->observation_log ? c->observation_log : "observations.csv";
    s->f_snap = c->snapshot_log ? c->snapshot_log : "snapshots.csv";
    s->precision = c->precision;
    if (s->eq_sample < 1 || s->ob_sample < 2 || s->eq_interval < 1 || s->ob_interval < 1) {
        delete s;
        return fail(LJ_E_INVALID, "invalid phase parameters");
    }
    s->pos0.resize(3 * (size_t)s->n);
    s->vel0.resize(3 * (size_t)s->n);
    s->box = lj_initial_condition(s->n, s->density, s->temperature, s->seed,
                                  LJ_LATTICE_FACE_CENTERED, s->pos0.data(), s->vel0.data());
    if (!(s->box > 0)) {
        delete s;
        return fail(LJ_E_INVALID, "invalid initial condition");
    }
    if (s->box < s->cutoff) {
        delete s;
        return fail(LJ_E_INVALID, "Simulation box size is less than the cutoff distance");
    }
    const double rc2 = s->cutoff * s->cutoff;
    const double rcm6 = 1 / (rc2 * rc2 * rc2);
    s->alpha = -4.0 * rcm6 * (rcm6 - 1.0);
    s->beta = 12.0 * rcm6 * (2.0 * rcm6 - 1.0);
    *out = s;
    return LJ_OK;
}

int lj_simulation_destroy(lj_simulation* sim)
{
    delete sim;
    return LJ_OK;
}

int lj_simulation_run(lj_simulation* sim, int echo)
{
    if (!sim) return fail(LJ_E_INVALID, "null simulation");
    return run_batch(&sim, 1, echo, nullptr);
}

int lj_simulation_run_batch(lj_simulation** sims, int nsim, int echo, int* status)
{
    if (!sims || nsim < 0) return fail(LJ_E_INVALID, "bad batch");
    for (int k = 0; k < nsim; ++k)
        if (!sims[k]) return fail(LJ_E_INVALID, "null simulation in batch");
    return run_batch(sims, nsim, echo, status);
}

// physics/lennard_jones_force.cpp:76-128 along z (physics/forces.hpp:80-94)
static void lj_along_z(const lj_simulation* s, double r, double* v, double* w, double* f)
{
    const double r2 = r * r, rc2 = s->cutoff * s->cutoff;
    if (!(r2 < rc2)) { *v = *w = *f = 0; return; }
    const double s6 = 1.0 / (r2 * r2 * r2);
    *v = 4.0 * s6 * (s6 - 1.0) + s->alpha + s->beta * (r2 / rc2 - 1.0);
    *w = 24.0 * s6 * (2.0 * s6 - 1.0) - 2.0 * s->beta * (r2 / rc2);
    *f = *w * r / r2;
}

double lj_simulation_potential(const lj_simulation* s, double d)
{
    double v, w, f;
    lj_along_z(s, d, &v, &w, &f);
    return v;
}

double lj_simulation_virial(const lj_simulation* s, double d)
{
    double v, w, f;
    lj_along_z(s, d, &v, &w, &f);
    return w;
}

double lj_simulation_force(const lj_simulation* s, double d)
{
    double v, w, f;
    lj_along_z(s, d, &v, &w, &f);
    return f;
}

}  // extern "C"

extern "C" int lj_thermo_observation(const lj_thermo* records, int count, int sample_size,
                                     double density, int particle_count, double* out)
{
    if (!records || !out || count < 1 || sample_size < 1 || particle_count < 1)
        return fail(LJ_E_INVALID, "lj_thermo_observation: bad arguments");
    Sample temps(sample_size), energies(sample_size), virials(sample_size);
    Sample2 msd(sample_size);
    for (int k = 0; k < count; ++k) {
        const lj_thermo& r = records[k];
        temps.push(r.temperature);
        energies.push(r.total_energy);
        virials.push(r.virial);
        msd.push({r.time, r.mean_square_displacement});
    }
    const Observation o = analyze(temps, energies, virials, msd, density, particle_count);
    const double v[6] = {o.temperature, o.density, o.total_energy, o.pressure, o.specific_heat,
                         o.diffusion_coefficient};
    std::copy(v, v + 6, out);
    return LJ_OK;
}
