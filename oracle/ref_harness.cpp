// Test harness around the compiled Kessler_Sim reference core.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled together with the
// reference's own sources (src/core/{orbital,collision,breakup,threads}.cpp,
// read in place from the reference checkout) into oracle/_ref/libkessler_ref.so
// by oracle/Makefile. It adds nothing to the algorithm: every function below
// forwards to the reference's public C++ API (orbital.hpp, collision.hpp,
// collision_filter.hpp, breakup.hpp) and exposes it with plain C types so
// that tests/golden/make_golden.py and bench.py's CPU-baseline leg can call
// it through ctypes.
//
// The product never links or loads this library.

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "breakup.hpp"
#include "collision.hpp"
#include "collision_filter.hpp"
#include "orbital.hpp"
#include "threads.hpp"

using namespace kessler;

extern "C" {

struct RefEvent {
    int32_t i, j;
    int64_t step;
    double t, miss_km, v_rel_km_s;
    int32_t collision, pad;
};

struct RefStats {
    int64_t candidate_pairs, refined_pairs, co_orbiting;
    double cell_km;
    int32_t max_cell_objects, pad;
};

int ref_configure_threads(int n) { return configure_threads(n); }

void ref_moon_position(double jd, double* out) {
    Vec3 m = moon_position(jd);
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
}

void ref_acceleration(const double* r, const double* rm, double* out) {
    Vec3 a = acceleration({r[0], r[1], r[2]}, {rm[0], rm[1], rm[2]});
    out[0] = a.x; out[1] = a.y; out[2] = a.z;
}

void ref_rk4(const double* s, double dt, const double* m0, const double* m1, const double* m2,
             double* out) {
    State st{{s[0], s[1], s[2]}, {s[3], s[4], s[5]}};
    State r = rk4(st, dt, {m0[0], m0[1], m0[2]}, {m1[0], m1[1], m1[2]}, {m2[0], m2[1], m2[2]});
    out[0] = r.r.x; out[1] = r.r.y; out[2] = r.r.z;
    out[3] = r.v.x; out[4] = r.v.y; out[5] = r.v.z;
}

int ref_filter_pair(const double* r, const double* v, double dt, double thr, double vmin2) {
    return filter_pair(r[0], r[1], r[2], v[0], v[1], v[2], dt, thr, vmin2);
}

int ref_prefilter_keep(const float* r, const float* v, float dt, float thr, float vmin2) {
    return prefilter_keep(r[0], r[1], r[2], v[0], v[1], v[2], dt, thr, vmin2) ? 1 : 0;
}

void ref_interpolate_state(const double* a, const double* b, double dt, double s, double* out) {
    State sa{{a[0], a[1], a[2]}, {a[3], a[4], a[5]}};
    State sb{{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
    State r = interpolate_state(sa, sb, dt, s);
    out[0] = r.r.x; out[1] = r.r.y; out[2] = r.r.z;
    out[3] = r.v.x; out[4] = r.v.y; out[5] = r.v.z;
}

static void fill_objects(std::vector<Object>& objs, int n, const double* states, const uint8_t* alive,
                         const double* size_m) {
    objs.resize(n);
    for (int i = 0; i < n; ++i) {
        const double* s = states + 6 * static_cast<size_t>(i);
        objs[i].id = i;
        objs[i].s = State{{s[0], s[1], s[2]}, {s[3], s[4], s[5]}};
        objs[i].alive = alive ? alive[i] != 0 : true;
        objs[i].size_m = size_m ? size_m[i] : 0.8;
    }
}

static void copy_events(const std::vector<Conjunction>& found, long step, RefEvent* out, int64_t cap,
                        int64_t& count) {
    for (const Conjunction& c : found) {
        if (out && count < cap) {
            RefEvent& e = out[count];
            e.i = c.i; e.j = c.j; e.step = step; e.t = c.t; e.miss_km = c.miss_km;
            e.v_rel_km_s = c.v_rel_km_s; e.collision = c.collision ? 1 : 0; e.pad = 0;
        }
        ++count;
    }
}

static void copy_stats(const ScreeningStats& st, RefStats* out) {
    out->candidate_pairs = st.candidate_pairs;
    out->refined_pairs = st.refined_pairs;
    out->co_orbiting = st.co_orbiting;
    out->cell_km = st.cell_km;
    out->max_cell_objects = st.max_cell_objects;
    out->pad = 0;
}

// One call of screen_step on explicit before / after states (CPU path).
int64_t ref_screen(int n, const double* before, const double* after, const uint8_t* alive_after,
                   const double* size_m, double t0, double dt, double thr, double radius_scale,
                   double vmin_km_s, RefEvent* out, int64_t cap, RefStats* stats) {
    std::vector<State> b(n);
    for (int i = 0; i < n; ++i) {
        const double* s = before + 6 * static_cast<size_t>(i);
        b[i] = State{{s[0], s[1], s[2]}, {s[3], s[4], s[5]}};
    }
    std::vector<Object> a;
    fill_objects(a, n, after, alive_after, size_m);
    ScreeningConfig cfg;
    cfg.threshold_km = thr;
    cfg.radius_scale = radius_scale;
    cfg.min_encounter_speed_km_s = vmin_km_s;
    ScreeningStats st;
    std::vector<Conjunction> found = screen_step(b, a, t0, dt, cfg, &st);
    if (stats) copy_stats(st, stats);
    int64_t count = 0;
    copy_events(found, 0, out, cap, count);
    return count;
}

// The main.cpp step loop (main.cpp:346-444) without breakup and without file
// output: copy `before`, Screener::begin, Simulation::step, Screener::finish.
// States and alive flags are updated in place. Returns the number of events
// (all of them, even beyond `cap`). `loop_s` receives the loop wall time.
int64_t ref_run(int n, double* states, uint8_t* alive, const double* size_m, double epoch_jd,
                double t_start, double dt, int64_t nsteps, int screen, double thr,
                double radius_scale, double vmin_km_s, RefEvent* out, int64_t cap,
                RefStats* per_step, double* loop_s) {
    Simulation sim;
    fill_objects(sim.objects, n, states, alive, size_m);
    sim.epoch_jd = epoch_jd;
    sim.t = t_start;
    sim.dt = dt;
    ScreeningConfig cfg;
    cfg.threshold_km = thr;
    cfg.radius_scale = radius_scale;
    cfg.min_encounter_speed_km_s = vmin_km_s;
    Screener screener(cfg);
    std::vector<State> before;
    int64_t count = 0;
    auto t0c = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < nsteps; ++k) {
        if (screen) {
            const int nobj = static_cast<int>(sim.objects.size());
            before.resize(nobj);
#pragma omp parallel for schedule(static)
            for (int i = 0; i < nobj; ++i) before[i] = sim.objects[i].s;
            screener.begin(before, sim.objects, sim.dt);
        }
        const double t0 = sim.t;
        sim.step();
        if (screen) {
            ScreeningStats st;
            std::vector<Conjunction> found = screener.finish(before, sim.objects, t0, sim.dt, &st);
            if (per_step) copy_stats(st, per_step + k);
            copy_events(found, static_cast<long>(k), out, cap, count);
        }
    }
    if (loop_s)
        *loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0c).count();
    for (int i = 0; i < n; ++i) {
        const State& s = sim.objects[i].s;
        double* d = states + 6 * static_cast<size_t>(i);
        d[0] = s.r.x; d[1] = s.r.y; d[2] = s.r.z; d[3] = s.v.x; d[4] = s.v.y; d[5] = s.v.z;
        alive[i] = sim.objects[i].alive ? 1 : 0;
    }
    return count;
}

// load_catalog + read_epoch_jd. Two-call protocol: with null buffers only the
// count is returned. rcs_code: 0 SMALL, 1 MEDIUM, 2 LARGE, 3 other.
int ref_load_catalog(const char* path, int cap, int32_t* ids, double* states, double* size_m,
                     int32_t* rcs_code, int32_t* type_code, double* epoch_jd, int* dropped) {
    int bad = 0;
    std::vector<Object> objs = load_catalog(path, &bad);
    if (dropped) *dropped = bad;
    if (epoch_jd) *epoch_jd = read_epoch_jd(path);
    const int n = static_cast<int>(objs.size());
    for (int i = 0; i < n && i < cap; ++i) {
        const Object& o = objs[i];
        if (ids) ids[i] = o.id;
        if (states) {
            double* d = states + 6 * static_cast<size_t>(i);
            d[0] = o.s.r.x; d[1] = o.s.r.y; d[2] = o.s.r.z;
            d[3] = o.s.v.x; d[4] = o.s.v.y; d[5] = o.s.v.z;
        }
        if (size_m) size_m[i] = o.size_m;
        if (rcs_code) rcs_code[i] = o.rcs == "SMALL" ? 0 : o.rcs == "MEDIUM" ? 1 : o.rcs == "LARGE" ? 2 : 3;
        if (type_code) type_code[i] = static_cast<int32_t>(o.type);
    }
    return n;
}

// PortableRng draws: kind 0 uniform, 1 normal, 2 poisson(mean), 3 raw
// mt19937_64 outputs (as doubles of the integer, exact below 2^53 only: use
// kind 4 for the raw bits written into out as uint64).
void ref_rng(uint64_t seed, int kind, double mean, int64_t n, double* out) {
    PortableRng rng(seed);
    std::mt19937_64 eng(seed);
    for (int64_t k = 0; k < n; ++k) {
        if (kind == 0) out[k] = rng.uniform();
        else if (kind == 1) out[k] = rng.normal();
        else if (kind == 2) out[k] = static_cast<double>(rng.poisson(mean));
        else {
            uint64_t x = eng();
            std::memcpy(out + k, &x, sizeof(x));
        }
    }
}

struct RefObject {
    int32_t id, event;
    int32_t alive, type;
    double state[6];
    double size_m, mass_kg, area_to_mass;
};

struct RefBreakup {
    int32_t id, target, projectile, catastrophic;
    double t, v_rel_km_s, energy_j_per_g, target_mass_kg, projectile_mass_kg, sbm_mass_kg,
        expected_fragments, fragments_mass_kg;
    int32_t fragments, escaped;
};

// Applies forced impacts (Fragmentation::impact, breakup.cpp:340-360) in
// sequence on one Fragmentation object, then returns every object. Objects
// 0..n-1 are given with their masses; fragments are appended after them.
// Two-call protocol on the returned object count.
int64_t ref_impacts(int n, const double* states, const double* size_m, const double* mass_kg,
                    double epoch_jd, double t, int nimp, const int32_t* targets,
                    const double* prj_mass, const double* vrel, uint64_t seed, double lc_m,
                    int64_t cap, RefObject* out, RefBreakup* events) {
    Simulation sim;
    fill_objects(sim.objects, n, states, nullptr, size_m);
    for (int i = 0; i < n; ++i) {
        sim.objects[i].mass_kg = mass_kg[i];
        sim.objects[i].id = 100000 + i;
        sim.objects[i].name = "OBJ " + std::to_string(i);
    }
    sim.epoch_jd = epoch_jd;
    sim.t = t;
    BreakupConfig cfg;
    cfg.seed = seed;
    cfg.lc_m = lc_m;
    Fragmentation frag(cfg);
    for (int k = 0; k < nimp; ++k) {
        BreakupEvent ev = frag.impact(sim, targets[k], prj_mass[k], vrel[k]);
        if (events) {
            RefBreakup& b = events[k];
            b.id = ev.id; b.target = ev.target; b.projectile = ev.projectile;
            b.catastrophic = ev.catastrophic ? 1 : 0; b.t = ev.t; b.v_rel_km_s = ev.v_rel_km_s;
            b.energy_j_per_g = ev.energy_j_per_g; b.target_mass_kg = ev.target_mass_kg;
            b.projectile_mass_kg = ev.projectile_mass_kg; b.sbm_mass_kg = ev.sbm_mass_kg;
            b.expected_fragments = ev.expected_fragments; b.fragments_mass_kg = ev.fragments_mass_kg;
            b.fragments = ev.fragments; b.escaped = ev.escaped;
        }
    }
    const int64_t total = static_cast<int64_t>(sim.objects.size());
    for (int64_t i = 0; i < total && i < cap && out; ++i) {
        const Object& o = sim.objects[i];
        RefObject& r = out[i];
        r.id = o.id; r.event = o.event; r.alive = o.alive ? 1 : 0; r.type = static_cast<int32_t>(o.type);
        r.state[0] = o.s.r.x; r.state[1] = o.s.r.y; r.state[2] = o.s.r.z;
        r.state[3] = o.s.v.x; r.state[4] = o.s.v.y; r.state[5] = o.s.v.z;
        r.size_m = o.size_m; r.mass_kg = o.mass_kg; r.area_to_mass = o.area_to_mass;
    }
    return total;
}

// Collisions applied by Fragmentation::apply (breakup.cpp:316-338) for one
// step's conjunction list; same output protocol as ref_impacts.
int64_t ref_apply(int n, const double* before, const double* after, const double* size_m,
                  const double* mass_kg, double epoch_jd, double t0, double dt, int nconj,
                  const RefEvent* conj, uint64_t seed, double lc_m, int64_t cap, RefObject* out,
                  RefBreakup* events, int* nevents) {
    Simulation sim;
    fill_objects(sim.objects, n, after, nullptr, size_m);
    std::vector<State> b(n);
    for (int i = 0; i < n; ++i) {
        const double* s = before + 6 * static_cast<size_t>(i);
        b[i] = State{{s[0], s[1], s[2]}, {s[3], s[4], s[5]}};
        sim.objects[i].mass_kg = mass_kg[i];
        sim.objects[i].id = 100000 + i;
    }
    sim.epoch_jd = epoch_jd;
    sim.t = t0 + dt;
    sim.dt = dt;
    std::vector<Conjunction> cs(nconj);
    for (int k = 0; k < nconj; ++k) {
        cs[k].i = conj[k].i; cs[k].j = conj[k].j; cs[k].t = conj[k].t;
        cs[k].miss_km = conj[k].miss_km; cs[k].v_rel_km_s = conj[k].v_rel_km_s;
        cs[k].collision = conj[k].collision != 0;
    }
    BreakupConfig cfg;
    cfg.seed = seed;
    cfg.lc_m = lc_m;
    Fragmentation frag(cfg);
    std::vector<BreakupEvent> evs = frag.apply(sim, b, t0, dt, cs);
    if (nevents) *nevents = static_cast<int>(evs.size());
    for (size_t k = 0; k < evs.size() && events; ++k) {
        const BreakupEvent& ev = evs[k];
        RefBreakup& r = events[k];
        r.id = ev.id; r.target = ev.target; r.projectile = ev.projectile;
        r.catastrophic = ev.catastrophic ? 1 : 0; r.t = ev.t; r.v_rel_km_s = ev.v_rel_km_s;
        r.energy_j_per_g = ev.energy_j_per_g; r.target_mass_kg = ev.target_mass_kg;
        r.projectile_mass_kg = ev.projectile_mass_kg; r.sbm_mass_kg = ev.sbm_mass_kg;
        r.expected_fragments = ev.expected_fragments; r.fragments_mass_kg = ev.fragments_mass_kg;
        r.fragments = ev.fragments; r.escaped = ev.escaped;
    }
    const int64_t total = static_cast<int64_t>(sim.objects.size());
    for (int64_t i = 0; i < total && i < cap && out; ++i) {
        const Object& o = sim.objects[i];
        RefObject& r = out[i];
        r.id = o.id; r.event = o.event; r.alive = o.alive ? 1 : 0; r.type = static_cast<int32_t>(o.type);
        r.state[0] = o.s.r.x; r.state[1] = o.s.r.y; r.state[2] = o.s.r.z;
        r.state[3] = o.s.v.x; r.state[4] = o.s.v.y; r.state[5] = o.s.v.z;
        r.size_m = o.size_m; r.mass_kg = o.mass_kg; r.area_to_mass = o.area_to_mass;
    }
    return total;
}

// write_snapshot of n objects (ids, states, alive flags) at time t.
void ref_write_snapshot(const char* path, int n, const int* ids, const double* states, const uint8_t* alive,
                        double t) {
    std::vector<Object> objs;
    fill_objects(objs, n, states, alive, nullptr);
    for (int i = 0; i < n; ++i) objs[i].id = ids[i];
    write_snapshot(path, objs, t);
}

// read_snapshot into a list of n objects (ids, fragment flags as event >= 0):
// the count after the replay (the first cap objects' ids, states and alive
// flags written back), or -1 where read_snapshot returns false.
int64_t ref_read_snapshot(const char* path, int n, int* ids, const int* events, int64_t cap, double* states,
                          uint8_t* alive, double* t) {
    std::vector<Object> objs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        objs[i].id = ids[i];
        objs[i].event = events[i];
    }
    if (!read_snapshot(path, objs, t)) return -1;
    const int64_t total = static_cast<int64_t>(objs.size());
    for (int64_t i = 0; i < total && i < cap; ++i) {
        const Object& o = objs[static_cast<size_t>(i)];
        ids[i] = o.id;
        double* d = states + 6 * static_cast<size_t>(i);
        d[0] = o.s.r.x; d[1] = o.s.r.y; d[2] = o.s.r.z; d[3] = o.s.v.x; d[4] = o.s.v.y; d[5] = o.s.v.z;
        alive[i] = o.alive ? 1 : 0;
    }
    return total;
}

}  // extern "C"
