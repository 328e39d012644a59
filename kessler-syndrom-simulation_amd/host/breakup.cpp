// NASA Standard Breakup Model (Johnson et al., Adv. Space Res. 28(9), 2001).
//
// Restates breakup.cpp:47-360 of the reference: the same draws in the same
// order and the same libm calls, so that a seed reproduces its fragments
// exactly (pinned by tests/test_breakup.py against the compiled reference).
// Built with -ffp-contract=off.

#include "breakup.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

namespace kessler {

namespace {

constexpr int NTYPES = static_cast<int>(ObjectType::Count);
// Typical radar cross-section per RCS class (SMALL, MEDIUM, LARGE, NONE), m^2.
constexpr double CLASS_AREA_M2[4] = {0.05, 0.5, 5.0, 0.5};
constexpr double FALLBACK_MASS_KG = 100.0;
// Bulk density of a fictitious projectile, only to bound its fragments' size.
constexpr double PROJECTILE_DENSITY_KG_M3 = 300.0;
constexpr long long MAX_FRAGMENTS_PER_EVENT = 200000;

double mass_or_fallback(const Object& o) { return o.mass_kg > 0 ? o.mass_kg : FALLBACK_MASS_KG; }

// Piecewise linear, constant outside [x0, x1].
double ramp(double x, double x0, double y0, double x1, double y1) {
    if (x <= x0) return y0;
    if (x >= x1) return y1;
    return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
}

int rcs_class(const std::string& rcs) {
    return rcs == "SMALL" ? 0 : rcs == "MEDIUM" ? 1 : rcs == "LARGE" ? 2 : 3;
}

}  // namespace

MassTable::MassTable() {
    // Rows SMALL, MEDIUM, LARGE, NONE; columns in ObjectType order.
    static const double defaults[4][NTYPES] = {
        {5.0, 0.5, 50.0, 1.0},
        {50.0, 5.0, 300.0, 10.0},
        {500.0, 50.0, 2000.0, 100.0},
        {100.0, 5.0, 500.0, 10.0},
    };
    for (int r = 0; r < 4; ++r)
        for (int t = 0; t < NTYPES; ++t) kg_[r][t] = defaults[r][t];
}

int MassTable::rcs_index(const std::string& rcs) { return rcs_class(rcs); }

double MassTable::mass_for(const std::string& rcs, ObjectType type) const {
    return kg_[rcs_index(rcs)][static_cast<int>(type)];
}

bool MassTable::load(const std::string& path, std::string* error) {
    std::ifstream in(path);
    if (!in) {
        if (error) *error = "cannot open " + path;
        return false;
    }
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#' || line.rfind("rcs,", 0) == 0) continue;
        std::stringstream ss(line);
        std::string rcs, type, mass;
        std::getline(ss, rcs, ',');
        std::getline(ss, type, ',');
        std::getline(ss, mass, ',');
        int t = -1;
        for (int k = 0; k < NTYPES; ++k)
            if (type == object_type_name(static_cast<ObjectType>(k))) t = k;
        const bool rcs_ok = rcs == "SMALL" || rcs == "MEDIUM" || rcs == "LARGE" || rcs == "NONE";
        const double kg = std::atof(mass.c_str());
        if (t < 0 || !rcs_ok || kg <= 0) {
            if (error) *error = path + ", line " + std::to_string(lineno) + " invalid: " + line;
            return false;
        }
        kg_[rcs_index(rcs)][t] = kg;
    }
    return true;
}

void assign_masses(std::vector<Object>& objs, const MassTable& table) {
    for (Object& o : objs) {
        if (o.is_fragment()) continue;
        o.mass_kg = table.mass_for(o.rcs, o.type);
        o.area_to_mass = CLASS_AREA_M2[rcs_class(o.rcs)] / o.mass_kg;
    }
}

double PortableRng::uniform() { return static_cast<double>(engine_() >> 11) * 0x1.0p-53; }

double PortableRng::normal() {
    const double u1 = 1.0 - uniform();  // (0, 1]: finite log
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
}

long long PortableRng::poisson(double mean) {
    if (mean <= 0.0) return 0;
    if (mean < 10.0) {
        const double limit = std::exp(-mean);
        long long k = 0;
        double prod = uniform();
        while (prod > limit) {
            ++k;
            prod *= uniform();
        }
        return k;
    }
    // Transformed rejection with squeeze (Hormann 1993, "PTRS").
    const double slam = std::sqrt(mean), loglam = std::log(mean);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double u = uniform() - 0.5;
        const double v = uniform();
        const double us = 0.5 - std::fabs(u);
        if (us <= 0.0 || v <= 0.0) continue;
        const double k = std::floor((2.0 * a / us + b) * u + mean + 0.43);
        if (us >= 0.07 && v <= vr) return static_cast<long long>(k);
        if (k < 0.0 || (us < 0.013 && v > us)) continue;
        if (std::log(v) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
            -mean + k * loglam - std::lgamma(k + 1.0))
            return static_cast<long long>(k);
    }
}

Fragmentation::Fragmentation(const BreakupConfig& cfg) : cfg_(cfg), rng_(cfg.seed) {}

// log10(A/m) of a fragment of size 10^log_l (SBM area-to-mass distribution).
double Fragmentation::sample_area_to_mass_log(double log_l) {
    const double l = std::pow(10.0, log_l);
    auto bimodal = [&]() {  // above 11 cm
        const double alpha = ramp(log_l, -1.95, 0.0, 0.55, 1.0);
        const double mu1 = ramp(log_l, -1.1, -0.6, 0.0, -0.95);
        const double s1 = ramp(log_l, -1.3, 0.1, -0.3, 0.3);
        const double mu2 = ramp(log_l, -0.7, -1.2, -0.1, -2.0);
        const double s2 = ramp(log_l, -0.5, 0.5, -0.3, 0.3);
        if (rng_.uniform() < alpha) return mu1 + s1 * rng_.normal();
        return mu2 + s2 * rng_.normal();
    };
    auto single = [&]() {  // below 8 cm
        const double mu = ramp(log_l, -1.75, -0.3, -1.25, -1.0);
        const double s = log_l <= -3.5 ? 0.2 : 0.2 + 0.1333 * (log_l + 3.5);
        return mu + s * rng_.normal();
    };
    if (l >= 0.11) return bimodal();
    if (l <= 0.08) return single();
    // Bridge: one model or the other, with a probability linear in size.
    return rng_.uniform() < (l - 0.08) / 0.03 ? bimodal() : single();
}

BreakupEvent Fragmentation::fragment(Simulation& sim, const Parent& tgt, const Parent& prj, double t,
                                     double remaining_s) {
    BreakupEvent ev;
    ev.id = next_event_++;
    ev.t = t;
    ev.target = tgt.index;
    ev.projectile = prj.index;
    ev.target_name = tgt.name;
    ev.projectile_name = prj.name;
    ev.target_mass_kg = tgt.mass_kg;
    ev.projectile_mass_kg = prj.mass_kg;
    ev.position = tgt.s.r;

    const double vrel = norm(tgt.s.v - prj.s.v);
    // One set of Moon samples for every fragment of the event.
    const double jd0 = sim.epoch_jd + t / 86400.0;
    const Vec3 moon0 = moon_position(jd0);
    const Vec3 moon1 = moon_position(jd0 + remaining_s / 2 / 86400.0);
    const Vec3 moon2 = moon_position(jd0 + remaining_s / 86400.0);
    ev.v_rel_km_s = vrel;
    ev.energy_j_per_g = 500.0 * prj.mass_kg * vrel * vrel / tgt.mass_kg;
    ev.catastrophic = ev.energy_j_per_g >= cfg_.catastrophic_j_per_g;

    double budget[2];  // mass each parent gives to fragments: 0 target, 1 projectile
    if (ev.catastrophic) {
        budget[0] = tgt.mass_kg;
        budget[1] = prj.mass_kg;
    } else {
        budget[1] = prj.mass_kg;
        budget[0] = std::max(0.0, prj.mass_kg * vrel * vrel - prj.mass_kg);
    }
    const double M = budget[0] + budget[1];
    ev.sbm_mass_kg = M;

    if (prj.index >= 0) sim.objects[prj.index].alive = false;
    if (tgt.index >= 0) {
        if (ev.catastrophic) sim.objects[tgt.index].alive = false;
        else sim.objects[tgt.index].mass_kg = std::max(tgt.mass_kg - budget[0], 0.0);
    }

    const double lc = cfg_.lc_m;
    const double slope = 1.71;
    ev.expected_fragments = M > 0 ? 0.1 * std::pow(M, 0.75) * std::pow(lc, -slope) : 0.0;
    long long count = 0;
    if (ev.expected_fragments > 0)
        count = std::min(rng_.poisson(ev.expected_fragments), MAX_FRAGMENTS_PER_EVENT);

    const Parent* parents[2] = {&tgt, &prj};
    double used[2] = {0.0, 0.0};
    const double p_target = M > 0 ? budget[0] / M : 0.0;
    sim.objects.reserve(sim.objects.size() + static_cast<size_t>(count));

    for (long long k = 0; k < count; ++k) {
        const int side = rng_.uniform() < p_target ? 0 : 1;
        const Parent& par = *parents[side];
        // Size by inverting the truncated power law on [L_c, parent size].
        const double lmax = par.size_m;
        if (lmax <= lc) continue;
        const double r = std::pow(lmax / lc, -slope);
        const double L = lc * std::pow(1.0 - rng_.uniform() * (1.0 - r), -1.0 / slope);
        const double chi = sample_area_to_mass_log(std::log10(L));
        const double area = 0.540424 * L * L;
        const double m = area / std::pow(10.0, chi);
        if (used[side] + m > budget[side]) continue;  // parent's mass exhausted
        used[side] += m;
        // Delta-v: log-normal magnitude capped at the impact speed, isotropic.
        double dv_ms = std::pow(10.0, 0.9 * chi + 2.9 + 0.4 * rng_.normal());
        dv_ms = std::min(dv_ms, vrel * 1000.0);
        const double cz = 2.0 * rng_.uniform() - 1.0;
        const double phi = 2.0 * PI * rng_.uniform();
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        const Vec3 dir{sz * std::cos(phi), sz * std::sin(phi), cz};
        const State birth{par.s.r, par.s.v + dir * (dv_ms * 1e-3)};
        // Hyperbolic fragments leave Earth orbit: dropped.
        if (0.5 * dot(birth.v, birth.v) - MU_EARTH / norm(birth.r) >= 0.0) {
            ++ev.escaped;
            continue;
        }
        Object f;
        f.id = next_id_++;
        f.name = "FRAG " + std::to_string(ev.id) + "-" + std::to_string(ev.fragments);
        f.type = ObjectType::Debris;
        f.rcs = area < 0.1 ? "SMALL" : area < 1.0 ? "MEDIUM" : "LARGE";
        f.size_m = L;
        f.mass_kg = m;
        f.area_to_mass = std::pow(10.0, chi);
        f.event = ev.id;
        f.s = remaining_s > 0 ? rk4(birth, remaining_s, moon0, moon1, moon2) : birth;
        f.alive = norm(f.s.r) >= R_EARTH;
        sim.objects.push_back(std::move(f));
        ++ev.fragments;
        ev.fragments_mass_kg += m;
    }
    fragments_created_ += ev.fragments;
    return ev;
}

std::vector<BreakupEvent> Fragmentation::apply(Simulation& sim, const std::vector<State>& before,
                                               double t0, double dt,
                                               const std::vector<ConjunctionRecord>& conj) {
    std::vector<BreakupEvent> out;
    for (const ConjunctionRecord& c : conj) {  // sorted by time
        if (!c.collision) continue;
        if (!sim.objects[c.i].alive || !sim.objects[c.j].alive) continue;
        const double s = (c.t - t0) / dt;
        auto parent = [&](int idx) {
            const Object& o = sim.objects[idx];
            // Hermite state at the collision instant (collision.cpp:289-292).
            const State& a = before[idx];
            const State& b = o.s;
            const double s2 = s * s, s3 = s2 * s;
            const Vec3 m0 = a.v * dt, m1 = b.v * dt;
            const Vec3 p = a.r * (2 * s3 - 3 * s2 + 1) + m0 * (s3 - 2 * s2 + s) + b.r * (-2 * s3 + 3 * s2) +
                           m1 * (s3 - s2);
            const Vec3 w = a.r * (6 * s2 - 6 * s) + m0 * (3 * s2 - 4 * s + 1) + b.r * (-6 * s2 + 6 * s) +
                           m1 * (3 * s2 - 2 * s);
            return Parent{idx, State{p, w * (1.0 / dt)}, mass_or_fallback(o), o.size_m, o.name};
        };
        const Parent a = parent(c.i), b = parent(c.j);
        const bool a_heavier = a.mass_kg >= b.mass_kg;
        out.push_back(fragment(sim, a_heavier ? a : b, a_heavier ? b : a, c.t, t0 + dt - c.t));
    }
    return out;
}

BreakupEvent Fragmentation::impact(Simulation& sim, int target, double projectile_mass_kg,
                                   double v_rel_km_s) {
    const Object& o = sim.objects[target];
    const Vec3 r = o.s.r, v = o.s.v;
    // Projectile with the target's horizontal speed rotated about the local
    // vertical by theta: |v_p - v| = 2 |v_h| sin(theta / 2) (Rodrigues).
    const Vec3 k = r * (1.0 / norm(r));
    const double vh = norm(v - k * dot(v, k));
    double theta = v_rel_km_s >= 2.0 * vh ? PI : 2.0 * std::asin(v_rel_km_s / (2.0 * vh));
    if (rng_.uniform() < 0.5) theta = -theta;
    const Vec3 vp = v * std::cos(theta) + cross(k, v) * std::sin(theta) +
                    k * (dot(k, v) * (1.0 - std::cos(theta)));
    const Parent tgt{target, o.s, mass_or_fallback(o), o.size_m, o.name};
    const Parent prj{-1, State{r, vp}, projectile_mass_kg,
                     std::cbrt(projectile_mass_kg / PROJECTILE_DENSITY_KG_M3), "PROJECTILE"};
    if (prj.mass_kg > tgt.mass_kg) return fragment(sim, prj, tgt, sim.t, 0.0);
    return fragment(sim, tgt, prj, sim.t, 0.0);
}

}  // namespace kessler
