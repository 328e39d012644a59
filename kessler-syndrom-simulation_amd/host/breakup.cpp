// Host-side NASA Standard Breakup Model: the seeded forced impacts that
// build cascade catalogues, and the collisions the CLI's cascade mode
// applies after each step (breakup.hpp surface, reference breakup.hpp:42-153).
//
// Model: Johnson, Krisko, Liou & Anz-Meador, "NASA's new breakup model of
// EVOLVE 4.0", Adv. Space Res. 28(9), 2001 -- catastrophic threshold on the
// specific impact energy, characteristic mass M, fragment count
// N(> L_c) = 0.1 M^0.75 L_c^-1.71, the spacecraft area-to-mass distribution
// (two normals in log10 A/m above 11 cm, one below 8 cm, a size-linear
// choice between them), A = 0.540424 L^2 and the log-normal delta-v law
// mu = 0.9 chi + 2.9, sigma = 0.4. Poisson counts above a mean of 10 use
// Hormann's transformed rejection with squeeze (PTRS; "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Math. and Econ. 12, 1993).
//
// Organisation: an event is first planned (plan_event: energy, regime, mass
// budgets, expected count), its parents settled, then fragments are drawn
// one at a time by FragmentDraws, which owns the order in which the draws
// consume the generator. A seed reproduces the reference's fragments
// exactly (breakup.cpp:121-360), which fixes three things kept here and
// nothing else: the sequence of generator draws, the operation order of
// each floating-point expression (the file is built with -ffp-contract=off),
// and the libm function behind each step. The CLI tests compare
// breakups.csv, population.csv and final_state.csv of a forced-impact
// cascade byte for byte with the reference binary's, and
// tests/test_breakup_host.py compares khost_impacts with the compiled
// reference's Fragmentation::impact.

#include "breakup.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

#include "collision.hpp"

namespace kessler {

namespace {

constexpr int kTypes = static_cast<int>(ObjectType::Count);

// Mass-table defaults (config/masses.csv of the reference), kg: rows are
// the RCS classes SMALL, MEDIUM, LARGE, NONE; columns ObjectType order.
constexpr double kDefaultMassKg[4][kTypes] = {
    {5.0, 0.5, 50.0, 1.0},
    {50.0, 5.0, 300.0, 10.0},
    {500.0, 50.0, 2000.0, 100.0},
    {100.0, 5.0, 500.0, 10.0},
};
// Radar cross-section of a catalogue object of each RCS class, m^2 (its
// area-to-mass ratio is this over its mass).
constexpr double kClassAreaM2[4] = {0.05, 0.5, 5.0, 0.5};
constexpr double kUnknownMassKg = 100.0;      // parent without a mass
constexpr double kProjectileDensity = 300.0;  // kg/m^3, sizes a fictitious projectile
constexpr long long kMaxFragments = 200000;   // per event, whatever L_c
constexpr double kSizeExponent = 1.71;        // N(> L) ~ L^-1.71

int rcs_class_of(const std::string& rcs) {
    static const char* const names[3] = {"SMALL", "MEDIUM", "LARGE"};
    for (int k = 0; k < 3; ++k)
        if (rcs == names[k]) return k;
    return 3;  // NONE and anything unrecognised
}

double parent_mass(const Object& o) { return o.mass_kg > 0 ? o.mass_kg : kUnknownMassKg; }

// An SBM coefficient that moves linearly in lambda = log10(L) between two
// knots and holds its end values outside them.
struct Knots {
    double x0, y0, x1, y1;
    double at(double x) const {
        if (x <= x0) return y0;
        if (x >= x1) return y1;
        return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
    }
};

// Spacecraft area-to-mass law (Johnson et al. 2001, section 3.3).
// Above 11 cm: weight alpha of N(mu1, sigma1), else N(mu2, sigma2).
constexpr Knots kAlpha{-1.95, 0.0, 0.55, 1.0};
constexpr Knots kMu1{-1.1, -0.6, 0.0, -0.95};
constexpr Knots kSigma1{-1.3, 0.1, -0.3, 0.3};
constexpr Knots kMu2{-0.7, -1.2, -0.1, -2.0};
constexpr Knots kSigma2{-0.5, 0.5, -0.3, 0.3};
// Below 8 cm: one normal.
constexpr Knots kMuSmall{-1.75, -0.3, -1.25, -1.0};
double sigma_small(double lam) { return lam <= -3.5 ? 0.2 : 0.2 + 0.1333 * (lam + 3.5); }

// Knuth's product of uniforms, for small means.
long long poisson_product(PortableRng& g, double mean) {
    const double stop = std::exp(-mean);
    long long count = 0;
    for (double prod = g.uniform(); prod > stop; prod *= g.uniform()) ++count;
    return count;
}

// Hormann's PTRS for means of 10 and above: a transformed-rejection
// proposal with a fast acceptance box and an exact log-density test.
long long poisson_ptrs(PortableRng& g, double mean) {
    const double sq = std::sqrt(mean);
    const double log_mean = std::log(mean);
    const double b = 0.931 + 2.53 * sq;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double v_box = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double u = g.uniform() - 0.5;
        const double v = g.uniform();
        const double us = 0.5 - std::fabs(u);
        if (us <= 0.0 || v <= 0.0) continue;
        const double k = std::floor((2.0 * a / us + b) * u + mean + 0.43);
        const bool in_box = us >= 0.07 && v <= v_box;
        if (in_box) return static_cast<long long>(k);
        const bool hopeless = k < 0.0 || (us < 0.013 && v > us);
        if (hopeless) continue;
        const double lhs = std::log(v) + std::log(inv_alpha) - std::log(a / (us * us) + b);
        const double rhs = -mean + k * log_mean - std::lgamma(k + 1.0);
        if (lhs <= rhs) return static_cast<long long>(k);
    }
}

// The per-fragment draws of one event, in the order they consume the
// generator: parent side, size (only when the parent can shed a tracked
// fragment), area-to-mass, then -- for fragments within their parent's
// mass budget -- delta-v magnitude and direction.
class FragmentDraws {
public:
    FragmentDraws(PortableRng& g, double lc) : g_(g), lc_(lc) {}

    int side(double p_target) { return g_.uniform() < p_target ? 0 : 1; }

    // Inverse CDF of the power law truncated to [L_c, lmax].
    double size(double lmax) {
        const double tail = std::pow(lmax / lc_, -kSizeExponent);
        return lc_ * std::pow(1.0 - g_.uniform() * (1.0 - tail), -1.0 / kSizeExponent);
    }

    // chi = log10(A/m) for a fragment of size 10^lam.
    double log_area_to_mass(double lam) {
        const double l = std::pow(10.0, lam);
        if (l >= 0.11) return bimodal(lam);
        if (l <= 0.08) return single(lam);
        const bool take_large = g_.uniform() < (l - 0.08) / 0.03;
        return take_large ? bimodal(lam) : single(lam);
    }

    // Log-normal magnitude (m/s), capped at the impact speed.
    double dv_ms(double chi, double vrel_km_s) {
        const double dv = std::pow(10.0, 0.9 * chi + 2.9 + 0.4 * g_.normal());
        return std::min(dv, vrel_km_s * 1000.0);
    }

    // Isotropic unit vector: uniform cos(polar angle), uniform azimuth.
    Vec3 direction() {
        const double cz = 2.0 * g_.uniform() - 1.0;
        const double az = 2.0 * PI * g_.uniform();
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        return Vec3{sz * std::cos(az), sz * std::sin(az), cz};
    }

private:
    double bimodal(double lam) {
        const double alpha = kAlpha.at(lam);
        const double mu1 = kMu1.at(lam), s1 = kSigma1.at(lam);
        const double mu2 = kMu2.at(lam), s2 = kSigma2.at(lam);
        if (g_.uniform() < alpha) return mu1 + s1 * g_.normal();
        return mu2 + s2 * g_.normal();
    }
    double single(double lam) {
        const double mu = kMuSmall.at(lam);
        const double s = sigma_small(lam);
        return mu + s * g_.normal();
    }

    PortableRng& g_;
    double lc_;
};

// Energy, regime and mass budgets of one collision (the heavier body is
// the target). Non-catastrophic: the projectile plus the ejecta it knocks
// off the target, M = m_p v^2 with v in km/s.
struct EventPlan {
    double vrel_km_s, energy_j_per_g, expected;
    bool catastrophic;
    double budget[2];  // mass each parent may give: 0 target, 1 projectile
    double total() const { return budget[0] + budget[1]; }
};

EventPlan plan_event(const State& tgt_s, double tgt_kg, const State& prj_s, double prj_kg,
                     const BreakupConfig& cfg) {
    EventPlan p;
    p.vrel_km_s = norm(tgt_s.v - prj_s.v);
    // (1/2) m_p (1000 v)^2 J over 1000 m_t grams.
    p.energy_j_per_g = 500.0 * prj_kg * p.vrel_km_s * p.vrel_km_s / tgt_kg;
    p.catastrophic = p.energy_j_per_g >= cfg.catastrophic_j_per_g;
    if (p.catastrophic) {
        p.budget[0] = tgt_kg;
        p.budget[1] = prj_kg;
    } else {
        p.budget[1] = prj_kg;
        p.budget[0] = std::max(0.0, prj_kg * p.vrel_km_s * p.vrel_km_s - prj_kg);
    }
    const double m = p.total();
    p.expected = m > 0 ? 0.1 * std::pow(m, 0.75) * std::pow(cfg.lc_m, -kSizeExponent) : 0.0;
    return p;
}

const char* rcs_of_area(double area_m2) { return area_m2 < 0.1 ? "SMALL" : area_m2 < 1.0 ? "MEDIUM" : "LARGE"; }

}  // namespace

// ---- Masses ---------------------------------------------------------------

MassTable::MassTable() {
    for (int r = 0; r < 4; ++r)
        for (int t = 0; t < kTypes; ++t) kg_[r][t] = kDefaultMassKg[r][t];
}

int MassTable::rcs_index(const std::string& rcs) { return rcs_class_of(rcs); }

double MassTable::mass_for(const std::string& rcs, ObjectType type) const {
    return kg_[rcs_index(rcs)][static_cast<int>(type)];
}

// CSV rows "rcs,type,mass_kg"; '#' comments and a "rcs," header skipped.
bool MassTable::load(const std::string& path, std::string* error) {
    std::ifstream in(path);
    if (!in) {
        if (error) *error = "cannot open " + path;
        return false;
    }
    auto bad = [&](int lineno, const std::string& line) {
        if (error) *error = path + ", line " + std::to_string(lineno) + " invalid: " + line;
        return false;
    };
    std::string line;
    for (int lineno = 1; std::getline(in, line); ++lineno) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#' || line.compare(0, 4, "rcs,") == 0) continue;
        std::string field[3];
        std::stringstream row(line);
        for (std::string& f : field) std::getline(row, f, ',');
        int type = -1;
        for (int k = 0; k < kTypes && type < 0; ++k)
            if (field[1] == object_type_name(static_cast<ObjectType>(k))) type = k;
        const bool known_rcs = field[0] == "NONE" || rcs_class_of(field[0]) < 3;
        const double kg = std::atof(field[2].c_str());
        if (type < 0 || !known_rcs || kg <= 0) return bad(lineno, line);
        kg_[rcs_index(field[0])][type] = kg;
    }
    return true;
}

void assign_masses(std::vector<Object>& objs, const MassTable& table) {
    for (Object& o : objs) {
        if (o.is_fragment()) continue;  // fragments carry their SBM mass
        o.mass_kg = table.mass_for(o.rcs, o.type);
        o.area_to_mass = kClassAreaM2[rcs_class_of(o.rcs)] / o.mass_kg;
    }
}

// ---- Portable draws -------------------------------------------------------

double PortableRng::uniform() { return static_cast<double>(engine_() >> 11) * 0x1.0p-53; }

double PortableRng::normal() {
    // Box-Muller on (0, 1] x [0, 1).
    const double u_log = 1.0 - uniform();
    const double u_ang = uniform();
    return std::sqrt(-2.0 * std::log(u_log)) * std::cos(2.0 * PI * u_ang);
}

long long PortableRng::poisson(double mean) {
    if (mean <= 0.0) return 0;
    return mean < 10.0 ? poisson_product(*this, mean) : poisson_ptrs(*this, mean);
}

// ---- Fragmentation --------------------------------------------------------

Fragmentation::Fragmentation(const BreakupConfig& cfg) : cfg_(cfg), rng_(cfg.seed) {}

double Fragmentation::sample_area_to_mass_log(double log_l) {
    return FragmentDraws(rng_, cfg_.lc_m).log_area_to_mass(log_l);
}

BreakupEvent Fragmentation::fragment(Simulation& sim, const Parent& tgt, const Parent& prj, double t,
                                     double remaining_s) {
    const EventPlan plan = plan_event(tgt.s, tgt.mass_kg, prj.s, prj.mass_kg, cfg_);
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
    ev.v_rel_km_s = plan.vrel_km_s;
    ev.energy_j_per_g = plan.energy_j_per_g;
    ev.catastrophic = plan.catastrophic;
    ev.sbm_mass_kg = plan.total();
    ev.expected_fragments = plan.expected;

    // Fragments of the event share its birth time and the rest of the step,
    // so one set of Moon samples serves every fragment's RK4 step.
    const double jd0 = sim.epoch_jd + t / 86400.0;
    const Vec3 moon[3] = {moon_position(jd0), moon_position(jd0 + remaining_s / 2 / 86400.0),
                          moon_position(jd0 + remaining_s / 86400.0)};

    // Parents first (appending fragments may move the object array).
    if (prj.index >= 0) sim.objects[prj.index].alive = false;
    if (tgt.index >= 0) {
        Object& target = sim.objects[tgt.index];
        if (plan.catastrophic) target.alive = false;
        else target.mass_kg = std::max(tgt.mass_kg - plan.budget[0], 0.0);
    }

    const double lc = cfg_.lc_m;
    const long long count = plan.expected > 0 ? std::min(rng_.poisson(plan.expected), kMaxFragments) : 0;
    const double p_target = plan.total() > 0 ? plan.budget[0] / plan.total() : 0.0;
    const Parent* side_parent[2] = {&tgt, &prj};
    double given[2] = {0.0, 0.0};
    FragmentDraws draw(rng_, lc);
    sim.objects.reserve(sim.objects.size() + static_cast<size_t>(count));

    for (long long k = 0; k < count; ++k) {
        const int side = draw.side(p_target);
        const Parent& par = *side_parent[side];
        if (par.size_m <= lc) continue;  // nothing tracked comes off a parent below L_c
        const double size = draw.size(par.size_m);
        const double chi = draw.log_area_to_mass(std::log10(size));
        const double area = 0.540424 * size * size;
        const double mass = area / std::pow(10.0, chi);
        // Mass rule, in draw order: a fragment its parent can no longer
        // supply is skipped; smaller ones after it may still fit.
        if (given[side] + mass > plan.budget[side]) continue;
        given[side] += mass;
        const double dv_ms = draw.dv_ms(chi, plan.vrel_km_s);
        const Vec3 dir = draw.direction();
        const State birth{par.s.r, par.s.v + dir * (dv_ms * 1e-3)};
        // Unbound fragments leave Earth orbit for good: dropped.
        const bool unbound = 0.5 * dot(birth.v, birth.v) - MU_EARTH / norm(birth.r) >= 0.0;
        if (unbound) {
            ++ev.escaped;
            continue;
        }
        Object f;
        f.id = next_id_++;
        f.name = "FRAG " + std::to_string(ev.id) + "-" + std::to_string(ev.fragments);
        f.type = ObjectType::Debris;
        f.rcs = rcs_of_area(area);
        f.size_m = size;
        f.mass_kg = mass;
        f.area_to_mass = std::pow(10.0, chi);
        f.event = ev.id;
        f.s = remaining_s > 0 ? rk4(birth, remaining_s, moon[0], moon[1], moon[2]) : birth;
        f.alive = norm(f.s.r) >= R_EARTH;
        sim.objects.push_back(std::move(f));
        ++ev.fragments;
        ev.fragments_mass_kg += mass;
    }
    fragments_created_ += ev.fragments;
    return ev;
}

std::vector<BreakupEvent> Fragmentation::apply(Simulation& sim, const std::vector<State>& before,
                                               double t0, double dt,
                                               const std::vector<ConjunctionRecord>& conj) {
    std::vector<BreakupEvent> out;
    for (const ConjunctionRecord& c : conj) {  // in (t, i, j) order
        if (!c.collision) continue;
        // A body destroyed earlier in the step collides no more.
        if (!sim.objects[c.i].alive || !sim.objects[c.j].alive) continue;
        const double frac = (c.t - t0) / dt;
        auto at_impact = [&](int idx) {
            const Object& o = sim.objects[idx];
            return Parent{idx, interpolate_state(before[idx], o.s, dt, frac), parent_mass(o), o.size_m, o.name};
        };
        const Parent pi = at_impact(c.i), pj = at_impact(c.j);
        const bool i_heavier = pi.mass_kg >= pj.mass_kg;
        out.push_back(fragment(sim, i_heavier ? pi : pj, i_heavier ? pj : pi, c.t, t0 + dt - c.t));
    }
    return out;
}

BreakupEvent Fragmentation::impact(Simulation& sim, int target, double projectile_mass_kg,
                                   double v_rel_km_s) {
    const Object& o = sim.objects[target];
    const Vec3 r = o.s.r, v = o.s.v;
    // The projectile moves at the target's speed, its velocity turned about
    // the local vertical k by theta, so that |v_p - v| = 2 |v_h| sin(theta/2)
    // (v_h: the horizontal part of v); Rodrigues' rotation formula.
    const Vec3 up = r * (1.0 / norm(r));
    const double v_horizontal = norm(v - up * dot(v, up));
    double theta = v_rel_km_s >= 2.0 * v_horizontal ? PI : 2.0 * std::asin(v_rel_km_s / (2.0 * v_horizontal));
    if (rng_.uniform() < 0.5) theta = -theta;  // crossing from either side
    const double cos_t = std::cos(theta);
    const Vec3 vp = v * cos_t + cross(up, v) * std::sin(theta) + up * (dot(up, v) * (1.0 - cos_t));
    const Parent body{target, o.s, parent_mass(o), o.size_m, o.name};
    const Parent bullet{-1, State{r, vp}, projectile_mass_kg, std::cbrt(projectile_mass_kg / kProjectileDensity),
                        "PROJECTILE"};
    // The SBM's target is the heavier body.
    return bullet.mass_kg > body.mass_kg ? fragment(sim, bullet, body, sim.t, 0.0)
                                         : fragment(sim, body, bullet, sim.t, 0.0);
}

}  // namespace kessler
