// NASA Standard Breakup Model fragmentation, host side.
//
// Keeps the reference's breakup.hpp surface (breakup.hpp:42-153): MassTable,
// assign_masses, PortableRng, BreakupConfig, BreakupEvent, Fragmentation
// with the same names and meaning. Used to build seeded cascade catalogues
// (forced impacts at t = 0) and to apply collisions found by the engine.
// The draws, their order and every libm call follow breakup.cpp:121-360, so
// a given seed gives the reference's fragments bit for bit.

#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "orbital.hpp"

namespace kessler {

struct Conjunction;

class MassTable {
public:
    MassTable();  // defaults of config/masses.csv
    bool load(const std::string& path, std::string* error = nullptr);
    double mass_for(const std::string& rcs, ObjectType type) const;

private:
    static int rcs_index(const std::string& rcs);  // SMALL, MEDIUM, LARGE, NONE
    double kg_[4][static_cast<int>(ObjectType::Count)];
};

void assign_masses(std::vector<Object>& objs, const MassTable& table);

// Draws reproducible across standard libraries: the mt19937_64 engine is
// fixed bit for bit by the C++ standard; the distributions are written out.
class PortableRng {
public:
    explicit PortableRng(std::uint64_t seed) : engine_(seed) {}
    double uniform();                // [0, 1), 53 bits
    double normal();                 // N(0, 1), Box-Muller
    long long poisson(double mean);  // Knuth below 10, Hormann's PTRS above

private:
    std::mt19937_64 engine_;
};

struct BreakupConfig {
    double lc_m = 0.10;                  // smallest tracked fragment (m)
    double catastrophic_j_per_g = 40.0;  // specific-energy threshold
    std::uint64_t seed = 1;
};

struct BreakupEvent {
    int id = 0;
    double t = 0.0;
    int target = -1, projectile = -1;  // indices; projectile -1 when fictitious
    std::string target_name, projectile_name;
    bool catastrophic = false;
    double v_rel_km_s = 0.0;
    double energy_j_per_g = 0.0;
    double target_mass_kg = 0.0, projectile_mass_kg = 0.0;
    double sbm_mass_kg = 0.0;
    double expected_fragments = 0.0;
    int fragments = 0;
    double fragments_mass_kg = 0.0;
    int escaped = 0;
    Vec3 position;
};

struct ConjunctionRecord {
    int i = 0, j = 0;
    double t = 0.0, miss_km = 0.0, v_rel_km_s = 0.0;
    bool collision = false;
};

class Fragmentation {
public:
    explicit Fragmentation(const BreakupConfig& cfg = {});

    BreakupConfig& config() { return cfg_; }
    const BreakupConfig& config() const { return cfg_; }

    // Collisions found in the step [t0, t0 + dt]: parents destroyed or
    // damaged, fragments appended to sim.objects already propagated to the
    // end of the step (breakup.cpp:316-338).
    std::vector<BreakupEvent> apply(Simulation& sim, const std::vector<State>& before, double t0,
                                    double dt, const std::vector<ConjunctionRecord>& conjunctions);

    // Impact of a fictitious projectile on `target` at the current time
    // (breakup.cpp:340-360).
    BreakupEvent impact(Simulation& sim, int target, double projectile_mass_kg, double v_rel_km_s);

    int events() const { return next_event_; }
    long long fragments_created() const { return fragments_created_; }

private:
    struct Parent {
        int index;
        State s;
        double mass_kg;
        double size_m;
        std::string name;
    };
    BreakupEvent fragment(Simulation& sim, const Parent& target, const Parent& projectile, double t,
                          double remaining_s);
    double sample_area_to_mass_log(double log_l);

    BreakupConfig cfg_;
    PortableRng rng_;
    int next_event_ = 0;
    int next_id_ = FRAGMENT_ID_BASE;
    long long fragments_created_ = 0;
};

}  // namespace kessler
