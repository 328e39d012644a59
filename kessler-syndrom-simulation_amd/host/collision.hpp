// Close approaches and collisions: the reference's collision.hpp surface
// (collision.hpp:24-128) over the gfx950 engine.
//
// Screening runs on the GPU only. Two ways to drive it, as in the reference:
//  * one-shot, on host-held states: screen() / screen_step() (khip_screen);
//  * in the step loop: begin() before Simulation::step(), finish() after.
//    With a Simulation attached (attach()), the step and its screening are
//    one device call and `before` is not read: the start-of-step states
//    never leave HBM.

#pragma once

#include <memory>
#include <string>
#include <vector>

#include "orbital.hpp"

namespace kessler {

struct ScreeningConfig {
    double threshold_km = 5.0;
    double radius_scale = 1.0;
    double min_encounter_speed_km_s = 0.01;
};

struct Conjunction {
    int i = 0, j = 0;       // indices into Simulation::objects, i < j
    double t = 0.0;         // TCA, s since epoch
    double miss_km = 0.0;
    double v_rel_km_s = 0.0;
    bool collision = false;
};

struct ScreeningStats {
    long long candidate_pairs = 0;
    long long refined_pairs = 0;
    long long co_orbiting = 0;
    double cell_km = 0.0;
    double grid_s = 0.0;
    double pairs_s = 0.0;
    double pairs_busy_s = 0.0;
    double refine_s = 0.0;
    double transfer_s = 0.0;
    double wait_s = 0.0;
    int threads = 0;
    int max_cell_objects = 0;
    long long max_cell_pairs = 0;
    bool gpu = true;
};

double collision_radius_km(const Object& o, double scale = 1.0);
State interpolate_state(const State& a, const State& b, double dt, double s);

class Screener {
public:
    explicit Screener(const ScreeningConfig& cfg = {});
    ~Screener();
    Screener(Screener&&) noexcept;
    Screener& operator=(Screener&&) noexcept;
    Screener(const Screener&) = delete;
    Screener& operator=(const Screener&) = delete;

    ScreeningConfig& config() { return cfg_; }
    const ScreeningConfig& config() const { return cfg_; }

    // Always true: this build is the GPU backend.
    static bool gpu_compiled();
    // Creates (on) or releases (off) the device engine. Returns false with
    // the reason in `message` when no device can be used; `message` gets the
    // device name on success.
    bool set_gpu(bool on, std::string* message = nullptr);
    bool gpu() const;

    // Device-resident loop: the simulation's steps are screened on the
    // device (shares this screener's engine with `sim`).
    void attach(Simulation& sim);

    std::vector<Conjunction> screen(const std::vector<State>& before, const std::vector<Object>& after,
                                    double t0, double dt, ScreeningStats* stats = nullptr);
    void begin(const std::vector<State>& before, const std::vector<Object>& objects, double dt);
    std::vector<Conjunction> finish(const std::vector<State>& before, const std::vector<Object>& after,
                                    double t0, double dt, ScreeningStats* stats = nullptr);

private:
    struct Impl;
    ScreeningConfig cfg_;
    std::unique_ptr<Impl> impl_;
};

// One-shot screening on the GPU (a throwaway Screener). Throws
// std::runtime_error when no device is available.
std::vector<Conjunction> screen_step(const std::vector<State>& before, const std::vector<Object>& after,
                                     double t0, double dt, const ScreeningConfig& cfg,
                                     ScreeningStats* stats = nullptr);

}  // namespace kessler
