// Screener over the gfx950 engine (collision.cpp:667-759 replaced): no CPU
// grid behind it. set_gpu keeps its role (collision.cpp:680-696): it is where
// a missing device is reported.

#include "collision.hpp"

#include <stdexcept>

#include "device_engine.hpp"

namespace kessler {

double collision_radius_km(const Object& o, double scale) { return 0.5 * o.size_m * 1e-3 * scale; }

State interpolate_state(const State& a, const State& b, double dt, double s) {
    const double s2 = s * s, s3 = s2 * s;
    const Vec3 m0 = a.v * dt, m1 = b.v * dt;
    const Vec3 p = a.r * (2 * s3 - 3 * s2 + 1) + m0 * (s3 - 2 * s2 + s) + b.r * (-2 * s3 + 3 * s2) +
                   m1 * (s3 - s2);
    const Vec3 w = a.r * (6 * s2 - 6 * s) + m0 * (3 * s2 - 4 * s + 1) + b.r * (-6 * s2 + 6 * s) +
                   m1 * (3 * s2 - 2 * s);
    return {p, w * (1.0 / dt)};
}

struct Screener::Impl {
    std::shared_ptr<DeviceEngine> engine;
    Simulation* attached = nullptr;
};

namespace {

Conjunction to_conjunction(const khip_event& e) {
    Conjunction c;
    c.i = e.i;
    c.j = e.j;
    c.t = e.t;
    c.miss_km = e.miss_km;
    c.v_rel_km_s = e.v_rel_km_s;
    c.collision = e.collision != 0;
    return c;
}

void fill_stats(const khip_stats& k, ScreeningStats* st) {
    if (!st) return;
    *st = ScreeningStats{};
    st->candidate_pairs = k.candidate_pairs;
    st->refined_pairs = k.refined_pairs;
    st->co_orbiting = k.co_orbiting;
    st->cell_km = k.cell_km;
    st->max_cell_objects = k.max_cell_objects;
    st->max_cell_pairs = static_cast<long long>(k.max_cell_objects) * (k.max_cell_objects - 1) / 2;
    st->gpu = true;
}

}  // namespace

Screener::Screener(const ScreeningConfig& cfg) : cfg_(cfg), impl_(new Impl) {}
Screener::~Screener() = default;
Screener::Screener(Screener&&) noexcept = default;
Screener& Screener::operator=(Screener&&) noexcept = default;

bool Screener::gpu_compiled() { return true; }

bool Screener::set_gpu(bool on, std::string* message) {
    if (!on) {
        impl_->engine.reset();
        return true;
    }
    if (impl_->engine) {
        if (message) *message = impl_->engine->name();
        return true;
    }
    impl_->engine = DeviceEngine::create(0, message);
    return impl_->engine != nullptr;
}

bool Screener::gpu() const { return impl_->engine != nullptr; }

void Screener::attach(Simulation& sim) {
    if (!impl_->engine) {
        std::string msg;
        if (!set_gpu(true, &msg)) throw std::runtime_error("no device engine: " + msg);
    }
    sim.engine = impl_->engine;
    sim.device_dirty = true;
    impl_->attached = &sim;
}

std::vector<Conjunction> Screener::screen(const std::vector<State>& before,
                                          const std::vector<Object>& after, double t0, double dt,
                                          ScreeningStats* stats) {
    if (!impl_->engine) throw std::runtime_error("Screener: no device engine (set_gpu(true) failed)");
    DeviceEngine& e = *impl_->engine;
    const size_t n = before.size();
    std::vector<double> b(6 * n), a(6 * n), size(n);
    std::vector<uint8_t> alive(n);
    for (size_t i = 0; i < n; ++i) {
        const State& s = before[i];
        const State& s2 = after[i].s;
        double* d = b.data() + 6 * i;
        d[0] = s.r.x; d[1] = s.r.y; d[2] = s.r.z; d[3] = s.v.x; d[4] = s.v.y; d[5] = s.v.z;
        double* d2 = a.data() + 6 * i;
        d2[0] = s2.r.x; d2[1] = s2.r.y; d2[2] = s2.r.z; d2[3] = s2.v.x; d2[4] = s2.v.y; d2[5] = s2.v.z;
        alive[i] = after[i].alive ? 1 : 0;
        size[i] = after[i].size_m;
    }
    const khip_screen_config cfg{cfg_.threshold_km, cfg_.radius_scale, cfg_.min_encounter_speed_km_s};
    e.check(khip_set_config(e.ctx(), &cfg), "khip_set_config");
    int64_t ne = 0;
    khip_stats ks{};
    e.check(khip_screen(e.ctx(), static_cast<int64_t>(n), b.data(), a.data(), alive.data(), size.data(),
                        t0, dt, nullptr, 0, &ne, &ks),
            "khip_screen");
    std::vector<khip_event> ev(static_cast<size_t>(ne));
    e.check(khip_screen(e.ctx(), static_cast<int64_t>(n), b.data(), a.data(), alive.data(), size.data(),
                        t0, dt, ev.data(), ne, &ne, &ks),
            "khip_screen");
    if (impl_->attached) impl_->attached->device_dirty = true;  // the population was replaced
    fill_stats(ks, stats);
    std::vector<Conjunction> out;
    out.reserve(ev.size());
    for (const khip_event& x : ev) out.push_back(to_conjunction(x));
    return out;
}

void Screener::begin(const std::vector<State>& before, const std::vector<Object>& objects, double dt) {
    (void)before;
    (void)objects;
    (void)dt;
    if (!impl_->attached || !impl_->engine) return;
    if (impl_->engine->cfg.threshold_km != cfg_.threshold_km ||
        impl_->engine->cfg.radius_scale != cfg_.radius_scale ||
        impl_->engine->cfg.min_encounter_speed_km_s != cfg_.min_encounter_speed_km_s) {
        impl_->engine->cfg = cfg_;
        impl_->attached->device_dirty = true;
    }
    impl_->engine->screen_next = true;
}

std::vector<Conjunction> Screener::finish(const std::vector<State>& before,
                                          const std::vector<Object>& after, double t0, double dt,
                                          ScreeningStats* stats) {
    if (impl_->attached && impl_->engine && impl_->engine->have_results) {
        DeviceEngine& e = *impl_->engine;
        e.have_results = false;
        fill_stats(e.stats, stats);
        std::vector<Conjunction> out;
        out.reserve(e.events.size());
        for (const khip_event& x : e.events) out.push_back(to_conjunction(x));
        return out;
    }
    return screen(before, after, t0, dt, stats);
}

std::vector<Conjunction> screen_step(const std::vector<State>& before, const std::vector<Object>& after,
                                     double t0, double dt, const ScreeningConfig& cfg,
                                     ScreeningStats* stats) {
    Screener s(cfg);
    std::string msg;
    if (!s.set_gpu(true, &msg)) throw std::runtime_error("screen_step: no device engine: " + msg);
    return s.screen(before, after, t0, dt, stats);
}

}  // namespace kessler
