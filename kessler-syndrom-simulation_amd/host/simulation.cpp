// Simulation::step on the device (orbital.cpp:286-300 replaced): the host
// vector `objects` is mirrored in HBM; steps run there, and the host copy is
// refreshed only when asked (sync_objects).

#include <algorithm>
#include <cstdio>
#include <vector>

#include "device_engine.hpp"
#include "orbital.hpp"

namespace kessler {

std::shared_ptr<DeviceEngine> DeviceEngine::create(int device, std::string* message) {
    char buf[512] = {0};
    khip_ctx* ctx = nullptr;
    const int rc = khip_create(device, &ctx, buf, sizeof(buf));
    if (message) *message = buf;
    if (rc != KHIP_OK) return nullptr;
    std::shared_ptr<DeviceEngine> e(new DeviceEngine());
    e->ctx_ = ctx;
    e->name_ = buf;
    return e;
}

DeviceEngine::~DeviceEngine() { khip_destroy(ctx_); }

void DeviceEngine::check(int rc, const char* what) const {
    if (rc != KHIP_OK) throw std::runtime_error(std::string(what) + ": " + khip_last_error(ctx_));
}

void Simulation::ensure_device() {
    if (!engine) {
        std::string msg;
        engine = DeviceEngine::create(0, &msg);
        if (!engine) throw std::runtime_error("no device engine: " + msg);
        device_dirty = true;
    }
    if (!device_dirty) return;
    const size_t n = objects.size();
    std::vector<double> states(6 * n), size(n);
    std::vector<uint8_t> alive(n);
    for (size_t i = 0; i < n; ++i) {
        const State& s = objects[i].s;
        double* d = states.data() + 6 * i;
        d[0] = s.r.x; d[1] = s.r.y; d[2] = s.r.z; d[3] = s.v.x; d[4] = s.v.y; d[5] = s.v.z;
        alive[i] = objects[i].alive ? 1 : 0;
        size[i] = objects[i].size_m;
    }
    khip_ctx* c = engine->ctx();
    const khip_screen_config cfg{engine->cfg.threshold_km, engine->cfg.radius_scale,
                                 engine->cfg.min_encounter_speed_km_s};
    engine->check(khip_set_config(c, &cfg), "khip_set_config");
    engine->check(khip_set_objects(c, static_cast<int64_t>(n), states.data(), alive.data(), size.data()),
                  "khip_set_objects");
    engine->check(khip_set_time(c, epoch_jd, t, dt), "khip_set_time");
    device_dirty = false;
    host_stale = false;
}

void Simulation::step() {
    ensure_device();
    khip_ctx* c = engine->ctx();
    double t_dev = 0.0;
    int64_t s = 0;
    engine->check(khip_get_time(c, &t_dev, &s), "khip_get_time");
    if (t_dev != t) engine->check(khip_set_time(c, epoch_jd, t, dt), "khip_set_time");
    const bool screen = engine->screen_next;
    engine->screen_next = false;
    engine->check(khip_run(c, 1, screen ? 1 : 0), "khip_run");
    engine->check(khip_get_time(c, &t, &s), "khip_get_time");
    host_stale = true;
    if (screen) {
        int64_t ne = 0, ns = 0;
        engine->check(khip_take_events(c, nullptr, 0, &ne), "khip_take_events");
        engine->events.resize(static_cast<size_t>(ne));
        engine->check(khip_take_events(c, engine->events.data(), ne, &ne), "khip_take_events");
        engine->check(khip_take_stats(c, &engine->stats, 1, &ns), "khip_take_stats");
        engine->have_results = true;
    }
}

void Simulation::advance(long n) {
    ensure_device();
    if (n <= 0) return;
    khip_ctx* c = engine->ctx();
    engine->check(khip_run(c, n, 0), "khip_run");
    int64_t s = 0;
    engine->check(khip_get_time(c, &t, &s), "khip_get_time");
    host_stale = true;
}

void Simulation::sync_objects() {
    if (!engine || !host_stale) return;
    const size_t n = objects.size();
    std::vector<double> states(6 * n);
    std::vector<uint8_t> alive(n);
    engine->check(khip_get_objects(engine->ctx(), states.data(), alive.data()), "khip_get_objects");
    for (size_t i = 0; i < n; ++i) {
        const double* d = states.data() + 6 * i;
        objects[i].s = State{{d[0], d[1], d[2]}, {d[3], d[4], d[5]}};
        objects[i].alive = alive[i] != 0;
    }
    host_stale = false;
}

}  // namespace kessler
