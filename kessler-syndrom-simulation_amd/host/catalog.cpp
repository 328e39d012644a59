// Synthetic catalogues, seeded cascades and catalogue I/O (C ABI of
// include/kessler_host.h). Host-only code: it builds the populations the
// engine runs on.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/kessler_host.h"
#include "breakup.hpp"
#include "orbital.hpp"

using namespace kessler;

namespace {

// Own draws for the generator (independent of the breakup model's stream).
class Draw {
public:
    explicit Draw(uint64_t seed) : eng_(seed) {}
    double uniform() { return static_cast<double>(eng_() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
    }
    double exponential(double mean) { return -mean * std::log(1.0 - uniform()); }
    uint64_t below(uint64_t n) { return eng_() % n; }

private:
    std::mt19937_64 eng_;
};

struct Kepler {
    double a, e, inc, raan, argp, M;
};

State to_eci(const Kepler& k) {
    double E = k.e < 0.8 ? k.M : PI;
    for (int it = 0; it < 100; ++it) {
        const double f = E - k.e * std::sin(E) - k.M;
        const double d = f / (1.0 - k.e * std::cos(E));
        E -= d;
        if (std::fabs(d) < 1e-15) break;
    }
    const double nu = 2.0 * std::atan2(std::sqrt(1.0 + k.e) * std::sin(E / 2),
                                       std::sqrt(1.0 - k.e) * std::cos(E / 2));
    const double p = k.a * (1.0 - k.e * k.e);
    const double r = k.a * (1.0 - k.e * std::cos(E));
    const double vf = std::sqrt(MU_EARTH / p);
    const double px = r * std::cos(nu), py = r * std::sin(nu);
    const double vx = -vf * std::sin(nu), vy = vf * (k.e + std::cos(nu));
    const double cO = std::cos(k.raan), sO = std::sin(k.raan);
    const double ci = std::cos(k.inc), si = std::sin(k.inc);
    const double cw = std::cos(k.argp), sw = std::sin(k.argp);
    const double r11 = cO * cw - sO * sw * ci, r12 = -cO * sw - sO * cw * ci;
    const double r21 = sO * cw + cO * sw * ci, r22 = -sO * sw + cO * cw * ci;
    const double r31 = sw * si, r32 = cw * si;
    return State{{r11 * px + r12 * py, r21 * px + r22 * py, r31 * px + r32 * py},
                 {r11 * vx + r12 * vy, r21 * vx + r22 * vy, r31 * vx + r32 * vy}};
}

double clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

Kepler draw_leo(Draw& d, double boost) {
    // Altitude: shells at 550 and 800 km over a broad 300-2000 km background.
    const double w550 = 0.45 * boost, w800 = 0.30 * boost, wbg = 0.25;
    const double u = d.uniform() * (w550 + w800 + wbg);
    double alt;
    if (u < w550) alt = 550.0 + 15.0 * d.normal();
    else if (u < w550 + w800) alt = 800.0 + 40.0 * d.normal();
    else alt = 300.0 * std::pow(2000.0 / 300.0, d.uniform());
    alt = clamp(alt, 300.0, 2000.0);
    const double v = d.uniform();
    double inc;
    if (v < 0.4) inc = 53.0 + 0.3 * d.normal();
    else if (v < 0.7) inc = d.uniform(97.0, 99.0);
    else inc = d.uniform(0.0, 100.0);
    double e = d.exponential(0.005);
    const double a = R_EARTH + alt;
    e = std::min(e, 1.0 - (R_EARTH + 200.0) / a);  // perigee above 200 km
    e = std::max(e, 0.0);
    return {a, e, inc * DEG, d.uniform(0, 2 * PI), d.uniform(0, 2 * PI), d.uniform(0, 2 * PI)};
}

Kepler draw_heo(Draw& d) {
    const double hp = d.uniform(200.0, 600.0);
    const double e = d.uniform(0.6, 0.75);
    const double a = (R_EARTH + hp) / (1.0 - e);
    const double inc = d.uniform() < 0.6 ? d.uniform(0.0, 28.5) : 63.4 + 0.5 * d.normal();
    return {a, e, inc * DEG, d.uniform(0, 2 * PI), d.uniform(0, 2 * PI), d.uniform(0, 2 * PI)};
}

Kepler draw_meo(Draw& d) {
    const double a = R_EARTH + 20200.0 + 30.0 * d.normal();
    return {a, std::min(d.exponential(0.003), 0.05), (55.0 + 0.5 * d.normal()) * DEG,
            d.uniform(0, 2 * PI), d.uniform(0, 2 * PI), d.uniform(0, 2 * PI)};
}

Kepler draw_geo(Draw& d) {
    const double a = R_EARTH + 35786.0 + 20.0 * d.normal();
    return {a, std::min(d.exponential(0.0005), 0.01), d.uniform(0.0, 15.0) * DEG,
            d.uniform(0, 2 * PI), d.uniform(0, 2 * PI), d.uniform(0, 2 * PI)};
}

const char* rcs_name(double size_m) {
    return size_m <= 0.2 ? "SMALL" : (size_m >= 4.0 ? "LARGE" : "MEDIUM");
}

Object to_object(const khost_object& k) {
    Object o;
    o.id = k.id;
    o.event = k.event;
    o.alive = k.alive != 0;
    o.type = static_cast<ObjectType>(k.type);
    o.s = State{{k.state[0], k.state[1], k.state[2]}, {k.state[3], k.state[4], k.state[5]}};
    o.size_m = k.size_m;
    o.mass_kg = k.mass_kg;
    o.area_to_mass = k.area_to_mass;
    o.rcs = rcs_name(k.size_m);
    o.name = "OBJ " + std::to_string(k.id);
    return o;
}

khost_object from_object(const Object& o) {
    khost_object k;
    k.id = o.id;
    k.event = o.event;
    k.alive = o.alive ? 1 : 0;
    k.type = static_cast<int32_t>(o.type);
    k.state[0] = o.s.r.x; k.state[1] = o.s.r.y; k.state[2] = o.s.r.z;
    k.state[3] = o.s.v.x; k.state[4] = o.s.v.y; k.state[5] = o.s.v.z;
    k.size_m = o.size_m;
    k.mass_kg = o.mass_kg;
    k.area_to_mass = o.area_to_mass;
    return k;
}

}  // namespace

extern "C" {

int khost_generate(int64_t n, uint64_t seed, const khost_mix* mix, khost_object* out) {
    if (n < 0 || (n > 0 && !out)) return -3;
    khost_mix m{0.7, 0.1, 0.1, 0.1, 1.0};
    if (mix) m = *mix;
    const double total = m.frac_leo + m.frac_heo + m.frac_meo + m.frac_geo;
    if (!(total > 0)) return -3;
    Draw d(seed);
    const MassTable masses;
    for (int64_t i = 0; i < n; ++i) {
        const double u = d.uniform() * total;
        Kepler k;
        if (u < m.frac_leo) k = draw_leo(d, m.leo_shell_boost);
        else if (u < m.frac_leo + m.frac_heo) k = draw_heo(d);
        else if (u < m.frac_leo + m.frac_heo + m.frac_meo) k = draw_meo(d);
        else k = draw_geo(d);
        const State s = to_eci(k);
        const double c = d.uniform();
        const double size = c < 0.5 ? 0.2 : (c < 0.8 ? 0.8 : 4.0);
        const ObjectType type = size >= 4.0 ? ObjectType::Payload : ObjectType::Debris;
        khost_object& o = out[i];
        o.id = static_cast<int32_t>(i + 1);
        o.event = -1;
        o.alive = 1;
        o.type = static_cast<int32_t>(type);
        o.state[0] = s.r.x; o.state[1] = s.r.y; o.state[2] = s.r.z;
        o.state[3] = s.v.x; o.state[4] = s.v.y; o.state[5] = s.v.z;
        o.size_m = size;
        o.mass_kg = masses.mass_for(rcs_name(size), type);
        o.area_to_mass = 0.0;
    }
    return 0;
}

void khost_kepler_state(const double* el, double* state6) {
    const State s = to_eci(Kepler{el[0], el[1], el[2], el[3], el[4], el[5]});
    state6[0] = s.r.x; state6[1] = s.r.y; state6[2] = s.r.z;
    state6[3] = s.v.x; state6[4] = s.v.y; state6[5] = s.v.z;
}

int64_t khost_impacts(int64_t n, const khost_object* in, double epoch_jd, double t, int nimp,
                      const int32_t* targets, const double* projectile_mass_kg,
                      const double* v_rel_km_s, uint64_t seed, double lc_m, int64_t cap,
                      khost_object* out, int32_t* events) {
    if (n < 0 || nimp < 0) return -3;
    Simulation sim;
    sim.objects.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) sim.objects.push_back(to_object(in[i]));
    sim.epoch_jd = epoch_jd;
    sim.t = t;
    BreakupConfig cfg;
    cfg.seed = seed;
    cfg.lc_m = lc_m;
    Fragmentation frag(cfg);
    for (int k = 0; k < nimp; ++k) {
        if (targets[k] < 0 || targets[k] >= static_cast<int32_t>(sim.objects.size())) return -3;
        const BreakupEvent ev = frag.impact(sim, targets[k], projectile_mass_kg[k], v_rel_km_s[k]);
        if (events) {
            events[2 * k] = ev.fragments;
            events[2 * k + 1] = ev.escaped;
        }
    }
    const int64_t total = static_cast<int64_t>(sim.objects.size());
    for (int64_t i = 0; i < total && i < cap && out; ++i) out[i] = from_object(sim.objects[i]);
    return total;
}

int64_t khost_cascade(int64_t n_catalog, uint64_t seed, int n_events, double lc_m, double mass_lo,
                      double mass_hi, double vrel_lo, double vrel_hi, int64_t cap,
                      khost_object* out) {
    if (n_catalog <= 0 || n_events < 0) return -3;
    std::vector<khost_object> cat(static_cast<size_t>(n_catalog));
    if (khost_generate(n_catalog, seed, nullptr, cat.data()) != 0) return -3;
    // Targets: distinct LARGE objects in LEO (perigee and apogee under 2,100 km).
    std::vector<int32_t> pool;
    for (int64_t i = 0; i < n_catalog; ++i) {
        const khost_object& o = cat[i];
        const double r = std::sqrt(o.state[0] * o.state[0] + o.state[1] * o.state[1] + o.state[2] * o.state[2]);
        if (o.size_m >= 4.0 && r < R_EARTH + 2100.0) pool.push_back(static_cast<int32_t>(i));
    }
    if (static_cast<int64_t>(pool.size()) < n_events) return -3;
    Draw d(seed ^ 0x9e3779b97f4a7c15ULL);
    std::vector<int32_t> targets;
    std::vector<double> pm, vr;
    for (int k = 0; k < n_events; ++k) {
        const uint64_t pick = d.below(pool.size() - k);
        std::swap(pool[k], pool[k + pick]);
        targets.push_back(pool[k]);
        cat[pool[k]].mass_kg = d.uniform(mass_lo, mass_hi);
        pm.push_back(d.uniform(mass_lo, mass_hi));
        vr.push_back(d.uniform(vrel_lo, vrel_hi));
    }
    return khost_impacts(n_catalog, cat.data(), JD_DEFAULT_EPOCH, 0.0, n_events, targets.data(),
                         pm.data(), vr.data(), seed, lc_m, cap, out, nullptr);
}

int khost_write_catalog(const char* path, int64_t n, const khost_object* objs, const char* epoch_iso) {
    std::FILE* f = std::fopen(path, "w");
    if (!f) return -3;
    std::fprintf(f, "# Synthetic catalogue (kessler-mi355x generator)\n");
    std::fprintf(f, "# Date de reference (epoch commun) : %s\n",
                 epoch_iso ? epoch_iso : "2026-08-01T10:00:00+00:00");
    std::fprintf(f, "# Colonnes : NORAD_ID | NOM | OBJECT_TYPE | INTL_DESIGNATOR | COUNTRY | RCS_SIZE | "
                    "RCS_VALUE_m2 | TLE_SOURCE_EPOCH | LAT_deg | LON_deg | ALT_km | X_km | Y_km | Z_km | "
                    "VX_km_s | VY_km_s | VZ_km_s | size_m\n");
    // Lines are formatted in parallel, 4,096 per chunk, and written in order.
    auto format = [&](const khost_object& o, char* buf, size_t cap) {
        const double* s = o.state;
        const double alt = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) - R_EARTH;
        return std::snprintf(buf, cap,
                             "%d | %s %d | %s | - | - | %s | 0 | - | 0 | 0 | %.3f | %.17g | %.17g | %.17g | "
                             "%.17g | %.17g | %.17g | size_m=%.17g\n",
                             o.id, o.event >= 0 ? "FRAG" : "OBJ", o.id,
                             object_type_name(static_cast<ObjectType>(o.type)), rcs_name(o.size_m), alt, s[0],
                             s[1], s[2], s[3], s[4], s[5], o.size_m);
    };
    constexpr int64_t CHUNK_LINES = 4096;
    const int64_t nchunks = (n + CHUNK_LINES - 1) / CHUNK_LINES;
    std::vector<std::string> text(static_cast<size_t>(nchunks > 0 ? nchunks : 0));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nchunks; ++c) {
        std::string& out = text[static_cast<size_t>(c)];
        out.reserve(static_cast<size_t>(CHUNK_LINES) * 240);
        char line[512];
        const int64_t end = std::min(n, (c + 1) * CHUNK_LINES);
        for (int64_t i = c * CHUNK_LINES; i < end; ++i) {
            const int len = format(objs[i], line, sizeof line);
            if (len < 0) continue;
            if (static_cast<size_t>(len) < sizeof line) {
                out.append(line, static_cast<size_t>(len));
            } else {
                std::string big(static_cast<size_t>(len) + 1, '\0');
                format(objs[i], &big[0], big.size());
                out.append(big.data(), static_cast<size_t>(len));
            }
        }
    }
    bool ok = true;
    for (const std::string& t : text) ok = ok && std::fwrite(t.data(), 1, t.size(), f) == t.size();
    return std::fclose(f) == 0 && ok ? 0 : -3;
}

int64_t khost_load_catalog(const char* path, int64_t cap, khost_object* out, double* epoch_jd,
                           int* dropped) {
    // The two-call protocol (count, then fill) parses the file once: the
    // last result is kept while the file's size and modification time are
    // unchanged.
    static std::mutex mu;
    static std::string last_path;
    static off_t last_size = -1;
    static struct timespec last_mtime {};
    static std::vector<khost_object> last;
    static double last_epoch = 0.0;
    static int last_bad = 0;
    struct stat st {};
    const bool have = ::stat(path, &st) == 0;
    std::lock_guard<std::mutex> lock(mu);
    if (!(have && last_path == path && last_size == st.st_size && last_mtime.tv_sec == st.st_mtim.tv_sec &&
          last_mtime.tv_nsec == st.st_mtim.tv_nsec)) {
        int bad = 0;
        std::vector<Object> objs = load_catalog(path, &bad);
        const MassTable masses;
        assign_masses(objs, masses);
        last.resize(objs.size());
        for (size_t i = 0; i < objs.size(); ++i) last[i] = from_object(objs[i]);
        last_epoch = read_epoch_jd(path);
        last_bad = bad;
        last_path = have ? path : std::string();
        last_size = have ? st.st_size : -1;
        last_mtime = have ? st.st_mtim : timespec{};
    }
    if (dropped) *dropped = last_bad;
    if (epoch_jd) *epoch_jd = last_epoch;
    const int64_t n = static_cast<int64_t>(last.size());
    for (int64_t i = 0; i < n && i < cap && out; ++i) out[i] = last[static_cast<size_t>(i)];
    return n;
}

int khost_write_snapshot(const char* path, int64_t n, const khost_object* objs, double t, int exact) {
    if (!path || n < 0 || (n > 0 && !objs)) return -3;
    std::vector<Object> v;
    v.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) v.push_back(to_object(objs[i]));
    if (exact) write_snapshot_exact(path, v, t);
    else write_snapshot(path, v, t);
    return 0;
}

int64_t khost_read_snapshot(const char* path, int64_t n, int64_t cap, khost_object* objs, double* t) {
    if (!path || n < 0 || (n > 0 && !objs)) return -3;
    std::vector<Object> v;
    v.reserve(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) v.push_back(to_object(objs[i]));
    double ts = 0.0;
    if (!read_snapshot(path, v, &ts)) return -1;
    if (t) *t = ts;
    const int64_t total = static_cast<int64_t>(v.size());
    for (int64_t i = 0; i < total && i < cap && objs; ++i) objs[i] = from_object(v[static_cast<size_t>(i)]);
    return total;
}

void khost_moon_position(double jd, double* out3) {
    const Vec3 m = moon_position(jd);
    out3[0] = m.x;
    out3[1] = m.y;
    out3[2] = m.z;
}

}  // extern "C"
