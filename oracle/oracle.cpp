// CPU oracle of the Kessler_Sim propagate-and-screen loop (C++17 + OpenMP).
//
// TEST INFRASTRUCTURE ONLY: the checker, never the product. Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
//
// Each function restates the reference algorithm with the same IEEE double
// operations in the same order, so that results are bit-identical to the
// compiled reference (pinned by tests/golden/ fixtures written from it).
// Built with -ffp-contract=off and no -march: no FMA contraction.
//
//   moon_position        orbital.cpp:53-85
//   acceleration         orbital.cpp:91-105
//   rk4                  orbital.cpp:107-118
//   step loop            orbital.cpp:286-300, main.cpp:346-444
//   filter_pair          collision_filter.hpp:37-66
//   prefilter_keep       collision_filter.hpp:85-114
//   cell_key / FORWARD   collision_filter.hpp:118-140
//   Hermite / golden     collision.cpp:77-105
//   refine_pair          collision.cpp:109-132
//   cell_size            collision.cpp:339-356
//   grid + enumeration   collision.cpp:360-563 (candidate set, counters)
//   event order          collision.cpp:739-747
//   loader / epoch       orbital.cpp:13-25, 136-188
//
// The grid here groups objects with a parallel sort instead of the
// reference's hash table / radix sort: grouping never changes which pairs
// are candidates, only how they are found.

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <omp.h>

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double DEG = PI / 180.0;
constexpr double MU_EARTH = 398600.4418;
constexpr double MU_MOON = 4902.800066;
constexpr double R_EARTH = 6378.137;
constexpr double J2 = 1.08262668e-3;
constexpr double JD_DEFAULT_EPOCH = 2461253.916666667;
constexpr double GRAVITY_GRADIENT = 2e-6;
constexpr float DR_ERR_KM = 5e-4f;
constexpr float DV_ERR_KM_S = 5e-6f;
constexpr long long KEY_OFFSET = 1LL << 20;

struct V3 {
    double x, y, z;
};
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }

struct St {
    V3 r, v;
};

inline St load_state(const double* p) { return {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}}; }
inline void store_state(double* p, const St& s) {
    p[0] = s.r.x; p[1] = s.r.y; p[2] = s.r.z; p[3] = s.v.x; p[4] = s.v.y; p[5] = s.v.z;
}

double wrap360(double d) { return std::fmod(std::fmod(d, 360.0) + 360.0, 360.0); }

V3 moon(double jd) {
    const double T = (jd - 2451545.0) / 36525.0;
    const double Lp = wrap360(218.3164477 + 481267.88123421 * T);
    double D = wrap360(297.8501921 + 445267.1114034 * T);
    double M = wrap360(357.5291092 + 35999.0502909 * T);
    double Mp = wrap360(134.9633964 + 477198.8675055 * T);
    double F = wrap360(93.2720950 + 483202.0175233 * T);
    D *= DEG; M *= DEG; Mp *= DEG; F *= DEG;
    // Longitude, latitude (deg) and distance (km): truncated Meeus ch. 47.
    const double lon = Lp + 6.288774 * std::sin(Mp) + 1.274027 * std::sin(2 * D - Mp)
                     + 0.658314 * std::sin(2 * D) + 0.213618 * std::sin(2 * Mp)
                     - 0.185116 * std::sin(M) - 0.114332 * std::sin(2 * F)
                     + 0.058793 * std::sin(2 * D - 2 * Mp) + 0.057066 * std::sin(2 * D - M - Mp)
                     + 0.053322 * std::sin(2 * D + Mp) + 0.045758 * std::sin(2 * D - M);
    const double lat = 5.128122 * std::sin(F) + 0.280602 * std::sin(Mp + F)
                     + 0.277693 * std::sin(Mp - F) + 0.173237 * std::sin(2 * D - F)
                     + 0.055413 * std::sin(2 * D - Mp + F) + 0.046271 * std::sin(2 * D - Mp - F);
    const double dist = 385000.56 - 20905.355 * std::cos(Mp) - 3699.111 * std::cos(2 * D - Mp)
                      - 2955.968 * std::cos(2 * D) - 569.925 * std::cos(2 * Mp)
                      + 48.888 * std::cos(M) - 3.149 * std::cos(2 * F)
                      + 246.158 * std::cos(2 * D - 2 * Mp) - 152.138 * std::cos(2 * D - M - Mp)
                      - 170.733 * std::cos(2 * D + Mp) - 204.586 * std::cos(2 * D - M);
    const double l = lon * DEG, b = lat * DEG;
    const double eps = (23.439291 - 0.0130042 * T) * DEG;
    const double xe = dist * std::cos(b) * std::cos(l);
    const double ye = dist * std::cos(b) * std::sin(l);
    const double ze = dist * std::sin(b);
    return {xe, ye * std::cos(eps) - ze * std::sin(eps), ye * std::sin(eps) + ze * std::cos(eps)};
}

V3 accel(V3 r, V3 rm) {
    const double r2 = dot(r, r), rn = std::sqrt(r2);
    V3 a = scale(r, -MU_EARTH / (r2 * rn));
    const double k = 1.5 * J2 * MU_EARTH * R_EARTH * R_EARTH / (r2 * r2 * rn);
    const double z2 = 5.0 * r.z * r.z / r2;
    a = add(a, V3{k * r.x * (z2 - 1.0), k * r.y * (z2 - 1.0), k * r.z * (z2 - 3.0)});
    const V3 d = sub(rm, r);
    const double dn = len(d), mn = len(rm);
    a = sub(add(a, scale(d, MU_MOON / (dn * dn * dn))), scale(rm, MU_MOON / (mn * mn * mn)));
    return a;
}

St rk4_step(const St& s, double dt, V3 m0, V3 m1, V3 m2) {
    const double h = dt / 2;
    const V3 k1r = s.v, k1v = accel(s.r, m0);
    const V3 y2r = add(s.r, scale(k1r, h)), k2r = add(s.v, scale(k1v, h));
    const V3 k2v = accel(y2r, m1);
    const V3 y3r = add(s.r, scale(k2r, h)), k3r = add(s.v, scale(k2v, h));
    const V3 k3v = accel(y3r, m1);
    const V3 y4r = add(s.r, scale(k3r, dt)), k4r = add(s.v, scale(k3v, dt));
    const V3 k4v = accel(y4r, m2);
    const double w = dt / 6;
    const V3 sr = add(add(add(k1r, scale(k2r, 2)), scale(k3r, 2)), k4r);
    const V3 sv = add(add(add(k1v, scale(k2v, 2)), scale(k3v, 2)), k4v);
    return {add(s.r, scale(sr, w)), add(s.v, scale(sv, w))};
}

int filter(double rx, double ry, double rz, double vx, double vy, double vz, double dt,
           double thr, double vmin2) {
    const double dv2 = vx * vx + vy * vy + vz * vz;
    const double rv = rx * vx + ry * vy + rz * vz;
    double ts = dv2 > 0 ? -rv / dv2 : 0.0;
    if (ts < 0.0) ts = 0.0;
    if (ts > dt) ts = dt;
    const double px = rx + vx * ts, py = ry + vy * ts, pz = rz + vz * ts;
    const double gate = thr + GRAVITY_GRADIENT * std::sqrt(dv2) * dt * dt * dt + 1e-3;
    if (std::sqrt(px * px + py * py + pz * pz) > gate) return 0;
    if (dv2 < vmin2) return 2;
    if (rv >= 0.0 && dv2 > GRAVITY_GRADIENT * (rx * rx + ry * ry + rz * rz)) return 0;
    return 1;
}

bool prefilter(float rx, float ry, float rz, float vx, float vy, float vz, float dt, float thr,
               float vmin2) {
    const float dv2 = vx * vx + vy * vy + vz * vz;
    const float rv = rx * vx + ry * vy + rz * vz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float dvn = std::sqrt(dv2), rn = std::sqrt(r2);
    float ts = dv2 > 0.0f ? -rv / dv2 : 0.0f;
    if (ts < 0.0f) ts = 0.0f;
    if (ts > dt) ts = dt;
    const float px = rx + vx * ts, py = ry + vy * ts, pz = rz + vz * ts;
    const float d = std::sqrt(px * px + py * py + pz * pz);
    const float gate = thr + static_cast<float>(GRAVITY_GRADIENT) * dvn * dt * dt * dt + 1e-3f;
    const float margin = 1e-5f * (rn + dvn * dt) + DR_ERR_KM + DV_ERR_KM_S * dt;
    if (d > gate + margin) return false;
    const float rv_tol = 1e-5f * rn * dvn + DR_ERR_KM * dvn + DV_ERR_KM_S * rn;
    if (dv2 > 1.1f * vmin2 && rv > rv_tol && dv2 > 1.01f * static_cast<float>(GRAVITY_GRADIENT) * r2)
        return false;
    return true;
}

uint64_t key_of(long long ix, long long iy, long long iz) {
    return (static_cast<uint64_t>(ix + KEY_OFFSET) << 42) |
           (static_cast<uint64_t>(iy + KEY_OFFSET) << 21) | static_cast<uint64_t>(iz + KEY_OFFSET);
}

void coords_of(uint64_t key, long long& ix, long long& iy, long long& iz) {
    const uint64_t mask = (1ULL << 21) - 1;
    ix = static_cast<long long>((key >> 42) & mask) - KEY_OFFSET;
    iy = static_cast<long long>((key >> 21) & mask) - KEY_OFFSET;
    iz = static_cast<long long>(key & mask) - KEY_OFFSET;
}

const int FWD[13][3] = {{1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0},
                        {1, 0, 1},   {1, 1, -1}, {1, 1, 0},  {1, 1, 1},  {0, 1, -1},
                        {0, 1, 0},   {0, 1, 1},  {0, 0, 1}};

// Cubic Hermite of the relative motion over one step, s in [0, 1].
struct Herm {
    V3 p0, m0, p1, m1;
    V3 pos(double s) const {
        const double s2 = s * s, s3 = s2 * s;
        return add(add(add(scale(p0, 2 * s3 - 3 * s2 + 1), scale(m0, s3 - 2 * s2 + s)),
                       scale(p1, -2 * s3 + 3 * s2)),
                   scale(m1, s3 - s2));
    }
    V3 vel(double s) const {
        const double s2 = s * s;
        return add(add(add(scale(p0, 6 * s2 - 6 * s), scale(m0, 3 * s2 - 4 * s + 1)),
                       scale(p1, -6 * s2 + 6 * s)),
                   scale(m1, 3 * s2 - 2 * s));
    }
};

double golden_min(const Herm& h) {
    auto f = [&](double s) { const V3 p = h.pos(s); return dot(p, p); };
    const double g = 0.6180339887498949;
    double a = 0.0, b = 1.0;
    double c = b - g * (b - a), d = a + g * (b - a);
    double fc = f(c), fd = f(d);
    for (int it = 0; it < 60; ++it) {
        if (fc < fd) {
            b = d; d = c; fd = fc; c = b - g * (b - a); fc = f(c);
        } else {
            a = c; c = d; fc = fd; d = a + g * (b - a); fd = f(d);
        }
    }
    return 0.5 * (a + b);
}

double radius_km(double size_m, double scale_) { return 0.5 * size_m * 1e-3 * scale_; }

struct Problem {
    int n;
    const double* before;
    const double* after;
    const uint8_t* alive;
    const double* size_m;
    double t0, dt, thr, rscale, vmin2;
};

bool refine(const Problem& P, int i, int j, orc_event& out) {
    const St bi = load_state(P.before + 6 * (size_t)i), bj = load_state(P.before + 6 * (size_t)j);
    const St ai = load_state(P.after + 6 * (size_t)i), aj = load_state(P.after + 6 * (size_t)j);
    const Herm h{sub(bj.r, bi.r), scale(sub(bj.v, bi.v), P.dt), sub(aj.r, ai.r),
                 scale(sub(aj.v, ai.v), P.dt)};
    const double s = golden_min(h);
    if (s <= 1e-9 || s >= 1.0 - 1e-9) return false;
    const double miss = len(h.pos(s));
    if (miss > P.thr) return false;
    out.i = i;
    out.j = j;
    out.t = P.t0 + s * P.dt;
    out.miss_km = miss;
    out.v_rel_km_s = len(h.vel(s)) / P.dt;
    out.collision = miss < radius_km(P.size_m[i], P.rscale) + radius_km(P.size_m[j], P.rscale) ? 1 : 0;
    out.pad = 0;
    return true;
}

double cell_size_of(int n, const double* before, const uint8_t* alive, double dt, double thr) {
    double vmax = 0.0;
#pragma omp parallel for reduction(max : vmax) schedule(static)
    for (int i = 0; i < n; ++i)
        if (alive[i]) {
            const double* p = before + 6 * (size_t)i;
            vmax = std::max(vmax, len(V3{p[3], p[4], p[5]}));
        }
    const double vrel_max = 2.0 * vmax;
    return std::max(1.0, vrel_max * dt + thr + GRAVITY_GRADIENT * vrel_max * dt * dt * dt);
}

struct Entry {
    uint64_t key;
    int idx;
};

struct CellRange {
    uint64_t key;
    int begin, end;
};

// Screens one step; events sorted by (t, i, j).
std::vector<orc_event> screen_once(const Problem& P, orc_stats& st) {
    const double cell = cell_size_of(P.n, P.before, P.alive, P.dt, P.thr);
    std::vector<Entry> ent;
    ent.reserve(P.n);
    for (int i = 0; i < P.n; ++i) {
        if (!P.alive[i]) continue;
        const double* r = P.before + 6 * (size_t)i;
        ent.push_back({key_of(static_cast<long long>(std::floor(r[0] / cell)),
                              static_cast<long long>(std::floor(r[1] / cell)),
                              static_cast<long long>(std::floor(r[2] / cell))),
                       i});
    }
    std::sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) {
        return a.key != b.key ? a.key < b.key : a.idx < b.idx;
    });
    std::vector<CellRange> cells;
    for (int k = 0; k < (int)ent.size(); ++k)
        if (k == 0 || ent[k].key != ent[k - 1].key) cells.push_back({ent[k].key, k, k});
    for (size_t c = 0; c < cells.size(); ++c)
        cells[c].end = c + 1 < cells.size() ? cells[c + 1].begin : (int)ent.size();
    auto find_cell = [&](uint64_t key) -> int {
        auto it = std::lower_bound(cells.begin(), cells.end(), key,
                                   [](const CellRange& c, uint64_t k) { return c.key < k; });
        return (it != cells.end() && it->key == key) ? (int)(it - cells.begin()) : -1;
    };

    const int nt = omp_get_max_threads();
    std::vector<std::vector<orc_event>> found(nt);
    long long cand = 0, refined = 0, co = 0;
    int max_cell = 0;
    const int ncells = (int)cells.size();
#pragma omp parallel reduction(+ : cand, refined, co) reduction(max : max_cell)
    {
        std::vector<orc_event>& out = found[omp_get_thread_num()];
        auto visit = [&](int a, int b) {
            ++cand;
            const int i = std::min(a, b), j = std::max(a, b);
            const double* si = P.before + 6 * (size_t)i;
            const double* sj = P.before + 6 * (size_t)j;
            const int r = filter(sj[0] - si[0], sj[1] - si[1], sj[2] - si[2], sj[3] - si[3],
                                 sj[4] - si[4], sj[5] - si[5], P.dt, P.thr, P.vmin2);
            if (r == 2) { ++co; return; }
            if (r != 1) return;
            ++refined;
            orc_event e;
            if (refine(P, i, j, e)) out.push_back(e);
        };
#pragma omp for schedule(dynamic, 16)
        for (int c = 0; c < ncells; ++c) {
            const CellRange& cc = cells[c];
            max_cell = std::max(max_cell, cc.end - cc.begin);
            for (int p = cc.begin; p < cc.end; ++p)
                for (int q = p + 1; q < cc.end; ++q) visit(ent[p].idx, ent[q].idx);
            long long x, y, z;
            coords_of(cc.key, x, y, z);
            for (const auto& o : FWD) {
                const int nc = find_cell(key_of(x + o[0], y + o[1], z + o[2]));
                if (nc < 0) continue;
                for (int p = cc.begin; p < cc.end; ++p)
                    for (int q = cells[nc].begin; q < cells[nc].end; ++q)
                        visit(ent[p].idx, ent[q].idx);
            }
        }
    }
    std::vector<orc_event> all;
    for (auto& f : found) all.insert(all.end(), f.begin(), f.end());
    std::sort(all.begin(), all.end(), [](const orc_event& a, const orc_event& b) {
        if (a.t != b.t) return a.t < b.t;
        if (a.i != b.i) return a.i < b.i;
        return a.j < b.j;
    });
    st.candidate_pairs = cand;
    st.refined_pairs = refined;
    st.co_orbiting = co;
    st.cell_km = cell;
    st.max_cell_objects = max_cell;
    st.pad = 0;
    return all;
}

std::string trim(const std::string& s) {
    const size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

double civil_to_jd(int y, int m, double day) {
    if (m <= 2) { y -= 1; m += 12; }
    const int A = y / 100;
    const int B = 2 - A + A / 4;
    return std::floor(365.25 * (y + 4716)) + std::floor(30.6001 * (m + 1)) + day + B - 1524.5;
}

double epoch_of(const char* path) {
    std::ifstream in(path);
    if (!in) return JD_DEFAULT_EPOCH;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] != '#') break;
        const size_t pos = line.find("epoch commun) :");
        if (pos == std::string::npos) continue;
        int y, mo, d, h, mi, s;
        if (std::sscanf(line.c_str() + pos, "epoch commun) : %d-%d-%dT%d:%d:%d", &y, &mo, &d, &h,
                        &mi, &s) == 6)
            return civil_to_jd(y, mo, d + (h + mi / 60.0 + s / 3600.0) / 24.0);
    }
    return JD_DEFAULT_EPOCH;
}

}  // namespace

extern "C" {

int orc_configure_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void orc_moon_position(double jd, double* out) {
    const V3 m = moon(jd);
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
}

void orc_acceleration(const double* r, const double* rm, double* out) {
    const V3 a = accel({r[0], r[1], r[2]}, {rm[0], rm[1], rm[2]});
    out[0] = a.x; out[1] = a.y; out[2] = a.z;
}

void orc_rk4(const double* s, double dt, const double* m0, const double* m1, const double* m2,
             double* out) {
    store_state(out, rk4_step(load_state(s), dt, {m0[0], m0[1], m0[2]}, {m1[0], m1[1], m1[2]},
                              {m2[0], m2[1], m2[2]}));
}

int orc_filter_pair(const double* r, const double* v, double dt, double thr, double vmin2) {
    return filter(r[0], r[1], r[2], v[0], v[1], v[2], dt, thr, vmin2);
}

int orc_prefilter_keep(const float* r, const float* v, float dt, float thr, float vmin2) {
    return prefilter(r[0], r[1], r[2], v[0], v[1], v[2], dt, thr, vmin2) ? 1 : 0;
}

void orc_interpolate_state(const double* a, const double* b, double dt, double s, double* out) {
    const St sa = load_state(a), sb = load_state(b);
    const Herm h{sa.r, scale(sa.v, dt), sb.r, scale(sb.v, dt)};
    const V3 p = h.pos(s), v = scale(h.vel(s), 1.0 / dt);
    out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = v.x; out[4] = v.y; out[5] = v.z;
}

double orc_cell_size(int n, const double* before, const uint8_t* alive, double dt, double thr) {
    return cell_size_of(n, before, alive, dt, thr);
}

uint64_t orc_cell_key(int64_t ix, int64_t iy, int64_t iz) { return key_of(ix, iy, iz); }

int64_t orc_screen(int n, const double* before, const double* after, const uint8_t* alive_after,
                   const double* size_m, double t0, double dt, double thr, double radius_scale,
                   double vmin_km_s, orc_event* out, int64_t cap, orc_stats* stats) {
    const Problem P{n, before, after, alive_after, size_m, t0, dt, thr, radius_scale,
                    vmin_km_s * vmin_km_s};
    orc_stats st;
    const std::vector<orc_event> ev = screen_once(P, st);
    if (stats) *stats = st;
    for (size_t k = 0; k < ev.size() && out && (int64_t)k < cap; ++k) out[k] = ev[k];
    return (int64_t)ev.size();
}

int64_t orc_run(int n, double* states, uint8_t* alive, const double* size_m, double epoch_jd,
                double t_start, double dt, int64_t nsteps, int screen, double thr,
                double radius_scale, double vmin_km_s, orc_event* out, int64_t cap,
                orc_stats* per_step, double* loop_s) {
    std::vector<double> before(6 * (size_t)n);
    double t = t_start;
    int64_t count = 0;
    const auto c0 = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < nsteps; ++k) {
        std::memcpy(before.data(), states, sizeof(double) * 6 * (size_t)n);
        // Simulation::step: Moon at t, t + dt/2, t + dt from the same
        // expressions as the reference.
        const V3 m0 = moon(epoch_jd + t / 86400.0);
        const V3 m1 = moon(epoch_jd + (t + dt / 2) / 86400.0);
        const V3 m2 = moon(epoch_jd + (t + dt) / 86400.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            if (!alive[i]) continue;
            double* p = states + 6 * (size_t)i;
            const St s = rk4_step(load_state(p), dt, m0, m1, m2);
            store_state(p, s);
            if (len(s.r) < R_EARTH) alive[i] = 0;
        }
        const double t0 = t;
        t += dt;
        if (!screen) continue;
        const Problem P{n, before.data(), states, alive, size_m, t0, dt, thr, radius_scale,
                        vmin_km_s * vmin_km_s};
        orc_stats st;
        const std::vector<orc_event> ev = screen_once(P, st);
        if (per_step) per_step[k] = st;
        for (const orc_event& e0 : ev) {
            if (out && count < cap) {
                out[count] = e0;
                out[count].step = k;
            }
            ++count;
        }
    }
    if (loop_s)
        *loop_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
    return count;
}

int64_t orc_bruteforce_candidates(int n, const double* before, const uint8_t* alive, double cell) {
    std::vector<long long> cx(n), cy(n), cz(n);
    for (int i = 0; i < n; ++i) {
        const double* r = before + 6 * (size_t)i;
        cx[i] = (long long)std::floor(r[0] / cell);
        cy[i] = (long long)std::floor(r[1] / cell);
        cz[i] = (long long)std::floor(r[2] / cell);
    }
    int64_t count = 0;
    for (int i = 0; i < n; ++i) {
        if (!alive[i]) continue;
        for (int j = i + 1; j < n; ++j) {
            if (!alive[j]) continue;
            if (std::llabs(cx[i] - cx[j]) <= 1 && std::llabs(cy[i] - cy[j]) <= 1 &&
                std::llabs(cz[i] - cz[j]) <= 1)
                ++count;
        }
    }
    return count;
}

int orc_load_catalog(const char* path, int cap, int32_t* ids, double* states, double* size_m,
                     int32_t* rcs_code, int32_t* type_code, double* epoch_jd, int* dropped) {
    if (epoch_jd) *epoch_jd = epoch_of(path);
    std::ifstream in(path);
    int bad = 0, n = 0;
    if (!in) {
        if (dropped) *dropped = 0;
        return 0;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::vector<std::string> f;
        std::stringstream ss(line);
        std::string tok;
        while (std::getline(ss, tok, '|')) f.push_back(trim(tok));
        if (f.size() < 17) { ++bad; continue; }
        try {
            const int id = std::stoi(f[0]);
            const std::string& rcs = f[5];
            const std::string& type = f[2];
            double s[6];
            for (int k = 0; k < 6; ++k) s[k] = std::stod(f[11 + k]);
            if (n < cap) {
                if (ids) ids[n] = id;
                if (states) std::memcpy(states + 6 * (size_t)n, s, sizeof(s));
                if (size_m) size_m[n] = rcs == "SMALL" ? 0.2 : rcs == "LARGE" ? 4.0 : 0.8;
                if (rcs_code) rcs_code[n] = rcs == "SMALL" ? 0 : rcs == "MEDIUM" ? 1 : rcs == "LARGE" ? 2 : 3;
                if (type_code)
                    type_code[n] = type == "PAYLOAD" ? 0 : type == "DEBRIS" ? 1 : type == "ROCKET BODY" ? 2 : 3;
            }
            ++n;
        } catch (...) {
            ++bad;
        }
    }
    if (dropped) *dropped = bad;
    return n;
}

}  // extern "C"
