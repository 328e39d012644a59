// Host side of the orbital core: the lunar ephemeris (evaluated on the host
// once per half step and broadcast to the RK4 kernel), the host RK4 used for
// fragments born inside a step, orbital elements and catalogue I/O.
//
// Operation order follows the reference (orbital.cpp:53-130) so that the
// Moon table and the host RK4 are bit-identical to it; built with
// -ffp-contract=off.

#include "orbital.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace kessler {

namespace {

std::string strip(const std::string& s) {
    const size_t b = s.find_first_not_of(" \t\r\n");
    const size_t e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

// Civil date (UTC) to Julian day, Meeus ch. 7 (orbital.cpp:19-25).
double julian_day(int y, int m, double day) {
    if (m <= 2) {
        y -= 1;
        m += 12;
    }
    const int century = y / 100;
    const int gregorian = 2 - century + century / 4;
    return std::floor(365.25 * (y + 4716)) + std::floor(30.6001 * (m + 1)) + day + gregorian - 1524.5;
}

double wrap_deg(double d) { return std::fmod(std::fmod(d, 360.0) + 360.0, 360.0); }

}  // namespace

ObjectType parse_object_type(const std::string& s) {
    if (s == "PAYLOAD") return ObjectType::Payload;
    if (s == "DEBRIS") return ObjectType::Debris;
    if (s == "ROCKET BODY") return ObjectType::RocketBody;
    return ObjectType::Unknown;
}

const char* object_type_name(ObjectType t) {
    switch (t) {
        case ObjectType::Payload: return "PAYLOAD";
        case ObjectType::Debris: return "DEBRIS";
        case ObjectType::RocketBody: return "ROCKET BODY";
        default: return "UNKNOWN";
    }
}

Vec3 moon_position(double jd) {
    const double T = (jd - 2451545.0) / 36525.0;
    const double Lp = wrap_deg(218.3164477 + 481267.88123421 * T);  // mean longitude
    double D = wrap_deg(297.8501921 + 445267.1114034 * T);          // mean elongation
    double M = wrap_deg(357.5291092 + 35999.0502909 * T);           // Sun mean anomaly
    double Mp = wrap_deg(134.9633964 + 477198.8675055 * T);         // Moon mean anomaly
    double F = wrap_deg(93.2720950 + 483202.0175233 * T);           // argument of latitude
    D *= DEG;
    M *= DEG;
    Mp *= DEG;
    F *= DEG;
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
    const double l = lon * DEG;
    const double b = lat * DEG;
    const double eps = (23.439291 - 0.0130042 * T) * DEG;  // obliquity
    const double xe = dist * std::cos(b) * std::cos(l);
    const double ye = dist * std::cos(b) * std::sin(l);
    const double ze = dist * std::sin(b);
    return {xe, ye * std::cos(eps) - ze * std::sin(eps), ye * std::sin(eps) + ze * std::cos(eps)};
}

Vec3 acceleration(const Vec3& r, const Vec3& rm) {
    const double r2 = dot(r, r);
    const double rn = std::sqrt(r2);
    Vec3 a = r * (-MU_EARTH / (r2 * rn));
    const double k = 1.5 * J2 * MU_EARTH * R_EARTH * R_EARTH / (r2 * r2 * rn);
    const double z2 = 5.0 * r.z * r.z / r2;
    a = a + Vec3{k * r.x * (z2 - 1.0), k * r.y * (z2 - 1.0), k * r.z * (z2 - 3.0)};
    const Vec3 d = rm - r;
    const double dn = norm(d);
    const double mn = norm(rm);
    return a + d * (MU_MOON / (dn * dn * dn)) - rm * (MU_MOON / (mn * mn * mn));
}

State rk4(const State& s, double dt, const Vec3& rm0, const Vec3& rm1, const Vec3& rm2) {
    const double h = dt / 2;
    const Vec3 k1r = s.v, k1v = acceleration(s.r, rm0);
    const Vec3 k2r = s.v + k1v * h, k2v = acceleration(s.r + k1r * h, rm1);
    const Vec3 k3r = s.v + k2v * h, k3v = acceleration(s.r + k2r * h, rm1);
    const Vec3 k4r = s.v + k3v * dt, k4v = acceleration(s.r + k3r * dt, rm2);
    const double w = dt / 6;
    return {s.r + (k1r + k2r * 2 + k3r * 2 + k4r) * w, s.v + (k1v + k2v * 2 + k3v * 2 + k4v) * w};
}

Elements elements(const State& s) {
    const double r = norm(s.r), v2 = dot(s.v, s.v);
    const double energy = v2 / 2 - MU_EARTH / r;
    const double a = -MU_EARTH / (2 * energy);
    const Vec3 h = cross(s.r, s.v);
    const Vec3 ev = cross(s.v, h) * (1.0 / MU_EARTH) - s.r * (1.0 / r);
    const double e = norm(ev);
    const double period = a > 0 ? 2 * PI * std::sqrt(a * a * a / MU_EARTH) / 60.0 : 0.0;
    return {a, e, std::acos(h.z / norm(h)) / DEG, a * (1 - e) - R_EARTH, a * (1 + e) - R_EARTH, period};
}

namespace {

// One catalogue line: 0 skipped (empty or comment), 1 object, 2 malformed.
struct ParsedLine {
    int status = 0;
    Object o;
};

// The fields of one line split as std::getline(stream, token, '|') splits
// them (no empty token after a final '|'), each stripped of blanks.
void parse_line(const char* b, const char* e, ParsedLine& out) {
    const std::string line(b, e);
    if (line.empty() || line[0] == '#') {
        out.status = 0;
        return;
    }
    std::vector<std::string> field;
    field.reserve(20);
    size_t start = 0;
    while (start < line.size()) {
        const size_t bar = line.find('|', start);
        const size_t stop = bar == std::string::npos ? line.size() : bar;
        field.push_back(strip(line.substr(start, stop - start)));
        if (bar == std::string::npos) break;
        start = bar + 1;
    }
    if (field.size() < 17) {
        out.status = 2;
        return;
    }
    try {
        Object& o = out.o;
        o.id = std::stoi(field[0]);
        o.name = field[1];
        o.type = parse_object_type(field[2]);
        o.country = field[4];
        o.rcs = field[5];
        // Size from the radar cross-section class: collision radii
        // 0.1 / 0.4 / 2 m for SMALL / MEDIUM (and unknown) / LARGE.
        o.size_m = o.rcs == "SMALL" ? 0.2 : (o.rcs == "LARGE" ? 4.0 : 0.8);
        o.s.r = {std::stod(field[11]), std::stod(field[12]), std::stod(field[13])};
        o.s.v = {std::stod(field[14]), std::stod(field[15]), std::stod(field[16])};
        // Extension: an 18th field "size_m=<L>" (written for fragments,
        // whose size is not an RCS class) overrides the class size.
        if (field.size() > 17 && field[17].rfind("size_m=", 0) == 0) o.size_m = std::stod(field[17].substr(7));
        out.status = 1;
    } catch (...) {
        out.status = 2;
    }
}

}  // namespace

// load_catalog (orbital.cpp:150-170): the whole file is read at once, lines
// are parsed in parallel (each independently, as the reference parses them
// one by one) and the objects kept in file order.
std::vector<Object> load_catalog(const std::string& path, int* dropped) {
    std::vector<Object> objs;
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::cerr << "cannot open " << path << "\n";
        if (dropped) *dropped = 0;
        return objs;
    }
    std::string buf;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size > 0) {
        buf.resize(static_cast<size_t>(size));
        buf.resize(std::fread(&buf[0], 1, buf.size(), f));
    }
    std::fclose(f);
    // Line boundaries ('\n' only, as std::getline); a last line without
    // '\n' counts, an empty tail after a final '\n' does not.
    std::vector<size_t> begin;
    begin.reserve(buf.size() / 64 + 1);
    for (size_t p = 0; p < buf.size();) {
        begin.push_back(p);
        const void* nl = std::memchr(buf.data() + p, '\n', buf.size() - p);
        p = nl ? static_cast<size_t>(static_cast<const char*>(nl) - buf.data()) + 1 : buf.size();
    }
    const int64_t nlines = static_cast<int64_t>(begin.size());
    std::vector<ParsedLine> parsed(static_cast<size_t>(nlines));
#pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t k = 0; k < nlines; ++k) {
        const size_t b = begin[static_cast<size_t>(k)];
        size_t e = k + 1 < nlines ? begin[static_cast<size_t>(k + 1)] : buf.size();
        if (e > b && buf[e - 1] == '\n') --e;
        parse_line(buf.data() + b, buf.data() + e, parsed[static_cast<size_t>(k)]);
    }
    int bad = 0;
    size_t keep = 0;
    for (const ParsedLine& pl : parsed) {
        keep += pl.status == 1 ? 1 : 0;
        bad += pl.status == 2 ? 1 : 0;
    }
    objs.reserve(keep);
    for (ParsedLine& pl : parsed)
        if (pl.status == 1) objs.push_back(std::move(pl.o));
    if (dropped) *dropped = bad;
    return objs;
}

double read_epoch_jd(const std::string& path) {
    std::ifstream in(path);
    if (!in) return JD_DEFAULT_EPOCH;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] != '#') break;
        const size_t pos = line.find("epoch commun) :");
        if (pos == std::string::npos) continue;
        int y, mo, d, h, mi, s;
        if (std::sscanf(line.c_str() + pos, "epoch commun) : %d-%d-%dT%d:%d:%d", &y, &mo, &d, &h, &mi,
                        &s) == 6)
            return julian_day(y, mo, d + (h + mi / 60.0 + s / 3600.0) / 24.0);
    }
    return JD_DEFAULT_EPOCH;
}

static void write_snapshot_impl(const std::string& path, const std::vector<Object>& objs, double t,
                                int precision) {
    std::ofstream out(path);
    out << "# t = " << t << " s apres l'epoch\nid,x_km,y_km,z_km,vx,vy,vz,alive\n";
    out.precision(precision);
    for (const Object& o : objs)
        out << o.id << ',' << o.s.r.x << ',' << o.s.r.y << ',' << o.s.r.z << ',' << o.s.v.x << ','
            << o.s.v.y << ',' << o.s.v.z << ',' << o.alive << '\n';
}

void write_snapshot(const std::string& path, const std::vector<Object>& objs, double t) {
    write_snapshot_impl(path, objs, t, 9);
}

void write_snapshot_exact(const std::string& path, const std::vector<Object>& objs, double t) {
    write_snapshot_impl(path, objs, t, 17);
}

// read_snapshot (orbital.hpp:133, orbital.cpp:199-249): comment lines carry
// "t = <seconds>", the first other line names the columns, then one row per
// object in the order written. A row must match the loaded object at its
// position by id; rows past the list are fragments born during the run
// (ids from FRAGMENT_ID_BASE), appended with generic properties; loaded
// fragments past the last row are dropped. A catalogue object without a row,
// a row out of order or a malformed row make the file not match (false).
bool read_snapshot(const std::string& path, std::vector<Object>& objs, double* t) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    size_t row = 0;
    bool columns_seen = false;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            const size_t at = line.find("t =");
            if (t && at != std::string::npos) *t = std::atof(line.c_str() + at + 3);
            continue;
        }
        if (!columns_seen) {
            columns_seen = true;
            continue;
        }
        // Comma-separated fields, split as std::getline(stream, token, ',').
        std::vector<std::string> f;
        size_t start = 0;
        while (start < line.size()) {
            const size_t comma = line.find(',', start);
            const size_t stop = comma == std::string::npos ? line.size() : comma;
            f.push_back(line.substr(start, stop - start));
            if (comma == std::string::npos) break;
            start = comma + 1;
        }
        if (f.size() < 8) return false;
        try {
            const int id = std::stoi(f[0]);
            if (row >= objs.size()) {
                if (id < FRAGMENT_ID_BASE) return false;
                Object o;
                o.id = id;
                o.name = "FRAG " + std::to_string(id - FRAGMENT_ID_BASE);
                o.type = ObjectType::Debris;
                o.rcs = "SMALL";
                o.size_m = 0.1;
                o.event = 0;
                objs.push_back(std::move(o));
            } else if (objs[row].id != id) {
                return false;
            }
            Object& o = objs[row];
            o.s.r = {std::stod(f[1]), std::stod(f[2]), std::stod(f[3])};
            o.s.v = {std::stod(f[4]), std::stod(f[5]), std::stod(f[6])};
            o.alive = std::stoi(f[7]) != 0;
        } catch (...) {
            return false;
        }
        ++row;
    }
    if (row < objs.size()) {
        for (size_t k = row; k < objs.size(); ++k)
            if (!objs[k].is_fragment()) return false;
        objs.resize(row);
    }
    return true;
}

std::string jd_to_utc_string(double jd) {
    const double z = std::floor(jd + 0.5);
    const double frac = jd + 0.5 - z;
    double a = z;
    if (z >= 2299161) {
        const double alpha = std::floor((z - 1867216.25) / 36524.25);
        a = z + 1 + alpha - std::floor(alpha / 4);
    }
    const double b = a + 1524;
    const double c = std::floor((b - 122.1) / 365.25);
    const double d = std::floor(365.25 * c);
    const double e = std::floor((b - d) / 30.6001);
    const int day = static_cast<int>(b - d - std::floor(30.6001 * e));
    const int month = static_cast<int>(e < 14 ? e - 1 : e - 13);
    const int year = static_cast<int>(month > 2 ? c - 4716 : c - 4715);
    const double hours = frac * 24.0;
    int hh = static_cast<int>(hours);
    int mm = static_cast<int>((hours - hh) * 60.0);
    int sec = static_cast<int>(((hours - hh) * 60.0 - mm) * 60.0 + 0.5);
    if (sec >= 60) { sec -= 60; ++mm; }
    if (mm >= 60) { mm -= 60; ++hh; }
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%04d-%02d-%02d %02d:%02d:%02d UTC", year, month, day, hh, mm, sec);
    return buf;
}

State Simulation::propagate_state(const State& s, double t_from, double duration) const {
    if (duration <= 0) return s;
    const double jd0 = epoch_jd + t_from / 86400.0;
    return rk4(s, duration, moon_position(jd0), moon_position(jd0 + duration / 2 / 86400.0),
               moon_position(jd0 + duration / 86400.0));
}

}  // namespace kessler
