// Orbital core of the MI355X build: constants, vector and object types, the
// analytic Moon, the dynamics, catalogue I/O and the Simulation driver.
//
// Keeps the reference's orbital.hpp surface (orbital.hpp:13-163) with the
// same names and meaning. Units: km, s, ECI frame of the input file.
// Simulation::step runs on the GPU through the engine's C ABI
// (include/kessler_hip.h): object states stay resident in HBM between steps
// and are copied back to `objects` only when the host asks (sync_objects).

#pragma once

#include <cmath>
#include <memory>
#include <string>
#include <vector>

namespace kessler {

inline constexpr double PI = 3.14159265358979323846;
inline constexpr double DEG = PI / 180.0;

inline constexpr double MU_EARTH = 398600.4418;  // km^3/s^2
inline constexpr double MU_MOON = 4902.800066;   // km^3/s^2
inline constexpr double R_EARTH = 6378.137;      // km, equatorial
inline constexpr double R_MOON = 1737.4;         // km
inline constexpr double J2 = 1.08262668e-3;

// 2026-08-01 10:00:00 UTC, used when the catalogue header carries no date.
inline constexpr double JD_DEFAULT_EPOCH = 2461253.916666667;

inline const std::string OUT_DIR = "output";

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
};

inline double dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }
inline Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct State {
    Vec3 r, v;
};

enum class ObjectType { Payload = 0, Debris, RocketBody, Unknown, Count };

ObjectType parse_object_type(const std::string& s);
const char* object_type_name(ObjectType t);

// Ids given to fragments created during a run, far above NORAD numbers.
inline constexpr int FRAGMENT_ID_BASE = 10000000;

struct Object {
    int id = 0;
    std::string name, country, rcs;
    ObjectType type = ObjectType::Unknown;
    State s;
    bool alive = true;
    double size_m = 0.8;        // characteristic size L (m); collision radius L/2
    double mass_kg = 0.0;       // 0 = unknown
    double area_to_mass = 0.0;  // m^2/kg
    int event = -1;             // breakup that created it, -1 for catalogue objects
    bool is_fragment() const { return event >= 0; }
};

// Truncated Meeus ch. 47 lunar series, equatorial ECI (orbital.cpp:53-85).
Vec3 moon_position(double jd);

// Earth point mass + J2 + Moon direct minus indirect term (orbital.cpp:91-105).
Vec3 acceleration(const Vec3& r, const Vec3& rm);

// Classical RK4 with the Moon at t, t + dt/2 (k2 and k3) and t + dt
// (orbital.cpp:107-118).
State rk4(const State& s, double dt, const Vec3& rm0, const Vec3& rm1, const Vec3& rm2);

struct Elements {
    double a = 0, e = 0, inc = 0, perigee_alt = 0, apogee_alt = 0, period_min = 0;
};
Elements elements(const State& s);

std::vector<Object> load_catalog(const std::string& path, int* dropped = nullptr);
double read_epoch_jd(const std::string& path);
void write_snapshot(const std::string& path, const std::vector<Object>& objs, double t);
// Full-precision variant (%.17g): exact round trip of every double.
void write_snapshot_exact(const std::string& path, const std::vector<Object>& objs, double t);
// Replays a snapshot into `objs` (orbital.hpp:133); false if unreadable or not
// matching the loaded list. `t` receives the snapshot time (s).
bool read_snapshot(const std::string& path, std::vector<Object>& objs, double* t = nullptr);
std::string jd_to_utc_string(double jd);

class DeviceEngine;

// Propagation state shared by the CLI and the tests. step() runs on the GPU;
// there is no CPU integrator behind it.
struct Simulation {
    std::vector<Object> objects;
    double epoch_jd = JD_DEFAULT_EPOCH;
    double t = 0.0;
    double dt = 10.0;

    // Advances one step on the device. Throws std::runtime_error when no
    // device engine can be created.
    void step();
    void advance(long n);

    // Single RK4 step over `duration` from t_from (orbital.cpp:302-308), on
    // the host: used for fragments born inside a step.
    State propagate_state(const State& s, double t_from, double duration) const;

    double jd() const { return epoch_jd + t / 86400.0; }
    Vec3 moon() const { return moon_position(jd()); }

    // Device plumbing: the engine owns the resident copy of `objects`.
    std::shared_ptr<DeviceEngine> engine;
    bool device_dirty = true;   // host objects changed since the last upload
    bool host_stale = false;    // device advanced since the last download
    void sync_objects();        // device -> host (states and alive flags)
    void ensure_device();       // host -> device when dirty
};

}  // namespace kessler
