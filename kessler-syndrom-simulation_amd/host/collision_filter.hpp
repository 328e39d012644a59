// Pair filter shared by the host code and the gfx950 kernels.
//
// Keeps the reference's collision_filter.hpp surface (collision_filter.hpp:
// 26-140): filter_pair, prefilter_keep, cell_key, cell_coords, the forward
// neighbour table and their constants, with the same names and meaning.
// The arithmetic is IEEE double / float in a fixed order; every translation
// unit that includes this file is built with -ffp-contract=off, so host and
// device give bit-identical decisions.

#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define KESSLER_HD __host__ __device__ inline
#else
#define KESSLER_HD inline
#endif

namespace kessler {

// Bound on the relative acceleration of two nearby objects per km of
// separation (s^-2): mu / r^3 is 1.4e-6 at the top of the atmosphere. Over a
// step dt at relative speed v the path leaves the straight line by at most
// GRAVITY_GRADIENT * v * dt^3.
constexpr double GRAVITY_GRADIENT = 2e-6;

enum FilterResult : int {
    FILTER_REJECT = 0,       // cannot come under the threshold in this step
    FILTER_REFINE = 1,       // needs the Hermite closest-approach solve
    FILTER_CO_ORBITING = 2,  // under the threshold but too slow to be an encounter
};

// Exact FP64 decision from the start-of-step relative state (object j minus
// object i): linear closest approach clamped to the step, distance gate,
// co-orbiting test, receding-and-convex rejection.
KESSLER_HD int filter_pair(double rx, double ry, double rz, double vx, double vy, double vz,
                           double dt, double threshold_km, double vmin2) {
    const double dv2 = vx * vx + vy * vy + vz * vz;
    const double rv = rx * vx + ry * vy + rz * vz;
    double ts = 0.0;
    if (dv2 > 0) ts = -rv / dv2;
    ts = ts < 0.0 ? 0.0 : ts;
    ts = ts > dt ? dt : ts;
    const double cx = rx + vx * ts;
    const double cy = ry + vy * ts;
    const double cz = rz + vz * ts;
    const double gate = threshold_km + GRAVITY_GRADIENT * sqrt(dv2) * dt * dt * dt + 1e-3;
    if (sqrt(cx * cx + cy * cy + cz * cz) > gate) return FILTER_REJECT;
    if (dv2 < vmin2) return FILTER_CO_ORBITING;
    // Receding at the start of the step with |p|^2 convex over the step
    // (2|dv|^2 + 2 dr.a > 0 whenever |dv|^2 > G |dr|^2): the minimum is at the
    // start, i.e. it belonged to the previous step.
    const double r2 = rx * rx + ry * ry + rz * rz;
    if (rv >= 0.0 && dv2 > GRAVITY_GRADIENT * r2) return FILTER_REJECT;
    return FILTER_REFINE;
}

// Absolute error bounds of the FP32 inputs of prefilter_keep: positions
// relative to the cell corner rounded to float (km), velocities rounded to
// float (km/s).
constexpr float DR_ERR_KM = 5e-4f;
constexpr float DV_ERR_KM_S = 5e-6f;

// Conservative FP32 screen: returns false only where filter_pair would
// return FILTER_REJECT for the exact inputs, given the error bounds above.
KESSLER_HD bool prefilter_keep(float rx, float ry, float rz, float vx, float vy, float vz,
                               float dt, float threshold_km, float vmin2) {
    const float dv2 = vx * vx + vy * vy + vz * vz;
    const float rv = rx * vx + ry * vy + rz * vz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float dvn = sqrtf(dv2);
    const float rn = sqrtf(r2);
    float ts = 0.0f;
    if (dv2 > 0.0f) ts = -rv / dv2;
    ts = ts < 0.0f ? 0.0f : ts;
    ts = ts > dt ? dt : ts;
    const float cx = rx + vx * ts;
    const float cy = ry + vy * ts;
    const float cz = rz + vz * ts;
    const float d = sqrtf(cx * cx + cy * cy + cz * cz);
    const float gate = threshold_km + static_cast<float>(GRAVITY_GRADIENT) * dvn * dt * dt * dt + 1e-3f;
    const float margin = 1e-5f * (rn + dvn * dt) + DR_ERR_KM + DV_ERR_KM_S * dt;
    if (d > gate + margin) return false;
    const float rv_tol = 1e-5f * rn * dvn + DR_ERR_KM * dvn + DV_ERR_KM_S * rn;
    const bool fast = dv2 > 1.1f * vmin2;
    const bool receding = rv > rv_tol;
    const bool convex = dv2 > 1.01f * static_cast<float>(GRAVITY_GRADIENT) * r2;
    return !(fast && receding && convex);
}

// 3 x 21-bit cell keys, coordinates biased by 2^20.
constexpr long long KEY_OFFSET = 1LL << 20;
constexpr unsigned long long NO_CELL = ~0ULL;

KESSLER_HD unsigned long long cell_key(long long ix, long long iy, long long iz) {
    const unsigned long long ux = static_cast<unsigned long long>(ix + KEY_OFFSET);
    const unsigned long long uy = static_cast<unsigned long long>(iy + KEY_OFFSET);
    const unsigned long long uz = static_cast<unsigned long long>(iz + KEY_OFFSET);
    return (ux << 42) | (uy << 21) | uz;
}

KESSLER_HD void cell_coords(unsigned long long key, long long& ix, long long& iy, long long& iz) {
    constexpr unsigned long long field = (1ULL << 21) - 1;
    ix = static_cast<long long>(key >> 42 & field) - KEY_OFFSET;
    iy = static_cast<long long>(key >> 21 & field) - KEY_OFFSET;
    iz = static_cast<long long>(key & field) - KEY_OFFSET;
}

// The 13 neighbour cells that come after a cell in lexicographic (x, y, z)
// order. Intra-cell pairs plus these visit every adjacent cell pair once.
#define KESSLER_FORWARD_NEIGHBOURS                                                              \
    {{1, -1, -1}, {1, -1, 0}, {1, -1, 1}, {1, 0, -1}, {1, 0, 0}, {1, 0, 1}, {1, 1, -1},         \
     {1, 1, 0},   {1, 1, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1}, {0, 0, 1}}

}  // namespace kessler
