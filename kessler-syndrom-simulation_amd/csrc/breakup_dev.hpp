// In-run breakup of colliding objects on the device (NASA Standard Breakup
// Model, Johnson et al., Adv. Space Res. 28(9), 2001), following the model
// of breakup.cpp:121-338: specific energy and the catastrophic threshold,
// mass budgets, N = Poisson(0.1 M^0.75 L_c^-1.71) capped at 200,000, sizes by
// inverting the truncated power law, the bi-normal / normal area-to-mass
// distribution, log-normal isotropic delta-v capped at the impact speed,
// hyperbolic fragments dropped, fragments stepped by RK4 to the end of the
// step.
//
// What differs from the host restatement, on purpose: the draws come from a
// counter-based generator (splitmix64 of seed, event and fragment indices)
// instead of one mt19937_64 sequence, so every fragment is sampled by its
// own lane; the Moon for the remaining part of the step is the quadratic
// through the step's three samples. Results are therefore statistically,
// not bitwise, the host model's (tests check the distributions against the
// model's closed forms).
//
// Per step, after the exact pair kernel has appended the step's events:
//   k_bu_events   one block: the step's collisions in (t, i, j) order;
//                 parents alive -> target / projectile by mass, energy,
//                 budgets, parents killed or damaged, fragment count; one
//                 job per collision and one log record
//   k_bu_sample   one lane per drawn fragment: side, size, area-to-mass,
//                 mass, delta-v vector
//   k_bu_budget   one lane per job, in fragment order: the reference's
//                 sequential mass rule (a fragment that would exceed its
//                 parent's budget is skipped), hyperbolic check, ranks
//   k_bu_bases    one lane: job offsets in the reserved slots
//   k_bu_write    one lane per kept fragment: birth state, RK4 to the end of
//                 the step, object arrays
// New objects go to slots [n, n + reserve) that were launched dead; the
// chunk is re-run with more room if any buffer is short.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace kessler {
namespace bu {

using dev::D3;
using dev::MoonSample;

constexpr int MAX_JOBS = 512;             // collisions handled per step
constexpr int MAX_FRAGS_PER_EVENT = 200000;
constexpr double FALLBACK_MASS_KG = 100.0;
constexpr double SLOPE = 1.71;
constexpr double PI_D = 3.14159265358979323846;

struct BreakupCfg {
    double lc_m;
    double catastrophic_j_per_g;
    double radius_scale;
    unsigned long long seed;
    int reserve;          // slots available past the live population
    int sample_cap;       // fragment samples per step
};

// Device-side counters of one chunk (reset per chunk).
struct BreakupState {
    int njobs, nsamples;
    int frag_used;        // reserved slots taken in this chunk
    int jobs_short, samples_short, reserve_short;
    int next_event;       // breakup event ids (persist across chunks)
    int nlog;             // log records of this chunk
    int pad;
};

struct Job {
    double t, remaining, vrel, p_target;
    D3 r[2], v[2];
    double size[2], budget[2];
    int count, event, sample_base, scen;
    int kept, escaped, base, log;
    double used_mass;
};

struct Sample {
    D3 dv;
    double L, chi, m;
    int side;  // 0 target, 1 projectile, -1 no fragment
    int rank;  // slot rank within the job, -1 not kept
};

// One breakup record (the BreakupEvent of breakup.hpp).
struct LogRecord {
    double t, v_rel_km_s, energy_j_per_g, target_mass_kg, projectile_mass_kg, sbm_mass_kg;
    double expected_fragments, fragments_mass_kg;
    double px, py, pz;
    int id, target, projectile, catastrophic;
    int drawn, fragments, escaped;
    int first;  // first fragment slot, relative to the chunk's reserved room
};

__host__ __device__ inline uint64_t mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Counter-based stream: draw c of stream (a, b).
struct Rng {
    uint64_t key;
    uint64_t c;
    __device__ Rng(uint64_t seed, uint64_t a, uint64_t b) : key(mix(seed ^ mix(a * 0x100000001b3ULL + b))), c(0) {}
    __device__ double uniform() { return static_cast<double>(mix(key + (c++)) >> 11) * 0x1.0p-53; }
    __device__ double normal() {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * PI_D * u2);
    }
    // Knuth's product method below 10, Hormann's transformed rejection
    // (PTRS, 1993) above: the algorithms of the host restatement.
    __device__ long long poisson(double mean) {
        if (mean <= 0.0) return 0;
        if (mean < 10.0) {
            const double limit = exp(-mean);
            long long k = 0;
            double prod = uniform();
            while (prod > limit) {
                ++k;
                prod *= uniform();
            }
            return k;
        }
        const double slam = sqrt(mean), loglam = log(mean);
        const double b = 0.931 + 2.53 * slam;
        const double a = -0.059 + 0.02483 * b;
        const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
        const double vr = 0.9277 - 3.6224 / (b - 2.0);
        for (int guard = 0; guard < 1000; ++guard) {
            const double u = uniform() - 0.5;
            const double v = uniform();
            const double us = 0.5 - fabs(u);
            if (us <= 0.0 || v <= 0.0) continue;
            const double k = floor((2.0 * a / us + b) * u + mean + 0.43);
            if (us >= 0.07 && v <= vr) return static_cast<long long>(k);
            if (k < 0.0 || (us < 0.013 && v > us)) continue;
            if (log(v) + log(inv_alpha) - log(a / (us * us) + b) <= -mean + k * loglam - lgamma(k + 1.0))
                return static_cast<long long>(k);
        }
        return static_cast<long long>(mean + 0.5);
    }
};

__device__ __forceinline__ double ramp(double x, double x0, double y0, double x1, double y1) {
    if (x <= x0) return y0;
    if (x >= x1) return y1;
    return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
}

// log10(A/m) of a fragment of size 10^log_l.
__device__ __forceinline__ double area_to_mass_log(Rng& g, double log_l) {
    const double l = pow(10.0, log_l);
    auto bimodal = [&]() {
        const double alpha = ramp(log_l, -1.95, 0.0, 0.55, 1.0);
        const double mu1 = ramp(log_l, -1.1, -0.6, 0.0, -0.95);
        const double s1 = ramp(log_l, -1.3, 0.1, -0.3, 0.3);
        const double mu2 = ramp(log_l, -0.7, -1.2, -0.1, -2.0);
        const double s2 = ramp(log_l, -0.5, 0.5, -0.3, 0.3);
        if (g.uniform() < alpha) return mu1 + s1 * g.normal();
        return mu2 + s2 * g.normal();
    };
    auto single = [&]() {
        const double mu = ramp(log_l, -1.75, -0.3, -1.25, -1.0);
        const double s = log_l <= -3.5 ? 0.2 : 0.2 + 0.1333 * (log_l + 3.5);
        return mu + s * g.normal();
    };
    if (l >= 0.11) return bimodal();
    if (l <= 0.08) return single();
    return g.uniform() < (l - 0.08) / 0.03 ? bimodal() : single();
}

// Quadratic through the Moon samples at 0, dt/2 and dt of the step.
__device__ __forceinline__ MoonSample moon_at(const MoonSample* m3, double f) {
    const double l0 = 2.0 * (f - 0.5) * (f - 1.0), l1 = -4.0 * f * (f - 1.0), l2 = 2.0 * f * (f - 0.5);
    MoonSample o;
    for (int k = 0; k < 3; ++k) {
        o.m[k] = l0 * m3[0].m[k] + l1 * m3[1].m[k] + l2 * m3[2].m[k];
        o.ind[k] = l0 * m3[0].ind[k] + l1 * m3[1].ind[k] + l2 * m3[2].ind[k];
    }
    return o;
}

}  // namespace bu
}  // namespace kessler
