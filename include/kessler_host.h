/* C ABI of the host-side helpers in libkessler_hip.so: synthetic catalogue
 * generation, seeded breakup cascades and catalogue I/O. Pure CPU code (no
 * device call); it builds the inputs the engine (kessler_hip.h) runs on.
 *
 * Reference interfaces replaced (file:line in the Kessler_Sim reference):
 *   khost_load_catalog   load_catalog + read_epoch_jd (orbital.hpp:119-123,
 *                        orbital.cpp:136-188)
 *   khost_write_catalog  the initial-state text format written by
 *                        scripts/spacetrack_export.py:293-381 and read by
 *                        orbital.cpp:150-164
 *   khost_impacts        Fragmentation::impact in sequence (breakup.hpp:128-129,
 *                        breakup.cpp:340-360)
 *   khost_moon_position  moon_position (orbital.hpp:92, orbital.cpp:53-85)
 *   khost_generate       no counterpart: the seeded synthetic catalogue of
 *                        the benchmark configurations
 *   khost_kepler_state   no counterpart: the Kepler-elements -> ECI
 *                        conversion the generator uses (the inverse of the
 *                        reference's elements(), orbital.cpp:120-130)
 */
#ifndef KESSLER_HOST_H
#define KESSLER_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One object (the reference's Object, orbital.hpp:67-84, without strings). */
typedef struct {
    int32_t id;
    int32_t event;  /* breakup that created it, -1 for catalogue objects */
    int32_t alive;
    int32_t type;   /* 0 payload, 1 debris, 2 rocket body, 3 unknown */
    double state[6];
    double size_m;
    double mass_kg;
    double area_to_mass;
} khost_object;

/* Orbit-mix parameters of the synthetic catalogue. */
typedef struct {
    double frac_leo, frac_heo, frac_meo, frac_geo; /* 0.7, 0.1, 0.1, 0.1 */
    double leo_shell_boost; /* > 1 concentrates more LEO objects in the 550 / 800 km shells */
} khost_mix;

/* Writes n synthetic objects (seeded) into out[0..n). Orbits: LEO with an
 * altitude mixture peaked at 550 and 800 km over 300-2000 km, inclinations
 * 53 deg / 97-99 deg SSO / uniform 0-100 deg, exponential eccentricity of
 * mean 0.005; GTO/HEO (perigee 200-600 km, e 0.6-0.75); MEO near 20,200 km
 * at 55 deg; GEO at 35,786 km, i 0-15 deg. RAAN, argument of perigee and
 * mean anomaly uniform. Size classes SMALL / MEDIUM / LARGE at 50/30/20 %.
 * mix may be NULL for the defaults. */
int khost_generate(int64_t n, uint64_t seed, const khost_mix* mix, khost_object* out);

/* ECI state (km, km/s) of the two-body orbit with elements el = {a (km), e,
 * inclination, RAAN, argument of perigee, mean anomaly} (angles in radians,
 * 0 <= e < 1): Kepler's equation by Newton's method, perifocal position and
 * velocity, rotation R3(-RAAN) R1(-i) R3(-argp). The conversion
 * khost_generate applies to every object. */
void khost_kepler_state(const double* el, double* state6);

/* Applies forced impacts in sequence on one Fragmentation object seeded with
 * `seed` (L_c = lc_m). Objects 0..n-1 are the inputs (masses as given);
 * fragments are appended. Returns the total object count; objects are
 * written to out only up to cap (call with cap 0 to size the buffer).
 * events (nimp records, may be NULL) receive each breakup's summary:
 * {fragments, escaped} as two int32 per event. */
int64_t khost_impacts(int64_t n, const khost_object* in, double epoch_jd, double t, int nimp,
                      const int32_t* targets, const double* projectile_mass_kg,
                      const double* v_rel_km_s, uint64_t seed, double lc_m, int64_t cap,
                      khost_object* out, int32_t* events);

/* Seeded cascade catalogue: khost_generate(n_catalog, seed) plus n_events
 * catastrophic forced impacts at t = 0 on distinct LEO objects of the LARGE
 * class, with target and projectile masses uniform in [mass_lo, mass_hi] kg
 * and impact speeds uniform in [vrel_lo, vrel_hi] km/s. Same output protocol
 * as khost_impacts. */
int64_t khost_cascade(int64_t n_catalog, uint64_t seed, int n_events, double lc_m, double mass_lo,
                      double mass_hi, double vrel_lo, double vrel_hi, int64_t cap,
                      khost_object* out);

/* Writes objects in the initial-state text format (17 '|' fields, state at
 * %.17g) with an 18th field, size_m, that the reference's loader ignores. */
int khost_write_catalog(const char* path, int64_t n, const khost_object* objs,
                        const char* epoch_iso);

/* Reads a catalogue file (reference loader semantics; an 18th numeric field,
 * when present, overrides the RCS-class size). Two-call protocol on cap. */
int64_t khost_load_catalog(const char* path, int64_t cap, khost_object* out, double* epoch_jd,
                           int* dropped);

void khost_moon_position(double jd, double* out3);

/* write_snapshot (orbital.hpp:125, orbital.cpp:190-197): the snapshot CSV of
 * objs at time t (s after the epoch); exact != 0 writes 17 significant
 * digits instead of the reference's 9. 0 on success. */
int khost_write_snapshot(const char* path, int64_t n, const khost_object* objs, double t, int exact);

/* read_snapshot (orbital.hpp:133, orbital.cpp:199-249): replays `path` into
 * objs[0..n), in and out. Returns the object count after the replay
 * (fragments appended or dropped; the first `cap` are written back), or -1
 * where the reference's function returns false (unreadable, or not matching
 * the list). *t receives the snapshot time. */
int64_t khost_read_snapshot(const char* path, int64_t n, int64_t cap, khost_object* objs, double* t);

#ifdef __cplusplus
}
#endif

#endif
