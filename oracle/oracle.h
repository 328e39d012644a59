/* CPU oracle of the Kessler_Sim propagate-and-screen loop.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared
 * against. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load liboracle.so. The product (libkessler_hip.so, kessler_sim) never
 * links it and has no CPU fallback through it.
 *
 * C ABI, plain types. States are arrays of 6 doubles per object
 * (x, y, z [km], vx, vy, vz [km/s]), ECI, as the reference's State
 * (orbital.hpp:51). Record layouts match the harness of the compiled
 * reference (oracle/ref_harness.cpp), so that tests can diff the two.
 *
 * Parity pins (see tests/test_oracle.py): every function is checked
 * against golden vectors written by the compiled reference
 * (tests/golden/), and against closed forms / invariants that do not need
 * the reference (two-body conservation, RK4 order 4, brute-force pair
 * enumeration, analytic linear TCA, prefilter conservativeness).
 */
#ifndef KESSLER_ORACLE_H
#define KESSLER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t i, j;      /* object indices, i < j */
    int64_t step;      /* step index within the run */
    double t;          /* TCA, s since epoch */
    double miss_km;
    double v_rel_km_s;
    int32_t collision; /* miss < R_i + R_j */
    int32_t pad;
} orc_event;

typedef struct {
    int64_t candidate_pairs, refined_pairs, co_orbiting;
    double cell_km;
    int32_t max_cell_objects, pad;
} orc_stats;

int orc_configure_threads(int n);

void orc_moon_position(double jd, double* out3);
void orc_acceleration(const double* r3, const double* rm3, double* out3);
void orc_rk4(const double* s6, double dt, const double* m0, const double* m1, const double* m2,
             double* out6);
int orc_filter_pair(const double* r3, const double* v3, double dt, double thr, double vmin2);
int orc_prefilter_keep(const float* r3, const float* v3, float dt, float thr, float vmin2);
void orc_interpolate_state(const double* a6, const double* b6, double dt, double s, double* out6);
double orc_cell_size(int n, const double* before, const uint8_t* alive_after, double dt,
                     double thr);
uint64_t orc_cell_key(int64_t ix, int64_t iy, int64_t iz);

/* One screening step (Screener::finish, CPU path). Returns the event count. */
int64_t orc_screen(int n, const double* before, const double* after, const uint8_t* alive_after,
                   const double* size_m, double t0, double dt, double thr, double radius_scale,
                   double vmin_km_s, orc_event* out, int64_t cap, orc_stats* stats);

/* nsteps of the main.cpp loop (no breakup), states/alive updated in place. */
int64_t orc_run(int n, double* states, uint8_t* alive, const double* size_m, double epoch_jd,
                double t_start, double dt, int64_t nsteps, int screen, double thr,
                double radius_scale, double vmin_km_s, orc_event* out, int64_t cap,
                orc_stats* per_step, double* loop_s);

/* Brute-force O(n^2) candidate enumeration for tiny inputs: every unordered
 * pair of end-of-step-alive objects whose start-of-step cells are equal or
 * adjacent (Chebyshev distance <= 1). Returns the pair count. */
int64_t orc_bruteforce_candidates(int n, const double* before, const uint8_t* alive_after,
                                  double cell);

int orc_load_catalog(const char* path, int cap, int32_t* ids, double* states, double* size_m,
                     int32_t* rcs_code, int32_t* type_code, double* epoch_jd, int* dropped);

#ifdef __cplusplus
}
#endif

#endif
