/* C ABI of the gfx950 propagate-and-screen engine (libkessler_hip.so).
 *
 * One engine context per host thread; one HIP stream per context; not
 * re-entrant. Every function returns 0 on success or a negative status; the
 * message of the last failure is kept in the context (khip_last_error) or,
 * for khip_create, written to the caller's buffer. Nothing throws across
 * this boundary.
 *
 * Object state is device-resident between calls: the host uploads it once
 * (khip_set_objects), runs any number of steps (khip_run), and reads back
 * states (khip_get_objects), events (khip_take_events) and per-step
 * statistics (khip_take_stats) only when it needs them.
 *
 * Which reference interface each entry point replaces (file:line in the
 * Kessler_Sim reference):
 *
 *   khip_create / khip_destroy     kessler::gpu::create / destroy
 *                                  (collision_gpu.hpp:47-48) and
 *                                  Screener::set_gpu (collision.hpp:93)
 *   khip_set_objects               gpu::staging + the per-step H2D upload
 *                                  (collision_gpu.hpp:53, collision.cpp:574-583);
 *                                  here once per run, not once per step
 *   khip_set_config                ScreeningConfig (collision.hpp:26-37)
 *   khip_set_time                  Simulation::{epoch_jd, t, dt} (orbital.hpp:145-148)
 *   khip_run                       the main.cpp step loop body (main.cpp:352-373):
 *                                  Simulation::step (orbital.cpp:286-300) and
 *                                  Screener::begin/finish (collision.cpp:706-750)
 *   khip_screen                    Screener::screen / screen_step on host-held
 *                                  before/after states (collision.hpp:100-102, 124-128)
 *   khip_take_events               gpu::collect (collision_gpu.hpp:62) + the
 *                                  exact filter/refine and (t, i, j) sort of
 *                                  collision.cpp:622-647, 739-747
 *   khip_take_stats                ScreeningStats (collision.hpp:47-62)
 *   khip_get_objects               Simulation::objects[i].{s, alive} read-back
 *                                  (orbital.hpp:145, 67-84)
 */
#ifndef KESSLER_HIP_H
#define KESSLER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KHIP_OK 0
#define KHIP_ERR_NO_DEVICE (-1)
#define KHIP_ERR_HIP (-2)
#define KHIP_ERR_ARG (-3)
#define KHIP_ERR_STATE (-4)

typedef struct khip_ctx khip_ctx;

/* ScreeningConfig (collision.hpp:26-37). */
typedef struct {
    double threshold_km;             /* default 5 */
    double radius_scale;             /* default 1 */
    double min_encounter_speed_km_s; /* default 0.01 */
} khip_screen_config;

/* Conjunction (collision.hpp:39-45) plus the step index it was found in. */
typedef struct {
    int32_t i, j;      /* object indices, i < j */
    int64_t step;      /* global step index since khip_set_time */
    double t;          /* TCA, s since epoch */
    double miss_km;
    double v_rel_km_s;
    int32_t collision; /* miss_km < R_i + R_j */
    int32_t pad;
} khip_event;

/* The ScreeningStats fields that the device computes (collision.hpp:47-62). */
typedef struct {
    int64_t candidate_pairs;
    int64_t refined_pairs;
    int64_t co_orbiting;
    double cell_km;
    int32_t max_cell_objects;
    int32_t cell_doublings; /* > 0 only where the grid extent needed a coarser cell (see khip_run) */
    int64_t propagated;     /* objects alive at the start of the step (0 for khip_screen) */
    int64_t dense_candidates; /* part of candidate_pairs screened by dense-cell work items */
    int64_t survivors;      /* candidate pairs the FP32 pre-filter kept for the FP64 filter
                               (no reference counterpart: the reference filters in FP64 only) */
} khip_stats;

/* Device time per kernel family, summed over the steps since the last
 * khip_reset_timing, measured with HIP events on the engine's stream when
 * timing is enabled (khip_set_timing). */
typedef struct {
    double propagate_ms;    /* k_propagate: fused RK4 step */
    double grid_ms;         /* keys, radix sort, cell table, neighbour ranges, work items */
    double pairs_ms;        /* the three kernels below */
    double pairs_sparse_ms; /* k_pairs_lanes: objects of sparse cells, FP32 pre-filter */
    double pairs_dense_ms;  /* k_pairs_dense: dense-cell work items, FP32 pre-filter */
    double exact_ms;        /* k_exact: FP64 filter, TCA, radius test, event append */
    double total_ms;        /* whole steps, launch gaps included */
    int64_t steps;
    int64_t reruns;         /* chunks run again after a device buffer grew */
    int64_t rerun_reasons;  /* OR of KHIP_RERUN_* over those re-runs */
    double reverify_ms;     /* learned form: k_reverify, the plain re-solve of events */
} khip_timing;

#define KHIP_RERUN_PASSES 1    /* the cell-key extent needed more radix passes */
#define KHIP_RERUN_EVENTS 2    /* event buffer */
#define KHIP_RERUN_ITEMS 4     /* dense-cell work items */
#define KHIP_RERUN_SURVIVORS 8 /* pre-filter survivor queue */
#define KHIP_RERUN_BREAKUPS 16 /* breakup jobs, fragment samples or reserved object slots */
#define KHIP_RERUN_ENTRIES 32  /* radial-shell entries of the learned form's screening */

/* Number of visible HIP devices (0 when none or the runtime fails). */
int khip_device_count(void);

/* Creates a context on `device`. On failure *out is NULL and `msg` (if
 * given) receives the reason; on success `msg` receives the device name. */
int khip_create(int device, khip_ctx** out, char* msg, size_t msg_cap);
void khip_destroy(khip_ctx* ctx);
const char* khip_last_error(const khip_ctx* ctx);

/* Uploads n objects: states is n x 6 doubles (x, y, z km; vx, vy, vz km/s),
 * alive n bytes (NULL = all alive), size_m n doubles (characteristic size L,
 * collision radius L/2). Replaces any previous population. */
int khip_set_objects(khip_ctx* ctx, int64_t n, const double* states, const uint8_t* alive,
                     const double* size_m);
int64_t khip_object_count(const khip_ctx* ctx);
/* Splits the uploaded population into nscen independent scenarios: objects
 * [offsets[k], offsets[k + 1]) form scenario k (offsets[0] = 0,
 * offsets[nscen] = n). Screening pairs objects of the same scenario only;
 * everything else runs over the concatenated arrays in the same launches.
 * khip_set_objects resets to one scenario. No reference counterpart (the
 * reference runs one population per process). */
int khip_set_scenarios(khip_ctx* ctx, int32_t nscen, const int64_t* offsets);
/* Reads back states (n x 6) and alive flags; either pointer may be NULL. */
int khip_get_objects(khip_ctx* ctx, double* states, uint8_t* alive);

/* Start-of-step states (n x 6) of the last step run, for the objects listed
 * in idx: the `before` states that Fragmentation::apply interpolates
 * (breakup.cpp:326-331, main.cpp:352-358), without a per-step host copy. */
int khip_get_previous(khip_ctx* ctx, int64_t n, const int32_t* idx, double* states);

int khip_set_config(khip_ctx* ctx, const khip_screen_config* cfg);
/* Epoch (Julian day), elapsed time t (s since epoch) and step dt (s).
 * Resets the global step counter to 0. */
int khip_set_time(khip_ctx* ctx, double epoch_jd, double t, double dt);
int khip_get_time(const khip_ctx* ctx, double* t, int64_t* step);

/* Runs nsteps steps of the main loop: FP64 RK4 of every alive object
 * (Earth + J2 + Moon) and, if `screen`, the screening of each step
 * (grid on start-of-step positions of the objects alive at the end of the
 * step, FP32 pre-filter, FP64 filter_pair, FP64 Hermite/golden-section TCA,
 * radius test). Events and per-step statistics accumulate on the host side
 * of the context until taken. */
int khip_run(khip_ctx* ctx, int64_t nsteps, int screen);

/* One screening of host-held states: `before` (start of step) and `after`
 * (end of step) are n x 6 doubles, alive_after n bytes, size_m n doubles.
 * Uses (and overwrites) the context's device population. Events and stats
 * are returned directly; *nevents receives the full count even beyond cap. */
int khip_screen(khip_ctx* ctx, int64_t n, const double* before, const double* after,
                const uint8_t* alive_after, const double* size_m, double t0, double dt,
                khip_event* events, int64_t cap, int64_t* nevents, khip_stats* stats);

/* khip_screen without device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Every
 * device buffer it uses (staged inputs, grid, pre-filter survivors, events)
 * is a region of one fixed array of about 27 MB that each GPU holds from the
 * library's load to the end of the process; the call makes no hipMalloc or
 * hipFree. Limits per call: n <= 32,768 objects (else KHIP_ERR_ARG), at most
 * 131,072 events, 1,048,576 pre-filter survivors and 65,536 dense-cell work
 * items (else KHIP_ERR_STATE, with a message). `events` must hold every
 * event: when cap is smaller, *nevents receives the count and nothing is
 * copied (KHIP_ERR_ARG). The context's own population is left untouched.
 * Calls on one GPU share its array, whatever the context: calls that
 * overlap on one GPU race. Events and stats equal khip_screen's. */
int khip_screen_static(khip_ctx* ctx, int64_t n, const double* before, const double* after,
                       const uint8_t* alive_after, const double* size_m, double t0, double dt,
                       khip_event* events, int64_t cap, int64_t* nevents, khip_stats* stats);

/* Moves up to cap accumulated events (sorted by step, t, i, j) to `out`;
 * *n receives the number moved. Call with out == NULL to query the count. */
int khip_take_events(khip_ctx* ctx, khip_event* out, int64_t cap, int64_t* n);
/* Same for per-step statistics, one record per screened step. */
int khip_take_stats(khip_ctx* ctx, khip_stats* out, int64_t cap, int64_t* n);

int khip_set_timing(khip_ctx* ctx, int enabled);
int khip_get_timing(const khip_ctx* ctx, khip_timing* out);
int khip_reset_timing(khip_ctx* ctx);

/* Propagation arithmetic. KHIP_PRECISION_FP64 (default) is the reference's
 * FP64 RK4, bit for bit. KHIP_PRECISION_MIXED keeps the FP64 state and
 * central term and evaluates J2 and the Moon in FP32 (the precision form);
 * screening is unchanged (FP32 pre-filter, FP64 filter, TCA and radius
 * test). No reference counterpart: the reference has one arithmetic. */
#define KHIP_PRECISION_FP64 0
#define KHIP_PRECISION_MIXED 1
/* KHIP_PRECISION_LEARNED: steps of tau (khip_set_time's dt, which must equal
 * the corrector's tau). Each object steps by one of three methods, chosen
 * at check steps: one mixed RK4 step of tau plus the corrector (near-circular
 * orbits below a_max_km), one mixed RK4 step of tau alone (near-circular
 * orbits above a_max_km), or `substeps` FP64 RK4 steps of tau / substeps
 * (the plain method; everything else). A check step (every check_every
 * steps, and the first step after khip_set_objects / khip_set_precision /
 * khip_set_learned) runs the plain method for every object and keeps the
 * cheaper method for an object only if it stayed within tol_km of the plain
 * result in position and within tol_a_km in semi-major axis. Screening
 * covers each window of tau / windows on the cubic Hermite dense output of
 * the step, with a grid on midpoints and the threshold raised by margin_km.
 * With shell_km > 0 each object is also binned by the radius range its
 * cubic covers in the window (fixed radial shells, shell_km wide below
 * 2,000 km altitude, coarser above; an object enters every shell its range
 * widened by half the threshold meets), and a pair is enumerated only
 * inside a shared shell, once: two objects closer than the threshold have
 * radii closer than the threshold. Every event found is then solved again on
 * plain FP64 trajectories from the step's start (filter_pair, Hermite +
 * golden section, radius test, in the sub-step of tau / substeps holding
 * the TCA) at the configured threshold plus the drift band, and kept with
 * the plain values when the plain solve accepts it; an event whose plain
 * minimum leaves the learned step keeps its learned values if its learned
 * miss distance is within that (events_unverified). The drift band
 * (margin_factor) covers the distance the learned trajectories have drifted
 * from plain ones since the run started: with it, every event of a plain run
 * from the same states is reported, together with events whose miss lies
 * within the band above the threshold (miss_km > threshold_km marks them;
 * the plain run's miss for those may lie on either side). Likewise the
 * velocity band lowers the minimum encounter speed (relative speeds drift
 * too). The search threshold is threshold + margin_km + band. Event step
 * indices count
 * sub-steps, as a plain run at dt = tau / substeps does; statistics have one
 * record per sub-step, a window's counts in the record of its first
 * sub-step and zero counts in the others. */
#define KHIP_PRECISION_LEARNED 2
int khip_set_precision(khip_ctx* ctx, int mode);

typedef struct {
    int32_t substeps;    /* plain steps per learned step (divides 12; default 6) */
    int32_t check_every; /* learned steps between checks (default 10) */
    double tol_km;       /* deviation from the plain step tolerated at a check (default 1e-5) */
    double margin_km;    /* screening threshold raise (default 0.025) */
    double e_max;        /* corrector / base-step domain: eccentricity below (default 0.1) */
    double a_max_km;     /* corrector domain: semi-major axis below (default 12000) */
    double tol_a_km;     /* semi-major-axis deviation tolerated at a check (default 5e-8: a
                            systematic 5e-8 km per 60 s step drifts ~5 m along track in 24 h) */
    int32_t windows;     /* screening windows per learned step (divides substeps; 0: two with
                            shells when substeps is even, else substeps / 2 when even, else
                            substeps). Fewer, longer windows mean fewer grid builds and more
                            candidate pairs per build. */
    double margin_factor; /* drift band from the measured error (> 0; 0: no band): every chunk the
                             band is margin_factor x D x ((T + L) / T)^2, with D the largest
                             distance between learned and plain trajectories so far over a sample
                             of up to 8,192 objects (evenly spaced) that the plain method follows
                             alongside from the last khip_set_objects / khip_set_time /
                             khip_set_precision change / khip_set_learned, T the time since then
                             and L the chunk's span (along-track drift grows quadratically in
                             time), rounded up to a ladder of 2^(1/4) steps. Default 4: both
                             objects of a pair drift, and the sample's maximum stands for the
                             population's. */
    double shell_km;     /* radial shell width of the screening (>= 1 km, with at most 4 windows):
                            objects are binned by their radius range over the window as well
                            as by cell, and pairs are enumerated within shells only (see
                            KHIP_PRECISION_LEARNED); 0 screens on cells alone. Default 16. */
} khip_learned_config;

typedef struct {
    int64_t learned, base, plain; /* objects per method after the last check */
    int64_t rejected;             /* domain candidates sent to the plain method, all checks */
    double dev_max_km;            /* largest deviation accepted at a check */
    int64_t checks, steps;
    int64_t events_verified;      /* events confirmed by the plain solve */
    int64_t events_dropped;       /* events the plain solve rejected (not reported) */
    int64_t events_unverified;    /* plain minimum outside the learned step, learned miss within the
                                     threshold (learned values kept) */
    double sample_dev_max_km;     /* D: largest learned-vs-plain distance over the followed sample */
    double sample_dev_v_max_km_s; /* the same in velocity */
    double band_km;               /* the drift band of the last chunk */
    double band_v_km_s;           /* its velocity band: relative speeds down to
                                     min_encounter_speed_km_s - band_v_km_s are kept */
    int64_t sample_objects;       /* size of the followed sample */
} khip_learned_report;

int khip_set_learned(khip_ctx* ctx, const khip_learned_config* cfg);
/* When enabled (and timing is off), each screened chunk of steps is
 * captured once into a HIP graph and replayed for every later chunk of the
 * same shape (length, check pattern, radix passes, buffers); per-step times
 * and indices come from a device table, so replays need no re-capture. All
 * forms, in-run breakups included (a chunk after the population grew is
 * captured again). Results are identical to direct launches. No reference
 * counterpart. */
int khip_set_graph(khip_ctx* ctx, int enabled);

/* ---- In-run breakups (cascade mode) ---------------------------------------
 *
 * Replaces Fragmentation::apply after each step's screening (main.cpp:408-415,
 * breakup.cpp:316-338) on the device: collisions of the step in (t, i, j)
 * order, parents alive -> the NASA Standard Breakup Model (target = heavier
 * parent; specific energy against the catastrophic threshold; mass budgets;
 * N = Poisson(0.1 M^0.75 L_c^-1.71), at most 200,000; sizes, area-to-mass,
 * log-normal isotropic delta-v; hyperbolic fragments dropped; RK4 to the end
 * of the step). New objects are appended after the population (indices
 * n, n + 1, ...) without a host round trip. The draws come from a
 * counter-based generator, not the reference's mt19937_64 sequence: the
 * fragments follow the model's distributions, not the reference's exact
 * draws (the host restatement in kessler_host.h keeps those). All three
 * forms: in the learned form the collisions are those the plain
 * re-verification confirms, the parents come from the 60 s step's cubic,
 * and each fragment starts on the plain method until the next check. */
typedef struct {
    int32_t enabled;
    int32_t reserve;             /* object slots reserved per chunk (grown when short) */
    double lc_m;                 /* smallest fragment (default 0.1 m) */
    double catastrophic_j_per_g; /* threshold (default 40 J/g) */
    uint64_t seed;
} khip_breakup_config;

/* BreakupEvent (breakup.hpp:86-104) of one breakup. */
typedef struct {
    double t, v_rel_km_s, energy_j_per_g, target_mass_kg, projectile_mass_kg, sbm_mass_kg;
    double expected_fragments, fragments_mass_kg;
    double position[3];          /* target at the collision (km) */
    int32_t id, target, projectile, catastrophic;
    int32_t drawn;               /* fragments drawn from the Poisson law */
    int32_t fragments;           /* fragments created (objects first_index ..) */
    int32_t escaped;             /* hyperbolic fragments dropped */
    int32_t first_index;
} khip_breakup_event;

int khip_set_breakup(khip_ctx* ctx, const khip_breakup_config* cfg);
/* Object masses (kg, n = khip_object_count; 0 = the model's 100 kg fallback),
 * Object::mass_kg (orbital.hpp:67-84). Damaged targets lose mass. */
int khip_set_masses(khip_ctx* ctx, int64_t n, const double* mass_kg);
int khip_get_masses(khip_ctx* ctx, int64_t n, double* mass_kg);
/* Characteristic sizes L (m), Object::size_m, of the n current objects. */
int khip_get_sizes(const khip_ctx* ctx, int64_t n, double* size_m);
/* Moves up to cap breakup records (in order) to out; out == NULL queries. */
int khip_take_breakups(khip_ctx* ctx, khip_breakup_event* out, int64_t cap, int64_t* n);

/* Launch and tiling parameters of the screening kernels. None changes which
 * conjunctions are found; cell_mult > 1 enlarges the grid cell beyond the
 * reference's (larger cells never change results, collision.cpp:337-338)
 * and so changes the candidate counts in the statistics. */
typedef struct {
    double cell_mult;     /* >= 1 (default 1: the reference's cell) */
    int32_t dense_min;    /* cells with at least this many objects use work items (128) */
    int32_t long_range;   /* sparse-cell ranges longer than this use work items (256; 1 to 16384) */
    int32_t seg;          /* partners per work item (256) */
    int32_t dense_blocks; /* persistent grid of the dense-cell pair kernel (1024, <= 4096) */
    int32_t exact_blocks; /* grid of the exact FP64 kernel (512, <= 2048) */
    int32_t sparse_kernel; /* sparse-cell pair kernel: 0 (default) one partner per lane against the
                              cell's own objects in LDS (catalogues); 1 one own object per lane
                              walking its partners (the crowded cells of fragment clouds) */
} khip_tuning;
int khip_set_tuning(khip_ctx* ctx, const khip_tuning* t);
int khip_get_tuning(const khip_ctx* ctx, khip_tuning* t);
/* Autotuner: times `steps` screened steps of the current population and
 * configuration for candidate tunings (coordinate descent over the fields;
 * cell_mult only if allow_cell), each from the same saved state, keeps the
 * fastest and restores the population, time, events and statistics. */
int khip_autotune(khip_ctx* ctx, int64_t steps, int allow_cell, khip_tuning* best, double* base_ms,
                  double* best_ms);
int khip_get_graph_counts(const khip_ctx* ctx, int64_t* captures, int64_t* launches);
/* Tallies since the last khip_set_objects. No reference counterpart. */
int khip_get_learned_report(const khip_ctx* ctx, khip_learned_report* out);

/* ---- Learned step corrector (no reference counterpart) ----------------
 *
 * A 10-128-128-128-6 perceptron (SiLU, BF16 weights on the matrix cores,
 * FP32 accumulation) that predicts, from ten inputs describing the state in
 * its local orbital frame, the difference between six FP64 RK4 steps of
 * tau/6 and one mixed-precision RK4 step of tau. Trained on the device by
 * the library's own forward and backward kernels (Adam on FP32 master
 * parameters). Parameter layout: W1[128][10], b1[128], W2[128][128],
 * b2[128], W3[128][128], b3[128], W4[6][128], b4[6] (row-major [out][in]). */
#define KHIP_NET_PARAMS 35206
#define KHIP_NET_INPUTS 10
#define KHIP_NET_OUTPUTS 6

/* Standardisation: input k = (feature_k - fmu[k]) / fsig[k]; frame
 * component k of the correction = (output_k * osig[k] + omu[k]) * scale. */
typedef struct {
    float fmu[10], fsig[10];
    float omu[6], osig[6];
    double tau; /* step (s) the corrector is trained for */
} khip_net_norm;

typedef struct {
    int32_t samples_per_orbit; /* samples emitted along each orbit */
    int32_t spacing_steps;     /* tau-steps between two samples of one orbit */
    int32_t iterations;        /* Adam steps */
    int32_t batch;             /* columns per step, a multiple of 1024 */
    double lr_max;             /* peak learning rate (5% warm-up, cosine decay) */
    double e_max;              /* learned domain: osculating eccentricity below e_max, */
    double a_max_km;           /* semi-major axis below a_max_km */
    double tau;                /* step the corrector is trained for (s) */
    uint64_t seed;             /* parameter initialisation and batch sampling */
} khip_train_config;

typedef struct {
    int64_t samples;   /* valid training samples */
    double loss_first; /* mean loss of the first 100 iterations */
    double loss_last;  /* mean loss of the last 100 iterations */
    double gen_ms;     /* device time of the sample generation */
    double train_ms;   /* device time of the training loop */
} khip_train_report;

/* Generates the training set on the device from n seeded orbits (n x 6
 * states at the context's epoch and time), then trains the corrector.
 * curve (may be NULL) receives the mean loss of every iteration. */
int khip_net_train(khip_ctx* ctx, int64_t n, const double* orbits, const khip_train_config* cfg,
                   khip_train_report* report, float* curve);
/* Loads / reads the parameters (KHIP_NET_PARAMS floats) and standardisation. */
int khip_net_set(khip_ctx* ctx, const float* params, const khip_net_norm* norm);
int khip_net_get(khip_ctx* ctx, float* params, khip_net_norm* norm);
/* Network outputs (n x 6) for standardised inputs (n x 10). */
int khip_net_infer(khip_ctx* ctx, int64_t n, const float* x, float* out);
/* Gradient of the training loss mean_b (1/6) sum_k (out_bk - y_bk)^2 over
 * the batch (n a multiple of 1024) with respect to the parameters, and the
 * loss itself, as the training step computes them. */
int khip_net_gradient(khip_ctx* ctx, int64_t n, const float* x, const float* y, float* grad, float* loss);

/* Steps per host synchronisation in khip_run (default 256). */
int khip_set_chunk(khip_ctx* ctx, int steps);

/* FP64 arithmetic self-check: out[k] = {sqrt(a[k]), a[k] / b[k],
 * floor(a[k] / b[k])} computed on the device, for comparison with the host's
 * IEEE results. n <= 1 << 20. */
int khip_fp64_selfcheck(khip_ctx* ctx, int64_t n, const double* a, const double* b, double* out);

#ifdef __cplusplus
}
#endif

#endif
