// gfx950 propagate-and-screen engine and its C ABI (include/kessler_hip.h).
//
// Device-resident design. Object state lives in HBM as structure-of-arrays
// FP64 (x, y, z, vx, vy, vz planes) in two buffers that swap every step, so
// one step reads the start-of-step buffer and writes the end-of-step one;
// the screening of that step then has both at hand with no host copy.
//
// One plain step is a fixed sequence of kernels on one stream:
//
//   k_propagate   fused four-stage RK4, one object per lane, coalesced SoA
//                 loads/stores; the Moon samples of the step are uniform
//                 (scalar) loads; per-block partials (max |v|, bounding box
//                 of the start-of-step positions, alive counts) for the
//                 next single-block kernel, no shared-address atomics.
//   k_grid_setup  combines the partials: cell size (collision.cpp:339-356)
//                 and a packed cell key (per-axis offsets from the bounding
//                 box, as few bits as it needs: 27 at 1M objects, so the
//                 radix sort runs three 9-bit passes); also writes the
//                 previous step's (or window's) statistics record and resets
//                 its accumulators.
//   k_keys        packed cell key per object (dead objects get a sentinel
//                 that sorts last), the object's compact FP32 record, and
//                 the first pass's digit histogram.
//   k_hist / k_digit_scan / k_scatter
//                 stable LSD radix sort of (key, index): per-tile digit
//                 histograms, per-digit scans, then a scatter whose in-tile
//                 ranks come from wave64 ballots.
//   k_cell_count / k_cell_write
//                 cell boundaries from wave ballots of the head flags; the
//                 cell table, the column table and the records gathered in
//                 sorted order.
//   k_neighbours  per cell: its five neighbour runs (the 13 forward cells of
//                 collision_filter.hpp:137-140 as z-runs of four columns plus
//                 the cell's own tail), found through the column table;
//                 dense cells registered for work items.
//   k_items       work items of dense cells and long runs.
//   k_pairs_lanes / k_pairs_dense
//                 the FP32 pre-filter on every candidate pair (sparse cells:
//                 a lane per partner against the own cell's records in LDS;
//                 dense cells: a wave per item); survivors go to a queue.
//   k_exact       the queue: FP64 filter_pair, Hermite + golden-section TCA,
//                 radius test; events appended with wave-aggregated atomics.
//   k_finish      the statistics record of a chunk's last step (earlier
//                 steps' records are written by the next k_grid_setup).
//
// The learned form (KHIP_PRECISION_LEARNED) replaces k_propagate by the
// method kernels over per-method lists (k_fp64_methods: the plain sub-steps
// and the mixed 60 s base steps at full occupancy; k_net_step: the corrector
// on MFMA; at checks k_check_plain for everyone, then k_check_base,
// k_net_step<true> on the corrector's domain and k_check_verdict), follows a
// sample of objects on plain twins beside the screening (k_sample_step, side
// stream), screens each step in windows on the step's cubic dense output
// (k_windows: window boundary states, reductions and, with radial shells,
// each object's shells; k_keys_shell: one sorted element per (object,
// shell)), window q on a second stream with its own grid buffers when q is
// odd, and solves every event again on plain FP64 trajectories (k_reverify,
// each stream its own events and its last window's statistics record,
// k_merge_events joining them). In-run breakups (k_bu_*) follow a step's
// screening in every form.
//
// The host only enqueues; it synchronises once per chunk of steps to drain
// events (into pinned memory) and statistics, and orders a chunk's events
// while the next chunk runs.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kessler_hip.h"
#include "../host/orbital.hpp"
#include "device_common.hpp"
#include "mlp.hpp"
#include "breakup_dev.hpp"

using namespace kessler;
using namespace kessler::dev;

namespace {

constexpr int ITEMS = 8;               // elements per thread per tile (BLOCK threads)
constexpr int TILE = BLOCK * ITEMS;    // 2048 elements per object tile
// The tile kernels of the grid build (keys, radix passes, cell table) take a
// tile with TBLOCK threads, TI elements each (a template argument: one to
// three, chosen per screening so that the tiles fit one round of blocks;
// TITEMS, two, at 1M objects): 16 waves per tile keep enough loads in
// flight with a few hundred tiles on 256 CUs.
constexpr int TBLOCK = 1024;
constexpr int TITEMS = TILE / TBLOCK;
constexpr int TWAVES = TBLOCK / 64;
constexpr int COL_BITS = 24;  // column table index width (see k_cell_write)
// Shell screening (learned form): a sorted element is one (object, radial
// shell) entry; its index word holds the object index in the low OBJ_BITS
// and NOT_FIRST for every entry but the object's innermost shell.
constexpr int OBJ_BITS = 30;
constexpr int OBJ_MASK = (1 << OBJ_BITS) - 1;
constexpr int NOT_FIRST = 1 << OBJ_BITS;
// LSD radix sort of the packed cell keys, 9-bit digits: at 1M objects the
// key has 27 bits (about 400 cells of 200 km per axis), which three passes
// cover.
constexpr int RADIX_BITS = 9;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr uint64_t DIGIT_MASK = RADIX - 1;
constexpr int KEY_BITS_MAX = 63;       // packed key width; the dead sentinel is all ones
constexpr uint64_t DEAD_KEY = ~0ULL;
constexpr int MAX_PASSES = 8;
// Cells with at least DENSE_MIN objects (fresh fragment clouds, crowded
// shells) are screened by work items: CHUNK own objects (one per lane)
// against up to SEG partners read by the whole wave at once.
constexpr int DENSE_MIN = 128;  // defaults; the tuned values live in GridState
// A sparse cell's range longer than LONG_RANGE (a neighbouring cloud) is
// also handed to the work items instead of being walked lane by lane.
// (Measured at 1M objects, 2 km: the sparse-cell kernel beats the work items
// for everything but real clouds.)
constexpr int LONG_RANGE = 256;
constexpr int CHUNK = 64;
constexpr int SEG = 256;
constexpr int DENSE_BLOCKS = 1024;
constexpr int EXACT_BLOCKS = 512;
constexpr int MAX_DENSE_BLOCKS = 4096;  // partial-record capacity for tuned grids
constexpr int MAX_EXACT_BLOCKS = 2048;

// Launch and tiling parameters the autotuner chooses. None of them changes
// which pairs are found; cell_mult > 1 (larger cells, collision.cpp:337-338)
// changes the candidate count only.
struct TileCfg {
    double cell_mult;
    int dense_min, long_range, seg, dense_blocks, exact_blocks;
    int sparse;  // sparse-cell pair kernel: 0 partners on the lanes, 1 one own object per lane
};

// Device-resident grid parameters and per-step accumulators.
struct GridState {
    // Accumulated by k_propagate / k_reduce over end-of-step-alive objects.
    unsigned long long vmax_bits;   // max |v| as raw double bits (non-negative)
    unsigned long long lo[3];       // ordered bits of min position
    unsigned long long hi[3];       // ordered bits of max position
    unsigned int alive;             // objects alive at the end of the step
    unsigned int propagated;        // objects alive at the start of the step
    // Derived by k_grid_setup.
    double cell;                    // cell edge (km)
    double cell_ref;                // the reference's cell size before any doubling
    long long imin[3];              // smallest cell coordinate per axis
    int bits[3];                    // packed bits per axis (x, y, z)
    int sbits;                      // scenario-id bits above x (0: one scenario)
    int dense_min, long_range, seg; // pair tiling of this screening (TileCfg)
    int npasses;                    // radix passes needed (0: nothing alive)
    int doublings;                  // times the cell was doubled to fit the key
    unsigned int ncells;
    int passes_max;                 // max npasses over the chunk (reset per chunk)
    int passes_short;               // a step needed more passes than were launched
    unsigned int n_items;           // dense work items of this step
    unsigned int n_dense;           // dense cells of this step
    unsigned int items_max;         // max n_items over the chunk
    int items_short;                // a step had more items than the buffer holds
    unsigned long long n_surv;      // pre-filter survivors of this step
    unsigned long long surv_max;    // max n_surv over the chunk
    int surv_short;                 // a step had more survivors than the queue holds
    int pad2;
    // Accumulated by the pair kernels.
    unsigned long long cand, refined, co;
    unsigned int max_cell;
    unsigned int pad1;
    unsigned long long events;      // events appended since the last drain
    unsigned long long ev_mark;     // events before the current learned step
    int stamp;                      // screening counter (column-table entries of this screening)
    int bbits;                      // radial-shell bits between y and z (shell screening; 0 otherwise)
    unsigned int nsort;             // elements the radix passes sort (objects, or shell entries)
    unsigned int nlive;             // live sorted elements (alive objects, or every shell entry)
    unsigned int bmin;              // innermost shell of this screening (shell screening)
    int entries_short;              // a screening had more shell entries than were launched
    unsigned int entries_max;       // max shell entries over the chunk
    int pad4;
};

struct DevEvent {
    int32_t i, j;
    int64_t step;
    double t, miss_km, v_rel_km_s;
    int32_t collision, pad;
};
static_assert(sizeof(DevEvent) == sizeof(khip_event), "event layout");

struct DevStats {
    long long cand, refined, co;
    double cell_km;
    int max_cell, doublings;
    long long propagated;
    long long dense_cand;
    long long survivors;
};
static_assert(sizeof(DevStats) == sizeof(khip_stats), "stats layout");

// Per-block partial reductions, combined by the next single-block kernel
// (no same-address atomics from thousands of waves).
struct StepPartial {
    unsigned long long vmax_bits;  // max |v| (raw bits of a non-negative double)
    unsigned long long lo[3];      // ordered bits of min start-of-step position
    unsigned long long hi[3];      // ordered bits of max start-of-step position
    unsigned int alive;            // alive at the end of the step
    unsigned int propagated;       // alive at the start of the step
    unsigned int bmin, bmax;       // shell range of the entries (shell screening)
};

struct PairPartial {
    unsigned long long cand, refined, co;
    unsigned int max_cell, pad;
};

// A dense work item: lanes take own objects [own_begin, own_begin +
// own_count) of one cell; the wave walks partners [part_begin, part_begin +
// part_count) of neighbour range `kind` (0-4, see k_neighbours); in range 0
// a lane pairs only with positions after its own.
struct PairItem {
    int own_begin, own_count, part_begin, part_count;
    int kind, pad0, pad1, pad2;
};

struct DenseCell {
    int cell, item_base;
};

// SoA view of one state buffer.
struct Planes {
    double *x, *y, *z, *vx, *vy, *vz;
};

__device__ __forceinline__ StepPartial empty_step_partial() {
    StepPartial p;
    p.vmax_bits = 0ULL;
    for (int a = 0; a < 3; ++a) {
        p.lo[a] = ~0ULL;
        p.hi[a] = 0ULL;
    }
    p.alive = 0;
    p.propagated = 0;
    p.bmin = ~0u;
    p.bmax = 0u;
    return p;
}

__device__ __forceinline__ void combine(StepPartial& a, const StepPartial& b) {
    a.vmax_bits = b.vmax_bits > a.vmax_bits ? b.vmax_bits : a.vmax_bits;
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = b.lo[k] < a.lo[k] ? b.lo[k] : a.lo[k];
        a.hi[k] = b.hi[k] > a.hi[k] ? b.hi[k] : a.hi[k];
    }
    a.alive += b.alive;
    a.propagated += b.propagated;
    a.bmin = b.bmin < a.bmin ? b.bmin : a.bmin;
    a.bmax = b.bmax > a.bmax ? b.bmax : a.bmax;
}

__device__ __forceinline__ StepPartial wave_reduce(StepPartial p) {
    p.vmax_bits = wave_max_u64(p.vmax_bits);
    for (int k = 0; k < 3; ++k) {
        p.lo[k] = wave_min_u64(p.lo[k]);
        p.hi[k] = wave_max_u64(p.hi[k]);
    }
    p.alive = wave_sum<unsigned int>(p.alive);
    p.propagated = wave_sum<unsigned int>(p.propagated);
    p.bmin = static_cast<unsigned int>(wave_min_u64(p.bmin));
    p.bmax = static_cast<unsigned int>(wave_max_u64(p.bmax));
    return p;
}

// Block-wide combine; thread 0 returns the block's total.
template <int NW = WAVES_PER_BLOCK>
__device__ __forceinline__ StepPartial block_reduce(StepPartial p, StepPartial* s_rows) {
    p = wave_reduce(p);
    const int wave = threadIdx.x / WAVE;
    if (__lane_id() == 0) s_rows[wave] = p;
    __syncthreads();
    StepPartial t = s_rows[0];
    for (int w = 1; w < NW; ++w) combine(t, s_rows[w]);
    return t;
}

__device__ __forceinline__ StepPartial object_partial(bool on, bool was, double vn, double rx,
                                                      double ry, double rz) {
    StepPartial p = empty_step_partial();
    if (on) {
        p.vmax_bits = static_cast<unsigned long long>(__double_as_longlong(vn));
        p.lo[0] = p.hi[0] = ordered_bits(rx);
        p.lo[1] = p.hi[1] = ordered_bits(ry);
        p.lo[2] = p.hi[2] = ordered_bits(rz);
        p.alive = 1;
    }
    p.propagated = was ? 1 : 0;
    return p;
}

// Block-wide exclusive scan of one int per thread (NW waves); returns the total.
template <int NW = WAVES_PER_BLOCK>
__device__ __forceinline__ int block_exclusive_scan(int v, int& total, int* s_wave) {
    const int lane = __lane_id();
    const int wave = threadIdx.x / WAVE;
    int incl = v;
    for (int off = 1; off < WAVE; off <<= 1) {
        const int o = __shfl_up(incl, off, WAVE);
        if (lane >= off) incl += o;
    }
    if (lane == WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < NW; ++w) {
        if (w < wave) before += s_wave[w];
        sum += s_wave[w];
    }
    __syncthreads();
    total = sum;
    return before + incl - v;
}

// ---------------------------------------------------------------------------
// Propagation
// ---------------------------------------------------------------------------

template <class Acc>
__global__ void __launch_bounds__(BLOCK) k_propagate(Planes cur, Planes nxt, uint8_t* alive, int n,
                                                      const MoonSample* __restrict__ moon3,
                                                      double dt, StepPartial* partials) {
    __shared__ StepPartial s_rows[WAVES_PER_BLOCK];
    StepPartial acc = empty_step_partial();
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        bool on = false, was = false;
        double vn = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
        D3 r{cur.x[i], cur.y[i], cur.z[i]};
        D3 v{cur.vx[i], cur.vy[i], cur.vz[i]};
        rx = r.x; ry = r.y; rz = r.z;
        vn = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        bool a = alive[i] != 0;
        was = a;
        if (a) {
            const MoonSample m0 = moon3[0], m1 = moon3[1], m2 = moon3[2];
            rk4(r, v, dt, m0, m1, m2, Acc());
            if (sqrt(r.x * r.x + r.y * r.y + r.z * r.z) < dev::R_EARTH) {
                a = false;
                alive[i] = 0;
            }
        }
        nxt.x[i] = r.x; nxt.y[i] = r.y; nxt.z[i] = r.z;
        nxt.vx[i] = v.x; nxt.vy[i] = v.y; nxt.vz[i] = v.z;
        on = a;
        combine(acc, object_partial(on, was, vn, rx, ry, rz));
    }
    const StepPartial t = block_reduce(acc, s_rows);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Reductions only (screening of host-given states).
__global__ void __launch_bounds__(BLOCK) k_reduce(Planes cur, const uint8_t* alive, int n,
                                                   StepPartial* partials) {
    __shared__ StepPartial s_rows[WAVES_PER_BLOCK];
    StepPartial acc = empty_step_partial();
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        if (!alive[i]) continue;
        const double vx = cur.vx[i], vy = cur.vy[i], vz = cur.vz[i];
        combine(acc, object_partial(true, false, sqrt(vx * vx + vy * vy + vz * vz), cur.x[i], cur.y[i], cur.z[i]));
    }
    const StepPartial t = block_reduce(acc, s_rows);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Learned form: steps of tau with a per-object method
// ---------------------------------------------------------------------------
//
// Every object steps by one of three methods, chosen at check steps:
//   CLASS_LEARNED  one mixed-precision RK4 step of tau + the corrector;
//   CLASS_BASE     one mixed-precision RK4 step of tau (MEO, GEO);
//   CLASS_PLAIN    `substeps` FP64 RK4 steps of tau / substeps (the plain
//                  method; the fallback).
// A check step runs the plain method for every object (the state it writes
// is the plain result) and, for objects whose orbit lies in the learned or
// base domain, also that method; an object keeps the cheaper method until
// the next check only if it stayed within tol_km of the plain result.
// Between checks each method runs on its own list of objects, so waves
// are uniform.

struct LearnedParams {
    double tau;
    int substeps;
    int pad;
    double e_max, a_max, tol_km, tol_a_km;
};

// Semi-major axis from the orbital energy (bound orbits; large otherwise).
__device__ __forceinline__ double semi_major(D3 r, D3 v) {
    const double en = 0.5 * (v.x * v.x + v.y * v.y + v.z * v.z) - dev::MU_EARTH / sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    return en < 0.0 ? -dev::MU_EARTH / (2.0 * en) : 1e30;
}

struct CheckStats {
    unsigned int count[3];        // objects per method after the check
    unsigned int rejected;        // domain candidates sent to the plain method
    unsigned long long dev_max;   // largest accepted deviation (km), double bits
};

// Followed sample: up to SAMPLE_MAX objects, evenly spaced over the
// population, that the plain method carries alongside a learned run from
// the last change of population, time, form or configuration. Their largest
// distance from the learned states is the measured drift of the learned
// trajectories, and the search margin of the next chunk follows it.
constexpr int SAMPLE_MAX = 8192;

struct SampleObj {
    double r[3], v[3];
    int idx;  // object index
    int on;   // 0: dropped (the object or its plain twin died)
};

// The thresholds of one learned step's screening, from the sample's drift
// so far (k_windows, or k_bands when the windows are not fused, writes them
// once per step on the device; the screening kernels read them, so captured
// graphs and chunk boundaries do not depend on the values).
struct Bands {
    double band, band_v;              // drift bands (km, km/s)
    double thr_search, thr_report;    // search: threshold + margin + band; report: threshold + band
    double vmin2_search, vmin2_report;
    float thr_f, base_f, vmin2_f, pad;
};

// Per-step values read on the device (the step table of a chunk).
struct StepInfo {
    double t;         // start of the step (s since epoch)
    long long macro;  // global learned-step index
};

struct BandCfg {
    double factor;     // margin_factor (0: no band)
    double threshold;  // configured threshold (km)
    double margin;     // configured search margin (km)
    double vmin;       // configured minimum encounter speed (km/s)
    double window;     // screening window (s)
    double dt;         // learned step (s)
    double t0;         // time the sample started
};

// The drift band for the step starting at t: the sample's drift D (position
// or velocity) measured up to t, carried to the step's end as
// D ((T + dt) / T)^2 with T = t - t0 (an error in the semi-major axis grows
// the along-track drift quadratically in time), times `factor` (both objects
// of a pair drift, and the sample's maximum stands for the population's).
// The search's minimum speed is lowered further by the change of a pair's
// relative velocity over a window of w seconds: (2 mu / r^3) d w with
// 2 mu / r^3 <= 3.1e-6 s^-2 above the surface and d <= thr + vmin w for a
// pair that closes to thr in the window (the plain solve tests the speed at
// its sub-step's start, the search at its window's).
__host__ __device__ inline Bands make_bands(const BandCfg& c, double dev, double dev_v, double t) {
    Bands b{};
    const double T = t - c.t0;
    double g2 = 0.0;
    if (c.factor > 0.0 && T > 0.0) {
        const double g = (T + c.dt) / T;
        g2 = g * g;
    }
    b.band = c.factor * dev * g2;
    b.band_v = c.factor * dev_v * g2;
    b.thr_report = c.threshold + b.band;
    b.thr_search = b.thr_report + c.margin;
    const double vmin = c.vmin - b.band_v > 0.0 ? c.vmin - b.band_v : 0.0;
    b.vmin2_report = vmin * vmin;
    double vs = vmin - 3.1e-6 * (b.thr_search + vmin * c.window) * c.window;
    vs = vs > 0.0 ? vs : 0.0;
    b.vmin2_search = vs * vs;
    b.thr_f = static_cast<float>(b.thr_search);
    b.vmin2_f = static_cast<float>(b.vmin2_search);
    b.base_f = b.thr_f + 1e-3f + DR_ERR_KM + DV_ERR_KM_S * static_cast<float>(c.window);
    return b;
}

// One object through `substeps` FP64 RK4 steps; moon13 holds the Moon every
// tau / 12 over the step. Stops at the sub-step whose end is under the
// surface (the object dies there, as a plain step of tau / substeps would).
__device__ __forceinline__ bool plain_substeps(D3& r, D3& v, const MoonSample* m, double h, int sub) {
    const int stride = 12 / sub;
    for (int k = 0; k < sub; ++k) {
        rk4(r, v, h, m[stride * k], m[stride * k + stride / 2], m[stride * (k + 1)]);
        if (sqrt(r.x * r.x + r.y * r.y + r.z * r.z) < dev::R_EARTH) return false;
    }
    return true;
}

__device__ __forceinline__ void store_state(Planes p, int i, D3 r, D3 v) {
    p.x[i] = r.x; p.y[i] = r.y; p.z[i] = r.z;
    p.vx[i] = v.x; p.vy[i] = v.y; p.vz[i] = v.z;
}

// An object dies in a learned step (under the surface): its final state
// goes into the step's start planes too (the end planes already hold it), so
// both buffers keep it whether or not a later step's kernels visit the
// object (a check drops dead objects from the method lists). Nothing reads
// a dead object's start state later in the step.
__device__ __forceinline__ void kill(Planes cur, uint8_t* alive, int i, D3 r, D3 v) {
    store_state(cur, i, r, v);
    alive[i] = 0;
}

// Corrector on the 64 objects of a wave (lane = object). x: the lane's
// standardised inputs; returns the lane's six outputs.
__device__ __forceinline__ void net_wave(const float x[net::NIN], const uint16_t* frags, const float* P,
                                         float o[net::NOUT]) {
    const int lane = static_cast<int>(__lane_id());
    const int j = lane & 15, g = lane >> 4;
    // Packed BF16 words of the lane's own object, then per tile the words
    // of column j (object 16 t + j).
    uint32_t w[10];
    net::input_words(x, w);
    net::u32x4 xin[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t wt[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) wt[k] = static_cast<uint32_t>(__shfl(static_cast<int>(w[k]), 16 * t + j, WAVE));
        xin[t] = net::input_operand_words(wt, g);
    }
    // Two passes of two 16-object tiles: the hidden activations of four tiles
    // at once would hold one wave per SIMD (over 400 registers); per tile the
    // arithmetic is the same, so the outputs are too.
    net::f32x4 out[4];
    {
        const net::u32x4 x0[2] = {xin[0], xin[1]};
        net::forward_tiles<2>(x0, frags, P, lane, out);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the passes apart (register pressure)
    {
        const net::u32x4 x1[2] = {xin[2], xin[3]};
        net::forward_tiles<2>(x1, frags, P, lane, out + 2);
    }
    // Output 4g' + e of column j of tile t sits in lane j + 16 g', element e.
    const int mt = lane >> 4, mj = lane & 15;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float a0 = __shfl(out[t].x, mj, WAVE), a1 = __shfl(out[t].y, mj, WAVE);
        const float a2 = __shfl(out[t].z, mj, WAVE), a3 = __shfl(out[t].w, mj, WAVE);
        const float a4 = __shfl(out[t].x, mj + 16, WAVE), a5 = __shfl(out[t].y, mj + 16, WAVE);
        if (mt == t) {
            o[0] = a0; o[1] = a1; o[2] = a2; o[3] = a3; o[4] = a4; o[5] = a5;
        }
    }
}

// The same with NT tiles per pass (4 / NT passes): fewer tiles per pass
// hold fewer registers (more waves per SIMD), more tiles per pass read each
// weight fragment from LDS for more MFMAs.
template <int NT>
__device__ __forceinline__ void net_wave_nt(const float x[net::NIN], const uint16_t* frags, const float* P,
                                            float o[net::NOUT]) {
    const int lane = static_cast<int>(__lane_id());
    const int j = lane & 15, g = lane >> 4;
    uint32_t w[10];
    net::input_words(x, w);
    net::u32x4 xin[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uint32_t wt[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) wt[k] = static_cast<uint32_t>(__shfl(static_cast<int>(w[k]), 16 * t + j, WAVE));
        xin[t] = net::input_operand_words(wt, g);
    }
    net::f32x4 out[4];
#pragma unroll
    for (int t0 = 0; t0 < 4; t0 += NT) {
        net::forward_tiles<NT>(xin + t0, frags, P, lane, out + t0);
        __builtin_amdgcn_sched_barrier(0);  // keep the passes apart (register pressure)
    }
    const int mt = lane >> 4, mj = lane & 15;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float a0 = __shfl(out[t].x, mj, WAVE), a1 = __shfl(out[t].y, mj, WAVE);
        const float a2 = __shfl(out[t].z, mj, WAVE), a3 = __shfl(out[t].w, mj, WAVE);
        const float a4 = __shfl(out[t].x, mj + 16, WAVE), a5 = __shfl(out[t].y, mj + 16, WAVE);
        if (mt == t) {
            o[0] = a0; o[1] = a1; o[2] = a2; o[3] = a3; o[4] = a4; o[5] = a5;
        }
    }
}

// KIND: CLASS_LEARNED / CLASS_BASE / CLASS_PLAIN list kernels, or KIND_CHECK
// over every object. Persistent grid, waves stride over tiles of 64.
constexpr int KIND_CHECK = 3;

template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_learned(Planes cur, Planes nxt, uint8_t* alive, uint8_t* mode,
                                                    const int* __restrict__ list, const int* __restrict__ counts,
                                                    int n, const MoonSample* __restrict__ moon13,
                                                    const uint16_t* __restrict__ ffwd, const float* __restrict__ P,
                                                    net::NetNorm nn, LearnedParams lp, CheckStats* cs) {
    constexpr bool USES_NET = KIND == net::CLASS_LEARNED || KIND == KIND_CHECK;
    __shared__ uint4 s_frag[USES_NET ? net::NFRAG_FWD * net::FRAG_HALVES / 8 : 1];
    if (USES_NET) {
        const uint4* src = reinterpret_cast<const uint4*>(ffwd);
        for (int k = threadIdx.x; k < net::NFRAG_FWD * net::FRAG_HALVES / 8; k += BLOCK) s_frag[k] = src[k];
        __syncthreads();
    }
    const uint16_t* frags = reinterpret_cast<const uint16_t*>(s_frag);
    const int count = KIND == KIND_CHECK ? n : counts[KIND];
    const int lane = static_cast<int>(__lane_id());
    const int wave = threadIdx.x / WAVE;
    const double h = lp.tau / lp.substeps;
    unsigned int cnt[3] = {0, 0, 0}, rejected = 0;
    double dev_max = 0.0;
    for (int tile = blockIdx.x * WAVES_PER_BLOCK + wave; tile * WAVE < count; tile += gridDim.x * WAVES_PER_BLOCK) {
        const int k = tile * WAVE + lane;
        const bool active = k < count;
        const int i = active ? (KIND == KIND_CHECK ? k : list[k]) : 0;
        bool on = false;
        D3 r{0.0, 0.0, 0.0}, v{0.0, 0.0, 0.0};
        if (active) {
            r = D3{cur.x[i], cur.y[i], cur.z[i]};
            v = D3{cur.vx[i], cur.vy[i], cur.vz[i]};
            on = alive[i] != 0;
            if (!on) store_state(nxt, i, r, v);  // frozen
        }
        int cls = KIND;
        if (KIND == KIND_CHECK) cls = on ? net::domain_class(r, v, lp.e_max, lp.a_max) : net::CLASS_PLAIN;
        // Base step and corrector (all lanes of a wave take part in the MFMA).
        D3 rb = r, vb = v;
        const bool want_base = on && (cls == net::CLASS_LEARNED || cls == net::CLASS_BASE);
        if (want_base) rk4(rb, vb, lp.tau, moon13[0], moon13[6], moon13[12], AccelMixed());
        if (USES_NET) {
            const bool want_net = on && cls == net::CLASS_LEARNED;
            if (__ballot(want_net) != 0ULL) {
                float x[net::NIN];
                net::Frame fr;
                if (want_net) {
                    net::features(r, v, moon13[0], nn, x, fr);
                } else {
#pragma unroll
                    for (int q = 0; q < net::NIN; ++q) x[q] = 0.0f;
                }
                float o[net::NOUT];
                net_wave(x, frags, P, o);
                if (want_net) net::apply_correction(rb, vb, o, nn, fr);
            }
        }
        if (!on) continue;
        if (KIND == net::CLASS_LEARNED || KIND == net::CLASS_BASE) {
            store_state(nxt, i, rb, vb);
            if (sqrt(rb.x * rb.x + rb.y * rb.y + rb.z * rb.z) < dev::R_EARTH) kill(cur, alive, i, rb, vb);
        } else {
            D3 rp = r, vp = v;
            const bool survives = plain_substeps(rp, vp, moon13, h, lp.substeps);
            store_state(nxt, i, rp, vp);
            if (!survives) kill(cur, alive, i, rp, vp);
            if (KIND == KIND_CHECK) {
                int m = net::CLASS_PLAIN;
                if (survives && cls != net::CLASS_PLAIN) {
                    const double dx = rb.x - rp.x, dy = rb.y - rp.y, dz = rb.z - rp.z;
                    const double d = sqrt(dx * dx + dy * dy + dz * dz);
                    // A per-step error in the semi-major axis grows into an
                    // along-track drift quadratic in time, so it is bounded
                    // separately from the position.
                    const double da = fabs(semi_major(rb, vb) - semi_major(rp, vp));
                    if (d <= lp.tol_km && da <= lp.tol_a_km) {
                        m = cls;
                        dev_max = d > dev_max ? d : dev_max;
                    } else {
                        ++rejected;
                    }
                }
                mode[i] = static_cast<uint8_t>(m);
                if (survives) ++cnt[m];
            }
        }
    }
    if (KIND == KIND_CHECK) {
        for (int q = 0; q < 3; ++q) {
            const unsigned int s = wave_sum(cnt[q]);
            if (lane == 0 && s) atomicAdd(&cs->count[q], s);
        }
        const unsigned int rj = wave_sum(rejected);
        if (lane == 0 && rj) atomicAdd(&cs->rejected, rj);
        const unsigned long long dm =
            wave_max_u64(static_cast<unsigned long long>(__double_as_longlong(dev_max)));
        if (lane == 0) atomicMax(&cs->dev_max, dm);
    }
}

// The same step split by occupancy: the FP64 work runs in a light kernel, one
// object per lane at full occupancy, and only the corrector (the MFMA work,
// whose weights fill LDS) in a kernel of its own. Arithmetic and results are
// those of k_learned, bit for bit.
//
// k_fp64_methods: the FP64 work of a step that is not a check, in one
// launch: the plain list's sub-steps and the mixed base step of tau of the
// corrected and base lists (into the next states; the base list with the
// re-entry test, the corrected objects get theirs after the correction),
// object for object k_learned's operations. The waves split between the plain list
// and the two base lists by cost (substeps RK4 steps per plain object, one
// per base step), each part striding over its tiles of 64; a plain object's
// chain of sub-steps then runs beside the base steps instead of after them.
// Lists: corrected at `lists`, plain at lists + c, base at lists + 2c.
__global__ void __launch_bounds__(BLOCK) k_fp64_methods(Planes cur, Planes nxt, uint8_t* alive,
                                                         const int* __restrict__ lists, size_t c,
                                                         const int* __restrict__ counts,
                                                         const MoonSample* __restrict__ moon13, LearnedParams lp) {
    const int nl = counts[net::CLASS_LEARNED], nb = counts[net::CLASS_BASE], np = counts[net::CLASS_PLAIN];
    const int lane = static_cast<int>(__lane_id());
    const int waves = static_cast<int>(gridDim.x) * WAVES_PER_BLOCK;
    const int w = static_cast<int>(blockIdx.x) * WAVES_PER_BLOCK + static_cast<int>(threadIdx.x) / WAVE;
    const int tp = (np + WAVE - 1) / WAVE, tl = (nl + WAVE - 1) / WAVE, tb = (nb + WAVE - 1) / WAVE;
    // Waves for the plain list: its share of the cost, at least one when it
    // has objects and at most all but one when the base lists have some.
    const long long cost_p = static_cast<long long>(lp.substeps) * np, cost_b = static_cast<long long>(nl) + nb;
    int wp = cost_p + cost_b > 0 ? static_cast<int>((waves * cost_p + cost_p + cost_b - 1) / (cost_p + cost_b)) : 0;
    if (tp > 0 && wp < 1) wp = 1;
    if (tl + tb > 0 && wp > waves - 1) wp = waves - 1;
    if (wp > tp) wp = tp;
    if (w < wp) {
        const int* list = lists + c;
        const double h = lp.tau / lp.substeps;
        for (int tile = w; tile < tp; tile += wp) {
            const int k = tile * WAVE + lane;
            if (k >= np) continue;
            const int i = list[k];
            D3 r{cur.x[i], cur.y[i], cur.z[i]}, v{cur.vx[i], cur.vy[i], cur.vz[i]};
            if (!alive[i]) {
                store_state(nxt, i, r, v);  // frozen
                continue;
            }
            const bool survives = plain_substeps(r, v, moon13, h, lp.substeps);
            store_state(nxt, i, r, v);
            if (!survives) kill(cur, alive, i, r, v);
        }
        return;
    }
    const int wb = waves - wp;
    for (int tile = w - wp; tile < tl + tb; tile += wb) {
        const bool learned = tile < tl;
        const int k = (learned ? tile : tile - tl) * WAVE + lane;
        if (k >= (learned ? nl : nb)) continue;
        const int i = learned ? lists[k] : lists[2 * c + k];
        D3 r{cur.x[i], cur.y[i], cur.z[i]}, v{cur.vx[i], cur.vy[i], cur.vz[i]};
        if (!alive[i]) {
            store_state(nxt, i, r, v);  // frozen
            continue;
        }
        rk4(r, v, lp.tau, moon13[0], moon13[6], moon13[12], AccelMixed());
        store_state(nxt, i, r, v);
        // The corrected objects' re-entry test follows their correction.
        if (!learned && sqrt(r.x * r.x + r.y * r.y + r.z * r.z) < dev::R_EARTH) kill(cur, alive, i, r, v);
    }
}

// k_check_plain: the plain method for every object (a check step's states),
// dead objects frozen; an object under the surface at a sub-step's end dies.
__global__ void __launch_bounds__(BLOCK) k_check_plain(Planes cur, Planes nxt, uint8_t* alive, int n,
                                                        const MoonSample* __restrict__ moon13, LearnedParams lp) {
    const double h = lp.tau / lp.substeps;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        D3 r{cur.x[i], cur.y[i], cur.z[i]}, v{cur.vx[i], cur.vy[i], cur.vz[i]};
        if (alive[i] && !plain_substeps(r, v, moon13, h, lp.substeps)) kill(cur, alive, i, r, v);
        store_state(nxt, i, r, v);
    }
}

// A check's verdict on one object of the corrector's or base step's domain
// (cls) from its candidate state (rb, vb: corrected or base) and its plain
// state (rp, vp): it keeps the cheaper method only within the position and
// semi-major-axis tolerances (the semi-major axis separately, since a
// per-step error in it grows into a quadratic along-track drift).
__device__ __forceinline__ int check_verdict(D3 rb, D3 vb, D3 rp, D3 vp, int cls, const LearnedParams& lp,
                                             double& dev_max, unsigned int& rejected) {
    const double dx = rb.x - rp.x, dy = rb.y - rp.y, dz = rb.z - rp.z;
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double da = fabs(semi_major(rb, vb) - semi_major(rp, vp));
    if (d <= lp.tol_km && da <= lp.tol_a_km) {
        dev_max = d > dev_max ? d : dev_max;
        return cls;
    }
    ++rejected;
    return net::CLASS_PLAIN;
}

// A block's check tallies, one record per block (k_build_lists combines
// them): atomics from every wave on the one CheckStats line serialize.
__device__ __forceinline__ void check_partial(const unsigned int (&cnt)[3], unsigned int rejected, double dev_max,
                                              CheckStats* out) {
    __shared__ CheckStats s_p[TWAVES];
    const int lane = static_cast<int>(__lane_id());
    const int wave = threadIdx.x / WAVE;
    CheckStats p;
    for (int q = 0; q < 3; ++q) p.count[q] = wave_sum(cnt[q]);
    p.rejected = wave_sum(rejected);
    p.dev_max = wave_max_u64(static_cast<unsigned long long>(__double_as_longlong(dev_max)));
    if (lane == 0) s_p[wave] = p;
    __syncthreads();
    if (threadIdx.x != 0) return;
    CheckStats t = s_p[0];
    for (int w = 1; w < static_cast<int>(blockDim.x) / WAVE; ++w) {
        for (int q = 0; q < 3; ++q) t.count[q] += s_p[w].count[q];
        t.rejected += s_p[w].rejected;
        t.dev_max = s_p[w].dev_max > t.dev_max ? s_p[w].dev_max : t.dev_max;
    }
    *out = t;
}

// A check's first pass, after k_check_plain (plain states in `nxt`, `alive`
// = alive and survived the plain sub-steps). Block b takes objects
// [b * chunk, (b + 1) * chunk): dead objects and those outside both domains
// take the plain method; the base step's domain is decided here (the mixed
// step against the plain state); the corrector's domain gets its mixed step
// written to `base` and is listed for k_net_step<true> (segment b of `list`,
// in object order, its length in segcnt[b]), so the corrector runs on full
// waves of its own objects only.
constexpr int CHECK_BLOCKS = 256;

__global__ void __launch_bounds__(TBLOCK) k_check_base(Planes cur, Planes nxt, Planes base, const uint8_t* alive,
                                                        uint8_t* mode, int n, int chunk,
                                                        const MoonSample* __restrict__ moon13, LearnedParams lp,
                                                        int* list, int* segcnt, CheckStats* parts) {
    __shared__ int s_wave[TWAVES];
    const int b0 = blockIdx.x * chunk;
    const int b1 = b0 + chunk < n ? b0 + chunk : n;
    unsigned int cnt[3] = {0, 0, 0}, rejected = 0;
    double dev_max = 0.0;
    int run = 0;
    for (int k0 = b0; k0 < b1; k0 += TBLOCK) {
        const int i = k0 + static_cast<int>(threadIdx.x);
        bool listed = false;
        if (i < b1) {
            const bool on = alive[i] != 0;
            D3 r{0.0, 0.0, 0.0}, v{0.0, 0.0, 0.0};
            if (on) {
                r = D3{cur.x[i], cur.y[i], cur.z[i]};
                v = D3{cur.vx[i], cur.vy[i], cur.vz[i]};
            }
            const int cls = on ? net::domain_class(r, v, lp.e_max, lp.a_max) : net::CLASS_PLAIN;
            if (cls == net::CLASS_PLAIN) {
                mode[i] = static_cast<uint8_t>(net::CLASS_PLAIN);
                if (on) ++cnt[net::CLASS_PLAIN];
            } else {
                rk4(r, v, lp.tau, moon13[0], moon13[6], moon13[12], AccelMixed());
                if (cls == net::CLASS_BASE) {
                    const D3 rp{nxt.x[i], nxt.y[i], nxt.z[i]}, vp{nxt.vx[i], nxt.vy[i], nxt.vz[i]};
                    const int m = check_verdict(r, v, rp, vp, cls, lp, dev_max, rejected);
                    mode[i] = static_cast<uint8_t>(m);
                    ++cnt[m];
                } else {
                    store_state(base, i, r, v);
                    listed = true;
                }
            }
        }
        int tot;
        const int ex = block_exclusive_scan<TWAVES>(listed ? 1 : 0, tot, s_wave);
        if (listed) list[b0 + run + ex] = i;
        run += tot;
    }
    if (threadIdx.x == 0) segcnt[blockIdx.x] = run;
    check_partial(cnt, rejected, dev_max, parts + blockIdx.x);
}

// A check's last pass over the corrector's domain: block b takes segment b
// of the list, the corrected states k_net_step<true> wrote to `base`, the
// plain ones in `nxt`; the verdict at full occupancy.
__global__ void __launch_bounds__(TBLOCK) k_check_verdict(Planes base, Planes nxt, uint8_t* mode,
                                                           const int* __restrict__ list,
                                                           const int* __restrict__ segcnt, int chunk,
                                                           LearnedParams lp, CheckStats* parts) {
    unsigned int cnt[3] = {0, 0, 0}, rejected = 0;
    double dev_max = 0.0;
    const int m_seg = segcnt[blockIdx.x];
    for (int j = threadIdx.x; j < m_seg; j += TBLOCK) {
        const int i = list[blockIdx.x * chunk + j];
        const D3 rb{base.x[i], base.y[i], base.z[i]}, vb{base.vx[i], base.vy[i], base.vz[i]};
        const D3 rp{nxt.x[i], nxt.y[i], nxt.z[i]}, vp{nxt.vx[i], nxt.vy[i], nxt.vz[i]};
        const int m = check_verdict(rb, vb, rp, vp, net::CLASS_LEARNED, lp, dev_max, rejected);
        mode[i] = static_cast<uint8_t>(m);
        ++cnt[m];
    }
    check_partial(cnt, rejected, dev_max, parts + CHECK_BLOCKS + blockIdx.x);
}

// k_net_step: the corrector on top of the base step's states. CHECK false:
// the corrected list, base states in `nxt`, corrected in place (re-entry
// test after the correction). CHECK true: the corrector's domain as
// k_check_base listed it (segments of `list`, lengths in segcnt, `chunk`
// apart), base states in `base`, corrected in place for k_check_verdict.
// With k_check_base and k_check_verdict the check's tallies and method
// flags are those k_learned<KIND_CHECK> writes.
constexpr int NET_NT = 2;          // tiles per pass in k_net_step
constexpr int NET_BLOCK = 256;     // two blocks share a CU's LDS: two waves per SIMD
constexpr int NET_WAVES = NET_BLOCK / WAVE;

template <bool CHECK>
__global__ void __launch_bounds__(NET_BLOCK) __attribute__((amdgpu_waves_per_eu(2))) k_net_step(Planes cur, Planes nxt, Planes base, uint8_t* alive,
                                                     uint8_t* mode, const int* __restrict__ list,
                                                     const int* __restrict__ counts, int n,
                                                     const MoonSample* __restrict__ moon13,
                                                     const uint16_t* __restrict__ ffwd, const float* __restrict__ P,
                                                     net::NetNorm nn, LearnedParams lp, CheckStats* cs,
                                                     const int* __restrict__ segcnt, int chunk) {
    __shared__ uint4 s_frag[net::NFRAG_FWD * net::FRAG_HALVES / 8];
    {
        const uint4* src = reinterpret_cast<const uint4*>(ffwd);
        for (int k = threadIdx.x; k < net::NFRAG_FWD * net::FRAG_HALVES / 8; k += NET_BLOCK) s_frag[k] = src[k];
        __syncthreads();
    }
    const uint16_t* frags = reinterpret_cast<const uint16_t*>(s_frag);
    const int lane = static_cast<int>(__lane_id());
    const int wave = threadIdx.x / WAVE;
    // At checks: the listed objects in tiles of 64, segment by segment
    // (segment b has ceil(segcnt[b] / 64) tiles). Every wave keeps the
    // tiles' exclusive prefix in registers, segments 4 lane .. 4 lane + 3 on
    // lane `lane` (no LDS beyond the fragments: two blocks per CU).
    constexpr int SEG_PER_LANE = CHECK_BLOCKS / WAVE;
    int tpre[SEG_PER_LANE];
    int count;
    if (CHECK) {
        int own = 0;
#pragma unroll
        for (int q = 0; q < SEG_PER_LANE; ++q) {
            tpre[q] = own;
            own += (segcnt[SEG_PER_LANE * lane + q] + WAVE - 1) / WAVE;
        }
        int incl = own;
        for (int off = 1; off < WAVE; off <<= 1) {
            const int o = __shfl_up(incl, off, WAVE);
            if (lane >= off) incl += o;
        }
#pragma unroll
        for (int q = 0; q < SEG_PER_LANE; ++q) tpre[q] += incl - own;
        count = __shfl(incl, WAVE - 1, WAVE) * WAVE;
    } else {
        count = counts[net::CLASS_LEARNED];
    }
    for (int tile = blockIdx.x * NET_WAVES + wave; tile * WAVE < count; tile += gridDim.x * NET_WAVES) {
        int k = tile * WAVE + lane;
        bool active = k < count;
        if (CHECK) {
            // The tile's segment: the last one whose first tile is at or
            // before it (wave-uniform), then the lane's place in it.
            const unsigned long long at = __ballot(tpre[0] <= tile);
            const int owner = 63 - __clzll(at);
            int seg = SEG_PER_LANE * owner, first = __shfl(tpre[0], owner, WAVE);
#pragma unroll
            for (int q = 1; q < SEG_PER_LANE; ++q) {
                const int f = __shfl(tpre[q], owner, WAVE);
                if (f <= tile) {
                    seg = SEG_PER_LANE * owner + q;
                    first = f;
                }
            }
            const int in_seg = (tile - first) * WAVE + lane;
            active = in_seg < segcnt[seg];
            k = seg * chunk + in_seg;
        }
        const int i = active ? list[k] : 0;
        const bool on = active && alive[i] != 0;
        D3 r{0.0, 0.0, 0.0}, v{0.0, 0.0, 0.0};
        if (on) {
            r = D3{cur.x[i], cur.y[i], cur.z[i]};
            v = D3{cur.vx[i], cur.vy[i], cur.vz[i]};
        }
        // At checks the list holds live objects of the corrector's domain only (k_check_base).
        const int cls = !on ? net::CLASS_PLAIN : net::CLASS_LEARNED;
        // The base state and the correction frame are taken after the
        // network (fewer registers live across it; the frame is the same
        // arithmetic on the same inputs).
        const bool want_net = cls == net::CLASS_LEARNED;
        float o[net::NOUT];
        const bool any_net = __ballot(want_net) != 0ULL;
        if (any_net) {
            float x[net::NIN];
            if (want_net) {
                net::Frame unused;
                net::features(r, v, moon13[0], nn, x, unused);
            } else {
#pragma unroll
                for (int q = 0; q < net::NIN; ++q) x[q] = 0.0f;
            }
            net_wave_nt<NET_NT>(x, frags, P, o);
        }
        const Planes src = CHECK ? base : nxt;
        D3 rb{0.0, 0.0, 0.0}, vb{0.0, 0.0, 0.0};
        if (cls != net::CLASS_PLAIN) {
            rb = D3{src.x[i], src.y[i], src.z[i]};
            vb = D3{src.vx[i], src.vy[i], src.vz[i]};
        }
        if (want_net) {
            float x[net::NIN];
            net::Frame fr;
            net::features(r, v, moon13[0], nn, x, fr);
            net::apply_correction(rb, vb, o, nn, fr);
        }
        if (!CHECK) {
            if (on) {
                store_state(nxt, i, rb, vb);
                if (sqrt(rb.x * rb.x + rb.y * rb.y + rb.z * rb.z) < dev::R_EARTH) kill(cur, alive, i, rb, vb);
            }
        } else if (on) {
            store_state(base, i, rb, vb);  // judged by k_check_verdict
        }
    }
}

// Per-method index lists of the objects alive after a check. Same-address
// atomics serialize (about 12 ns each on MI355X), so each block takes one
// contiguous chunk of objects (LIST_BLOCKS blocks at most): a first pass
// counts the chunk's objects per method, one atomic per method reserves the
// chunk's stretch of each list, and a second pass writes the indices in
// object order (wave ballots, running offsets).
constexpr int LIST_BLOCKS = 256;

__global__ void __launch_bounds__(BLOCK) k_build_lists(const uint8_t* alive, const uint8_t* mode, int n,
                                                        int* lists, int cap, int* counts,
                                                        const CheckStats* __restrict__ parts, CheckStats* cs) {
    __shared__ int s_base[3];
    __shared__ int s_wave[3][WAVES_PER_BLOCK];
    if (parts && blockIdx.x == 0) {
        // The check's tallies: k_check_base's and k_check_verdict's block records.
        unsigned int c[3] = {0, 0, 0}, rj = 0;
        unsigned long long dm = 0;
        for (int k = threadIdx.x; k < 2 * CHECK_BLOCKS; k += BLOCK) {
            for (int q = 0; q < 3; ++q) c[q] += parts[k].count[q];
            rj += parts[k].rejected;
            dm = parts[k].dev_max > dm ? parts[k].dev_max : dm;
        }
        CheckStats p;
        for (int q = 0; q < 3; ++q) p.count[q] = wave_sum(c[q]);
        p.rejected = wave_sum(rj);
        p.dev_max = wave_max_u64(dm);
        if (__lane_id() == 0) {
            for (int q = 0; q < 3; ++q) atomicAdd(&cs->count[q], p.count[q]);
            atomicAdd(&cs->rejected, p.rejected);
            atomicMax(&cs->dev_max, p.dev_max);
        }
    }
    const int chunk = ((n + gridDim.x - 1) / gridDim.x + BLOCK - 1) / BLOCK * BLOCK;
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    const int wave = threadIdx.x / WAVE;
    const unsigned long long below = lanes_below();
    int cnt[3] = {0, 0, 0};
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        const int m = alive[i] ? mode[i] : -1;
        for (int q = 0; q < 3; ++q) cnt[q] += m == q ? 1 : 0;
    }
    for (int q = 0; q < 3; ++q) {
        const int c = static_cast<int>(wave_sum<unsigned int>(static_cast<unsigned int>(cnt[q])));
        if (__lane_id() == 0) s_wave[q][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) {
            int t = 0;
            for (int w = 0; w < WAVES_PER_BLOCK; ++w) t += s_wave[q][w];
            s_base[q] = t ? atomicAdd(&counts[q], t) : 0;
        }
    }
    __syncthreads();
    int run[3] = {s_base[0], s_base[1], s_base[2]};
    __syncthreads();  // s_wave is reused below
    for (int b = i0; b < i1; b += BLOCK) {
        const int i = b + threadIdx.x;
        const int m = i < i1 && alive[i] ? mode[i] : -1;
        int rank[3];
        for (int q = 0; q < 3; ++q) {
            const unsigned long long bq = __ballot(m == q);
            rank[q] = __popcll(bq & below);
            if (__lane_id() == 0) s_wave[q][wave] = __popcll(bq);
        }
        __syncthreads();
        if (m >= 0) {
            int off = run[m] + rank[m];
            for (int w = 0; w < wave; ++w) off += s_wave[m][w];
            lists[m * cap + off] = i;
        }
        for (int q = 0; q < 3; ++q)
            for (int w = 0; w < WAVES_PER_BLOCK; ++w) run[q] += s_wave[q][w];
        __syncthreads();
    }
}

// Dense output: state at fraction s of the step from the cubic Hermite of an
// object's start and end states (positions and velocities).
__device__ __forceinline__ void hermite_state(const double x0[3], const double v0[3], const double x1[3],
                                              const double v1[3], double dt, double s, double p[3], double w[3]) {
    const double s2 = s * s, s3 = s2 * s;
    const double c0 = 2 * s3 - 3 * s2 + 1, c1 = s3 - 2 * s2 + s, c2 = -2 * s3 + 3 * s2, c3 = s3 - s2;
    const double d0 = 6 * s2 - 6 * s, d1 = 3 * s2 - 4 * s + 1, d2 = -6 * s2 + 6 * s, d3 = 3 * s2 - 2 * s;
    for (int k = 0; k < 3; ++k) {
        p[k] = x0[k] * c0 + v0[k] * dt * c1 + x1[k] * c2 + v1[k] * dt * c3;
        w[k] = (x0[k] * d0 + v0[k] * dt * d1 + x1[k] * d2 + v1[k] * dt * d3) / dt;
    }
}

// Sub-step q of sub of a learned step from a (start) to b (end): its start
// state (a itself for q = 0, else the dense output at q / sub, recomputed
// bit for bit as the previous sub-step wrote it) and end state (b itself
// for the last sub-step, else the dense output at (q + 1) / sub, written to
// out), and the reductions for its midpoint grid: max of |v| over both ends
// and the bounding box of the midpoints, of the objects alive at the end of
// the step. One pass over a and b.
__global__ void __launch_bounds__(BLOCK) k_substep(Planes a, Planes b, const uint8_t* alive, int n, double dt,
                                                    int q, int sub, Planes out, StepPartial* partials) {
    __shared__ StepPartial s_rows[WAVES_PER_BLOCK];
    StepPartial acc = empty_step_partial();
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        bool on = false;
        double vn = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
        const double x0[3] = {a.x[i], a.y[i], a.z[i]}, v0[3] = {a.vx[i], a.vy[i], a.vz[i]};
        const double x1[3] = {b.x[i], b.y[i], b.z[i]}, v1[3] = {b.vx[i], b.vy[i], b.vz[i]};
        double ps[3], ws[3], pe[3], we[3];
        if (q == 0) {
            for (int k = 0; k < 3; ++k) ps[k] = x0[k], ws[k] = v0[k];
        } else {
            hermite_state(x0, v0, x1, v1, dt, static_cast<double>(q) / sub, ps, ws);
        }
        if (q == sub - 1) {
            for (int k = 0; k < 3; ++k) pe[k] = x1[k], we[k] = v1[k];
        } else {
            hermite_state(x0, v0, x1, v1, dt, static_cast<double>(q + 1) / sub, pe, we);
            out.x[i] = pe[0]; out.y[i] = pe[1]; out.z[i] = pe[2];
            out.vx[i] = we[0]; out.vy[i] = we[1]; out.vz[i] = we[2];
        }
        if (alive[i]) {
            on = true;
            rx = 0.5 * (ps[0] + pe[0]);
            ry = 0.5 * (ps[1] + pe[1]);
            rz = 0.5 * (ps[2] + pe[2]);
            const double va = sqrt(ws[0] * ws[0] + ws[1] * ws[1] + ws[2] * ws[2]);
            const double vb = sqrt(we[0] * we[0] + we[1] * we[1] + we[2] * we[2]);
            vn = va > vb ? va : vb;
        }
        combine(acc, object_partial(on, on, vn, rx, ry, rz));
    }
    const StepPartial t = block_reduce(acc, s_rows);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Radial shells of the learned form's screening. Two objects closer than
// thr at some instant have radii closer than thr then, so the intervals
// [rho_lo - thr / 2, rho_hi + thr / 2] of their radius ranges over the
// window share a point, and the shell holding that point holds an entry of
// each. Every object gets one entry per shell its interval meets; a pair is
// screened in its innermost common shell only, the one where at least one
// of the two entries is its object's first (NOT_FIRST clear on one side).
// Shells are fixed intervals of the radius: w km wide from r0 to r1 = r0 +
// n1 w (the crowded low-Earth-orbit region), w2 km wide beyond, and the
// last one (index nmax) open-ended. shell_of is non-decreasing in rho.
struct ShellCfg {
    double r0, w, r1, w2;
    int n1, nmax;
    double half;  // half the search threshold plus rounding slack (km)
    float r0f, r1f, inv_w, inv_w2;  // FP32 copies for shell_of
};

// Shell index of radius rho, in FP32 (reciprocal widths): a radius within
// SHELL_SLACK of a boundary may land on either side, which the callers'
// widening by SHELL_SLACK covers.
__device__ __forceinline__ int shell_of(float rho, const ShellCfg& c) {
    if (rho < c.r1f) return rho <= c.r0f ? 0 : static_cast<int>((rho - c.r0f) * c.inv_w);
    const float k = (rho - c.r1f) * c.inv_w2;
    return k >= static_cast<float>(c.nmax - c.n1) ? c.nmax : c.n1 + static_cast<int>(k);
}

// FP32 error allowance of the radius bounds and shell indices below: 2e-6
// relative (the few FP32 operations err by well under 1e-6) plus 0.1 m.
__device__ __forceinline__ float shell_slack(float r) { return 2e-6f * r + 1e-4f; }

// Radius range of a window's cubic (the Hermite of its end states p0, v0
// and p1, v1 over h seconds). The curve lies in the convex hull of its
// Bezier control points P0 = p0, P1 = p0 + h v0 / 3, P2 = p1 - h v1 / 3,
// P3 = p1, so |p(t)| <= max |Pk|, and |p(t)| >= p(t).u >= min Pk.u for the
// unit vector u along P0 + P3. The control points are formed in FP64, the
// bounds in FP32 and widened by shell_slack.
__device__ __forceinline__ void radius_range(const double p0[3], const double v0[3], const double p1[3],
                                             const double v1[3], double h, float& lo, float& hi) {
    float P[4][3];
    for (int k = 0; k < 3; ++k) {
        P[0][k] = static_cast<float>(p0[k]);
        P[1][k] = static_cast<float>(p0[k] + h * v0[k] * (1.0 / 3.0));
        P[2][k] = static_cast<float>(p1[k] - h * v1[k] * (1.0 / 3.0));
        P[3][k] = static_cast<float>(p1[k]);
    }
    const float ux = P[0][0] + P[3][0], uy = P[0][1] + P[3][1], uz = P[0][2] + P[3][2];
    const float un2 = ux * ux + uy * uy + uz * uz;
    const float inv = un2 > 0.0f ? __builtin_amdgcn_rsqf(un2) : 0.0f;
    float h2 = 0.0f, dmin = 3.0e38f;
    for (int j = 0; j < 4; ++j) {
        const float r2 = P[j][0] * P[j][0] + P[j][1] * P[j][1] + P[j][2] * P[j][2];
        h2 = r2 > h2 ? r2 : h2;
        const float d = (P[j][0] * ux + P[j][1] * uy + P[j][2] * uz) * inv;
        dmin = d < dmin ? d : dmin;
    }
    hi = __builtin_sqrtf(h2);
    lo = un2 > 0.0f ? dmin : 0.0f;
    const float sl = shell_slack(hi);
    hi += sl;
    lo -= sl;
}

// Every screening window of a learned step in one pass over its start (a)
// and end (b) states: window q's end state (the dense output at (q + 1) / nw,
// written to the q-th dense planes of `dense`, or b itself for the last
// window) and its midpoint-grid reductions (partials + q * gridDim.x). Window
// q starts at window q - 1's end, the same dense output k_substep recomputes,
// so states and reductions equal k_substep's window by window. Also marks
// the events before the step's screening (no kernel appends events while
// this one runs), as k_mark_events does for the per-window path.
// With shells (shells != 0): each alive object's first shell and shell count
// per window (shell_lo / shell_n + q * cap), the shell range in the
// partials, and the entries per object tile of TILE objects added to
// tile_cnt + q * tstride (k_grid_setup reads and clears them).
constexpr int FUSED_WINDOWS_MAX = 4;

// NW: the windows (1 .. FUSED_WINDOWS_MAX), a template argument so that only
// NW windows' partial reductions are held in registers.
template <int NW>
__global__ void __launch_bounds__(BLOCK) k_windows(Planes a, Planes b, const uint8_t* alive, int n, double dt,
                                                    double* dense, size_t cap, StepPartial* partials,
                                                    GridState* gs, unsigned long long ev_cap, int shells,
                                                    ShellCfg sc, int* shell_lo, uint16_t* shell_n, int* tile_cnt,
                                                    int tstride, const unsigned long long* bdev,
                                                    const StepInfo* binfo, BandCfg bcfg, Bands* bands) {
    // With a drift band: the step's thresholds from the sample's drift so
    // far (every block computes them; block 0 stores them for the kernels
    // that follow).
    if (bands) {
        const double d = bdev ? __longlong_as_double(static_cast<long long>(bdev[0])) : 0.0;
        const double dv = bdev ? __longlong_as_double(static_cast<long long>(bdev[1])) : 0.0;
        const Bands bb = make_bands(bcfg, d, dv, binfo->t);
        sc.half = 0.5 * bb.thr_search + 1e-6;
        if (blockIdx.x == 0 && threadIdx.x == 0) *bands = bb;
    }
    __shared__ StepPartial s_rows[WAVES_PER_BLOCK];
    if (blockIdx.x == 0 && threadIdx.x == 0) gs->ev_mark = gs->events < ev_cap ? gs->events : ev_cap;
    constexpr int nw = NW;
    StepPartial acc[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) acc[q] = empty_step_partial();
    const double hw = dt / nw;
    // Block-uniform trip count: the per-wave tile sums below need every lane.
    for (int i0 = blockIdx.x * BLOCK; i0 < n; i0 += gridDim.x * BLOCK) {
        const int i = i0 + static_cast<int>(threadIdx.x);
        const bool in = i < n;
        const int ii = in ? i : n - 1;
        const double x0[3] = {a.x[ii], a.y[ii], a.z[ii]}, v0[3] = {a.vx[ii], a.vy[ii], a.vz[ii]};
        const double x1[3] = {b.x[ii], b.y[ii], b.z[ii]}, v1[3] = {b.vx[ii], b.vy[ii], b.vz[ii]};
        const bool on = in && alive[ii] != 0;
        double ps[3] = {x0[0], x0[1], x0[2]}, ws[3] = {v0[0], v0[1], v0[2]};
        // |v| at the window's start: the previous window's end value.
        double va = on ? sqrt(ws[0] * ws[0] + ws[1] * ws[1] + ws[2] * ws[2]) : 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            double pe[3], we[3];
            if (q == nw - 1) {
                for (int k = 0; k < 3; ++k) pe[k] = x1[k], we[k] = v1[k];
            } else {
                hermite_state(x0, v0, x1, v1, dt, static_cast<double>(q + 1) / nw, pe, we);
                if (in) {
                    double* o = dense + static_cast<size_t>(q) * 6 * cap;
                    o[i] = pe[0]; o[cap + i] = pe[1]; o[2 * cap + i] = pe[2];
                    o[3 * cap + i] = we[0]; o[4 * cap + i] = we[1]; o[5 * cap + i] = we[2];
                }
            }
            double vn = 0.0, rx = 0.0, ry = 0.0, rz = 0.0;
            if (on) {
                rx = 0.5 * (ps[0] + pe[0]);
                ry = 0.5 * (ps[1] + pe[1]);
                rz = 0.5 * (ps[2] + pe[2]);
                const double vb = sqrt(we[0] * we[0] + we[1] * we[1] + we[2] * we[2]);
                vn = va > vb ? va : vb;
                va = vb;
            }
            StepPartial op = object_partial(on, on, vn, rx, ry, rz);
            if (shells) {
                int nb = 0;
                if (on) {
                    float lo, hi;
                    radius_range(ps, ws, pe, we, hw, lo, hi);
                    const float half = static_cast<float>(sc.half);
                    const int s0 = shell_of(lo - half, sc), s1 = shell_of(hi + half, sc);
                    nb = s1 - s0 + 1;
                    shell_lo[static_cast<size_t>(q) * cap + i] = s0;
                    shell_n[static_cast<size_t>(q) * cap + i] = static_cast<uint16_t>(nb);
                    op.bmin = static_cast<unsigned int>(s0);
                    op.bmax = static_cast<unsigned int>(s1);
                } else if (in) {
                    shell_n[static_cast<size_t>(q) * cap + i] = 0;
                }
                // A wave's 64 objects lie in one tile (TILE is a multiple of 64).
                const int sum = wave_sum<int>(nb);
                if (__lane_id() == 0 && sum > 0) atomicAdd(&tile_cnt[q * tstride + i0 / TILE], sum);
            }
            combine(acc[q], op);
            for (int k = 0; k < 3; ++k) ps[k] = pe[k], ws[k] = we[k];
        }
    }
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        const StepPartial t = block_reduce(acc[q], s_rows);
        if (threadIdx.x == 0) partials[q * gridDim.x + blockIdx.x] = t;
        __syncthreads();  // s_rows is reused by the next window
    }
}

__global__ void k_reset(GridState* gs) {
    if (threadIdx.x != 0) return;
    gs->passes_max = 0;
    gs->passes_short = 0;
    gs->items_max = 0;
    gs->items_short = 0;
    gs->n_items = 0;
    gs->n_dense = 0;
    gs->n_surv = 0;
    gs->surv_max = 0;
    gs->surv_short = 0;
    gs->vmax_bits = 0ULL;
    for (int a = 0; a < 3; ++a) {
        gs->lo[a] = ~0ULL;
        gs->hi[a] = 0ULL;
    }
    gs->alive = 0;
    gs->propagated = 0;
    gs->cand = gs->refined = gs->co = 0ULL;
    gs->max_cell = 0;
    gs->ncells = 0;
    gs->npasses = 0;
    gs->entries_short = 0;
    gs->entries_max = 0;
}

__global__ void k_zero_events(GridState* gs) {
    if (threadIdx.x == 0) gs->events = 0ULL;
}

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------

// A screening's statistics record (k_finish), passed to the next screening's
// k_grid_setup to be written there, one launch fewer (out == nullptr: none).
struct FinishArgs {
    DevStats* out;
    const PairPartial* parts;
    int nparts, npair;
    const PairPartial* dparts;
    int ndparts;
    const PairPartial* eparts;
    int neparts, fill;
};

__device__ __forceinline__ void finish_window(GridState* gs, DevStats* out, int screened,
                                                   const PairPartial* parts, int nparts, int npair,
                                                   const PairPartial* dparts, int ndparts,
                                                   const PairPartial* eparts, int neparts, int fill) {
    __shared__ unsigned long long s_red[TWAVES][6];
    unsigned long long a = 0, b = 0, c = 0, m = 0, dn = 0, sv = 0;
    if (screened) {
        // max_cell from the k_neighbours blocks, candidates from the k_pairs
        // blocks. Loads go in batches of FIN_BATCH per thread, so a thread
        // waits once per batch rather than once per record.
        constexpr int FIN_BATCH = 8;
        const int nt = static_cast<int>(blockDim.x);
        const int np = nparts > npair ? nparts : npair;
        for (int k0 = threadIdx.x; k0 < np; k0 += nt * FIN_BATCH) {
            unsigned int mc[FIN_BATCH];
            unsigned long long cd[FIN_BATCH];
#pragma unroll
            for (int u = 0; u < FIN_BATCH; ++u) {
                const int k = k0 + u * nt;
                mc[u] = k < nparts ? parts[k].max_cell : 0u;
                cd[u] = k < npair ? parts[k].cand : 0ULL;
            }
#pragma unroll
            for (int u = 0; u < FIN_BATCH; ++u) {
                m = mc[u] > m ? mc[u] : m;
                a += cd[u];
            }
        }
        for (int k = threadIdx.x; k < ndparts; k += nt) dn += dparts[k].cand;
        for (int k = threadIdx.x; k < neparts; k += nt) {
            b += eparts[k].refined;
            c += eparts[k].co;
            sv += eparts[k].cand;
        }
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    m = wave_max_u64(m);
    dn = wave_sum(dn);
    sv = wave_sum(sv);
    const int wave = threadIdx.x / WAVE;
    if (__lane_id() == 0) {
        s_red[wave][0] = a;
        s_red[wave][1] = b;
        s_red[wave][2] = c;
        s_red[wave][3] = m;
        s_red[wave][4] = dn;
        s_red[wave][5] = sv;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int waves = static_cast<int>(blockDim.x) / WAVE;  // at most TWAVES
    for (int w = 1; w < waves; ++w) {
        s_red[0][0] += s_red[w][0];
        s_red[0][1] += s_red[w][1];
        s_red[0][2] += s_red[w][2];
        s_red[0][3] = s_red[w][3] > s_red[0][3] ? s_red[w][3] : s_red[0][3];
        s_red[0][4] += s_red[w][4];
        s_red[0][5] += s_red[w][5];
    }
    if (screened) {
        DevStats s;
        s.cand = static_cast<long long>(s_red[0][0] + s_red[0][4]);
        s.dense_cand = static_cast<long long>(s_red[0][4]);
        s.survivors = static_cast<long long>(s_red[0][5]);
        s.refined = static_cast<long long>(s_red[0][1]);
        s.co = static_cast<long long>(s_red[0][2]);
        gs->max_cell = static_cast<unsigned int>(s_red[0][3]);
        s.cell_km = gs->cell_ref;
        s.max_cell = static_cast<int>(gs->max_cell);
        s.doublings = gs->doublings;
        s.propagated = static_cast<long long>(gs->propagated);
        *out = s;
        // Records of the window's later sub-steps: zero counts, objects counted.
        for (int j = 1; j <= fill; ++j) {
            DevStats f{};
            f.cell_km = s.cell_km;
            f.propagated = s.propagated;
            out[j] = f;
        }
    }
    gs->propagated = 0;
    gs->vmax_bits = 0ULL;
    for (int a = 0; a < 3; ++a) {
        gs->lo[a] = ~0ULL;
        gs->hi[a] = 0ULL;
    }
    gs->alive = 0;
    gs->cand = gs->refined = gs->co = 0ULL;
    gs->max_cell = 0;
    gs->ncells = 0;
    gs->npasses = 0;
    gs->n_items = 0;
    gs->n_dense = 0;
    if (gs->n_surv > gs->surv_max) gs->surv_max = gs->n_surv;
    gs->n_surv = 0;
}

__device__ __forceinline__ int bits_for(long long range) {
    return range <= 0 ? 0 : 64 - __clzll(static_cast<unsigned long long>(range));
}

// Shell screening arguments of k_grid_setup and k_keys_shell (on == 0: the
// plain grid, one sorted element per object).
struct ShellScreen {
    int on;
    const int* tile_cnt;  // entries per object tile (k_windows), cleared here after use
    int* tile_base;       // their exclusive prefix: first entry of each object tile
    int ntiles;           // object tiles
    unsigned int ecap;    // entries the sort and cell kernels were launched for
};

// midpoint != 0: the grid is built on the midpoint of each object's start and
// end positions (learned form); a pair that comes within thr during the step
// then has midpoints closer than vmax dt + thr + G vmax dt^3 (each object is
// within vmax dt / 2 of its midpoint over the step), so cells of that size
// and the 13 forward neighbours enumerate every such pair.
// nsort: elements the plain grid sorts (every object slot, dead ones last).
// With shells, the sorted elements are the shell entries: the packed key
// gains the shell (offset from the innermost, bbits wide) between y and z,
// so neighbour cells stay z-runs of one shell, and the entry count comes
// from the prefix of the object tiles' counts.
__global__ void __launch_bounds__(TBLOCK) k_grid_setup(GridState* gs, const StepPartial* partials,
                                                       int nparts, double dt, double thr,
                                                       int launched_passes, int midpoint, int sbits,
                                                       TileCfg tc, FinishArgs fin, int nsort, ShellScreen sh,
                                                       const Bands* bands) {
    __shared__ StepPartial s_rows[TWAVES];
    if (bands) thr = bands->thr_search;
    __shared__ int s_wave[TWAVES];
    long long entries = 0;
    if (sh.on) {
        // Exclusive prefix of the object tiles' entry counts (every thread).
        long long carry = 0;
        for (int t0 = 0; t0 < sh.ntiles; t0 += TBLOCK) {
            const int t = t0 + static_cast<int>(threadIdx.x);
            const int v = t < sh.ntiles ? sh.tile_cnt[t] : 0;
            int sum;
            const int ex = block_exclusive_scan<TWAVES>(v, sum, s_wave);
            if (t < sh.ntiles) {
                const long long at = carry + ex;
                sh.tile_base[t] = at < static_cast<long long>(sh.ecap) ? static_cast<int>(at) : static_cast<int>(sh.ecap);
                const_cast<int*>(sh.tile_cnt)[t] = 0;
            }
            carry += sum;
        }
        entries = carry;
    }
    // The previous screening's record first: it reads that screening's grid
    // fields, which thread 0 then overwrites below.
    if (fin.out)
        finish_window(gs, fin.out, 1, fin.parts, fin.nparts, fin.npair, fin.dparts, fin.ndparts, fin.eparts,
                      fin.neparts, fin.fill);
    StepPartial acc = empty_step_partial();
    for (int b = threadIdx.x; b < nparts; b += TBLOCK) combine(acc, partials[b]);
    const StepPartial tot = block_reduce<TWAVES>(acc, s_rows);
    if (threadIdx.x != 0) return;
    gs->vmax_bits = tot.vmax_bits;
    for (int a = 0; a < 3; ++a) {
        gs->lo[a] = tot.lo[a];
        gs->hi[a] = tot.hi[a];
    }
    gs->alive = tot.alive;
    gs->propagated = tot.propagated;
    if (sh.on) {
        if (entries > static_cast<long long>(sh.ecap)) gs->entries_short = 1;
        const unsigned int e = static_cast<unsigned int>(entries < static_cast<long long>(sh.ecap) ? entries : sh.ecap);
        if (static_cast<unsigned int>(entries < 0xffffffffLL ? entries : 0xffffffffLL) > gs->entries_max)
            gs->entries_max = static_cast<unsigned int>(entries < 0xffffffffLL ? entries : 0xffffffffLL);
        gs->nsort = gs->nlive = e;
        gs->bmin = tot.bmin;
        gs->bbits = tot.alive ? bits_for(static_cast<long long>(tot.bmax) - static_cast<long long>(tot.bmin)) : 0;
    } else {
        gs->nsort = static_cast<unsigned int>(nsort);
        gs->nlive = tot.alive;
        gs->bmin = 0;
        gs->bbits = 0;
    }
    const double vmax = __longlong_as_double(static_cast<long long>(gs->vmax_bits));
    const double vrel_max = midpoint ? vmax : 2.0 * vmax;
    const double raw = vrel_max * dt + thr + GRAVITY_GRADIENT * vrel_max * dt * dt * dt;
    double cell = raw > 1.0 ? raw : 1.0;
    gs->cell_ref = cell;
    cell = cell * tc.cell_mult;
    gs->dense_min = tc.dense_min;
    gs->long_range = tc.long_range;
    gs->seg = tc.seg;
    gs->doublings = 0;
    gs->stamp += 1;
    if (gs->alive == 0) {
        gs->cell = cell;
        gs->npasses = 0;
        return;
    }
    const double lo[3] = {from_ordered_bits(gs->lo[0]), from_ordered_bits(gs->lo[1]),
                          from_ordered_bits(gs->lo[2])};
    const double hi[3] = {from_ordered_bits(gs->hi[0]), from_ordered_bits(gs->hi[1]),
                          from_ordered_bits(gs->hi[2])};
    // floor(x / cell) is monotone in x, so the extreme cells come from the
    // extreme coordinates. A cell larger than the reference's never changes
    // which pairs pass the filters (collision.cpp:337-338); it is doubled
    // only if the packed key would not fit.
    // Concatenated scenarios: the scenario id is the key's top field, so
    // cells (and the 13 forward neighbours) never mix scenarios.
    gs->sbits = sbits;
    for (;;) {
        int total = sbits + gs->bbits;
        for (int a = 0; a < 3; ++a) {
            const long long i0 = static_cast<long long>(floor(lo[a] / cell));
            const long long i1 = static_cast<long long>(floor(hi[a] / cell));
            gs->imin[a] = i0;
            gs->bits[a] = bits_for(i1 - i0);
            total += gs->bits[a];
        }
        if (total <= KEY_BITS_MAX) {
            // Dead objects carry the all-ones key and must sort after every
            // live one (the sort is stable, so equal keys would keep index
            // order). A live key is all ones in its low `total` bits only
            // if its top cell field is; while that field's largest value
            // stays below all ones, `total` sorted bits keep dead keys
            // last, otherwise one more bit is sorted. Shell entries are
            // all live: `total` bits.
            const long long top = static_cast<long long>(floor(hi[0] / cell)) - gs->imin[0];
            const bool top_full = !sh.on && top == (1LL << gs->bits[0]) - 1;
            gs->npasses = (total + (top_full ? 1 : 0) + RADIX_BITS - 1) / RADIX_BITS;
            break;
        }
        cell = cell * 2.0;
        gs->doublings += 1;
    }
    gs->cell = cell;
    if (gs->npasses > gs->passes_max) gs->passes_max = gs->npasses;
    if (gs->npasses > launched_passes) gs->passes_short = 1;
}

// Also writes, in object order (coalesced), each alive object's compact
// 32-byte FP32 record for the pair kernels: position relative to its cell's
// low corner (FP64 subtraction, then rounding: the error DR_ERR_KM bounds)
// with the object index in w, then velocity with the cell's packed z
// coordinate in w. k_cell_write gathers one record (one cache line) per
// sorted object.
template <int TI = TITEMS>
__global__ void __launch_bounds__(TBLOCK) k_keys(Planes cur, Planes end, int midpoint, const uint8_t* alive,
                                                 const int* __restrict__ scen, int n, const GridState* gs,
                                                 uint64_t* keys, int* idx, int* hist, int ntiles, float4* rec) {
    __shared__ int s_hist[RADIX];
    const GridState g = *gs;
    if (g.npasses == 0) return;
    if (threadIdx.x < RADIX) s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * (TI * TBLOCK);
    for (int r = 0; r < TI; ++r) {
        const int e = base + r * TBLOCK + threadIdx.x;
        if (e < n) {
            uint64_t k = DEAD_KEY;
            if (alive[e]) {
                const double x = cur.x[e], y = cur.y[e], z = cur.z[e];
                const double gx = midpoint ? 0.5 * (x + end.x[e]) : x;
                const double gy = midpoint ? 0.5 * (y + end.y[e]) : y;
                const double gz = midpoint ? 0.5 * (z + end.z[e]) : z;
                const double c = g.cell;
                const long long ix = static_cast<long long>(floor(gx / c));
                const long long iy = static_cast<long long>(floor(gy / c));
                const long long iz = static_cast<long long>(floor(gz / c));
                const uint64_t pz = static_cast<uint64_t>(iz - g.imin[2]);
                k = (static_cast<uint64_t>(ix - g.imin[0]) << (g.bits[1] + g.bits[2])) |
                    (static_cast<uint64_t>(iy - g.imin[1]) << g.bits[2]) | pz;
                if (g.sbits)
                    k |= static_cast<uint64_t>(scen[e]) << (g.bits[0] + g.bits[1] + g.bits[2]);
                rec[2 * e] = make_float4(static_cast<float>(x - static_cast<double>(ix) * c),
                                         static_cast<float>(y - static_cast<double>(iy) * c),
                                         static_cast<float>(z - static_cast<double>(iz) * c), __int_as_float(e));
                rec[2 * e + 1] = make_float4(static_cast<float>(cur.vx[e]), static_cast<float>(cur.vy[e]),
                                             static_cast<float>(cur.vz[e]), __int_as_float(static_cast<int>(pz)));
            }
            keys[e] = k;
            idx[e] = e;
            atomicAdd(&s_hist[static_cast<int>(k & DIGIT_MASK)], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < RADIX) hist[threadIdx.x * ntiles + blockIdx.x] = s_hist[threadIdx.x];
}

// Shell entries of one object tile (TILE objects, the tiles k_windows
// counted): each alive object's compact record in object order (as k_keys
// writes it) and one (key, index word) per shell of the object, at the
// tile's first entry (k_grid_setup) plus the exclusive prefix of the counts
// within the tile. The key is the cell's packed key with the shell offset
// between y and z; the index word is the object index, NOT_FIRST set on
// every entry but the first.
__global__ void __launch_bounds__(TBLOCK) k_keys_shell(Planes cur, Planes end, const uint8_t* alive,
                                                       const int* __restrict__ scen, int n, const GridState* gs,
                                                       const int* __restrict__ shell_lo,
                                                       const uint16_t* __restrict__ shell_n,
                                                       const int* __restrict__ tile_base, unsigned int ecap,
                                                       uint64_t* keys, int* idx, float4* rec) {
    __shared__ int s_wave[TWAVES];
    const GridState g = *gs;
    if (g.npasses == 0) return;
    const int base = blockIdx.x * TILE;
    const int bz = g.bits[2], bb = g.bbits;
    unsigned int off = static_cast<unsigned int>(tile_base[blockIdx.x]);
    for (int r = 0; r < TITEMS; ++r) {
        const int e = base + r * TBLOCK + threadIdx.x;
        int nb = 0, s0 = 0;
        uint64_t k = 0;
        if (e < n && alive[e]) {
            nb = shell_n[e];
            s0 = shell_lo[e];
            const double x = cur.x[e], y = cur.y[e], z = cur.z[e];
            const double gx = 0.5 * (x + end.x[e]), gy = 0.5 * (y + end.y[e]), gz = 0.5 * (z + end.z[e]);
            const double c = g.cell;
            const long long ix = static_cast<long long>(floor(gx / c));
            const long long iy = static_cast<long long>(floor(gy / c));
            const long long iz = static_cast<long long>(floor(gz / c));
            const uint64_t pz = static_cast<uint64_t>(iz - g.imin[2]);
            k = (static_cast<uint64_t>(ix - g.imin[0]) << (g.bits[1] + bb + bz)) |
                (static_cast<uint64_t>(iy - g.imin[1]) << (bb + bz)) | pz;
            if (g.sbits) k |= static_cast<uint64_t>(scen[e]) << (g.bits[0] + g.bits[1] + bb + bz);
            rec[2 * e] = make_float4(static_cast<float>(x - static_cast<double>(ix) * c),
                                     static_cast<float>(y - static_cast<double>(iy) * c),
                                     static_cast<float>(z - static_cast<double>(iz) * c), __int_as_float(e));
            rec[2 * e + 1] = make_float4(static_cast<float>(cur.vx[e]), static_cast<float>(cur.vy[e]),
                                         static_cast<float>(cur.vz[e]), __int_as_float(static_cast<int>(pz)));
        }
        int tot;
        const int ex = block_exclusive_scan<TWAVES>(nb, tot, s_wave);
        for (int j = 0; j < nb; ++j) {
            const unsigned int pos = off + static_cast<unsigned int>(ex + j);
            if (pos >= ecap) break;
            keys[pos] = k | (static_cast<uint64_t>(s0 + j - static_cast<int>(g.bmin)) << bz);
            idx[pos] = e | (j ? NOT_FIRST : 0);
        }
        off += static_cast<unsigned int>(tot);
    }
}

template <int TI = TITEMS>
__global__ void __launch_bounds__(TBLOCK) k_hist(const uint64_t* keys, int pass, const GridState* gs,
                                                 int* hist, int ntiles) {
    __shared__ int s_hist[RADIX];
    if (pass >= gs->npasses) return;
    const int n = static_cast<int>(gs->nsort);
    if (threadIdx.x < RADIX) s_hist[threadIdx.x] = 0;
    __syncthreads();
    const int shift = RADIX_BITS * pass;
    const int base = blockIdx.x * (TI * TBLOCK);
    for (int r = 0; r < TI; ++r) {
        const int e = base + r * TBLOCK + threadIdx.x;
        if (e < n) atomicAdd(&s_hist[static_cast<int>((keys[e] >> shift) & DIGIT_MASK)], 1);
    }
    __syncthreads();
    if (threadIdx.x < RADIX) hist[threadIdx.x * ntiles + blockIdx.x] = s_hist[threadIdx.x];
}

// One block per digit: exclusive scan of that digit's per-tile counts.
__global__ void __launch_bounds__(BLOCK) k_digit_scan(int* hist, int ntiles, int pass,
                                                       const GridState* gs, int* totals) {
    __shared__ int s_wave[WAVES_PER_BLOCK];
    if (pass >= gs->npasses) return;
    int* row = hist + blockIdx.x * ntiles;
    int carry = 0;
    for (int b0 = 0; b0 < ntiles; b0 += BLOCK) {
        const int b = b0 + threadIdx.x;
        const int v = b < ntiles ? row[b] : 0;
        int tot;
        const int ex = block_exclusive_scan(v, tot, s_wave);
        if (b < ntiles) row[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter of one radix pass. In-tile ranks: elements are taken in
// index order, TBLOCK per round; a wave finds the lanes holding the same
// digit with RADIX_BITS ballots, and per-wave digit counts in LDS give the
// cross-wave offsets.
template <int TI = TITEMS>
__global__ void __launch_bounds__(TBLOCK) k_scatter(const uint64_t* keys_in, const int* idx_in,
                                                     uint64_t* keys_out, int* idx_out,
                                                     int pass, const GridState* gs,
                                                     const int* hist, const int* totals,
                                                     int ntiles) {
    __shared__ int s_off[RADIX];
    __shared__ int s_wcnt[TWAVES][RADIX];
    __shared__ int s_wave[TWAVES];
    if (pass >= gs->npasses) return;
    const int n = static_cast<int>(gs->nsort);
    const int t = threadIdx.x;
    {
        int tot;
        const int digit_base = block_exclusive_scan<TWAVES>(t < RADIX ? totals[t] : 0, tot, s_wave);
        if (t < RADIX) {
            s_off[t] = digit_base + hist[t * ntiles + blockIdx.x];
            for (int w = 0; w < TWAVES; ++w) s_wcnt[w][t] = 0;
        }
    }
    __syncthreads();
    const int shift = RADIX_BITS * pass;
    const int wave = t / WAVE;
    const unsigned long long below = lanes_below();
    const int base = blockIdx.x * (TI * TBLOCK);
    uint64_t keys[TI];
    int ids[TI];
#pragma unroll
    for (int r = 0; r < TI; ++r) {  // all of the tile's loads in flight at once
        const int e = base + r * TBLOCK + t;
        keys[r] = e < n ? keys_in[e] : 0ULL;
        ids[r] = e < n ? idx_in[e] : 0;
    }
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        const int e = base + r * TBLOCK + t;
        const bool valid = e < n;
        const uint64_t key = keys[r];
        const int id = ids[r];
        const int d = static_cast<int>((key >> shift) & DIGIT_MASK);
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < RADIX_BITS; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) s_wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pre = 0;
            for (int w = 0; w < wave; ++w) pre += s_wcnt[w][d];
            const int pos = s_off[d] + pre + rank;
            keys_out[pos] = key;
            idx_out[pos] = id;
        }
        __syncthreads();
        if (t < RADIX) {
            int add = 0;
            for (int w = 0; w < TWAVES; ++w) {
                add += s_wcnt[w][t];
                s_wcnt[w][t] = 0;
            }
            s_off[t] += add;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ const uint64_t* sorted_keys(const GridState& g, const uint64_t* k0,
                                                       const uint64_t* k1) {
    return (g.npasses & 1) ? k1 : k0;
}
__device__ __forceinline__ const int* sorted_idx(const GridState& g, const int* i0, const int* i1) {
    return (g.npasses & 1) ? i1 : i0;
}

// Cell heads per tile over the first `nlive` sorted elements.
template <int TI = TITEMS>
__global__ void __launch_bounds__(TBLOCK) k_cell_count(const uint64_t* k0, const uint64_t* k1,
                                                        const GridState* gs, int* bcount) {
    __shared__ int s_wave[TWAVES];
    const GridState& g = *gs;
    const int A = static_cast<int>(g.nlive);
    const uint64_t* keys = sorted_keys(g, k0, k1);
    const int base = blockIdx.x * (TI * TBLOCK);
    int cnt = 0;
    if (base < A && g.npasses > 0) {
        for (int r = 0; r < TI; ++r) {
            const int e = base + r * TBLOCK + threadIdx.x;
            if (e < A && (e == 0 || keys[e] != keys[e - 1])) ++cnt;
        }
    }
    int tot;
    block_exclusive_scan<TWAVES>(cnt, tot, s_wave);
    if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

// Cell table and compact records in sorted order. Element e = tile +
// r * TBLOCK + thread (r < TITEMS), so every load and store is coalesced;
// cells are numbered from wave ballots of the head flags and one block-wide
// prefix over the tile's TITEMS x TWAVES wave segments (in element order).
// Each sorted object's record is two float4 in lrec (32 bytes, one place):
//   lrec[2e]:     position relative to its cell's low corner, w = object
//                 index bits;
//   lrec[2e + 1]: velocity, w = the cell's packed z coordinate (int bits),
//                 from which pairs get their z cell offset.
template <int TI = TITEMS>
__global__ void __launch_bounds__(TBLOCK) k_cell_write(const uint64_t* k0, const uint64_t* k1,
                                                       const int* i0, const int* i1,
                                                       GridState* gs, const int* bcount, int ntiles,
                                                       const float4* __restrict__ rec, uint64_t* cell_key,
                                                       int* cell_begin, int* cell_of,
                                                       float4* lrec, int4* __restrict__ coltab) {
    __shared__ int s_seg[TI * TWAVES];
    __shared__ int s_wave[TWAVES];
    const GridState& g = *gs;
    const int A = static_cast<int>(g.nlive);
    const int base = blockIdx.x * (TI * TBLOCK);
    if (base >= A || g.npasses == 0) return;
    // Cells of the earlier tiles: the sum of their head counts (k_cell_count),
    // a few loads per thread; block 0 sums every tile and writes the total.
    int run;
    {
        const int upto = blockIdx.x == 0 ? ntiles : static_cast<int>(blockIdx.x);
        int part = 0;
        for (int k = threadIdx.x; k < upto; k += TBLOCK) part += bcount[k];
        int tot;
        block_exclusive_scan<TWAVES>(part, tot, s_wave);
        run = blockIdx.x == 0 ? 0 : tot;
        if (blockIdx.x == 0 && threadIdx.x == 0) gs->ncells = static_cast<unsigned int>(tot);
    }
    const uint64_t* keys = sorted_keys(g, k0, k1);
    const int* sidx = sorted_idx(g, i0, i1);
    const int lane = static_cast<int>(__lane_id());
    const int wave = threadIdx.x / WAVE;
    const unsigned long long below = lanes_below();
    uint64_t key[TI], prv[TI];
    unsigned long long bal[TI];
    bool head[TI];
    float4 p[TI], v[TI];
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        const int e = base + r * TBLOCK + threadIdx.x;
        const bool valid = e < A;
        key[r] = valid ? keys[e] : 0ULL;
        uint64_t prev = __shfl_up(key[r], 1, WAVE);
        if (lane == 0) prev = valid && e > 0 ? keys[e - 1] : ~0ULL;
        prv[r] = prev;
        head[r] = valid && (e == 0 || key[r] != prev);
        bal[r] = __ballot(head[r]);
        if (lane == 0) s_seg[r * TWAVES + wave] = __popcll(bal[r]);
        if (valid) {
            // The index word (object index, and NOT_FIRST for shell entries)
            // replaces the record's w.
            const int id = sidx[e];
            p[r] = rec[2 * (id & OBJ_MASK)];
            v[r] = rec[2 * (id & OBJ_MASK) + 1];
            p[r].w = __int_as_float(id);
        }
    }
    __syncthreads();
    int pre[TI];
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        for (int w = 0; w < TWAVES; ++w) {
            if (w == wave) pre[r] = run;
            run += s_seg[r * TWAVES + w];
        }
    }
    // Column table (when the column index fits COL_BITS): the first cell of
    // a column opens its entry and closes the previous column's; the last
    // element closes its own column.
    const bool use_tab = coltab != nullptr && g.sbits + g.bits[0] + g.bits[1] <= COL_BITS;
    const int bz = g.bits[2] + g.bbits;  // a column: one (x, y) (and shell) run
#pragma unroll
    for (int r = 0; r < TI; ++r) {
        const int e = base + r * TBLOCK + threadIdx.x;
        if (e >= A) break;
        const int c = pre[r] + __popcll(bal[r] & below) + (head[r] ? 1 : 0) - 1;
        if (head[r]) {
            cell_key[c] = key[r];
            cell_begin[c] = e;
            const uint64_t col = key[r] >> bz;
            if (use_tab && (e == 0 || (prv[r] >> bz) != col)) {
                coltab[col].x = g.stamp;
                coltab[col].y = c;
                if (e > 0) coltab[prv[r] >> bz].z = c;
            }
        }
        if (use_tab && e == A - 1) coltab[key[r] >> bz].z = c + 1;
        cell_of[e] = c;
        lrec[2 * e] = p[r];
        lrec[2 * e + 1] = v[r];
    }
}

// ---------------------------------------------------------------------------
// Neighbour ranges
// ---------------------------------------------------------------------------
//
// Packed keys hold z in their lowest bits, so the cells (x', y', z-1),
// (x', y', z) and (x', y', z+1) that exist are consecutive in key order and
// their objects form one contiguous run of the sorted array. The 13 forward
// neighbours of cell (x, y, z) (collision_filter.hpp:137-140) are therefore
// four runs, for (dx, dy) = (1, -1), (1, 0), (1, 1), (0, 1), plus the single
// cell (0, 0, 1), which directly follows cell (x, y, z) in key order and
// merges with the cell's own tail. Each cell keeps NRANGES (begin, end)
// runs of sorted positions:
//   range 0   the cell itself and, if present, cell (x, y, z+1): a lane at
//             sorted position p pairs with positions q > p only;
//   range 1-4 the three-cell runs above; every object pairs.
// The candidate set is the reference's (intra-cell pairs + 13 forward
// cells), each pair once.

constexpr int NRANGES = 5;

__device__ __forceinline__ void range_offset(int r, int& dx, int& dy) {
    dx = r == 4 ? 0 : (r == 0 ? 0 : 1);
    dy = r == 0 ? 0 : (r == 4 ? 1 : r - 2);
}

__device__ __forceinline__ bool axis_fits(long long v, int bits) {
    return v >= 0 && (bits >= 62 || v < (1LL << bits));
}


__device__ __forceinline__ int segments(int n, int seg) { return (n + seg - 1) / seg; }

// Dense cell with k objects at sorted positions [b, b + k): per chunk of
// CHUNK own objects, the range-0 tail from the chunk start plus ranges 1-4,
// each cut into SEG-long segments.
// Items of a sparse cell (k < dense_min objects at [b, b + k)): its ranges
// longer than long_range objects (range 0 counted from the cell start), per
// chunk of CHUNK own objects (a wave's lanes hold one chunk), cut into
// seg-long segments; range 0 of chunk i starts at the chunk.
__device__ __forceinline__ int sparse_chunk_items(int b, int i, const int2* rng, int long_range, int seg) {
    int items = 0;
    for (int r = 0; r < NRANGES; ++r) {
        const int len = rng[r].y - rng[r].x;
        if (len > long_range) items += segments(r == 0 ? rng[0].y - (b + i * CHUNK) : len, seg);
    }
    return items;
}

__device__ __forceinline__ int long_range_items(int b, int k, const int2* rng, int long_range, int seg) {
    int items = 0;
    for (int i = 0; i * CHUNK < k; ++i) items += sparse_chunk_items(b, i, rng, long_range, seg);
    return items;
}

__device__ __forceinline__ int dense_items(int b, int k, const int2* rng, int seg) {
    int fixed = 0;
    for (int r = 1; r < NRANGES; ++r) fixed += segments(rng[r].y - rng[r].x, seg);
    int items = 0;
    for (int i = 0; i * CHUNK < k; ++i) items += segments(rng[0].y - (b + i * CHUNK), seg) + fixed;
    return items;
}

// Column table (written by k_cell_write): for every occupied (x, y) column
// of the packed key space, the run [begin, end) of its cells in the sorted
// cell table, tagged with the screening's stamp (entries of earlier
// screenings read as empty, so the table is never cleared). Used when the
// column index fits COL_BITS.


__global__ void __launch_bounds__(BLOCK) k_neighbours(GridState* gs, const uint64_t* cell_key,
                                                       const int* cell_begin, int* cell_count,
                                                       int2* ranges, PairPartial* parts,
                                                       DenseCell* dense, unsigned int item_cap,
                                                       const int4* __restrict__ coltab) {
    __shared__ unsigned int s_max[WAVES_PER_BLOCK];
    __shared__ int s_wave[WAVES_PER_BLOCK];
    __shared__ int s_base[2];
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    const int nc = static_cast<int>(gs->ncells);
    const int A = static_cast<int>(gs->nlive);
    unsigned int cnt = 0;
    int items = 0;
    if (c < nc) {
        const int b = cell_begin[c];
        const int e = c + 1 < nc ? cell_begin[c + 1] : A;
        cnt = static_cast<unsigned int>(e - b);
        cell_count[c] = static_cast<int>(cnt);
        const int bx = gs->bits[0], by = gs->bits[1], bz = gs->bits[2];
        // Low field: the shell (bbits, shell screening) above z; neighbour
        // cells share the shell, so their z-runs stay consecutive keys.
        const int zf = bz + gs->bbits;
        const bool use_tab = coltab != nullptr && gs->sbits + bx + by <= COL_BITS;
        const int stamp = gs->stamp;
        const uint64_t key = cell_key[c];
        const uint64_t mz = (1ULL << bz) - 1ULL;
        const uint64_t shell_part = key & (((1ULL << zf) - 1ULL) ^ mz);
        const uint64_t scen_part = gs->sbits ? (key >> (bx + by + zf)) << (bx + by + zf) : 0ULL;
        const long long x = static_cast<long long>((key >> (by + zf)) & ((1ULL << bx) - 1ULL));
        const long long y = static_cast<long long>((key >> zf) & ((1ULL << by) - 1ULL));
        const long long z = static_cast<long long>(key & mz);
        int2 rng[NRANGES];
        // Range 0: own cell, then (x, y, z+1) if it exists (the next key).
        int e0 = e;
        if (c + 1 < nc && cell_key[c + 1] == key + 1ULL && (key & mz) != mz)
            e0 = c + 2 < nc ? cell_begin[c + 2] : A;
        rng[0] = make_int2(b, e0);
        const long long zlo = z > 0 ? z - 1 : 0;
        const long long zhi = axis_fits(z + 1, bz) ? z + 1 : z;
        // Ranges 1-4: the cells (x+dx, y+dy, zlo..zhi), a run of at most three
        // consecutive cells of that column. The four lower-bound searches run
        // in lockstep (their loads are issued together); the run's end is
        // among the next three cells.
        uint64_t lo_key[NRANGES], hi_key[NRANGES];
        int base[NRANGES], len[NRANGES];
        bool valid[NRANGES];
#pragma unroll
        for (int r = 1; r < NRANGES; ++r) {
            int dx, dy;
            range_offset(r, dx, dy);
            const long long nx = x + dx, ny = y + dy;
            base[r] = 0;
            len[r] = 0;
            lo_key[r] = hi_key[r] = 0ULL;
            valid[r] = axis_fits(nx, bx) && axis_fits(ny, by);
            if (valid[r]) {
                const uint64_t row = scen_part | (static_cast<uint64_t>(nx) << (by + zf)) |
                                     (static_cast<uint64_t>(ny) << zf);
                lo_key[r] = row | shell_part | static_cast<uint64_t>(zlo);
                hi_key[r] = row | shell_part | static_cast<uint64_t>(zhi);
                len[r] = nc;
                if (use_tab) {
                    const int4 t = coltab[row >> zf];
                    valid[r] = t.x == stamp;  // the column has cells in this screening
                    base[r] = valid[r] ? t.y : 0;
                    len[r] = valid[r] ? t.z - t.y : 0;
                }
            }
        }
        for (;;) {
            bool more = false;
#pragma unroll
            for (int r = 1; r < NRANGES; ++r) more |= len[r] > 0;
            if (!more) break;
            uint64_t probe[NRANGES];
#pragma unroll
            for (int r = 1; r < NRANGES; ++r) probe[r] = len[r] > 0 ? cell_key[base[r] + (len[r] >> 1)] : 0ULL;
#pragma unroll
            for (int r = 1; r < NRANGES; ++r) {
                if (len[r] <= 0) continue;
                const int half = len[r] >> 1;
                if (probe[r] < lo_key[r]) {
                    base[r] += half + 1;
                    len[r] -= half + 1;
                } else {
                    len[r] = half;
                }
            }
        }
#pragma unroll
        for (int r = 1; r < NRANGES; ++r) {
            // base[r]: first cell >= (x+dx, y+dy, zlo) (or a cell of another column).
            const int i0 = base[r];
            int i1 = i0;
            if (valid[r]) {
                uint64_t k3[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) k3[m] = i0 + m < nc ? cell_key[i0 + m] : ~0ULL;
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    if (i1 == i0 + m && k3[m] >= lo_key[r] && k3[m] <= hi_key[r]) ++i1;
            }
            rng[r] = i0 < i1 ? make_int2(cell_begin[i0], i1 < nc ? cell_begin[i1] : A) : make_int2(0, 0);
        }
        for (int r = 0; r < NRANGES; ++r) ranges[c * NRANGES + r] = rng[r];
        items = static_cast<int>(cnt) >= gs->dense_min
                    ? dense_items(b, static_cast<int>(cnt), rng, gs->seg)
                    : long_range_items(b, static_cast<int>(cnt), rng, gs->long_range, gs->seg);
    }
    // Block-aggregated registration of dense cells: two atomics per block.
    int tot_items, tot_dense;
    const int ex_items = block_exclusive_scan(items, tot_items, s_wave);
    const int ex_dense = block_exclusive_scan(items > 0 ? 1 : 0, tot_dense, s_wave);
    if (threadIdx.x == 0) {
        s_base[0] = tot_items ? static_cast<int>(atomicAdd(&gs->n_items, static_cast<unsigned int>(tot_items))) : 0;
        s_base[1] = tot_dense ? static_cast<int>(atomicAdd(&gs->n_dense, static_cast<unsigned int>(tot_dense))) : 0;
    }
    __syncthreads();
    if (items > 0) {
        const int base = s_base[0] + ex_items;
        dense[s_base[1] + ex_dense] = DenseCell{c, base};
        if (static_cast<unsigned int>(base + items) > item_cap) gs->items_short = 1;
    }
    const unsigned long long m = wave_max_u64(cnt);
    if (__lane_id() == 0) s_max[threadIdx.x / WAVE] = static_cast<unsigned int>(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int mm = 0;
        for (int w = 0; w < WAVES_PER_BLOCK; ++w) mm = s_max[w] > mm ? s_max[w] : mm;
        parts[blockIdx.x].max_cell = mm;
    }
}

// Writes the work items of every dense cell: one block per dense cell
// (grid-stride), one thread per chunk.
__global__ void __launch_bounds__(BLOCK) k_items(GridState* gs, const DenseCell* dense,
                                                  const int* cell_begin, const int* cell_count,
                                                  const int2* ranges, PairItem* items,
                                                  unsigned int item_cap) {
    const int nd = static_cast<int>(gs->n_dense);
    const int seg = gs->seg, long_range = gs->long_range, dmin = gs->dense_min;
    if (blockIdx.x == 0 && threadIdx.x == 0 && gs->n_items > gs->items_max) gs->items_max = gs->n_items;
    for (int d = blockIdx.x; d < nd; d += gridDim.x) {
        const DenseCell dc = dense[d];
        const int c = dc.cell;
        const int b = cell_begin[c];
        const int k = cell_count[c];
        int2 rng[NRANGES];
        for (int r = 0; r < NRANGES; ++r) rng[r] = ranges[c * NRANGES + r];
        if (k < dmin) {
            // Sparse cell: only its long ranges, one thread per chunk of own objects.
            for (int i = threadIdx.x; i * CHUNK < k; i += BLOCK) {
                int off = dc.item_base;
                for (int j = 0; j < i; ++j) off += sparse_chunk_items(b, j, rng, long_range, seg);
                const int own_begin = b + i * CHUNK;
                const int own_count = min(CHUNK, k - i * CHUNK);
                for (int r = 0; r < NRANGES; ++r) {
                    if (rng[r].y - rng[r].x <= long_range) continue;
                    const int r0 = r == 0 ? own_begin : rng[r].x;
                    const int len = rng[r].y - r0;
                    for (int sgi = 0; sgi * seg < len; ++sgi, ++off) {
                        if (static_cast<unsigned int>(off) >= item_cap) continue;
                        items[off] = PairItem{own_begin, own_count, r0 + sgi * seg, min(seg, len - sgi * seg), r,
                                              0, 0, 0};
                    }
                }
            }
            continue;
        }
        int fixed = 0;
        for (int r = 1; r < NRANGES; ++r) fixed += segments(rng[r].y - rng[r].x, seg);
        const int nch = (k + CHUNK - 1) / CHUNK;
        for (int i = threadIdx.x; i < nch; i += BLOCK) {
            int off = dc.item_base + i * fixed;
            for (int j = 0; j < i; ++j) off += segments(rng[0].y - (b + j * CHUNK), seg);
            const int own_begin = b + i * CHUNK;
            const int own_count = min(CHUNK, k - i * CHUNK);
            for (int r = 0; r < NRANGES; ++r) {
                const int r0 = r == 0 ? own_begin : rng[r].x;
                const int len = rng[r].y - r0;
                for (int sgi = 0; sgi * seg < len; ++sgi, ++off) {
                    if (static_cast<unsigned int>(off) >= item_cap) continue;
                    items[off] = PairItem{own_begin, own_count, r0 + sgi * seg, min(seg, len - sgi * seg),
                                          r, 0, 0, 0};
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pairs
// ---------------------------------------------------------------------------

// Time and index of one learned step, read by the kernels from a per-chunk
// device table so that a captured chunk graph replays for later chunks.
struct PairParams {
    Planes before, after;
    const Bands* bands;     // learned form with a drift band: thresholds of the step (else null)
    const double* radius;   // collision radius (km) per object
    const int* cell_of;
    const int* cell_begin;
    const int* cell_count;
    const int2* ranges;
    const float4* lrec;     // sorted compact records: position, velocity (2 float4 each)
    DevEvent* events;
    unsigned long long cap;
    uint2* surv;                    // pre-filter survivors (object index pairs)
    unsigned long long surv_cap;
    long long step;
    double t0, dt, thr, vmin2;
    const StepInfo* info;   // learned form: step and t0 from here (window sub_q of sub_n)
    int sub_q, sub_n;
    int sub_per;            // learned form: plain sub-steps per screening window
    float dt_f, thr_f, vmin2_f;
    float gdt3_f;   // GRAVITY_GRADIENT * dt^3
    float base_f;   // threshold + 1e-3 + DR_ERR_KM + DV_ERR_KM_S * dt
};

// Upper bound of the pre-filter's lim^2 over every candidate pair of a
// screening: partners lie within two cell edges of the own object per axis
// (neighbour cells), so |dr|_1 <= 6 cell, and |dv|_1 <= sqrt(3) |dv| <=
// 2 sqrt(3) vmax (vmax: the largest speed of the velocities in the records);
// lim grows with both. Safety factors cover FP32 rounding.
// The step's thresholds from the device (learned form with a drift band).
__device__ __forceinline__ PairParams with_bands(PairParams P) {
    if (P.bands) {
        const Bands b = *P.bands;
        P.thr = b.thr_search;
        P.vmin2 = b.vmin2_search;
        P.thr_f = b.thr_f;
        P.vmin2_f = b.vmin2_f;
        P.base_f = b.base_f;
    }
    return P;
}

__device__ __forceinline__ float pair_lim_max2(const GridState* gs, const PairParams& P) {
    const double vmax = __longlong_as_double(static_cast<long long>(gs->vmax_bits));
    const double dvn = 2.0 * 1.7320508075688772 * vmax * 1.001;
    const double rn = 6.01 * gs->cell;
    const double lim = static_cast<double>(P.base_f) + static_cast<double>(P.gdt3_f) * dvn +
                       1e-5 * (rn + dvn * static_cast<double>(P.dt_f));
    return static_cast<float>(lim * lim * 1.001);
}

struct PairCounters {
    unsigned int cand, refined, co;
};

__device__ __forceinline__ void append_event(GridState* gs, const PairParams& P, const DevEvent& ev) {
    // Wave-aggregated append among the lanes that reach this point together.
    const unsigned long long m = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    const int rank = __popcll(m & lanes_below());
    unsigned long long base = 0;
    if (static_cast<int>(__lane_id()) == leader) base = atomicAdd(&gs->events, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, leader, WAVE);
    const unsigned long long slot = base + rank;
    if (slot < P.cap) P.events[slot] = ev;
}

__device__ __forceinline__ void exact_pair(GridState* gs, const PairParams& P, int a, int b,
                                           PairCounters& cnt) {
    const int i = a < b ? a : b, j = a < b ? b : a;
    const D3 ri{P.before.x[i], P.before.y[i], P.before.z[i]};
    const D3 rj{P.before.x[j], P.before.y[j], P.before.z[j]};
    const D3 vi{P.before.vx[i], P.before.vy[i], P.before.vz[i]};
    const D3 vj{P.before.vx[j], P.before.vy[j], P.before.vz[j]};
    const int res = filter_pair(rj.x - ri.x, rj.y - ri.y, rj.z - ri.z, vj.x - vi.x, vj.y - vi.y,
                                vj.z - vi.z, P.dt, P.thr, P.vmin2);
    if (res == FILTER_CO_ORBITING) {
        ++cnt.co;
        return;
    }
    if (res != FILTER_REFINE) return;
    ++cnt.refined;
    const D3 ai{P.after.x[i], P.after.y[i], P.after.z[i]};
    const D3 aj{P.after.x[j], P.after.y[j], P.after.z[j]};
    const D3 wi{P.after.vx[i], P.after.vy[i], P.after.vz[i]};
    const D3 wj{P.after.vx[j], P.after.vy[j], P.after.vz[j]};
    const double dt = P.dt;
    const Hermite h{{rj.x - ri.x, rj.y - ri.y, rj.z - ri.z},
                    {(vj.x - vi.x) * dt, (vj.y - vi.y) * dt, (vj.z - vi.z) * dt},
                    {aj.x - ai.x, aj.y - ai.y, aj.z - ai.z},
                    {(wj.x - wi.x) * dt, (wj.y - wi.y) * dt, (wj.z - wi.z) * dt}};
    const double s = golden_min(h);
    if (s <= 1e-9 || s >= 1.0 - 1e-9) return;
    const D3 p = h.pos(s);
    const double miss = sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (miss > P.thr) return;
    const D3 w = h.vel(s);
    DevEvent ev;
    ev.i = i;
    ev.j = j;
    long long step = P.step;
    double t0 = P.t0;
    if (P.info) {
        // The event's step is the plain sub-step holding the TCA.
        const int j = min(P.sub_per - 1, static_cast<int>(s * P.sub_per));
        step = (P.info->macro * P.sub_n + P.sub_q) * P.sub_per + j;
        t0 = P.info->t + static_cast<double>(P.sub_q) * dt;
    }
    ev.step = step;
    ev.t = t0 + s * dt;
    ev.miss_km = miss;
    ev.v_rel_km_s = sqrt(w.x * w.x + w.y * w.y + w.z * w.z) / dt;
    ev.collision = miss < P.radius[i] + P.radius[j] ? 1 : 0;
    ev.pad = 0;
    append_event(gs, P, ev);
}

// One candidate, in both pair kernels: FP32 relative state from the compact
// records (positions relative to the cell corner, cell offset (dx, dy, dz)
// in whole cells, built as in the reference's GPU path so that DR_ERR_KM
// bounds it), with the partner cell's (x, y) corner offset folded into the
// own position: (ox + (lb.x - la.x)) regrouped as lb.x - (la.x - ox), the
// same operands with one rounding each, inside DR_ERR_KM. Pairs the
// pre-filter keeps go to the survivor queue for the FP64 kernel.

// Appends a pre-filter survivor, wave-aggregated among the lanes that
// reach this point together (sparse cells, where survivors are few).
__device__ __forceinline__ void push_survivor(GridState* gs, const PairParams& P, int a, int b) {
    const unsigned long long m = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    const int rank = __popcll(m & lanes_below());
    unsigned long long base = 0;
    if (static_cast<int>(__lane_id()) == leader)
        base = atomicAdd(&gs->n_surv, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, leader, WAVE);
    const unsigned long long slot = base + rank;
    if (slot < P.surv_cap) P.surv[slot] = make_uint2(static_cast<unsigned int>(a), static_cast<unsigned int>(b));
}

__device__ __forceinline__ void store_counters(PairPartial* parts, const PairCounters& cnt,
                                               unsigned long long (*s_cnt)[3]) {
    const unsigned long long c0 = wave_sum<unsigned long long>(cnt.cand);
    const unsigned long long c1 = wave_sum<unsigned long long>(cnt.refined);
    const unsigned long long c2 = wave_sum<unsigned long long>(cnt.co);
    const int wave = threadIdx.x / WAVE;
    if (__lane_id() == 0) {
        s_cnt[wave][0] = c0;
        s_cnt[wave][1] = c1;
        s_cnt[wave][2] = c2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0, c = 0;
        for (int w = 0; w < WAVES_PER_BLOCK; ++w) {
            a += s_cnt[w][0];
            b += s_cnt[w][1];
            c += s_cnt[w][2];
        }
        parts[blockIdx.x].cand = a;
        parts[blockIdx.x].refined = b;
        parts[blockIdx.x].co = c;
    }
}

constexpr unsigned int SURV_INVALID = 0xffffffffu;

// Per-wave survivor slots: reservations that grow geometrically (64 ..
// 4096), so a fresh cloud whose every pair survives costs one counter
// atomic per thousands of survivors; unused reserved slots are marked
// SURV_INVALID. push() is called by the whole wave.
struct SurvivorSlots {
    unsigned long long base = 0;  // wave-uniform
    unsigned int left = 0, size = 32;

    __device__ __forceinline__ void push(GridState* gs, const PairParams& P, bool keep, int a, int b, int lane,
                                         unsigned long long below) {
        const unsigned long long m = __ballot(keep);
        if (m == 0ULL) return;
        const unsigned int cnt = static_cast<unsigned int>(__popcll(m));
        if (left < cnt) {
            if (static_cast<unsigned int>(lane) < left && base + lane < P.surv_cap)
                P.surv[base + lane] = make_uint2(SURV_INVALID, SURV_INVALID);
            size = size < 4096u ? 2u * size : 4096u;
            const unsigned int R = size > cnt ? size : cnt;
            unsigned long long r = 0;
            if (lane == 0) r = atomicAdd(&gs->n_surv, static_cast<unsigned long long>(R));
            base = __shfl(r, 0, WAVE);
            left = R;
        }
        if (keep) {
            const unsigned long long slot = base + __popcll(m & below);
            if (slot < P.surv_cap)
                P.surv[slot] = make_uint2(static_cast<unsigned int>(a), static_cast<unsigned int>(b));
        }
        base += cnt;
        left -= cnt;
    }

    __device__ __forceinline__ void close(const PairParams& P, int lane) {
        for (unsigned int k = static_cast<unsigned int>(lane); k < left; k += WAVE)
            if (base + k < P.surv_cap) P.surv[base + k] = make_uint2(SURV_INVALID, SURV_INVALID);
    }
};

// Sparse cells, one own object per lane (the tuning's sparse_kernel 1). A
// unit is one (own object, neighbour range) pair: the run
// of partners the object is tested against (range 0 from the position after
// its own). A block takes PAIR_GROUP consecutive sorted objects, ranks their
// units by length in LDS (counting sort, longest first) and hands them to
// its waves in rounds of 64, so the lanes of a round walk runs of nearly
// equal length. A lane loads PAIR_BATCH partner records at a time into the
// registers that test them (no copies of loads in flight, which would make
// every iteration wait for its own loads). Survivors are few here: they are
// appended one wave-aggregated atomic at a time (a per-wave reservation
// would need wave-uniform loops and leave unused slots for k_exact).
constexpr int PAIR_GROUP = BLOCK;   // sorted objects per block, one per thread
constexpr int UNIT_BUCKETS = 64;    // unit length classes, longest first (unit_bucket)
constexpr int UNIT_OBJ_BITS = 8;    // log2(PAIR_GROUP)
constexpr int PAIR_BATCH = 4;
static_assert((1 << UNIT_OBJ_BITS) == PAIR_GROUP && UNIT_BUCKETS == WAVE, "unit packing");

// Length class of a unit, longest first: lengths 1-32 one class each
// (buckets 63..32), then classes 8 long (33-40 -> 31, ..., 273-280 -> 1),
// longer -> 0.
__device__ __forceinline__ int unit_bucket(int l) {
    if (l <= 32) return UNIT_BUCKETS - l;
    const int c = 31 - ((l - 33) >> 3);
    return c > 0 ? c : 0;
}

// SHELL: the sorted elements are shell entries; a pair is tested in the
// shell where one of its two entries is its object's first.
template <bool SHELL>
__global__ void __launch_bounds__(BLOCK) k_pairs_units(GridState* gs, PairParams Pin, PairPartial* parts) {
    const PairParams P = with_bands(Pin);
    __shared__ unsigned long long s_cnt[WAVES_PER_BLOCK][3];
    __shared__ int s_hist[UNIT_BUCKETS];
    __shared__ int s_total;
    __shared__ int2 s_unit[PAIR_GROUP * NRANGES];
    const int A = static_cast<int>(gs->nlive);
    const int g0 = blockIdx.x * PAIR_GROUP;
    const int lane = static_cast<int>(__lane_id());
    const int wave = threadIdx.x / WAVE;
    unsigned int cand = 0;
    const float lim_max2 = pair_lim_max2(gs, P);
    if (gs->npasses > 0 && g0 < A) {  // uniform over the block
        if (threadIdx.x < UNIT_BUCKETS) s_hist[threadIdx.x] = 0;
        __syncthreads();
        const int long_range = gs->long_range, dmin = gs->dense_min;
        const int p = g0 + threadIdx.x;
        const int c = p < A ? P.cell_of[p] : 0;
        const bool own = p < A && P.cell_count[c] < dmin;
        int q0[NRANGES], len[NRANGES], rank[NRANGES];
#pragma unroll
        for (int r = 0; r < NRANGES; ++r) {
            int2 run = make_int2(0, 0);
            if (own) run = P.ranges[c * NRANGES + r];
            const int first = r == 0 ? p + 1 : run.x;
            const int l = run.y - run.x > long_range ? 0 : run.y - first;  // long ranges: work items
            q0[r] = first;
            len[r] = l > 0 ? l : 0;
            rank[r] = 0;
            if (l > 0) rank[r] = atomicAdd(&s_hist[unit_bucket(l)], 1);
        }
        __syncthreads();
        // Exclusive scan of the bucket counts (one wave, one bucket per lane).
        if (wave == 0) {
            const int v = s_hist[lane];
            int incl = v;
            for (int d = 1; d < WAVE; d <<= 1) {
                const int t = __shfl_up(incl, d, WAVE);
                if (lane >= d) incl += t;
            }
            s_hist[lane] = incl - v;
            if (lane == WAVE - 1) s_total = incl;
        }
        __syncthreads();
        const int total = s_total;
#pragma unroll
        for (int r = 0; r < NRANGES; ++r) {
            const int l = len[r];
            if (l == 0) continue;
            const int bucket = unit_bucket(l);
            s_unit[s_hist[bucket] + rank[r]] =
                make_int2(q0[r], static_cast<int>(threadIdx.x) | (r << UNIT_OBJ_BITS) | (l << (UNIT_OBJ_BITS + 3)));
        }
        __syncthreads();
        const float cellf = static_cast<float>(gs->cell);
        for (int u = wave * WAVE + lane; u - lane < total; u += BLOCK) {
            if (u >= total) continue;
            const int2 d = s_unit[u];
            const int po = g0 + (d.y & (PAIR_GROUP - 1));
            const int r = (d.y >> UNIT_OBJ_BITS) & 7;
            const int ql = d.x + (d.y >> (UNIT_OBJ_BITS + 3)) - 1;  // last partner
            int dx, dy;
            range_offset(r, dx, dy);
            const float4 la = P.lrec[2 * po], va = P.lrec[2 * po + 1];
            const float ax = la.x - static_cast<float>(dx) * cellf;
            const float ay = la.y - static_cast<float>(dy) * cellf;
            cand += static_cast<unsigned int>(ql - d.x + 1);
            for (int q = d.x; q <= ql; q += PAIR_BATCH) {
                float4 b[PAIR_BATCH], w[PAIR_BATCH];
#pragma unroll
                for (int k = 0; k < PAIR_BATCH; ++k) {
                    const int qq = q + k < ql ? q + k : ql;
                    b[k] = P.lrec[2 * qq];
                    w[k] = P.lrec[2 * qq + 1];
                }
#pragma unroll
                for (int k = 0; k < PAIR_BATCH; ++k) {
                    if (q + k > ql) continue;
                    const int ia = __float_as_int(la.w), ib = __float_as_int(b[k].w);
                    if (SHELL && (ia & ib & NOT_FIRST)) continue;  // tested in an inner common shell
                    const float oz = static_cast<float>(__float_as_int(w[k].w) - __float_as_int(va.w)) * cellf;
                    if (prefilter_keep_fast(b[k].x - ax, b[k].y - ay, oz + (b[k].z - la.z), w[k].x - va.x,
                                            w[k].y - va.y, w[k].z - va.z, P.dt_f, P.gdt3_f, P.base_f, P.vmin2_f, lim_max2))
                        push_survivor(gs, P, ia & OBJ_MASK, ib & OBJ_MASK);
                }
            }
        }
        __syncthreads();  // s_cnt is reused below
    }
    store_counters(parts, PairCounters{cand, 0, 0}, s_cnt);
}

// Sparse cells, partners on the lanes. A block takes PAIR_GROUP consecutive
// sorted elements; their compact records go to LDS. The elements of one cell
// inside the block form a segment, and a unit is (segment, neighbour range):
// every partner of the range against every own element of the segment
// (range 0: the partners after each own element, q > p). Each lane takes
// one partner of one unit, holds its record in registers (the partners of a
// unit are consecutive: coalesced loads) and walks the segment's own
// records from LDS, where all lanes of a unit read the same record at once.
// Per candidate that is one LDS broadcast instead of a partner-record load
// per lane. Units and lanes are numbered by (element, range), so the lane
// work is the same every run. Dense cells (dense_min objects or more) and
// ranges longer than long_range go to the work items (k_pairs_dense).
// Against one own object per lane walking its partners (global loads per
// candidate), this measured 30% faster at 1M objects, 2 km.
constexpr int MAX_UNITS = PAIR_GROUP * NRANGES;
constexpr int OWN_BATCH = 4;
constexpr int LANE_MAP = 2048;  // lanes mapped to their unit per round of the block
static_assert(PAIR_GROUP <= 256 && NRANGES <= 8, "lane map code: element | range << 8");

template <bool SHELL>
__global__ void __launch_bounds__(BLOCK) k_pairs_lanes(GridState* gs, PairParams Pin, PairPartial* parts) {
    const PairParams P = with_bands(Pin);
    __shared__ unsigned long long s_cnt[WAVES_PER_BLOCK][3];
    __shared__ float4 s_rec[2 * PAIR_GROUP];
    __shared__ int s_cell[PAIR_GROUP];
    __shared__ int s_own[PAIR_GROUP];
    __shared__ int s_first[MAX_UNITS], s_start[MAX_UNITS];
    __shared__ uint16_t s_map[LANE_MAP];
    __shared__ int s_wave[WAVES_PER_BLOCK];
    const int A = static_cast<int>(gs->nlive);
    const int g0 = blockIdx.x * PAIR_GROUP;
    const int t = threadIdx.x;
    unsigned int cand = 0;
    const float lim_max2 = pair_lim_max2(gs, P);
    if (gs->npasses > 0 && g0 < A) {  // uniform over the block
        const int gend = min(A, g0 + PAIR_GROUP);
        const int p = g0 + t;
        const int c = p < A ? P.cell_of[p] : -1;
        s_cell[t] = c;
        if (p < A) {
            s_rec[2 * t] = P.lrec[2 * p];
            s_rec[2 * t + 1] = P.lrec[2 * p + 1];
        }
        __syncthreads();
        // Segment heads: the block's first element of each cell.
        const bool head = p < A && (t == 0 || s_cell[t - 1] != c);
        int m = 0;
        int2 run[NRANGES];
        int lanes[NRANGES];
        int lane_sum = 0;
        if (head && P.cell_count[c] < gs->dense_min) {
            while (t + m < gend - g0 && s_cell[t + m] == c) ++m;
#pragma unroll
            for (int r = 0; r < NRANGES; ++r) {
                run[r] = P.ranges[c * NRANGES + r];
                int k = 0;
                if (run[r].y - run[r].x <= gs->long_range) k = r == 0 ? run[r].y - (p + 1) : run[r].y - run[r].x;
                lanes[r] = k > 0 ? k : 0;
                lane_sum += lanes[r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < NRANGES; ++r) {
                run[r] = make_int2(0, 0);
                lanes[r] = 0;
            }
        }
        int total;
        const int lane0 = block_exclusive_scan(lane_sum, total, s_wave);
        // Per (element, range) of a head: first partner and first lane; per
        // element its own count. Lanes find their unit through a lane map
        // (element | range << 8 per lane), LANE_MAP lanes at a time.
        if (head) {
            int st = lane0;
#pragma unroll
            for (int r = 0; r < NRANGES; ++r) {
                s_first[t * NRANGES + r] = r == 0 ? p + 1 : run[r].x;
                s_start[t * NRANGES + r] = st;
                st += lanes[r];
            }
        }
        s_own[t] = m;
        const float cellf = static_cast<float>(gs->cell);
        const int wave = t / WAVE;
        const int lane = static_cast<int>(__lane_id());
        for (int w0 = 0; w0 < total; w0 += LANE_MAP) {
            const int w1 = min(total, w0 + LANE_MAP);
            __syncthreads();  // the previous window's map is read out
            if (head && lane_sum > 0) {
                int st = lane0;
#pragma unroll
                for (int r = 0; r < NRANGES; ++r) {
                    const int a = max(st, w0), b = min(st + lanes[r], w1);
                    const uint16_t code = static_cast<uint16_t>(t | (r << 8));
                    for (int l = a; l < b; ++l) s_map[l - w0] = code;
                    st += lanes[r];
                }
            }
            __syncthreads();
            for (int l0 = w0 + wave * WAVE; l0 < w1; l0 += BLOCK) {
                const int l = l0 + lane;
                if (l >= w1) continue;
                const int code = s_map[l - w0];
                const int e = code & 0xff, r = code >> 8;
                const int q = s_first[e * NRANGES + r] + (l - s_start[e * NRANGES + r]);
                // Own elements of the segment [e, e + own) that pair with q.
                const int m_seg = s_own[e];
                const int own = r == 0 ? min(m_seg, q - (g0 + e)) : m_seg;
                int dx, dy;
                range_offset(r, dx, dy);
                const float4 b = P.lrec[2 * q], w = P.lrec[2 * q + 1];
                const int ib = __float_as_int(b.w);
                const int wz = __float_as_int(w.w);
                // The partner's position in the own cell's frame (one cell
                // edge per offset; one rounding, inside DR_ERR_KM).
                const float bx = b.x + static_cast<float>(dx) * cellf, by = b.y + static_cast<float>(dy) * cellf;
                cand += static_cast<unsigned int>(own);
                // OWN_BATCH own records read at a time (their LDS reads in
                // flight together), then tested.
                for (int k0 = 0; k0 < own; k0 += OWN_BATCH) {
                    float4 la[OWN_BATCH], va[OWN_BATCH];
#pragma unroll
                    for (int u = 0; u < OWN_BATCH; ++u) {
                        const int kk = e + min(k0 + u, own - 1);
                        la[u] = s_rec[2 * kk];
                        va[u] = s_rec[2 * kk + 1];
                    }
#pragma unroll
                    for (int u = 0; u < OWN_BATCH; ++u) {
                        if (k0 + u >= own) continue;
                        const int ia = __float_as_int(la[u].w);
                        if (SHELL && (ia & ib & NOT_FIRST)) continue;  // tested in an inner common shell
                        const float oz = static_cast<float>(wz - __float_as_int(va[u].w)) * cellf;
                        if (prefilter_keep_fast(bx - la[u].x, by - la[u].y, oz + (b.z - la[u].z), w.x - va[u].x,
                                                w.y - va[u].y, w.z - va[u].z, P.dt_f, P.gdt3_f, P.base_f,
                                                P.vmin2_f, lim_max2))
                            push_survivor(gs, P, ia & OBJ_MASK, ib & OBJ_MASK);
                    }
                }
            }
        }
        __syncthreads();  // s_cnt is reused below
    }
    store_counters(parts, PairCounters{cand, 0, 0}, s_cnt);
}

constexpr int PARTNER_BATCH = 4;

// Dense cells and long ranges: one wave per work item (grid-stride over
// the items), in one of two orientations, whichever takes fewer wave
// iterations:
//   own objects on the lanes: every lane tests its own object against the
//     same partner, whose records are wave-uniform scalar loads
//     (PARTNER_BATCH in flight);
//   partners on the lanes (few own objects, a long run: the long ranges of
//     sparse cells): 64 partners per lane load, the own objects' records
//     scalar loads.
// Both evaluate the same expression on the same operands, own object first.
template <bool SHELL>
__global__ void __launch_bounds__(BLOCK) k_pairs_dense(GridState* gs, PairParams Pin,
                                                        const PairItem* __restrict__ items,
                                                        const float4* __restrict__ lrec,
                                                        unsigned int item_cap, PairPartial* parts) {
    __shared__ unsigned long long s_cnt[WAVES_PER_BLOCK][3];
    const PairParams P = with_bands(Pin);
    const int lane = static_cast<int>(__lane_id());
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / WAVE);
    const unsigned int total_waves = gridDim.x * WAVES_PER_BLOCK;
    const unsigned int n_items = min(gs->n_items, item_cap);
    const float cellf = static_cast<float>(gs->cell);
    const unsigned long long below = lanes_below();
    unsigned int cand = 0;
    const float lim_max2 = pair_lim_max2(gs, P);
    SurvivorSlots slots;
    for (unsigned int it = blockIdx.x * WAVES_PER_BLOCK + wave; it < n_items; it += total_waves) {
        const PairItem item = items[it];
        int dx, dy;
        range_offset(item.kind, dx, dy);
        const float ox = static_cast<float>(dx) * cellf, oy = static_cast<float>(dy) * cellf;
        const int q1 = item.part_begin + item.part_count;
        const bool partners_on_lanes =
            item.own_count * ((item.part_count + WAVE - 1) / WAVE) < item.part_count;
        if (partners_on_lanes) {
            for (int q0 = item.part_begin; q0 < q1; q0 += WAVE) {
                const int q = q0 + lane;
                const bool has = q < q1;
                float4 lb = make_float4(0.f, 0.f, 0.f, 0.f), vb = lb;
                if (has) {
                    lb = lrec[2 * q];
                    vb = lrec[2 * q + 1];
                }
                for (int j0 = 0; j0 < item.own_count; j0 += PARTNER_BATCH) {
                    float4 ta[PARTNER_BATCH], tv[PARTNER_BATCH];
#pragma unroll
                    for (int j = 0; j < PARTNER_BATCH; ++j) {
                        const int pj = item.own_begin + (j0 + j < item.own_count ? j0 + j : item.own_count - 1);
                        ta[j] = lrec[2 * pj];
                        tv[j] = lrec[2 * pj + 1];
                    }
#pragma unroll
                    for (int j = 0; j < PARTNER_BATCH; ++j) {
                        const float4 la = ta[j], va = tv[j];
                        const int p = item.own_begin + j0 + j;
                        // Range 0 pairs only with later positions.
                        const int ia = __float_as_int(la.w), ib = __float_as_int(lb.w);
                        const bool test = has && j0 + j < item.own_count && (item.kind != 0 || q > p) &&
                                          !(SHELL && (ia & ib & NOT_FIRST));
                        bool keep = false;
                        if (test) {
                            ++cand;
                            const float ax = la.x - ox, ay = la.y - oy;
                            const float oz =
                                static_cast<float>(__float_as_int(vb.w) - __float_as_int(va.w)) * cellf;
                            keep = prefilter_keep_fast(lb.x - ax, lb.y - ay, oz + (lb.z - la.z), vb.x - va.x,
                                                       vb.y - va.y, vb.z - va.z, P.dt_f, P.gdt3_f, P.base_f,
                                                       P.vmin2_f, lim_max2);
                        }
                        slots.push(gs, P, keep, ia & OBJ_MASK, ib & OBJ_MASK, lane, below);
                    }
                }
            }
            continue;
        }
        const int p = item.own_begin + lane;
        const bool active = lane < item.own_count;
        float4 la = make_float4(0.f, 0.f, 0.f, 0.f), va = la;
        if (active) {
            la = lrec[2 * p];
            va = lrec[2 * p + 1];
        }
        // Cell-corner offsets folded into the own position: dr = lb - (la - o).
        const float ax = la.x - ox;
        const float ay = la.y - oy;
        const int za = __float_as_int(va.w);
        // Range 0 pairs only with later positions.
        const int first = item.kind == 0 ? p + 1 : -1;
        for (int q0 = item.part_begin; q0 < q1; q0 += PARTNER_BATCH) {
            float4 tb[PARTNER_BATCH], tw[PARTNER_BATCH];
#pragma unroll
            for (int j = 0; j < PARTNER_BATCH; ++j) {
                const int q = q0 + j < q1 ? q0 + j : q1 - 1;
                tb[j] = lrec[2 * q];
                tw[j] = lrec[2 * q + 1];
            }
#pragma unroll
            for (int j = 0; j < PARTNER_BATCH; ++j) {
                const float4 lb = tb[j];
                const float4 vb = tw[j];
                const int ia = __float_as_int(la.w), ib = __float_as_int(lb.w);
                const bool test = active && q0 + j < q1 && q0 + j >= first && !(SHELL && (ia & ib & NOT_FIRST));
                bool keep = false;
                if (test) {
                    ++cand;
                    const float oz = static_cast<float>(__float_as_int(vb.w) - za) * cellf;
                    // (ox + (lb.x - la.x)) regrouped as lb.x - (la.x - ox): the
                    // same operands, one rounding each, inside DR_ERR_KM.
                    keep = prefilter_keep_fast(lb.x - ax, lb.y - ay, oz + (lb.z - la.z), vb.x - va.x,
                                               vb.y - va.y, vb.z - va.z, P.dt_f, P.gdt3_f, P.base_f,
                                               P.vmin2_f, lim_max2);
                }
                slots.push(gs, P, keep, ia & OBJ_MASK, ib & OBJ_MASK, lane, below);
            }
        }
    }
    slots.close(P, lane);
    store_counters(parts, PairCounters{cand, 0, 0}, s_cnt);
}

// FP64 exact filter, TCA and radius test of the pre-filter survivors: one
// lane per survivor (grid-stride), coherent FP64 work.
__global__ void __launch_bounds__(BLOCK) k_exact(GridState* gs, PairParams Pin, PairPartial* parts) {
    const PairParams P = with_bands(Pin);
    __shared__ unsigned long long s_cnt[WAVES_PER_BLOCK][3];
    const unsigned long long n = gs->n_surv < P.surv_cap ? gs->n_surv : P.surv_cap;
    if (blockIdx.x == 0 && threadIdx.x == 0 && gs->n_surv > P.surv_cap) gs->surv_short = 1;
    PairCounters cnt{0, 0, 0};
    for (unsigned long long k = static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x; k < n;
         k += static_cast<unsigned long long>(gridDim.x) * BLOCK) {
        const uint2 pr = P.surv[k];
        if (pr.x == SURV_INVALID) continue;  // unused reserved slot
        ++cnt.cand;                          // survivors
        exact_pair(gs, P, static_cast<int>(pr.x), static_cast<int>(pr.y), cnt);
    }
    store_counters(parts, cnt, s_cnt);
}

// ---------------------------------------------------------------------------
// Learned form: plain FP64 re-verification of every reported event
// ---------------------------------------------------------------------------
//
// An event the learned form reports in sub-step k of a learned step is
// solved again on plain trajectories: both objects are stepped from the
// learned step's start state by FP64 RK4 sub-steps (the plain method) to the
// start and end of sub-step k, and the reference's exact test (filter_pair,
// Hermite + golden section, radius test; collision.cpp:109-132) runs on that
// pair of states, with the caller's threshold plus the drift band (the
// distance learned trajectories may have drifted from plain ones since the
// run began; the search margin above that serves the learned search only).
// A minimum on a sub-step edge is looked for in the neighbouring sub-step of
// the same learned step, and so is one of a pair the filter rejects in the
// event's own sub-step (receding there: the minimum may lie just before the
// boundary). The event's TCA, miss distance, speed and
// collision flag become the plain solve's; an event the plain solve rejects
// is dropped (pad = 1). One whose plain minimum leaves the learned step
// (a TCA within a fraction of a millisecond of the step's edge) keeps its
// learned values (pad = 2) if its learned miss distance is within the
// reporting threshold, and is dropped otherwise.

struct ReverifyStats {
    unsigned long long verified, dropped, kept_unverified;
};

// 0: event (s, miss, vrel); 1: rejected; 2: minimum at s ~ 0; 3: at s ~ 1;
// 4: rejected, receding at the start.
__device__ __forceinline__ int plain_refine(D3 ri, D3 vi, D3 rj, D3 vj, D3 ai, D3 wi, D3 aj, D3 wj, double dt,
                                            double thr, double vmin2, double& s, double& miss, double& vrel) {
    const double rx = rj.x - ri.x, ry = rj.y - ri.y, rz = rj.z - ri.z;
    const double ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
    const int res = filter_pair(rx, ry, rz, ux, uy, uz, dt, thr, vmin2);
    if (res == FILTER_REJECT && rx * ux + ry * uy + rz * uz >= 0.0) return 4;
    if (res != FILTER_REFINE) return 1;
    const Hermite h{{rj.x - ri.x, rj.y - ri.y, rj.z - ri.z},
                    {(vj.x - vi.x) * dt, (vj.y - vi.y) * dt, (vj.z - vi.z) * dt},
                    {aj.x - ai.x, aj.y - ai.y, aj.z - ai.z},
                    {(wj.x - wi.x) * dt, (wj.y - wi.y) * dt, (wj.z - wi.z) * dt}};
    s = golden_min(h);
    if (s <= 1e-9) return 2;
    if (s >= 1.0 - 1e-9) return 3;
    const D3 p = h.pos(s);
    miss = sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (miss > thr) return 1;
    const D3 w = h.vel(s);
    vrel = sqrt(w.x * w.x + w.y * w.y + w.z * w.z) / dt;
    return 0;
}

// States of objects i and j at the start of sub-steps k and k + 1. Both are
// stepped in the same loop body, so their two independent RK4 chains
// interleave (a lane's work here is one long dependent chain per object,
// and few waves run).
__device__ __forceinline__ void plain_at2(Planes s0, int i, int j, const MoonSample* m, double h, int sub, int k,
                                          D3& rai, D3& vai, D3& rbi, D3& vbi, D3& raj, D3& vaj, D3& rbj,
                                          D3& vbj) {
    const int stride = 12 / sub;
    D3 ri{s0.x[i], s0.y[i], s0.z[i]}, vi{s0.vx[i], s0.vy[i], s0.vz[i]};
    D3 rj{s0.x[j], s0.y[j], s0.z[j]}, vj{s0.vx[j], s0.vy[j], s0.vz[j]};
    for (int q = 0; q < k; ++q) {
        const MoonSample& m0 = m[stride * q];
        const MoonSample& m1 = m[stride * q + stride / 2];
        const MoonSample& m2 = m[stride * (q + 1)];
        rk4(ri, vi, h, m0, m1, m2);
        rk4(rj, vj, h, m0, m1, m2);
    }
    rai = ri;
    vai = vi;
    raj = rj;
    vaj = vj;
    const MoonSample& m0 = m[stride * k];
    const MoonSample& m1 = m[stride * k + stride / 2];
    const MoonSample& m2 = m[stride * (k + 1)];
    rk4(ri, vi, h, m0, m1, m2);
    rk4(rj, vj, h, m0, m1, m2);
    rbi = ri;
    vbi = vi;
    rbj = rj;
    vbj = vj;
}

__global__ void __launch_bounds__(BLOCK) k_reverify(GridState* gs, DevEvent* events, unsigned long long cap,
                                                     Planes s0, const double* radius,
                                                     const MoonSample* __restrict__ moon13, const StepInfo* info,
                                                     double h, int sub, double thr_report, double vmin2,
                                                     ReverifyStats* rs, const Bands* bands, FinishArgs fin) {
    if (bands) {
        thr_report = bands->thr_report;
        vmin2 = bands->vmin2_report;
    }
    const double t_step = info->t;
    const long long sub0 = info->macro * sub;
    const unsigned long long e0 = gs->ev_mark;
    const unsigned long long e1 = gs->events < cap ? gs->events : cap;
    unsigned long long ver = 0, drop = 0, kept = 0;
    for (unsigned long long e = e0 + static_cast<unsigned long long>(blockIdx.x) * BLOCK + threadIdx.x; e < e1;
         e += static_cast<unsigned long long>(gridDim.x) * BLOCK) {
        DevEvent ev = events[e];
        const int i = ev.i, j = ev.j;
        int k = static_cast<int>(ev.step - sub0);
        int code = 1, tries = 0, dir = 0;
        double s = 0.0, miss = 0.0, vrel = 0.0;
        while (tries < 2 && k >= 0 && k < sub) {
            D3 rai, vai, rbi, vbi, raj, vaj, rbj, vbj;
            plain_at2(s0, i, j, moon13, h, sub, k, rai, vai, rbi, vbi, raj, vaj, rbj, vbj);
            code = plain_refine(rai, vai, raj, vaj, rbi, vbi, rbj, vbj, h, thr_report, vmin2, s, miss, vrel);
            ++tries;
            // A minimum on the sub-step's start (code 2), or a pair already
            // receding there on the first try (code 4: the learned drift can
            // move a TCA from just after a sub-step boundary to just before
            // it), is looked for in the previous sub-step.
            if ((code == 2 || (code == 4 && tries == 1)) && dir <= 0) {
                code = 2;
                dir = -1;
                --k;
            } else if (code == 3 && dir >= 0) {
                dir = 1;
                ++k;
            } else {
                if (code == 4) code = 1;
                break;
            }
        }
        if (code == 0) {
            ev.step = sub0 + k;
            ev.t = t_step + static_cast<double>(k) * h + s * h;
            ev.miss_km = miss;
            ev.v_rel_km_s = vrel;
            ev.collision = miss < radius[i] + radius[j] ? 1 : 0;
            ev.pad = 0;
            ++ver;
        } else if ((code == 2 || code == 3) && (k < 0 || k >= sub) && ev.miss_km <= thr_report) {
            ev.pad = 2;  // the plain minimum lies in the neighbouring learned step
            ++kept;
        } else {
            ev.pad = 1;
            ++drop;
        }
        events[e] = ev;
    }
    ver = wave_sum(ver);
    drop = wave_sum(drop);
    kept = wave_sum(kept);
    if (__lane_id() == 0) {
        if (ver) atomicAdd(&rs->verified, ver);
        if (drop) atomicAdd(&rs->dropped, drop);
        if (kept) atomicAdd(&rs->kept_unverified, kept);
    }
    // The screening's statistics record (in place of a k_finish launch); it
    // touches no field read above.
    if (blockIdx.x == 0 && fin.out)
        finish_window(gs, fin.out, 1, fin.parts, fin.nparts, fin.npair, fin.dparts, fin.ndparts, fin.eparts,
                      fin.neparts, fin.fill);
}

__global__ void k_zero_reverify(ReverifyStats* rs) {
    if (threadIdx.x == 0) {
        rs->verified = 0;
        rs->dropped = 0;
        rs->kept_unverified = 0;
    }
}

// Appends the events a second screening stream found (its own buffer and
// counter) to the context's events, and empties the second buffer. The
// count is carried in full, so an overflow of either buffer shows as an
// overflow of the context's and the chunk runs again with both grown.
__global__ void __launch_bounds__(TBLOCK) k_merge_events(GridState* gs, DevEvent* events, unsigned long long cap,
                                                         GridState* gs2, const DevEvent* __restrict__ ev2,
                                                         unsigned long long cap2) {
    const unsigned long long base = gs->events;
    const unsigned long long n2 = gs2->events;
    const unsigned long long m = n2 < cap2 ? n2 : cap2;
    for (unsigned long long i = threadIdx.x; i < m; i += TBLOCK)
        if (base + i < cap) events[base + i] = ev2[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        gs->events = base + n2;
        gs2->events = 0ULL;
        gs2->ev_mark = 0ULL;
    }
}

__global__ void k_mark_events(GridState* gs, unsigned long long cap) {
    if (threadIdx.x == 0) gs->ev_mark = gs->events < cap ? gs->events : cap;
}

// ---------------------------------------------------------------------------
// In-run breakups (every form): collisions of the step -> jobs ->
// fragment samples -> mass budgets -> new objects in reserved slots
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(BLOCK) k_bu_events(const GridState* gs, const DevEvent* __restrict__ events,
                                                      unsigned long long cap, Planes pb, Planes pa, uint8_t* alive,
                                                      double* mass, const double* size_m, const int* scen, double t_arg,
                                                      const StepInfo* info, double dt, bu::BreakupCfg cfg,
                                                      bu::BreakupState* st, bu::Job* jobs, bu::LogRecord* logs,
                                                      int log_cap) {
    // The step's start time: from the step table when one is given (captured
    // chunks), else the argument.
    const double t0 = info ? info->t : t_arg;
    __shared__ int s_n;
    __shared__ unsigned long long s_idx[bu::MAX_JOBS];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const unsigned long long e0 = gs->ev_mark, e1 = gs->events < cap ? gs->events : cap;
    for (unsigned long long e = e0 + threadIdx.x; e < e1; e += BLOCK) {
        // pad 1: a learned-form event the plain re-verification rejected.
        if (!events[e].collision || events[e].pad == 1) continue;
        const int k = atomicAdd(&s_n, 1);
        if (k < bu::MAX_JOBS) s_idx[k] = e;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    st->njobs = 0;
    st->nsamples = 0;
    const int n = s_n < bu::MAX_JOBS ? s_n : bu::MAX_JOBS;
    if (s_n > bu::MAX_JOBS) st->jobs_short = 1;
    // (t, i, j) order, as the host applies a step's collisions.
    for (int a = 1; a < n; ++a) {
        const unsigned long long x = s_idx[a];
        int b = a - 1;
        while (b >= 0) {
            const DevEvent& p = events[s_idx[b]];
            const DevEvent& q = events[x];
            const bool after = p.t > q.t || (p.t == q.t && (p.i > q.i || (p.i == q.i && p.j > q.j)));
            if (!after) break;
            s_idx[b + 1] = s_idx[b];
            --b;
        }
        s_idx[b + 1] = x;
    }
    for (int c = 0; c < n; ++c) {
        const DevEvent ev = events[s_idx[c]];
        const int ii = ev.i, jj = ev.j;
        if (!alive[ii] || !alive[jj]) continue;
        // Hermite state of both objects at the collision instant (collision.cpp:289-292).
        const double s = (ev.t - t0) / dt;
        D3 r[2], v[2];
        const int idx[2] = {ii, jj};
        for (int q = 0; q < 2; ++q) {
            const int o = idx[q];
            const Hermite h{{pb.x[o], pb.y[o], pb.z[o]},
                            {pb.vx[o] * dt, pb.vy[o] * dt, pb.vz[o] * dt},
                            {pa.x[o], pa.y[o], pa.z[o]},
                            {pa.vx[o] * dt, pa.vy[o] * dt, pa.vz[o] * dt}};
            r[q] = h.pos(s);
            const D3 w = h.vel(s);
            v[q] = D3{w.x / dt, w.y / dt, w.z / dt};
        }
        const double mi = mass[ii] > 0 ? mass[ii] : bu::FALLBACK_MASS_KG;
        const double mj = mass[jj] > 0 ? mass[jj] : bu::FALLBACK_MASS_KG;
        const int tq = mi >= mj ? 0 : 1, pq = 1 - tq;
        const int tgt = idx[tq], prj = idx[pq];
        const double mt = tq == 0 ? mi : mj, mp = tq == 0 ? mj : mi;
        const double dvx = v[tq].x - v[pq].x, dvy = v[tq].y - v[pq].y, dvz = v[tq].z - v[pq].z;
        const double vrel = sqrt(dvx * dvx + dvy * dvy + dvz * dvz);
        const double energy = 500.0 * mp * vrel * vrel / mt;
        const bool cat = energy >= cfg.catastrophic_j_per_g;
        double b0, b1;
        if (cat) {
            b0 = mt;
            b1 = mp;
        } else {
            b1 = mp;
            b0 = fmax(0.0, mp * vrel * vrel - mp);
        }
        const double M = b0 + b1;
        alive[prj] = 0;
        if (cat) alive[tgt] = 0;
        else mass[tgt] = fmax(mt - b0, 0.0);
        const double expected = M > 0 ? 0.1 * pow(M, 0.75) * pow(cfg.lc_m, -bu::SLOPE) : 0.0;
        const int id = st->next_event++;
        bu::Rng g(cfg.seed, static_cast<uint64_t>(id), 0xffffffffULL);
        long long cnt = expected > 0 ? g.poisson(expected) : 0;
        if (cnt > bu::MAX_FRAGS_PER_EVENT) cnt = bu::MAX_FRAGS_PER_EVENT;
        if (st->nsamples + cnt > cfg.sample_cap) {
            st->samples_short = 1;
            cnt = 0;
        }
        bu::Job jb;
        jb.t = ev.t;
        jb.remaining = t0 + dt - ev.t;
        jb.vrel = vrel;
        jb.p_target = M > 0 ? b0 / M : 0.0;
        jb.r[0] = r[tq];
        jb.v[0] = v[tq];
        jb.r[1] = r[pq];
        jb.v[1] = v[pq];
        jb.size[0] = size_m[tgt];
        jb.size[1] = size_m[prj];
        jb.budget[0] = b0;
        jb.budget[1] = b1;
        jb.count = static_cast<int>(cnt);
        jb.event = id;
        jb.sample_base = st->nsamples;
        jb.scen = scen ? scen[tgt] : 0;
        jb.kept = jb.escaped = jb.base = 0;
        jb.used_mass = 0.0;
        jb.log = st->nlog < log_cap ? st->nlog : -1;
        if (jb.log >= 0) {
            bu::LogRecord& L = logs[jb.log];
            L.t = ev.t;
            L.v_rel_km_s = vrel;
            L.energy_j_per_g = energy;
            L.target_mass_kg = mt;
            L.projectile_mass_kg = mp;
            L.sbm_mass_kg = M;
            L.expected_fragments = expected;
            L.fragments_mass_kg = 0.0;
            L.px = r[tq].x;
            L.py = r[tq].y;
            L.pz = r[tq].z;
            L.id = id;
            L.target = tgt;
            L.projectile = prj;
            L.catastrophic = cat ? 1 : 0;
            L.drawn = jb.count;
            L.fragments = 0;
            L.escaped = 0;
            L.first = 0;
        }
        ++st->nlog;
        st->nsamples += jb.count;
        jobs[st->njobs++] = jb;
    }
}

// Last job whose first sample is <= f.
__device__ __forceinline__ int job_of(const bu::Job* jobs, int njobs, int f) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].sample_base <= f) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Fragment sample f (of job jb): its own counter-based stream.
__device__ void sample_one(const bu::Job& jb, int f, bu::Sample* samples, const bu::BreakupCfg& cfg) {
    const int k = f - jb.sample_base;
    bu::Rng g(cfg.seed, static_cast<uint64_t>(jb.event), static_cast<uint64_t>(k));
    bu::Sample sm;
    sm.rank = -1;
    sm.side = g.uniform() < jb.p_target ? 0 : 1;
    const double lmax = jb.size[sm.side];
    if (lmax <= cfg.lc_m) {
        sm.side = -1;
        samples[f] = sm;
        return;
    }
    // Size by inverting the truncated power law on [L_c, parent size].
    const double rr = pow(lmax / cfg.lc_m, -bu::SLOPE);
    const double L = cfg.lc_m * pow(1.0 - g.uniform() * (1.0 - rr), -1.0 / bu::SLOPE);
    const double chi = bu::area_to_mass_log(g, log10(L));
    const double area = 0.540424 * L * L;
    const double m = area / pow(10.0, chi);
    double dv_ms = pow(10.0, 0.9 * chi + 2.9 + 0.4 * g.normal());
    dv_ms = fmin(dv_ms, jb.vrel * 1000.0);
    const double cz = 2.0 * g.uniform() - 1.0;
    const double phi = 2.0 * bu::PI_D * g.uniform();
    const double sz = sqrt(fmax(0.0, 1.0 - cz * cz));
    const double a = dv_ms * 1e-3;
    sm.dv = D3{sz * cos(phi) * a, sz * sin(phi) * a, cz * a};
    sm.L = L;
    sm.chi = chi;
    sm.m = m;
    samples[f] = sm;
}

__global__ void __launch_bounds__(BLOCK) k_bu_sample(const bu::BreakupState* st, const bu::Job* __restrict__ jobs,
                                                      bu::Sample* samples, bu::BreakupCfg cfg) {
    const int ns = st->nsamples, nj = st->njobs;
    for (int f = blockIdx.x * BLOCK + threadIdx.x; f < ns; f += gridDim.x * BLOCK)
        sample_one(jobs[job_of(jobs, nj, f)], f, samples, cfg);
}

// One lane per job, fragments in order: the parent's mass budget (a fragment
// that would exceed it is skipped, later smaller ones may still fit), then
// the hyperbolic check; kept fragments get consecutive ranks.
__global__ void __launch_bounds__(64) k_bu_budget(const bu::BreakupState* st, bu::Job* jobs, bu::Sample* samples,
                                                   bu::LogRecord* logs) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= st->njobs) return;
    bu::Job& jb = jobs[j];
    double used[2] = {0.0, 0.0};
    int kept = 0, esc = 0;
    double mass_kept = 0.0;
    for (int k = 0; k < jb.count; ++k) {
        bu::Sample& sm = samples[jb.sample_base + k];
        if (sm.side < 0) continue;
        if (used[sm.side] + sm.m > jb.budget[sm.side]) continue;
        used[sm.side] += sm.m;
        const D3 r = jb.r[sm.side];
        const D3 v{jb.v[sm.side].x + sm.dv.x, jb.v[sm.side].y + sm.dv.y, jb.v[sm.side].z + sm.dv.z};
        if (0.5 * (v.x * v.x + v.y * v.y + v.z * v.z) - dev::MU_EARTH / sqrt(r.x * r.x + r.y * r.y + r.z * r.z) >=
            0.0) {
            ++esc;
            continue;
        }
        sm.rank = kept++;
        mass_kept += sm.m;
    }
    jb.kept = kept;
    jb.escaped = esc;
    jb.used_mass = mass_kept;
    if (jb.log >= 0) {
        logs[jb.log].fragments = kept;
        logs[jb.log].escaped = esc;
        logs[jb.log].fragments_mass_kg = mass_kept;
    }
}

__global__ void k_bu_bases(bu::BreakupState* st, bu::Job* jobs, bu::LogRecord* logs, int reserve) {
    if (threadIdx.x != 0) return;
    for (int j = 0; j < st->njobs; ++j) {
        jobs[j].base = st->frag_used;
        if (jobs[j].log >= 0) logs[jobs[j].log].first = st->frag_used;
        st->frag_used += jobs[j].kept;
    }
    if (st->frag_used > reserve) st->reserve_short = 1;
}

// New objects: birth at the parent's collision state plus delta-v, RK4 over
// the rest of the step (Moon from the step's samples), object arrays.
// moon: the Moon at the step's start, middle and end, `mstride` entries
// apart. Learned form (mode != nullptr): each fragment starts on the plain
// method, appended to the plain list (plain_list, counts[CLASS_PLAIN]) for
// the steps up to the next check, and its state goes to both state buffers
// (pb too): objects outside the lists, as a fragment born below the surface
// is, are not copied forward step by step.
__global__ void __launch_bounds__(BLOCK) k_bu_write(const bu::BreakupState* st, const bu::Job* __restrict__ jobs,
                                                     const bu::Sample* __restrict__ samples, int n0, int reserve,
                                                     Planes pa, uint8_t* alive, double* radius, double* size_m,
                                                     double* mass, int* scen, const MoonSample* __restrict__ moon,
                                                     int mstride, double t_arg, const StepInfo* info, double dt,
                                                     bu::BreakupCfg cfg, uint8_t* mode, int* plain_list, int* counts,
                                                     Planes pb) {
    const double t0 = info ? info->t : t_arg;  // as in k_bu_events
    const int ns = st->nsamples, nj = st->njobs;
    const MoonSample moon3[3] = {moon[0], moon[mstride], moon[2 * mstride]};
    for (int f = blockIdx.x * BLOCK + threadIdx.x; f < ns; f += gridDim.x * BLOCK) {
        const bu::Sample sm = samples[f];
        if (sm.rank < 0) continue;
        const bu::Job& jb = jobs[job_of(jobs, nj, f)];
        const int off = jb.base + sm.rank;
        if (off >= reserve) continue;
        const int slot = n0 + off;
        D3 r = jb.r[sm.side];
        D3 v{jb.v[sm.side].x + sm.dv.x, jb.v[sm.side].y + sm.dv.y, jb.v[sm.side].z + sm.dv.z};
        if (jb.remaining > 0.0) {
            const double f0 = (jb.t - t0) / dt;
            const MoonSample m0 = bu::moon_at(moon3, f0);
            const MoonSample m1 = bu::moon_at(moon3, f0 + 0.5 * jb.remaining / dt);
            const MoonSample m2 = bu::moon_at(moon3, 1.0);
            rk4(r, v, jb.remaining, m0, m1, m2);
        }
        pa.x[slot] = r.x;
        pa.y[slot] = r.y;
        pa.z[slot] = r.z;
        pa.vx[slot] = v.x;
        pa.vy[slot] = v.y;
        pa.vz[slot] = v.z;
        const bool up = sqrt(r.x * r.x + r.y * r.y + r.z * r.z) >= dev::R_EARTH;
        alive[slot] = up ? 1 : 0;
        radius[slot] = 0.5 * sm.L * 1e-3 * cfg.radius_scale;
        size_m[slot] = sm.L;
        mass[slot] = sm.m;
        if (scen) scen[slot] = jb.scen;
        if (mode) {
            mode[slot] = net::CLASS_PLAIN;
            if (up) plain_list[atomicAdd(&counts[net::CLASS_PLAIN], 1)] = slot;
            pb.x[slot] = r.x;
            pb.y[slot] = r.y;
            pb.z[slot] = r.z;
            pb.vx[slot] = v.x;
            pb.vy[slot] = v.y;
            pb.vz[slot] = v.z;
        }
    }
}

__global__ void __launch_bounds__(TBLOCK) k_finish(GridState* gs, DevStats* out, int screened,
                                                   const PairPartial* parts, int nparts, int npair,
                                                   const PairPartial* dparts, int ndparts,
                                                   const PairPartial* eparts, int neparts, int fill) {
    finish_window(gs, out, screened, parts, nparts, npair, dparts, ndparts, eparts, neparts, fill);
}

__global__ void k_fp64_check(int n, const double* a, const double* b, double* out) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const double x = a[k], y = b[k];
    out[3 * k + 0] = sqrt(x);
    out[3 * k + 1] = x / y;
    out[3 * k + 2] = floor(x / y);
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

// Per-screening device buffers of a second screening stream (swapped into
// the context while its kernels are enqueued; sizes follow the context's).
struct ScreenUnit {
    GridState* gs = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    int* idx[2] = {nullptr, nullptr};
    int* hist = nullptr;
    int* totals = nullptr;
    int* bcount = nullptr;
    uint64_t* cell_key = nullptr;
    int* cell_begin = nullptr;
    int* cell_count = nullptr;
    int* cell_of = nullptr;
    int2* ranges = nullptr;
    float4* lrec = nullptr;
    float4* rec = nullptr;
    int4* coltab = nullptr;
    PairPartial* pair_parts = nullptr;
    PairPartial* dense_parts = nullptr;
    PairPartial* exact_parts = nullptr;
    uint2* surv = nullptr;
    int64_t surv_cap = 0;
    DenseCell* dense = nullptr;
    PairItem* items = nullptr;
    int64_t item_cap = 0;
    DevEvent* events = nullptr;
    int64_t ev_cap = 0;
    int64_t ecap = 0, cap = 0;  // element and object capacities the buffers were sized for
    hipStream_t stream = nullptr;
    hipEvent_t ev_go = nullptr, ev_done = nullptr;
};

struct khip_ctx {
    int device = 0;
    std::string name, error;
    hipStream_t stream = nullptr;

    int64_t n = 0, cap = 0;
    double* state[2] = {nullptr, nullptr};  // SoA planes, 6 * cap doubles each
    int cur = 0;
    uint8_t* alive = nullptr;
    double* radius = nullptr;
    double* size_m = nullptr;
    std::vector<double> host_size;

    // Grid buffers.
    uint64_t* keys[2] = {nullptr, nullptr};
    int* idx[2] = {nullptr, nullptr};
    int* hist = nullptr;
    int* totals = nullptr;
    int* bcount = nullptr;
    uint64_t* cell_key = nullptr;
    int* cell_begin = nullptr;
    int* cell_count = nullptr;
    int* cell_of = nullptr;
    int2* ranges = nullptr;
    float4* lrec = nullptr;  // compact records in sorted order, 2 float4 each
    TileCfg tile{1.0, DENSE_MIN, LONG_RANGE, SEG, DENSE_BLOCKS, EXACT_BLOCKS, 0};

    // In-run breakups.
    khip_breakup_config bcfg{0, 65536, 0.1, 40.0, 1};
    int64_t n_launch = 0;          // objects the per-step kernels cover (n + reserved slots)
    double* mass = nullptr;        // kg per object (0: the model's fallback)
    double* snap_mass = nullptr;
    bu::BreakupState* bstate = nullptr;
    bu::BreakupState* snap_bstate = nullptr;
    bu::Job* jobs = nullptr;
    bu::Sample* samples = nullptr;
    int sample_cap = 0;
    bu::LogRecord* blogs = nullptr;
    int blog_cap = 0;
    int bu_reserve = 0;
    bool reset_breakups = true;    // breakup ids restart with a new population
    std::vector<khip_breakup_event> host_breakups;
    int4* coltab = nullptr;  // column table, 1 << COL_BITS entries
    int* scen = nullptr;     // scenario id per object (concatenated scenarios)
    int sbits = 0;
    std::vector<int64_t> scen_off;
    float4* rec = nullptr;   // the same records in object order, 2 float4 each (written by k_keys)
    GridState* gs = nullptr;
    StepPartial* step_parts = nullptr;
    StepPartial* win_parts = nullptr;  // learned form: FUSED_WINDOWS_MAX * PART_BLOCKS window reductions
    PairPartial* pair_parts = nullptr;
    PairPartial* dense_parts = nullptr;
    PairPartial* exact_parts = nullptr;
    uint2* surv = nullptr;
    int64_t surv_cap = 0;
    DenseCell* dense = nullptr;
    PairItem* items = nullptr;
    int64_t item_cap = 0;

    // Snapshot for re-running a chunk after an event-buffer overflow.
    double* snap_state = nullptr;
    uint8_t* snap_alive = nullptr;

    DevEvent* events = nullptr;
    int64_t ev_cap = 0;
    DevStats* stats = nullptr;
    int64_t stats_cap = 0;
    MoonSample* moon = nullptr;
    int64_t moon_cap = 0;

    khip_screen_config cfg{5.0, 1.0, 0.01};
    double epoch_jd = JD_DEFAULT_EPOCH, t = 0.0, dt = 10.0;
    int64_t step = 0;
    int chunk = 256;
    int passes = 3;   // radix passes launched per step (adapted per chunk)
    int precision = KHIP_PRECISION_FP64;

    // Learned corrector: FP32 master parameters, Adam moments, BF16
    // fragments (forward and backward order), standardisation.
    float* net_p = nullptr;
    float* net_m1 = nullptr;
    float* net_m2 = nullptr;
    uint16_t* net_ffwd = nullptr;
    uint16_t* net_fbwd = nullptr;
    net::NetNorm net_norm{};
    bool net_ready = false;

    // Learned form: per-object method, per-method lists, dense-output
    // states of the sub-step boundaries, check and re-verification tallies.
    khip_learned_config lcfg{6, 10, 1e-5, 0.025, 0.1, 12000.0, 5e-8, 0, 4.0, 16.0};
    int64_t lrn_cap = 0;
    uint8_t* mode = nullptr;
    uint8_t* snap_mode = nullptr;
    int* lists = nullptr;        // 3 * cap: learned, plain, base
    int* snap_lists = nullptr;
    int* counts = nullptr;       // 3 list sizes
    int* segcnt = nullptr;       // a check's corrector-domain list: objects per segment (k_check_base)
    CheckStats* cparts = nullptr;  // a check's per-block tallies (k_check_base, k_check_verdict)
    CheckStats* cstats = nullptr;
    ReverifyStats* rstats = nullptr;
    double* substates = nullptr; // (substeps - 1) * 6 * cap: dense-output states
    int dense_k = 0;
    // Shell screening: first shell and shell count per object and window,
    // entries per object tile and their prefix (FUSED_WINDOWS_MAX windows).
    int* shell_lo = nullptr;
    uint16_t* shell_n = nullptr;
    int* tile_cnt = nullptr;
    int* tile_base = nullptr;
    int tstride = 0;
    int64_t shell_cap = 0;
    int64_t ecap = 0;     // sorted-element capacity of the grid buffers (>= cap)
    int64_t elaunch = 0;  // shell entries the grid kernels are launched for
    bool force_check = true;
    bool reset_checks = true;    // zero the device check tallies before the next learned run
    khip_learned_report lrep{};
    // Followed sample (plain twins of up to SAMPLE_MAX objects) and its
    // largest drift; the twins step on a side stream beside the screening.
    SampleObj* sample = nullptr;
    SampleObj* snap_sample = nullptr;
    Bands* bands = nullptr;     // the current learned step's thresholds (drift band)
    unsigned long long* sample_dev = nullptr;
    int sample_n = 0;
    bool sample_reset = true;   // start the sample afresh at the next learned chunk
    double sample_t0 = 0.0;     // time the sample started
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int n_cus = 256;

    // Learned-form step table and the captured chunk graph.
    StepInfo* steptab = nullptr;
    int64_t steptab_cap = 0;
    bool use_graph = false;
    // Learned steps with the FP64 work in full-occupancy kernels and the
    // corrector alone in the MFMA kernel (KHIP_LEARNED_FUSED=1: one kernel
    // per method, for comparison; results are bit-identical).
    bool split_learned = true;


    // A second set of screening buffers and its stream: the learned form
    // screens its windows on two streams at once (window q on unit q % 2),
    // so one window's latency-bound grid kernels run beside the other's
    // pair kernels. KHIP_WINDOW_STREAMS=0: one stream.
    ScreenUnit unit2;
    bool window_streams = true;
    int tile_items = 0;  // KHIP_TILE_ITEMS: elements per thread of the tile kernels (0: by size)
    bool unit2_used = false;  // the chunk being drained screened on unit2
    bool in_unit2 = false;    // unit2's buffers are swapped in (no timing marks)

    uint64_t gen = 1;               // bumped by every (re)allocation and configuration change
    // Instantiated chunk graphs keyed by what they were captured for (most
    // recently used last; at most kGraphCache).
    std::vector<std::pair<std::vector<int64_t>, hipGraphExec_t>> graphs;
    int64_t graph_launches = 0, graph_captures = 0;
    int shrink_votes = 0;  // consecutive chunks that needed fewer passes
    int64_t last_rerun = 0;  // reasons of the last re-run (KHIP_RERUN_* bits)

    std::vector<khip_event> host_events;
    // One drained chunk's events as the device wrote them, in pinned host
    // memory (a full-speed copy), ordered into host_events later: the chunk
    // loops do that while the next chunk runs on the device.
    khip_event* ev_pinned = nullptr;
    int64_t ev_pinned_cap = 0;
    std::vector<khip_event> ev_pageable;  // the staging when pinned memory is refused
    struct {
        bool on = false;
        int64_t step0 = 0, steps = 0, n = 0;
    } pend;
    std::vector<khip_stats> host_stats;

    bool timing = false;
    unsigned char* scratch = nullptr;  // this device's copy of the allocation-free screening's array
    int static_passes = 3;             // radix passes launched by khip_screen_static (adapted)
    std::vector<hipEvent_t> ev_pool;   // 4 per step of a chunk
    khip_timing tm{};
    std::vector<MoonSample> moon_host;

    Planes planes(int which) const {
        double* b = state[which];
        const size_t c = static_cast<size_t>(cap);
        return Planes{b, b + c, b + 2 * c, b + 3 * c, b + 4 * c, b + 5 * c};
    }
    int ntiles() const { return static_cast<int>((n_launch + TILE - 1) / TILE); }
};

namespace {

bool fail(khip_ctx* ctx, const std::string& what, hipError_t e) {
    ctx->error = what + ": " + hipGetErrorString(e);
    return false;
}

#define KHIP_TRY(ctx, call)                                         \
    do {                                                            \
        hipError_t err_ = (call);                                   \
        if (err_ != hipSuccess) return fail((ctx), #call, err_);    \
    } while (0)

template <typename T>
bool dev_alloc(khip_ctx* ctx, T** p, size_t count) {
    ctx->gen += 1;  // pointers baked into a captured graph change
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    KHIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count));
    return true;
}

// Temporary device buffers of one call (freed when it returns).
struct TrainScratch {
    std::vector<void*> ptrs;
    template <typename T>
    bool get(khip_ctx* ctx, T** p, size_t n) {
        *p = nullptr;
        KHIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * (n ? n : 1)));
        ptrs.push_back(*p);
        return true;
    }
    ~TrainScratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

void free_all(khip_ctx* ctx) {
    void* ptrs[] = {ctx->state[0], ctx->state[1], ctx->alive, ctx->radius, ctx->size_m,
                    ctx->keys[0], ctx->keys[1], ctx->idx[0], ctx->idx[1], ctx->hist,
                    ctx->totals, ctx->bcount, ctx->cell_key, ctx->cell_begin, ctx->cell_count,
                    ctx->cell_of, ctx->ranges, ctx->lrec, ctx->rec, ctx->coltab, ctx->scen, ctx->steptab, ctx->mass, ctx->snap_mass,
                    ctx->bstate, ctx->snap_bstate, ctx->jobs, ctx->samples, ctx->blogs,
                    ctx->snap_state,
                    ctx->snap_alive, ctx->events, ctx->stats, ctx->moon, ctx->gs,
                    ctx->step_parts, ctx->win_parts, ctx->pair_parts, ctx->dense_parts, ctx->dense, ctx->items,
                    ctx->exact_parts, ctx->surv, ctx->net_p, ctx->net_m1, ctx->net_m2,
                    ctx->net_ffwd, ctx->net_fbwd, ctx->mode, ctx->snap_mode, ctx->lists,
                    ctx->snap_lists, ctx->counts, ctx->segcnt, ctx->cparts, ctx->cstats, ctx->rstats, ctx->substates,
                    ctx->shell_lo, ctx->shell_n, ctx->tile_cnt, ctx->tile_base, ctx->sample, ctx->snap_sample,
                    ctx->sample_dev, ctx->bands};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

bool ensure_items(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->item_cap && ctx->items) return true;
    if (!dev_alloc(ctx, &ctx->items, want)) return false;
    ctx->item_cap = want;
    return true;
}

bool ensure_surv(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->surv_cap && ctx->surv) return true;
    if (!dev_alloc(ctx, &ctx->surv, want)) return false;
    ctx->surv_cap = want;
    return true;
}

// Buffers indexed by sorted element or cell: objects for the plain grid,
// shell entries for the shell screening.
bool ensure_sort_capacity(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->ecap) return true;
    const int64_t cap = std::max<int64_t>(want, ctx->cap);
    const int64_t tiles = (cap + TBLOCK - 1) / TBLOCK;  // tiles of one element per thread at most
    ctx->ecap = cap;
    return dev_alloc(ctx, &ctx->keys[0], cap) && dev_alloc(ctx, &ctx->keys[1], cap) &&
           dev_alloc(ctx, &ctx->idx[0], cap) && dev_alloc(ctx, &ctx->idx[1], cap) &&
           dev_alloc(ctx, &ctx->hist, RADIX * tiles) && dev_alloc(ctx, &ctx->bcount, tiles) &&
           dev_alloc(ctx, &ctx->cell_key, cap) && dev_alloc(ctx, &ctx->cell_begin, cap) &&
           dev_alloc(ctx, &ctx->cell_count, cap) && dev_alloc(ctx, &ctx->cell_of, cap) &&
           dev_alloc(ctx, &ctx->ranges, NRANGES * static_cast<size_t>(cap)) &&
           dev_alloc(ctx, &ctx->lrec, 2 * static_cast<size_t>(cap)) &&
           dev_alloc(ctx, &ctx->pair_parts, (cap + BLOCK - 1) / BLOCK) && dev_alloc(ctx, &ctx->dense, cap);
}

bool ensure_capacity(khip_ctx* ctx, int64_t n) {
    if (n <= ctx->cap) return true;
    // Objects are re-uploaded by the caller after a growth.
    const int64_t cap = n + n / 2 + 1024;
    ctx->cap = cap;
    for (int b = 0; b < 2; ++b)
        if (!dev_alloc(ctx, &ctx->state[b], 6 * static_cast<size_t>(cap))) return false;
    return dev_alloc(ctx, &ctx->alive, cap) && dev_alloc(ctx, &ctx->radius, cap) &&
           dev_alloc(ctx, &ctx->size_m, cap) && dev_alloc(ctx, &ctx->mass, cap) &&
           dev_alloc(ctx, &ctx->snap_mass, cap) && ensure_sort_capacity(ctx, cap) &&
           dev_alloc(ctx, &ctx->totals, RADIX) &&
           dev_alloc(ctx, &ctx->rec, 2 * static_cast<size_t>(cap)) &&
           dev_alloc(ctx, &ctx->snap_state, 6 * static_cast<size_t>(cap)) &&
           dev_alloc(ctx, &ctx->snap_alive, cap) &&
           dev_alloc(ctx, &ctx->step_parts, (cap + BLOCK - 1) / BLOCK) &&
           dev_alloc(ctx, &ctx->dense_parts, MAX_DENSE_BLOCKS) &&
           dev_alloc(ctx, &ctx->exact_parts, MAX_EXACT_BLOCKS) &&
           ensure_surv(ctx, std::max<int64_t>(ctx->surv_cap, 1 << 24)) &&
           ensure_items(ctx, std::max<int64_t>(ctx->item_cap, 1 << 20));
}

bool ensure_events(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->ev_cap) return true;
    if (!dev_alloc(ctx, &ctx->events, want)) return false;
    ctx->ev_cap = want;
    return true;
}

bool ensure_stats(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->stats_cap) return true;
    if (!dev_alloc(ctx, &ctx->stats, want)) return false;
    ctx->stats_cap = want;
    return true;
}

// Exchanges the context's per-screening buffers (and its stream) with the
// second unit's: screenings enqueued in between run on unit2.
void swap_unit2(khip_ctx* c) {
    using std::swap;
    ScreenUnit& u = c->unit2;
    swap(c->gs, u.gs);
    swap(c->keys[0], u.keys[0]);
    swap(c->keys[1], u.keys[1]);
    swap(c->idx[0], u.idx[0]);
    swap(c->idx[1], u.idx[1]);
    swap(c->hist, u.hist);
    swap(c->totals, u.totals);
    swap(c->bcount, u.bcount);
    swap(c->cell_key, u.cell_key);
    swap(c->cell_begin, u.cell_begin);
    swap(c->cell_count, u.cell_count);
    swap(c->cell_of, u.cell_of);
    swap(c->ranges, u.ranges);
    swap(c->lrec, u.lrec);
    swap(c->rec, u.rec);
    swap(c->coltab, u.coltab);
    swap(c->pair_parts, u.pair_parts);
    swap(c->dense_parts, u.dense_parts);
    swap(c->exact_parts, u.exact_parts);
    swap(c->surv, u.surv);
    swap(c->surv_cap, u.surv_cap);
    swap(c->dense, u.dense);
    swap(c->items, u.items);
    swap(c->item_cap, u.item_cap);
    swap(c->events, u.events);
    swap(c->ev_cap, u.ev_cap);
    swap(c->stream, u.stream);
    c->in_unit2 = !c->in_unit2;
}

// Sizes unit2's buffers like the context's (after any growth of those).
bool sync_unit2(khip_ctx* ctx) {
    ScreenUnit& u = ctx->unit2;
    if (!u.gs) {
        KHIP_TRY(ctx, hipStreamCreateWithFlags(&u.stream, hipStreamNonBlocking));
        KHIP_TRY(ctx, hipEventCreateWithFlags(&u.ev_go, hipEventDisableTiming));
        KHIP_TRY(ctx, hipEventCreateWithFlags(&u.ev_done, hipEventDisableTiming));
        if (!dev_alloc(ctx, &u.gs, 1) || !dev_alloc(ctx, &u.coltab, size_t{1} << COL_BITS) ||
            !dev_alloc(ctx, &u.totals, RADIX) || !dev_alloc(ctx, &u.dense_parts, MAX_DENSE_BLOCKS) ||
            !dev_alloc(ctx, &u.exact_parts, MAX_EXACT_BLOCKS))
            return false;
        KHIP_TRY(ctx, hipMemset(u.gs, 0, sizeof(GridState)));
        KHIP_TRY(ctx, hipMemset(u.coltab, 0, sizeof(int4) << COL_BITS));
        KHIP_TRY(ctx, hipDeviceSynchronize());
    }
    if (u.ecap != ctx->ecap) {
        const int64_t cap = ctx->ecap;
        const int64_t tiles = (cap + TBLOCK - 1) / TBLOCK;  // tiles of one element per thread at most
        if (!dev_alloc(ctx, &u.keys[0], cap) || !dev_alloc(ctx, &u.keys[1], cap) || !dev_alloc(ctx, &u.idx[0], cap) ||
            !dev_alloc(ctx, &u.idx[1], cap) || !dev_alloc(ctx, &u.hist, RADIX * tiles) ||
            !dev_alloc(ctx, &u.bcount, tiles) || !dev_alloc(ctx, &u.cell_key, cap) ||
            !dev_alloc(ctx, &u.cell_begin, cap) || !dev_alloc(ctx, &u.cell_count, cap) ||
            !dev_alloc(ctx, &u.cell_of, cap) || !dev_alloc(ctx, &u.ranges, NRANGES * static_cast<size_t>(cap)) ||
            !dev_alloc(ctx, &u.lrec, 2 * static_cast<size_t>(cap)) ||
            !dev_alloc(ctx, &u.pair_parts, (cap + BLOCK - 1) / BLOCK) || !dev_alloc(ctx, &u.dense, cap))
            return false;
        u.ecap = cap;
    }
    if (u.cap != ctx->cap) {
        if (!dev_alloc(ctx, &u.rec, 2 * static_cast<size_t>(ctx->cap))) return false;
        u.cap = ctx->cap;
    }
    if (u.surv_cap != ctx->surv_cap) {
        if (!dev_alloc(ctx, &u.surv, ctx->surv_cap)) return false;
        u.surv_cap = ctx->surv_cap;
    }
    if (u.item_cap != ctx->item_cap) {
        if (!dev_alloc(ctx, &u.items, ctx->item_cap)) return false;
        u.item_cap = ctx->item_cap;
    }
    if (u.ev_cap != ctx->ev_cap) {
        if (!dev_alloc(ctx, &u.events, ctx->ev_cap)) return false;
        u.ev_cap = ctx->ev_cap;
    }
    return true;
}

void free_unit2(khip_ctx* ctx) {
    ScreenUnit& u = ctx->unit2;
    if (u.stream) (void)hipStreamSynchronize(u.stream);
    void* ptrs[] = {u.gs, u.keys[0], u.keys[1], u.idx[0], u.idx[1], u.hist, u.totals, u.bcount, u.cell_key,
                    u.cell_begin, u.cell_count, u.cell_of, u.ranges, u.lrec, u.rec, u.coltab, u.pair_parts,
                    u.dense_parts, u.exact_parts, u.surv, u.dense, u.items, u.events};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (u.ev_go) (void)hipEventDestroy(u.ev_go);
    if (u.ev_done) (void)hipEventDestroy(u.ev_done);
    if (u.stream) (void)hipStreamDestroy(u.stream);
    u = ScreenUnit{};
}

// Moon samples for steps [0, nsteps) starting at time t: entry 2k is the
// Moon at the start of step k, 2k+1 at its midpoint; entry 2k+2 (the end of
// step k) is the start of step k+1, bit for bit, since both are evaluated at
// the same double t + dt (orbital.cpp:287-289, 299).
bool upload_moon(khip_ctx* ctx, double t_start, int64_t nsteps) {
    const int64_t entries = 2 * nsteps + 1;
    if (entries > ctx->moon_cap) {
        if (!dev_alloc(ctx, &ctx->moon, entries)) return false;
        ctx->moon_cap = entries;
    }
    std::vector<MoonSample>& tab = ctx->moon_host;
    tab.resize(static_cast<size_t>(entries));
    double t = t_start;
    auto fill = [&](MoonSample& s, const Vec3& m) {
        s.m[0] = m.x; s.m[1] = m.y; s.m[2] = m.z;
        const double mn = norm(m);
        const double f = kessler::MU_MOON / (mn * mn * mn);
        s.ind[0] = m.x * f; s.ind[1] = m.y * f; s.ind[2] = m.z * f;
    };
    for (int64_t k = 0; k < nsteps; ++k) {
        fill(tab[2 * k], moon_position(ctx->epoch_jd + t / 86400.0));
        fill(tab[2 * k + 1], moon_position(ctx->epoch_jd + (t + ctx->dt / 2) / 86400.0));
        fill(tab[2 * k + 2], moon_position(ctx->epoch_jd + (t + ctx->dt) / 86400.0));
        t += ctx->dt;
    }
    KHIP_TRY(ctx, hipMemcpyAsync(ctx->moon, tab.data(), sizeof(MoonSample) * entries,
                                 hipMemcpyHostToDevice, ctx->stream));
    return true;
}

// Initial device event capacity (48 B each). Cascades with fresh fragment
// clouds report ~1,000 sibling conjunctions per step; a chunk of 256 steps
// fits without a re-run.
constexpr int64_t kInitialEvents = 1 << 22;

inline int grid_of(int64_t n) { return static_cast<int>((n + BLOCK - 1) / BLOCK); }
// Grid of the kernels that write per-block StepPartials (k_propagate,
// k_reduce, k_substep; grid-stride): at most PART_BLOCKS partials for
// k_grid_setup's single block to combine.
constexpr int PART_BLOCKS = 1024;
inline int part_grid(int64_t n) { return std::max(1, std::min(grid_of(n), PART_BLOCKS)); }

// Timing marks of step slot `slot` in the current chunk: 0 before the
// propagation, 1 after it, 2 before the pair kernels, 3 after k_pairs_lanes,
// 4 after k_pairs_dense, 5 after k_exact.
// The learned form uses one slot per sub-step: marks 1-5 around that
// sub-step's screening (1 before its reduction), and, in the slot of the
// step's first sub-step, 6 / 7 around the propagation; mark 0 of the last
// sub-step's slot follows the re-verification.
constexpr int MARKS = 8;
int g_slot = 0;

bool ensure_event_pool(khip_ctx* ctx, int64_t steps) {
    while (static_cast<int64_t>(ctx->ev_pool.size()) < MARKS * steps) {
        hipEvent_t e;
        KHIP_TRY(ctx, hipEventCreate(&e));
        ctx->ev_pool.push_back(e);
    }
    return true;
}

void mark(khip_ctx* ctx, int k) {
    if (ctx->timing && !ctx->in_unit2) (void)hipEventRecord(ctx->ev_pool[MARKS * g_slot + k], ctx->stream);
}

// Enqueues the screening kernels for the step whose start / end states are
// `before` / `after` (stats record `stat_slot`). `prev` (if set) is the
// previous screening's record, written by this screening's k_grid_setup;
// with `defer` set, this screening's record is returned there instead of
// launched (the next screening must then take it as `prev`).
// shell_q >= 0: shell screening of learned window shell_q (sorted elements
// are the shell entries k_windows counted; ctx->elaunch of them at most).
bool enqueue_screen(khip_ctx* ctx, Planes pb, Planes pa, double t0, int64_t step, int stat_slot, double dt,
                    double thr, int midpoint, const StepInfo* info = nullptr, int sub_q = 0, int sub_n = 1,
                    int sub_per = 1, const StepPartial* parts = nullptr, const FinishArgs* prev = nullptr,
                    FinishArgs* defer = nullptr, int shell_q = -1, const Bands* bands = nullptr) {
    const int n_obj = static_cast<int>(ctx->n_launch);
    const int obj_tiles = ctx->ntiles();
    const bool shells = shell_q >= 0;
    // Sorted elements the grid kernels are launched for.
    const int n = shells ? static_cast<int>(ctx->elaunch) : n_obj;
    // Tiles of the tile kernels: one to three elements per thread, the
    // fewest that keep the tiles within one round of blocks (two 16-wave
    // blocks per CU), so small populations still spread over the CUs. The
    // radix passes (and k_keys, which counts the first pass per tile) use
    // `ti`; the cell kernels at most two (at three k_cell_write needs 80
    // registers and holds one block per CU).
    const int64_t round = 2LL * ctx->n_cus * TBLOCK;
    const int ti = ctx->tile_items > 0 ? ctx->tile_items : n <= round ? 1 : (n <= 2 * round || !shells ? 2 : 3);
    const int ci = ti < 2 ? ti : 2;
    const int tiles = static_cast<int>((static_cast<int64_t>(n) + ti * TBLOCK - 1) / (ti * TBLOCK));
    const int ctiles = static_cast<int>((static_cast<int64_t>(n) + ci * TBLOCK - 1) / (ci * TBLOCK));
    hipStream_t s = ctx->stream;
    const TileCfg& tc = ctx->tile;
    FinishArgs fin{};
    if (prev) fin = *prev;
    ShellScreen sh{0, nullptr, nullptr, 0, 0};
    if (shells)
        sh = ShellScreen{1, ctx->tile_cnt + static_cast<size_t>(shell_q) * ctx->tstride,
                         ctx->tile_base + static_cast<size_t>(shell_q) * ctx->tstride, obj_tiles,
                         static_cast<unsigned int>(n)};
    k_grid_setup<<<1, TBLOCK, 0, s>>>(ctx->gs, parts ? parts : ctx->step_parts, part_grid(n_obj), dt, thr, ctx->passes,
                                     midpoint, ctx->sbits, tc, fin, n_obj, sh, bands);
    if (shells)
        k_keys_shell<<<obj_tiles, TBLOCK, 0, s>>>(
            pb, pa, ctx->alive, ctx->scen, n_obj, ctx->gs, ctx->shell_lo + static_cast<size_t>(shell_q) * ctx->cap,
            ctx->shell_n + static_cast<size_t>(shell_q) * ctx->cap, sh.tile_base, static_cast<unsigned int>(n),
            ctx->keys[0], ctx->idx[0], ctx->rec);
    else
        (ti == 1 ? k_keys<1> : ti == 2 ? k_keys<2> : k_keys<3>)<<<tiles, TBLOCK, 0, s>>>(
            pb, pa, midpoint, ctx->alive, ctx->scen, n, ctx->gs, ctx->keys[0], ctx->idx[0], ctx->hist, tiles,
            ctx->rec);
    auto hist_kernel = ti == 1 ? k_hist<1> : ti == 2 ? k_hist<2> : k_hist<3>;
    auto scatter_kernel = ti == 1 ? k_scatter<1> : ti == 2 ? k_scatter<2> : k_scatter<3>;
    for (int pass = 0; pass < ctx->passes; ++pass) {
        const int src = pass & 1, dst = src ^ 1;
        // k_keys counts the first pass's digits; shell entries are not in
        // object tiles, so their first pass is counted here too.
        if (pass > 0 || shells)
            hist_kernel<<<tiles, TBLOCK, 0, s>>>(ctx->keys[src], pass, ctx->gs, ctx->hist, tiles);
        k_digit_scan<<<RADIX, BLOCK, 0, s>>>(ctx->hist, tiles, pass, ctx->gs, ctx->totals);
        scatter_kernel<<<tiles, TBLOCK, 0, s>>>(ctx->keys[src], ctx->idx[src], ctx->keys[dst], ctx->idx[dst], pass,
                                               ctx->gs, ctx->hist, ctx->totals, tiles);
    }
    (ci == 1 ? k_cell_count<1> : k_cell_count<2>)<<<ctiles, TBLOCK, 0, s>>>(ctx->keys[0], ctx->keys[1], ctx->gs,
                                                                           ctx->bcount);
    (ci == 1 ? k_cell_write<1> : k_cell_write<2>)<<<ctiles, TBLOCK, 0, s>>>(
        ctx->keys[0], ctx->keys[1], ctx->idx[0], ctx->idx[1], ctx->gs, ctx->bcount, ctiles, ctx->rec,
        ctx->cell_key, ctx->cell_begin, ctx->cell_of, ctx->lrec, ctx->coltab);
    const unsigned int icap = static_cast<unsigned int>(std::min<int64_t>(ctx->item_cap, 0x7fffffff));
    k_neighbours<<<grid_of(n), BLOCK, 0, s>>>(ctx->gs, ctx->cell_key, ctx->cell_begin,
                                              ctx->cell_count, ctx->ranges, ctx->pair_parts, ctx->dense,
                                              icap, ctx->coltab);
    k_items<<<tc.dense_blocks, BLOCK, 0, s>>>(ctx->gs, ctx->dense, ctx->cell_begin, ctx->cell_count,
                                           ctx->ranges, ctx->items, icap);
    mark(ctx, 2);
    PairParams P;
    P.before = pb;
    P.after = pa;
    P.radius = ctx->radius;
    P.cell_of = ctx->cell_of;
    P.cell_begin = ctx->cell_begin;
    P.cell_count = ctx->cell_count;
    P.ranges = ctx->ranges;
    P.lrec = ctx->lrec;
    P.events = ctx->events;
    P.cap = static_cast<unsigned long long>(ctx->ev_cap);
    P.surv = ctx->surv;
    P.surv_cap = static_cast<unsigned long long>(ctx->surv_cap);
    P.step = step;
    P.t0 = t0;
    P.info = info;
    P.sub_q = sub_q;
    P.sub_n = sub_n;
    P.sub_per = sub_per;
    P.dt = dt;
    P.thr = thr;
    P.vmin2 = ctx->cfg.min_encounter_speed_km_s * ctx->cfg.min_encounter_speed_km_s;
    P.bands = bands;
    P.dt_f = static_cast<float>(P.dt);
    P.thr_f = static_cast<float>(P.thr);
    P.vmin2_f = static_cast<float>(P.vmin2);
    P.gdt3_f = static_cast<float>(GRAVITY_GRADIENT) * P.dt_f * P.dt_f * P.dt_f;
    P.base_f = P.thr_f + 1e-3f + DR_ERR_KM + DV_ERR_KM_S * P.dt_f;
    const int pair_blocks = std::max(1, (n + PAIR_GROUP - 1) / PAIR_GROUP);
    if (tc.sparse == 1) {
        if (shells)
            k_pairs_units<true><<<pair_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->pair_parts);
        else
            k_pairs_units<false><<<pair_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->pair_parts);
    } else if (shells) {
        k_pairs_lanes<true><<<pair_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->pair_parts);
    } else {
        k_pairs_lanes<false><<<pair_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->pair_parts);
    }
    mark(ctx, 3);
    if (shells)
        k_pairs_dense<true><<<tc.dense_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->items, ctx->lrec, icap,
                                                             ctx->dense_parts);
    else
        k_pairs_dense<false><<<tc.dense_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->items, ctx->lrec, icap,
                                                              ctx->dense_parts);
    mark(ctx, 4);
    k_exact<<<tc.exact_blocks, BLOCK, 0, s>>>(ctx->gs, P, ctx->exact_parts);
    mark(ctx, 5);
    const FinishArgs mine{ctx->stats + stat_slot, ctx->pair_parts, grid_of(n), pair_blocks, ctx->dense_parts,
                          tc.dense_blocks, ctx->exact_parts, tc.exact_blocks, info ? sub_per - 1 : 0};
    if (defer)
        *defer = mine;
    else
        k_finish<<<1, TBLOCK, 0, s>>>(ctx->gs, mine.out, 1, mine.parts, mine.nparts, mine.npair, mine.dparts,
                                     mine.ndparts, mine.eparts, mine.neparts, mine.fill);
    KHIP_TRY(ctx, hipGetLastError());
    return true;
}

// Called after the chunk's stream synchronisation.
void accumulate_timing(khip_ctx* ctx, int64_t steps, bool screened) {
    for (int64_t k = 0; k < steps; ++k) {
        hipEvent_t* e = ctx->ev_pool.data() + MARKS * k;
        float a = 0, b = 0, c = 0, d = 0, f = 0, g = 0;
        (void)hipEventElapsedTime(&a, e[0], e[1]);
        ctx->tm.propagate_ms += a;
        if (screened) {
            (void)hipEventElapsedTime(&b, e[1], e[2]);
            (void)hipEventElapsedTime(&c, e[2], e[3]);
            (void)hipEventElapsedTime(&f, e[3], e[4]);
            (void)hipEventElapsedTime(&g, e[4], e[5]);
            (void)hipEventElapsedTime(&d, e[0], e[5]);
            ctx->tm.grid_ms += b;
            ctx->tm.pairs_sparse_ms += c;
            ctx->tm.pairs_dense_ms += f;
            ctx->tm.exact_ms += g;
            ctx->tm.pairs_ms += c + f + g;
            ctx->tm.total_ms += d;
        } else {
            (void)hipEventElapsedTime(&d, e[0], e[5]);
            ctx->tm.total_ms += d;
        }
        ctx->tm.steps += 1;
    }
}

void accumulate_timing_learned(khip_ctx* ctx, int64_t steps, int sub, int nw, bool screened) {
    const int spw = sub / nw;
    for (int64_t k = 0; k < steps; ++k) {
        hipEvent_t* e0 = ctx->ev_pool.data() + MARKS * (k * sub);
        hipEvent_t* el = ctx->ev_pool.data() + MARKS * (k * sub + (nw - 1) * spw);
        float a = 0, d = 0;
        (void)hipEventElapsedTime(&a, e0[6], e0[7]);
        ctx->tm.propagate_ms += a;
        (void)hipEventElapsedTime(&d, e0[6], screened ? el[0] : e0[7]);
        ctx->tm.total_ms += d;
        if (screened) {
            for (int q = 0; q < nw; ++q) {
                hipEvent_t* e = ctx->ev_pool.data() + MARKS * (k * sub + q * spw);
                float b = 0, c = 0, f = 0, g = 0;
                (void)hipEventElapsedTime(&b, e[1], e[2]);
                (void)hipEventElapsedTime(&c, e[2], e[3]);
                (void)hipEventElapsedTime(&f, e[3], e[4]);
                (void)hipEventElapsedTime(&g, e[4], e[5]);
                ctx->tm.grid_ms += b;
                ctx->tm.pairs_sparse_ms += c;
                ctx->tm.pairs_dense_ms += f;
                ctx->tm.exact_ms += g;
                ctx->tm.pairs_ms += c + f + g;
            }
            float r = 0;
            (void)hipEventElapsedTime(&r, el[5], el[0]);
            ctx->tm.reverify_ms += r;
        }
        ctx->tm.steps += 1;
    }
}

// Waits for the chunk and moves its events and statistics to the host.
// `rerun` is set when the chunk must run again: its events overflowed the
// device buffer (grown here), or a step needed more radix passes than were
// launched (raised here). The pass count also shrinks to what the chunk
// needed, so the next chunk launches no idle passes.
// Orders a drained chunk's events (pinned staging) into host_events.
void flush_pending(khip_ctx* ctx) {
    if (!ctx->pend.on) return;
    ctx->pend.on = false;
    const int64_t step0 = ctx->pend.step0, steps = ctx->pend.steps;
    const khip_event* raw = ctx->ev_pinned ? ctx->ev_pinned : ctx->ev_pageable.data();
    const int64_t ne = ctx->pend.n;
    // (step, t, i, j) order, as Screener::finish returns each step's events
    // (collision.cpp:739-747): counting sort on the chunk's step index, then
    // each step's events sorted on their own, steps in parallel. Events the
    // plain re-verification of the learned form rejected (pad 1) are left out.
    std::vector<int64_t> start(static_cast<size_t>(steps) + 1, 0);
    int64_t nk = 0;
    for (int64_t e = 0; e < ne; ++e) {
        if (raw[e].pad == 1) continue;
        ++start[static_cast<size_t>(raw[e].step - step0) + 1];
        ++nk;
    }
    for (int64_t k = 0; k < steps; ++k) start[k + 1] += start[k];
    std::vector<khip_event>& host = ctx->host_events;
    const size_t out0 = host.size();
    if (out0 + static_cast<size_t>(nk) > host.capacity())
        host.reserve(std::max(out0 + static_cast<size_t>(nk), 2 * host.capacity()));
    host.resize(out0 + static_cast<size_t>(nk));
    khip_event* out = host.data() + out0;
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
        if (raw[e].pad != 1) out[fill[static_cast<size_t>(raw[e].step - step0)]++] = raw[e];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = 0; k < steps; ++k)
        std::sort(out + start[k], out + start[k + 1], [](const khip_event& a, const khip_event& b) {
            if (a.t != b.t) return a.t < b.t;
            if (a.i != b.i) return a.i < b.i;
            return a.j < b.j;
        });
}

// A drained chunk that has to run again: its staged events are dropped.
void discard_pending(khip_ctx* ctx) { ctx->pend.on = false; }

// Orders the staged events when the caller returns (every exit path).
struct FlushOnExit {
    khip_ctx* ctx;
    ~FlushOnExit() { flush_pending(ctx); }
};

bool drain_chunk(khip_ctx* ctx, int64_t step0, int64_t steps, int screen, bool& rerun, bool defer = false) {
    flush_pending(ctx);  // the staging buffer is about to be reused
    GridState g;
    KHIP_TRY(ctx, hipMemcpyAsync(&g, ctx->gs, sizeof(g), hipMemcpyDeviceToHost, ctx->stream));
    KHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->unit2_used) {
        // The second screening stream's needs count as the context's (the
        // two are sized alike; its events are already merged).
        GridState g2;
        KHIP_TRY(ctx, hipMemcpy(&g2, ctx->unit2.gs, sizeof(g2), hipMemcpyDeviceToHost));
        g.passes_short |= g2.passes_short;
        g.passes_max = std::max(g.passes_max, g2.passes_max);
        g.items_short |= g2.items_short;
        g.items_max = std::max(g.items_max, g2.items_max);
        g.surv_short |= g2.surv_short;
        g.surv_max = std::max(g.surv_max, g2.surv_max);
        g.entries_short |= g2.entries_short;
        g.entries_max = std::max(g.entries_max, g2.entries_max);
    }
    rerun = false;
    ctx->last_rerun = 0;
    if (screen && g.passes_short) {
        ctx->passes = std::min(MAX_PASSES, std::max(g.passes_max, ctx->passes + 1));
        ctx->last_rerun |= KHIP_RERUN_PASSES;
    }
    if (static_cast<int64_t>(g.events) > ctx->ev_cap) {
        if (!ensure_events(ctx, 2 * static_cast<int64_t>(g.events) + 1024)) return false;
        ctx->last_rerun |= KHIP_RERUN_EVENTS;
    }
    if (screen && g.items_short) {
        if (!ensure_items(ctx, 2 * static_cast<int64_t>(g.items_max) + 4096)) return false;
        ctx->last_rerun |= KHIP_RERUN_ITEMS;
    }
    if (screen && g.surv_short) {
        if (!ensure_surv(ctx, 2 * static_cast<int64_t>(g.surv_max) + 4096)) return false;
        ctx->last_rerun |= KHIP_RERUN_SURVIVORS;
    }
    if (screen && g.entries_short) {
        const int64_t want = static_cast<int64_t>(g.entries_max) + static_cast<int64_t>(g.entries_max) / 4 + TILE;
        if (want >= (1LL << 31) - TILE) {
            ctx->error = "shell screening: entries exceed the sorted-element index range";
            return false;
        }
        if (!ensure_sort_capacity(ctx, want)) return false;
        ctx->elaunch = want;
        ctx->last_rerun |= KHIP_RERUN_ENTRIES;
    }
    rerun = ctx->last_rerun != 0;
    if (rerun) return true;
    // Fewer passes only after four chunks in a row needed fewer: a bounding
    // box that hovers at a bit boundary must not cause repeated re-runs.
    if (screen && g.passes_max > 0 && g.passes_max < ctx->passes) {
        if (++ctx->shrink_votes >= 4) {
            ctx->passes = g.passes_max;
            ctx->shrink_votes = 0;
        }
    } else {
        ctx->shrink_votes = 0;
    }
    if (!screen) return true;
    const int64_t ne = static_cast<int64_t>(g.events);
    if (ne > ctx->ev_pinned_cap) {
        if (ctx->ev_pinned) (void)hipHostFree(ctx->ev_pinned);
        ctx->ev_pinned = nullptr;
        ctx->ev_pinned_cap = 0;
        const int64_t want = ne + ne / 4 + 1024;
        if (hipHostMalloc(reinterpret_cast<void**>(&ctx->ev_pinned), sizeof(khip_event) * want) == hipSuccess) {
            ctx->ev_pageable = std::vector<khip_event>();
        } else {
            // Pinned memory refused: pageable staging (a slower copy).
            (void)hipGetLastError();
            ctx->ev_pinned = nullptr;
            ctx->ev_pageable.resize(static_cast<size_t>(want));
        }
        ctx->ev_pinned_cap = want;
    }
    khip_event* stage = ctx->ev_pinned ? ctx->ev_pinned : ctx->ev_pageable.data();
    if (ne > 0) KHIP_TRY(ctx, hipMemcpy(stage, ctx->events, sizeof(khip_event) * ne, hipMemcpyDeviceToHost));
    std::vector<khip_stats> st(static_cast<size_t>(steps));
    KHIP_TRY(ctx, hipMemcpy(st.data(), ctx->stats, sizeof(khip_stats) * steps, hipMemcpyDeviceToHost));
    ctx->host_stats.insert(ctx->host_stats.end(), st.begin(), st.end());
    ctx->pend.on = true;
    ctx->pend.step0 = step0;
    ctx->pend.steps = steps;
    ctx->pend.n = ne;
    if (!defer) flush_pending(ctx);
    return true;
}

bool grow_preserve(khip_ctx* ctx, int64_t want);

__global__ void k_bu_chunk_begin(bu::BreakupState* st) {
    if (threadIdx.x != 0) return;
    st->njobs = st->nsamples = st->frag_used = 0;
    st->jobs_short = st->samples_short = st->reserve_short = 0;
    st->nlog = 0;
}

constexpr int kBreakupLogCap = 1 << 16;

bool ensure_breakup(khip_ctx* ctx) {
    if (!ctx->bstate) {
        if (!dev_alloc(ctx, &ctx->bstate, 1) || !dev_alloc(ctx, &ctx->snap_bstate, 1) ||
            !dev_alloc(ctx, &ctx->jobs, bu::MAX_JOBS) || !dev_alloc(ctx, &ctx->blogs, kBreakupLogCap))
            return false;
        KHIP_TRY(ctx, hipMemset(ctx->bstate, 0, sizeof(bu::BreakupState)));
        ctx->blog_cap = kBreakupLogCap;
    }
    if (ctx->sample_cap == 0) {
        if (!dev_alloc(ctx, &ctx->samples, 1 << 20)) return false;
        ctx->sample_cap = 1 << 20;
    }
    if (ctx->bu_reserve < ctx->bcfg.reserve) ctx->bu_reserve = ctx->bcfg.reserve;
    return true;
}

constexpr size_t kGraphCache = 8;

// Launches a chunk as a HIP graph: captured once per key (what the launches
// depend on beyond device memory: buffers, lengths, check pattern, radix
// passes) and replayed for every later chunk of the same key. Up to
// kGraphCache instantiated graphs are kept (most recently used last).
bool launch_graph(khip_ctx* ctx, const std::vector<int64_t>& key, const std::function<bool()>& enqueue) {
    hipStream_t s = ctx->stream;
    hipGraphExec_t exec = nullptr;
    for (size_t g = 0; g < ctx->graphs.size(); ++g) {
        if (ctx->graphs[g].first != key) continue;
        exec = ctx->graphs[g].second;
        std::rotate(ctx->graphs.begin() + g, ctx->graphs.begin() + g + 1, ctx->graphs.end());
        break;
    }
    if (!exec) {
        // Entries built for an older generation can never match again.
        for (size_t g = 0; g < ctx->graphs.size();) {
            if (ctx->graphs[g].first[0] != key[0] || ctx->graphs.size() >= kGraphCache) {
                (void)hipGraphExecDestroy(ctx->graphs[g].second);
                ctx->graphs.erase(ctx->graphs.begin() + g);
            } else {
                ++g;
            }
        }
        KHIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const bool ok = enqueue();
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (!ok) return false;
        if (ec != hipSuccess) return fail(ctx, "hipStreamEndCapture", ec);
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(ctx, "hipGraphInstantiate", ei);
        ctx->graphs.emplace_back(key, exec);
        ctx->graph_captures += 1;
    }
    KHIP_TRY(ctx, hipGraphLaunch(exec, s));
    ctx->graph_launches += 1;
    return true;
}

bool ensure_steptab(khip_ctx* ctx, int64_t len) {
    if (len <= ctx->steptab_cap) return true;
    if (!dev_alloc(ctx, &ctx->steptab, len)) return false;
    ctx->steptab_cap = len;
    return true;
}

bool run_steps(khip_ctx* ctx, int64_t nsteps, int screen) {
    FlushOnExit flush_guard{ctx};
    hipStream_t s = ctx->stream;
    if (ctx->n == 0) {
        for (int64_t k = 0; k < nsteps; ++k) {
            if (screen)
                ctx->host_stats.push_back(khip_stats{0, 0, 0, std::max(1.0, ctx->cfg.threshold_km), 0, 0, 0, 0});
            ctx->t += ctx->dt;
            ctx->step += 1;
        }
        return true;
    }
    // In-run breakups act on the collisions the screening finds.
    const bool bk = ctx->bcfg.enabled && screen;
    if (bk && !ensure_breakup(ctx)) return false;
    if (bk && ctx->reset_breakups) {
        KHIP_TRY(ctx, hipMemsetAsync(ctx->bstate, 0, sizeof(bu::BreakupState), s));
        ctx->reset_breakups = false;
    }
    int64_t done = 0;
    // Chunks start at one step and double up to ctx->chunk: a buffer that
    // must grow (events, work items, survivors, radix passes, breakup room)
    // costs the re-run of a short chunk only, usually right at the start.
    int64_t next_len = 1;
    while (done < nsteps) {
        const int64_t len = std::min<int64_t>(std::min<int64_t>(next_len, ctx->chunk), nsteps - done);
        next_len = std::min<int64_t>(2 * next_len, ctx->chunk);
        if (!ensure_stats(ctx, len) || !ensure_events(ctx, std::max<int64_t>(ctx->ev_cap, kInitialEvents)) ||
            (ctx->timing && !ensure_event_pool(ctx, len)))
            return false;
        if (!upload_moon(ctx, ctx->t, len)) return false;
        const double t_chunk = ctx->t;
        for (;;) {
            // Breakup room: the per-step kernels cover `reserve` dead slots past
            // the population; fragments of this chunk take them in order.
            const int64_t n0 = ctx->n;
            int reserve = 0;
            if (bk) {
                reserve = ctx->bu_reserve;
                if (!grow_preserve(ctx, n0 + reserve)) return false;
                if (n0 + reserve >= (1LL << 31) - TILE) {
                    ctx->error = "breakup room exceeds the object index range";
                    return false;
                }
                KHIP_TRY(ctx, hipMemsetAsync(ctx->alive + n0, 0, static_cast<size_t>(reserve), s));
                k_bu_chunk_begin<<<1, 64, 0, s>>>(ctx->bstate);
            }
            const int n = static_cast<int>(n0 + reserve);
            ctx->n_launch = n;
            // Snapshot of the chunk's starting population, for a re-run.
            const size_t c = static_cast<size_t>(ctx->cap);
            const int cur0 = ctx->cur;
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_state, ctx->state[cur0], sizeof(double) * 6 * c,
                                         hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_alive, ctx->alive, c, hipMemcpyDeviceToDevice, s));
            if (bk) {
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_mass, ctx->mass, sizeof(double) * c, hipMemcpyDeviceToDevice, s));
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_bstate, ctx->bstate, sizeof(bu::BreakupState),
                                             hipMemcpyDeviceToDevice, s));
            }
            bu::BreakupCfg bc;
            bc.lc_m = ctx->bcfg.lc_m;
            bc.catastrophic_j_per_g = ctx->bcfg.catastrophic_j_per_g;
            bc.radius_scale = ctx->cfg.radius_scale;
            bc.seed = ctx->bcfg.seed;
            bc.reserve = reserve;
            bc.sample_cap = ctx->sample_cap;
            // A chunk may run as a captured graph: the per-step time and
            // index then come from the step table (breakups included).
            const bool graphed = ctx->use_graph && !ctx->timing && screen;
            if (graphed) {
                if (!ensure_steptab(ctx, len)) return false;
                std::vector<StepInfo> tab(static_cast<size_t>(len));
                double tt = t_chunk;
                for (int64_t k = 0; k < len; ++k) {
                    tab[k].t = tt;
                    tab[k].macro = ctx->step + done + k;
                    tt += ctx->dt;
                }
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->steptab, tab.data(), sizeof(StepInfo) * len,
                                             hipMemcpyHostToDevice, s));
            }
            double t = t_chunk;
            int cur = cur0;
            auto enqueue_chunk = [&]() -> bool {
            k_reset<<<1, 64, 0, s>>>(ctx->gs);
            k_zero_events<<<1, 64, 0, s>>>(ctx->gs);
            t = t_chunk;
            cur = cur0;
            FinishArgs pending{};  // step k - 1's record, written by step k's grid setup
            for (int64_t k = 0; k < len; ++k) {
                const int nxt = cur ^ 1;
                g_slot = static_cast<int>(k);
                mark(ctx, 0);
                if (ctx->precision == KHIP_PRECISION_MIXED)
                    k_propagate<AccelMixed><<<part_grid(n), BLOCK, 0, s>>>(
                        ctx->planes(cur), ctx->planes(nxt), ctx->alive, n, ctx->moon + 2 * k, ctx->dt,
                        ctx->step_parts);
                else
                    k_propagate<AccelFp64><<<part_grid(n), BLOCK, 0, s>>>(
                        ctx->planes(cur), ctx->planes(nxt), ctx->alive, n, ctx->moon + 2 * k, ctx->dt,
                        ctx->step_parts);
                mark(ctx, 1);
                if (screen) {
                    if (bk) k_mark_events<<<1, 64, 0, s>>>(ctx->gs, static_cast<unsigned long long>(ctx->ev_cap));
                    FinishArgs mine{};
                    if (!enqueue_screen(ctx, ctx->planes(cur), ctx->planes(nxt), t, ctx->step + done + k,
                                        static_cast<int>(k), ctx->dt, ctx->cfg.threshold_km, 0,
                                        graphed ? ctx->steptab + k : nullptr, 0, 1, 1, nullptr,
                                        pending.out ? &pending : nullptr, k + 1 < len ? &mine : nullptr))
                        return false;
                    pending = mine;
                    if (bk) {
                        const Planes pb = ctx->planes(cur), pa = ctx->planes(nxt);
                        int* sc = ctx->sbits ? ctx->scen : nullptr;
                        const StepInfo* si = graphed ? ctx->steptab + k : nullptr;
                        k_bu_events<<<1, BLOCK, 0, s>>>(ctx->gs, ctx->events,
                                                        static_cast<unsigned long long>(ctx->ev_cap), pb, pa,
                                                        ctx->alive, ctx->mass, ctx->size_m, sc, t, si, ctx->dt, bc,
                                                        ctx->bstate, ctx->jobs, ctx->blogs, ctx->blog_cap);
                        k_bu_sample<<<EXACT_BLOCKS, BLOCK, 0, s>>>(ctx->bstate, ctx->jobs, ctx->samples, bc);
                        k_bu_budget<<<bu::MAX_JOBS / 64, 64, 0, s>>>(ctx->bstate, ctx->jobs, ctx->samples,
                                                                      ctx->blogs);
                        k_bu_bases<<<1, 64, 0, s>>>(ctx->bstate, ctx->jobs, ctx->blogs, reserve);
                        k_bu_write<<<EXACT_BLOCKS, BLOCK, 0, s>>>(ctx->bstate, ctx->jobs, ctx->samples,
                                                                  static_cast<int>(n0), reserve, pa, ctx->alive,
                                                                  ctx->radius, ctx->size_m, ctx->mass, sc,
                                                                  ctx->moon + 2 * k, 1, t, si, ctx->dt, bc, nullptr,
                                                                  nullptr, nullptr, pb);
                    }
                } else {
                    k_finish<<<1, TBLOCK, 0, s>>>(ctx->gs, ctx->stats, 0, ctx->pair_parts, 0, 0,
                                                 ctx->dense_parts, 0, ctx->exact_parts, 0, 0);
                    mark(ctx, 5);
                }
                t += ctx->dt;
                cur = nxt;
            }
            return true;
            };
            if (graphed) {
                const std::vector<int64_t> key = {static_cast<int64_t>(ctx->gen), len, cur0, screen, ctx->passes, n,
                                                  ctx->precision, -1, bk ? n0 : -1, bk ? reserve : -1,
                                                  bk ? ctx->sample_cap : -1};
                if (!launch_graph(ctx, key, enqueue_chunk)) return false;
            } else if (!enqueue_chunk()) {
                return false;
            }
            // The previous chunk's events are ordered while this one runs.
            flush_pending(ctx);
            // The chunk's end (a replayed graph runs no host code).
            t = t_chunk;
            for (int64_t k = 0; k < len; ++k) t += ctx->dt;
            cur = (len & 1) ? cur0 ^ 1 : cur0;
            KHIP_TRY(ctx, hipGetLastError());
            bool overflow = false;
            if (!drain_chunk(ctx, ctx->step + done, len, screen, overflow, true)) return false;
            bu::BreakupState bs{};
            if (bk) {
                KHIP_TRY(ctx, hipMemcpy(&bs, ctx->bstate, sizeof(bs), hipMemcpyDeviceToHost));
                if (bs.jobs_short || bs.samples_short || bs.reserve_short) {
                    if (!overflow) {
                        // drain_chunk kept this chunk's events and statistics: drop them.
                        discard_pending(ctx);
                        ctx->host_stats.resize(ctx->host_stats.size() - static_cast<size_t>(len));
                    }
                    overflow = true;
                }
            }
            if (!overflow) {
                if (ctx->timing) accumulate_timing(ctx, len, screen != 0);
                ctx->cur = cur;
                ctx->t = t;
                if (bk) {
                    const int used = std::min(bs.frag_used, reserve);
                    const int nl = std::min(bs.nlog, ctx->blog_cap);
                    std::vector<bu::LogRecord> lg(static_cast<size_t>(nl));
                    if (nl > 0)
                        KHIP_TRY(ctx, hipMemcpy(lg.data(), ctx->blogs, sizeof(bu::LogRecord) * nl,
                                                hipMemcpyDeviceToHost));
                    for (const bu::LogRecord& r : lg) {
                        khip_breakup_event e;
                        std::memcpy(&e, &r, sizeof(e));
                        e.first_index = static_cast<int32_t>(n0) + r.first;
                        ctx->host_breakups.push_back(e);
                    }
                    if (used > 0) {
                        std::vector<double> sz(static_cast<size_t>(used));
                        KHIP_TRY(ctx, hipMemcpy(sz.data(), ctx->size_m + n0, sizeof(double) * used,
                                                hipMemcpyDeviceToHost));
                        ctx->host_size.insert(ctx->host_size.end(), sz.begin(), sz.end());
                        ctx->n = n0 + used;
                    }
                    ctx->n_launch = ctx->n;
                }
                break;
            }
            // A device buffer was too small: restore the chunk's start, grow, re-run.
            ctx->tm.reruns += 1;
            ctx->tm.rerun_reasons |= ctx->last_rerun;
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->state[cur0], ctx->snap_state, sizeof(double) * 6 * c,
                                         hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->alive, ctx->snap_alive, c, hipMemcpyDeviceToDevice, s));
            if (bk) {
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->mass, ctx->snap_mass, sizeof(double) * c, hipMemcpyDeviceToDevice, s));
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->bstate, ctx->snap_bstate, sizeof(bu::BreakupState),
                                             hipMemcpyDeviceToDevice, s));
                if (bs.reserve_short) ctx->bu_reserve = std::max(2 * ctx->bu_reserve, bs.frag_used + 4096);
                if (bs.samples_short) {
                    const int want = std::max(2 * ctx->sample_cap, bs.nsamples + (1 << 16));
                    if (!dev_alloc(ctx, &ctx->samples, static_cast<size_t>(want))) return false;
                    ctx->sample_cap = want;
                }
                ctx->tm.rerun_reasons |= KHIP_RERUN_BREAKUPS;
                KHIP_TRY(ctx, hipStreamSynchronize(s));
            }
        }
        ctx->n_launch = ctx->n;
        done += len;
    }
    ctx->step += nsteps;
    return true;
}

// ---------------------------------------------------------------------------
// Learned form: host orchestration
// ---------------------------------------------------------------------------

// Screening windows per learned step: the configured count, or by default
// two (half a step each) with radial shells when the sub-steps split
// evenly, and half the sub-steps (two sub-steps per window) without them.
// Shells cut the candidates of a window so far that two windows beat three:
// at 1M objects, 2 km, three windows cost a third more grid builds than
// they save in pair tests.
int learned_windows(const khip_learned_config& L) {
    if (L.windows > 0) return L.windows;
    if (L.shell_km > 0 && L.substeps % 2 == 0) return 2;
    return L.substeps % 2 == 0 ? L.substeps / 2 : L.substeps;
}

bool ensure_learned(khip_ctx* ctx) {
    const int sub = ctx->lcfg.substeps;
    if (ctx->lrn_cap >= ctx->cap && ctx->mode && ctx->dense_k >= sub - 1) return true;
    const size_t c = static_cast<size_t>(ctx->cap);
    if (ctx->mode && ctx->lrn_cap < ctx->cap) {
        // The population outgrew the arrays (breakup room): the methods and
        // lists move to the new capacity as they are, so results do not
        // depend on when the room grew.
        const size_t c0 = static_cast<size_t>(ctx->lrn_cap);
        uint8_t* mode = nullptr;
        int* lists = nullptr;
        KHIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&mode), c));
        KHIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&lists), sizeof(int) * 3 * c));
        KHIP_TRY(ctx, hipMemset(mode, net::CLASS_PLAIN, c));
        KHIP_TRY(ctx, hipMemcpy(mode, ctx->mode, c0, hipMemcpyDeviceToDevice));
        for (int k = 0; k < 3; ++k)
            KHIP_TRY(ctx, hipMemcpy(lists + k * c, ctx->lists + k * c0, sizeof(int) * c0, hipMemcpyDeviceToDevice));
        (void)hipFree(ctx->mode);
        (void)hipFree(ctx->lists);
        ctx->mode = mode;
        ctx->lists = lists;
        ctx->gen += 1;
        if (!dev_alloc(ctx, &ctx->snap_mode, c) || !dev_alloc(ctx, &ctx->snap_lists, 3 * c + 3) ||
            !dev_alloc(ctx, &ctx->substates, static_cast<size_t>(std::max(1, sub - 1)) * 6 * c))
            return false;
        ctx->lrn_cap = ctx->cap;
        if (ctx->dense_k >= sub - 1) return true;
    }
    if (!dev_alloc(ctx, &ctx->mode, c) || !dev_alloc(ctx, &ctx->snap_mode, c) ||
        !dev_alloc(ctx, &ctx->lists, 3 * c) || !dev_alloc(ctx, &ctx->snap_lists, 3 * c + 3) ||
        !dev_alloc(ctx, &ctx->counts, 3) || !dev_alloc(ctx, &ctx->segcnt, CHECK_BLOCKS) ||
        !dev_alloc(ctx, &ctx->cparts, 2 * CHECK_BLOCKS) || !dev_alloc(ctx, &ctx->cstats, 1) || !dev_alloc(ctx, &ctx->rstats, 1) ||
        !dev_alloc(ctx, &ctx->substates, static_cast<size_t>(std::max(1, sub - 1)) * 6 * c) ||
        !dev_alloc(ctx, &ctx->win_parts, static_cast<size_t>(FUSED_WINDOWS_MAX) * PART_BLOCKS))
        return false;
    KHIP_TRY(ctx, hipMemset(ctx->mode, net::CLASS_PLAIN, c));
    KHIP_TRY(ctx, hipMemset(ctx->counts, 0, 3 * sizeof(int)));
    KHIP_TRY(ctx, hipMemset(ctx->cstats, 0, sizeof(CheckStats)));
    ctx->lrn_cap = ctx->cap;
    ctx->dense_k = sub - 1;
    ctx->force_check = true;
    return true;
}

// Shell screening buffers: first shell and shell count per object and
// window, and the per-object-tile entry counts (zeroed here once; each
// window's k_grid_setup clears the counts it consumed).
bool ensure_shells(khip_ctx* ctx) {
    const int tiles = static_cast<int>((ctx->cap + TILE - 1) / TILE);
    if (ctx->shell_cap >= ctx->cap && ctx->tile_cnt && ctx->tstride >= tiles) return true;
    const size_t c = static_cast<size_t>(ctx->cap);
    if (!dev_alloc(ctx, &ctx->shell_lo, FUSED_WINDOWS_MAX * c) || !dev_alloc(ctx, &ctx->shell_n, FUSED_WINDOWS_MAX * c) ||
        !dev_alloc(ctx, &ctx->tile_cnt, static_cast<size_t>(FUSED_WINDOWS_MAX) * tiles) ||
        !dev_alloc(ctx, &ctx->tile_base, static_cast<size_t>(FUSED_WINDOWS_MAX) * tiles))
        return false;
    KHIP_TRY(ctx, hipMemset(ctx->tile_cnt, 0, sizeof(int) * FUSED_WINDOWS_MAX * tiles));
    ctx->tstride = tiles;
    ctx->shell_cap = ctx->cap;
    return true;
}

// The radial shells for a search threshold thr: shell_km wide from 6,300 km
// to the first multiple past 8,400 km (about 2,000 km altitude), 256 km (or
// shell_km, if wider) beyond, the last one open from about 60,000 km.
ShellCfg shell_layout(double shell_km, double thr) {
    ShellCfg c;
    c.r0 = 6300.0;
    c.w = shell_km;
    c.n1 = static_cast<int>(std::ceil((8400.0 - c.r0) / c.w));
    c.r1 = c.r0 + c.n1 * c.w;
    c.w2 = std::max(256.0, shell_km);
    c.nmax = c.n1 + static_cast<int>(std::ceil(std::max(0.0, 60000.0 - c.r1) / c.w2));
    c.half = 0.5 * thr + 1e-6;
    c.r0f = static_cast<float>(c.r0);
    c.r1f = static_cast<float>(c.r1);
    c.inv_w = static_cast<float>(1.0 / c.w);
    c.inv_w2 = static_cast<float>(1.0 / c.w2);
    return c;
}

Planes dense_planes(khip_ctx* ctx, int k) {
    const size_t c = static_cast<size_t>(ctx->cap);
    double* b = ctx->substates + static_cast<size_t>(k) * 6 * c;
    return Planes{b, b + c, b + 2 * c, b + 3 * c, b + 4 * c, b + 5 * c};
}

// Moon every tau / 12 from t_start over nsteps learned steps, plus two
// entries past the end.
bool upload_moon12(khip_ctx* ctx, double t_start, int64_t nsteps) {
    const int64_t entries = 12 * nsteps + 3;
    if (entries > ctx->moon_cap) {
        if (!dev_alloc(ctx, &ctx->moon, entries)) return false;
        ctx->moon_cap = entries;
    }
    std::vector<MoonSample>& tab = ctx->moon_host;
    tab.resize(static_cast<size_t>(entries));
    const double q = ctx->dt / 12.0;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < entries; ++k) {
        const Vec3 m = moon_position(ctx->epoch_jd + (t_start + static_cast<double>(k) * q) / 86400.0);
        MoonSample& s = tab[static_cast<size_t>(k)];
        s.m[0] = m.x; s.m[1] = m.y; s.m[2] = m.z;
        const double mn = norm(m);
        const double f = kessler::MU_MOON / (mn * mn * mn);
        s.ind[0] = m.x * f; s.ind[1] = m.y * f; s.ind[2] = m.z * f;
    }
    KHIP_TRY(ctx, hipMemcpyAsync(ctx->moon, tab.data(), sizeof(MoonSample) * entries, hipMemcpyHostToDevice,
                                 ctx->stream));
    return true;
}

__global__ void k_check_begin(CheckStats* cs, int* counts) {
    if (threadIdx.x < 3) {
        cs->count[threadIdx.x] = 0;
        counts[threadIdx.x] = 0;
    }
}

__global__ void __launch_bounds__(BLOCK) k_sample_init(SampleObj* s, int ns, int stride, Planes cur,
                                                        const uint8_t* __restrict__ alive, unsigned long long* dev) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if (k == 0) {
        dev[0] = 0ULL;
        dev[1] = 0ULL;
    }
    if (k >= ns) return;
    const int i = k * stride;
    SampleObj o;
    o.r[0] = cur.x[i]; o.r[1] = cur.y[i]; o.r[2] = cur.z[i];
    o.v[0] = cur.vx[i]; o.v[1] = cur.vy[i]; o.v[2] = cur.vz[i];
    o.idx = i;
    o.on = alive[i] != 0;
    s[k] = o;
}

// One learned step of the sample: the plain sub-steps from the twin's own
// state, then its distances from the object's learned end-of-step state in
// position (dev[0]) and velocity (dev[1]) (wave maxima, atomicMax of the
// doubles' bits: distances are >= 0).
__global__ void __launch_bounds__(BLOCK) k_sample_step(SampleObj* s, int ns, Planes nxt,
                                                        const uint8_t* __restrict__ alive,
                                                        const MoonSample* __restrict__ moon13, double h, int sub,
                                                        unsigned long long* dev) {
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    double d = 0.0, dv = 0.0;
    if (k < ns && s[k].on) {
        SampleObj o = s[k];
        D3 r{o.r[0], o.r[1], o.r[2]}, v{o.v[0], o.v[1], o.v[2]};
        const int i = o.idx;
        if (!plain_substeps(r, v, moon13, h, sub) || !alive[i]) {
            s[k].on = 0;
        } else {
            o.r[0] = r.x; o.r[1] = r.y; o.r[2] = r.z;
            o.v[0] = v.x; o.v[1] = v.y; o.v[2] = v.z;
            s[k] = o;
            const double dx = nxt.x[i] - r.x, dy = nxt.y[i] - r.y, dz = nxt.z[i] - r.z;
            d = sqrt(dx * dx + dy * dy + dz * dz);
            const double ex = nxt.vx[i] - v.x, ey = nxt.vy[i] - v.y, ez = nxt.vz[i] - v.z;
            dv = sqrt(ex * ex + ey * ey + ez * ez);
        }
    }
    const unsigned long long m = wave_max_u64(static_cast<unsigned long long>(__double_as_longlong(d)));
    const unsigned long long mv = wave_max_u64(static_cast<unsigned long long>(__double_as_longlong(dv)));
    if (__lane_id() == 0 && m) atomicMax(dev, m);
    if (__lane_id() == 0 && mv) atomicMax(dev + 1, mv);
}

// One learned step's thresholds (one thread): read before the step's sample
// update is forked, so every run computes them from the same D.
__global__ void k_bands(const unsigned long long* dev, const StepInfo* info, BandCfg c, Bands* out) {
    if (threadIdx.x != 0) return;
    const double d = dev ? __longlong_as_double(static_cast<long long>(dev[0])) : 0.0;
    const double dv = dev ? __longlong_as_double(static_cast<long long>(dev[1])) : 0.0;
    *out = make_bands(c, d, dv, info->t);
}

bool ensure_sample(khip_ctx* ctx) {
    if (ctx->sample) return true;
    if (!ctx->side) KHIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    if (!ctx->ev_fork) KHIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    if (!ctx->ev_join) KHIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    return dev_alloc(ctx, &ctx->sample, SAMPLE_MAX) && dev_alloc(ctx, &ctx->snap_sample, SAMPLE_MAX) &&
           dev_alloc(ctx, &ctx->sample_dev, 2) && dev_alloc(ctx, &ctx->bands, 1);
}

bool run_steps_learned(khip_ctx* ctx, int64_t nsteps, int screen) {
    FlushOnExit flush_guard{ctx};
    if (!ctx->net_ready) {
        ctx->error = "learned form: no corrector loaded or trained";
        return false;
    }
    if (std::fabs(ctx->dt - ctx->net_norm.tau) > 1e-9 * ctx->net_norm.tau) {
        ctx->error = "learned form: the step (khip_set_time dt) must equal the corrector's tau";
        return false;
    }
    const khip_learned_config& L = ctx->lcfg;
    if (12 % L.substeps != 0) {
        ctx->error = "learned form: substeps must divide 12";
        return false;
    }
    // In-run breakups act on the collisions the screening confirms.
    const bool bk = ctx->bcfg.enabled && screen;
    int n = static_cast<int>(ctx->n);
    ctx->n_launch = ctx->n;
    hipStream_t s = ctx->stream;
    const int sub = L.substeps;
    const double h = ctx->dt / sub;
    if (n == 0) {
        for (int64_t k = 0; k < nsteps; ++k) {
            if (screen)
                for (int q = 0; q < sub; ++q)
                    ctx->host_stats.push_back(khip_stats{0, 0, 0, std::max(1.0, ctx->cfg.threshold_km), 0, 0, 0, 0});
            ctx->t += ctx->dt;
            ctx->step += 1;
        }
        return true;
    }
    if (!ensure_learned(ctx)) return false;
    const bool follow = L.margin_factor > 0;
    if (follow && !ensure_sample(ctx)) return false;
    if (ctx->reset_checks) {
        KHIP_TRY(ctx, hipMemsetAsync(ctx->cstats, 0, sizeof(CheckStats), s));
        ctx->reset_checks = false;
    }
    if (bk && !ensure_breakup(ctx)) return false;
    if (bk && ctx->reset_breakups) {
        KHIP_TRY(ctx, hipMemsetAsync(ctx->bstate, 0, sizeof(bu::BreakupState), s));
        ctx->reset_breakups = false;
    }
    const int nw = learned_windows(L);
    const int spw = sub / nw;
    const double hw = ctx->dt / nw;
    // Radial shells: with the fused window pass only (its reductions count
    // the entries), and within the index word's object range.
    const bool shells = screen && L.shell_km > 0 && nw <= FUSED_WINDOWS_MAX;
    LearnedParams lp;
    lp.tau = ctx->dt;
    lp.substeps = sub;
    lp.pad = 0;
    lp.e_max = L.e_max;
    lp.a_max = L.a_max_km;
    lp.tol_km = L.tol_km;
    lp.tol_a_km = L.tol_a_km;
    size_t c = static_cast<size_t>(ctx->cap);
    const int g_net = ctx->n_cus;
    int g_all = std::min(grid_of(n), 4 * ctx->n_cus);
    int64_t done = 0, next_len = 1;
    while (done < nsteps) {
        const int64_t len = std::min<int64_t>(std::min<int64_t>(next_len, ctx->chunk), nsteps - done);
        next_len = std::min<int64_t>(2 * next_len, ctx->chunk);
        if (!ensure_stats(ctx, len * sub) ||
            !ensure_events(ctx, std::max<int64_t>(ctx->ev_cap, kInitialEvents)) ||
            (ctx->timing && !ensure_event_pool(ctx, len * sub)) || !upload_moon12(ctx, ctx->t, len))
            return false;
        // Breakup room, as in the plain form: `reserve` dead slots past the
        // population, taken in order by this chunk's fragments. A grown
        // capacity carries the learned arrays over (ensure_learned).
        const int64_t n0 = ctx->n;
        int reserve = 0;
        if (bk) {
            reserve = ctx->bu_reserve;
            if (!grow_preserve(ctx, n0 + reserve)) return false;
            if (n0 + reserve >= (1LL << 31) - TILE) {
                ctx->error = "breakup room exceeds the object index range";
                return false;
            }
            if (!ensure_learned(ctx)) return false;
            KHIP_TRY(ctx, hipMemsetAsync(ctx->alive + n0, 0, static_cast<size_t>(reserve), s));
        }
        n = static_cast<int>(n0 + reserve);
        ctx->n_launch = n;
        c = static_cast<size_t>(ctx->cap);
        g_all = std::min(grid_of(n), 4 * ctx->n_cus);
        if (shells) {
            if (n >= (1 << OBJ_BITS)) {
                ctx->error = "shell screening: more objects than the entry index word holds";
                return false;
            }
            if (!ensure_shells(ctx)) return false;
            if (ctx->elaunch < n) {
                const int64_t want = n + n / 2 + TILE;
                if (!ensure_sort_capacity(ctx, want)) return false;
                ctx->elaunch = want;
            }
        }
        bu::BreakupCfg bc{};
        if (bk) {
            bc.lc_m = ctx->bcfg.lc_m;
            bc.catastrophic_j_per_g = ctx->bcfg.catastrophic_j_per_g;
            bc.radius_scale = ctx->cfg.radius_scale;
            bc.seed = ctx->bcfg.seed;
            bc.reserve = reserve;
            bc.sample_cap = ctx->sample_cap;
            k_bu_chunk_begin<<<1, 64, 0, s>>>(ctx->bstate);
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_mass, ctx->mass, sizeof(double) * c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_bstate, ctx->bstate, sizeof(bu::BreakupState),
                                         hipMemcpyDeviceToDevice, s));
        }
        // The followed sample: started afresh after a change of population,
        // time or configuration; saved with the chunk's other state.
        if (follow && ctx->sample_reset) {
            const int ns = static_cast<int>(std::min<int64_t>(SAMPLE_MAX, n0));
            const int stride = static_cast<int>(std::max<int64_t>(1, n0 / SAMPLE_MAX));
            k_sample_init<<<(ns + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(ctx->sample, ns, stride, ctx->planes(ctx->cur),
                                                                    ctx->alive, ctx->sample_dev);
            ctx->sample_n = ns;
            ctx->sample_t0 = ctx->t;
            ctx->sample_reset = false;
        }
        unsigned long long dev0[2] = {0, 0};
        if (follow) {
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_sample, ctx->sample, sizeof(SampleObj) * ctx->sample_n,
                                         hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(dev0, ctx->sample_dev, sizeof(dev0), hipMemcpyDeviceToHost, s));
        }
        bool retry_chunk = false;  // breakup room grown: the chunk runs again from the top
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_state, ctx->state[ctx->cur], sizeof(double) * 6 * c,
                                     hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_alive, ctx->alive, c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_mode, ctx->mode, c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_lists, ctx->lists, sizeof(int) * 3 * c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->snap_lists + 3 * c, ctx->counts, sizeof(int) * 3, hipMemcpyDeviceToDevice, s));
        CheckStats cs0;
        KHIP_TRY(ctx, hipMemcpyAsync(&cs0, ctx->cstats, sizeof(cs0), hipMemcpyDeviceToHost, s));
        KHIP_TRY(ctx, hipStreamSynchronize(s));
        const int cur0 = ctx->cur;
        const double t_chunk = ctx->t;
        const bool force0 = ctx->force_check;
        // With a drift band, each step's thresholds are computed on the
        // device (k_windows or k_bands) from the sample's drift; the host's
        // values below are then the band-free ones the layout starts from.
        const double thr = ctx->cfg.threshold_km + L.margin_km;
        const double vmin2 = ctx->cfg.min_encounter_speed_km_s * ctx->cfg.min_encounter_speed_km_s;
        const BandCfg bcfg{follow ? L.margin_factor : 0.0, ctx->cfg.threshold_km, L.margin_km,
                           ctx->cfg.min_encounter_speed_km_s, hw, ctx->dt, ctx->sample_t0};
        const Bands* bands = follow ? ctx->bands : nullptr;
        const ShellCfg shell_cfg = shell_layout(L.shell_km > 0 ? L.shell_km : 16.0, thr);
        // Windows on two streams (window q on unit q % 2) when their states
        // all come from the fused window pass.
        // (Not in a run with kernel marks: those time one stream.)
        const bool two = screen && ctx->window_streams && !ctx->timing && nw >= 2 && nw <= FUSED_WINDOWS_MAX;
        for (;;) {
            if (two && !sync_unit2(ctx)) return false;
            ctx->unit2_used = two;
            // Step table of the chunk and the host's copy of the check pattern.
            if (!ensure_steptab(ctx, len)) return false;
            std::vector<StepInfo> tab(static_cast<size_t>(len));
            std::vector<int64_t> pattern;
            int64_t checks = 0;
            bool force = force0;
            double t = t_chunk;
            for (int64_t k = 0; k < len; ++k) {
                tab[k].t = t;
                tab[k].macro = ctx->step + done + k;
                const bool chk = force || tab[k].macro % L.check_every == 0;
                pattern.push_back(chk ? 1 : 0);
                if (chk) {
                    force = false;
                    ++checks;
                }
                t += ctx->dt;
            }
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->steptab, tab.data(), sizeof(StepInfo) * len, hipMemcpyHostToDevice, s));
            const int cur_end = (len & 1) ? cur0 ^ 1 : cur0;
            // The chunk's launches: plain kernel arguments only (pointers,
            // sizes, configuration); the per-step time and index come from
            // the step table, so the same sequence serves every chunk of the
            // same shape.
            auto enqueue_chunk = [&]() -> bool {
                k_reset<<<1, 64, 0, s>>>(ctx->gs);
                k_zero_events<<<1, 64, 0, s>>>(ctx->gs);
                k_zero_reverify<<<1, 64, 0, s>>>(ctx->rstats);
                if (two) {
                    k_reset<<<1, 64, 0, s>>>(ctx->unit2.gs);
                    k_zero_events<<<1, 64, 0, s>>>(ctx->unit2.gs);
                }
                int cur = cur0;
                for (int64_t k = 0; k < len; ++k) {
                    const int nxt = cur ^ 1;
                    const MoonSample* m13 = ctx->moon + 12 * k;
                    const StepInfo* info = ctx->steptab + k;
                    g_slot = static_cast<int>(k * sub);
                    mark(ctx, 6);
                    const Planes pc = ctx->planes(cur), pn = ctx->planes(nxt);
                    if (ctx->split_learned && pattern[static_cast<size_t>(k)]) {
                        const Planes pw = dense_planes(ctx, 0);  // free until the screening's windows
                        // Plain sub-steps for everyone; then the base step's domain
                        // decided and the corrector's domain listed (in the learned
                        // list's room, rebuilt below); then the corrector on that list.
                        const int chunk = (n + CHECK_BLOCKS - 1) / CHECK_BLOCKS;
                        k_check_begin<<<1, 64, 0, s>>>(ctx->cstats, ctx->counts);
                        k_check_plain<<<grid_of(n), BLOCK, 0, s>>>(pc, pn, ctx->alive, n, m13, lp);
                        k_check_base<<<CHECK_BLOCKS, TBLOCK, 0, s>>>(pc, pn, pw, ctx->alive, ctx->mode, n, chunk, m13,
                                                                    lp, ctx->lists, ctx->segcnt, ctx->cparts);
                        k_net_step<true><<<2 * g_net, NET_BLOCK, 0, s>>>(
                            pc, pn, pw, ctx->alive, ctx->mode, ctx->lists, ctx->counts, n, m13, ctx->net_ffwd,
                            ctx->net_p, ctx->net_norm, lp, ctx->cstats, ctx->segcnt, chunk);
                        k_check_verdict<<<CHECK_BLOCKS, TBLOCK, 0, s>>>(pw, pn, ctx->mode, ctx->lists, ctx->segcnt,
                                                                        chunk, lp, ctx->cparts);
                        k_build_lists<<<std::max(1, std::min(grid_of(n), LIST_BLOCKS)), BLOCK, 0, s>>>(
                            ctx->alive, ctx->mode, n, ctx->lists, static_cast<int>(c), ctx->counts, ctx->cparts,
                            ctx->cstats);
                    } else if (ctx->split_learned) {
                        // The FP64 methods in one launch, then the corrector on the
                        // corrected list's base steps.
                        k_fp64_methods<<<g_all, BLOCK, 0, s>>>(pc, pn, ctx->alive, ctx->lists, c, ctx->counts, m13,
                                                               lp);
                        k_net_step<false><<<2 * g_net, NET_BLOCK, 0, s>>>(
                            pc, pn, pn, ctx->alive, ctx->mode, ctx->lists, ctx->counts, n, m13, ctx->net_ffwd,
                            ctx->net_p, ctx->net_norm, lp, ctx->cstats, nullptr, 0);
                    } else if (pattern[static_cast<size_t>(k)]) {
                        k_check_begin<<<1, 64, 0, s>>>(ctx->cstats, ctx->counts);
                        k_learned<KIND_CHECK><<<std::min(grid_of(n), g_net), BLOCK, 0, s>>>(
                            pc, pn, ctx->alive, ctx->mode, nullptr, ctx->counts, n, m13, ctx->net_ffwd, ctx->net_p,
                            ctx->net_norm, lp, ctx->cstats);
                        k_build_lists<<<std::max(1, std::min(grid_of(n), LIST_BLOCKS)), BLOCK, 0, s>>>(ctx->alive, ctx->mode, n, ctx->lists,
                                                                   static_cast<int>(c), ctx->counts, nullptr, ctx->cstats);
                    } else {
                        k_learned<net::CLASS_LEARNED><<<g_net, BLOCK, 0, s>>>(
                            pc, pn, ctx->alive, ctx->mode, ctx->lists, ctx->counts, n, m13, ctx->net_ffwd,
                            ctx->net_p, ctx->net_norm, lp, ctx->cstats);
                        k_learned<net::CLASS_BASE><<<g_all, BLOCK, 0, s>>>(
                            pc, pn, ctx->alive, ctx->mode, ctx->lists + 2 * c, ctx->counts, n, m13, ctx->net_ffwd,
                            ctx->net_p, ctx->net_norm, lp, ctx->cstats);
                        k_learned<net::CLASS_PLAIN><<<g_all, BLOCK, 0, s>>>(
                            pc, pn, ctx->alive, ctx->mode, ctx->lists + c, ctx->counts, n, m13, ctx->net_ffwd,
                            ctx->net_p, ctx->net_norm, lp, ctx->cstats);
                    }
                    mark(ctx, 7);
                    // The step's thresholds, from the drift so far (before
                    // this step's sample update is forked): computed by
                    // k_windows when the screening's windows come from it.
                    const bool wfused = screen && nw <= FUSED_WINDOWS_MAX;
                    // The sample's plain step runs on the side stream beside
                    // the screening (it reads the end-of-step states and
                    // flags, and writes only its own arrays); joined before
                    // anything writes the flags again.
                    auto fork_sample = [&](hipEvent_t ev) -> bool {
                        KHIP_TRY(ctx, hipStreamWaitEvent(ctx->side, ev, 0));
                        k_sample_step<<<(ctx->sample_n + BLOCK - 1) / BLOCK, BLOCK, 0, ctx->side>>>(
                            ctx->sample, ctx->sample_n, pn, ctx->alive, m13, h, sub, ctx->sample_dev);
                        KHIP_TRY(ctx, hipEventRecord(ctx->ev_join, ctx->side));
                        return true;
                    };
                    if (follow && !wfused) {
                        k_bands<<<1, 64, 0, s>>>(ctx->sample_dev, info, bcfg, ctx->bands);
                        KHIP_TRY(ctx, hipEventRecord(ctx->ev_fork, s));
                        if (!fork_sample(ctx->ev_fork)) return false;
                    }
                    if (screen) {
                        // Screening windows of tau / nw, each spw sub-steps long;
                        // up to FUSED_WINDOWS_MAX windows' states and reductions
                        // come from one pass over the step's start and end.
                        const bool fused = nw <= FUSED_WINDOWS_MAX;
                        // Per stream, the previous window's record, written by the next
                        // window's grid setup; the stream's last one by its k_reverify.
                        FinishArgs pending[2] = {};
                        if (!fused)
                            k_mark_events<<<1, 64, 0, s>>>(ctx->gs, static_cast<unsigned long long>(ctx->ev_cap));
                        for (int q = 0; q < nw; ++q) {
                            g_slot = static_cast<int>(k * sub + q * spw);
                            mark(ctx, 1);
                            const Planes pb = q == 0 ? pc : dense_planes(ctx, q - 1);
                            const Planes pa = q == nw - 1 ? pn : dense_planes(ctx, q);
                            if (!fused)
                                k_substep<<<part_grid(n), BLOCK, 0, s>>>(pc, pn, ctx->alive, n, ctx->dt, q, nw, pa,
                                                                       ctx->step_parts);
                            else if (q == 0) {
                                auto kw = nw == 1 ? k_windows<1> : nw == 2 ? k_windows<2> : nw == 3 ? k_windows<3>
                                                                                                   : k_windows<4>;
                                kw<<<part_grid(n), BLOCK, 0, s>>>(
                                    pc, pn, ctx->alive, n, ctx->dt, ctx->substates, c, ctx->win_parts, ctx->gs,
                                    static_cast<unsigned long long>(ctx->ev_cap), shells ? 1 : 0, shell_cfg,
                                    ctx->shell_lo, ctx->shell_n, ctx->tile_cnt, ctx->tstride,
                                    follow ? ctx->sample_dev : nullptr, info, bcfg, follow ? ctx->bands : nullptr);
                                // The second stream and the sample's step start once every
                                // window's states and the thresholds are written (one event).
                                const hipEvent_t go = follow ? ctx->ev_fork : (two ? ctx->unit2.ev_go : nullptr);
                                if (go) KHIP_TRY(ctx, hipEventRecord(go, s));
                                if (follow && !fork_sample(go)) return false;
                                if (two) KHIP_TRY(ctx, hipStreamWaitEvent(ctx->unit2.stream, go, 0));
                            }
                            // With two streams every window writes its own record
                            // (no record is passed between windows of different streams).
                            const bool on2 = two && (q & 1);
                            FinishArgs& prev = pending[on2 ? 1 : 0];
                            FinishArgs mine{};
                            if (on2) swap_unit2(ctx);
                            const bool ok = enqueue_screen(
                                ctx, pb, pa, 0.0, 0, static_cast<int>(k * sub + q * spw), hw, thr, 1, info, q, nw, spw,
                                fused ? ctx->win_parts + static_cast<size_t>(q) * part_grid(n) : nullptr,
                                prev.out ? &prev : nullptr, &mine, shells ? q : -1, bands);
                            if (on2) swap_unit2(ctx);
                            if (!ok) return false;
                            prev = mine;
                        }
                        // Every event found is solved again on plain trajectories;
                        // with two streams each stream its own events, then the
                        // second stream's go into the step's events.
                        auto reverify = [&](const FinishArgs& fin) {
                            k_reverify<<<EXACT_BLOCKS, BLOCK, 0, ctx->stream>>>(
                                ctx->gs, ctx->events, static_cast<unsigned long long>(ctx->ev_cap), pc, ctx->radius,
                                m13, info, h, sub, ctx->cfg.threshold_km, vmin2, ctx->rstats, bands, fin);
                        };
                        reverify(pending[0]);
                        if (two) {
                            swap_unit2(ctx);
                            reverify(pending[1]);
                            swap_unit2(ctx);
                            KHIP_TRY(ctx, hipEventRecord(ctx->unit2.ev_done, ctx->unit2.stream));
                            KHIP_TRY(ctx, hipStreamWaitEvent(s, ctx->unit2.ev_done, 0));
                            k_merge_events<<<1, TBLOCK, 0, s>>>(ctx->gs, ctx->events,
                                                                static_cast<unsigned long long>(ctx->ev_cap),
                                                                ctx->unit2.gs, ctx->unit2.events,
                                                                static_cast<unsigned long long>(ctx->unit2.ev_cap));
                        }
                        if (follow) KHIP_TRY(ctx, hipStreamWaitEvent(s, ctx->ev_join, 0));
                        if (bk) {
                            // The step's confirmed collisions, in (t, i, j) order:
                            // parents at the collision instant from the step's
                            // cubic, fragments stepped to its end (Moon: the
                            // step's start, middle and end samples, 6 entries
                            // apart), started on the plain method.
                            const double t_step = t_chunk + static_cast<double>(k) * ctx->dt;
                            int* sc = ctx->sbits ? ctx->scen : nullptr;
                            k_bu_events<<<1, BLOCK, 0, s>>>(ctx->gs, ctx->events,
                                                            static_cast<unsigned long long>(ctx->ev_cap), pc, pn,
                                                            ctx->alive, ctx->mass, ctx->size_m, sc, t_step, info,
                                                            ctx->dt, bc, ctx->bstate, ctx->jobs, ctx->blogs,
                                                            ctx->blog_cap);
                            k_bu_sample<<<EXACT_BLOCKS, BLOCK, 0, s>>>(ctx->bstate, ctx->jobs, ctx->samples, bc);
                            k_bu_budget<<<bu::MAX_JOBS / 64, 64, 0, s>>>(ctx->bstate, ctx->jobs, ctx->samples,
                                                                          ctx->blogs);
                            k_bu_bases<<<1, 64, 0, s>>>(ctx->bstate, ctx->jobs, ctx->blogs, reserve);
                            k_bu_write<<<EXACT_BLOCKS, BLOCK, 0, s>>>(
                                ctx->bstate, ctx->jobs, ctx->samples, static_cast<int>(n0), reserve, pn, ctx->alive,
                                ctx->radius, ctx->size_m, ctx->mass, sc, m13, 6, t_step, info, ctx->dt, bc,
                                ctx->mode, ctx->lists + c, ctx->counts, pc);
                        }
                        mark(ctx, 0);
                    } else {
                        k_finish<<<1, TBLOCK, 0, s>>>(ctx->gs, ctx->stats, 0, ctx->pair_parts, 0, 0, ctx->dense_parts,
                                                     0, ctx->exact_parts, 0, 0);
                        if (follow) KHIP_TRY(ctx, hipStreamWaitEvent(s, ctx->ev_join, 0));
                    }
                    cur = nxt;
                }
                return true;
            };
            if (ctx->use_graph && !ctx->timing) {
                std::vector<int64_t> key = {static_cast<int64_t>(ctx->gen), len, cur0, screen, ctx->passes, n,
                                            two ? 1 : 0, shells ? ctx->elaunch : 0, follow ? ctx->sample_n : -1,
                                            __builtin_bit_cast(int64_t, ctx->sample_t0), bk ? n0 : -1,
                                            bk ? reserve : -1, bk ? ctx->sample_cap : -1};
                key.insert(key.end(), pattern.begin(), pattern.end());
                if (!launch_graph(ctx, key, enqueue_chunk)) return false;
            } else if (!enqueue_chunk()) {
                return false;
            }
            // The previous chunk's events are ordered while this one runs.
            flush_pending(ctx);
            const int cur = cur_end;
            KHIP_TRY(ctx, hipGetLastError());
            bool overflow = false;
            const bool drained = drain_chunk(ctx, (ctx->step + done) * sub, len * sub, screen, overflow, true);
            ctx->unit2_used = false;
            if (!drained) return false;
            bu::BreakupState bs{};
            if (bk) {
                KHIP_TRY(ctx, hipMemcpy(&bs, ctx->bstate, sizeof(bs), hipMemcpyDeviceToHost));
                if (bs.jobs_short || bs.samples_short || bs.reserve_short) {
                    if (!overflow) {
                        // drain_chunk kept this chunk's events and statistics: drop them.
                        discard_pending(ctx);
                        ctx->host_stats.resize(ctx->host_stats.size() - static_cast<size_t>(len * sub));
                    }
                    overflow = true;
                    retry_chunk = true;
                }
            }
            if (!overflow) {
                if (bk) {
                    const int used = std::min(bs.frag_used, reserve);
                    const int nl = std::min(bs.nlog, ctx->blog_cap);
                    std::vector<bu::LogRecord> lg(static_cast<size_t>(nl));
                    if (nl > 0)
                        KHIP_TRY(ctx, hipMemcpy(lg.data(), ctx->blogs, sizeof(bu::LogRecord) * nl,
                                                hipMemcpyDeviceToHost));
                    for (const bu::LogRecord& r : lg) {
                        khip_breakup_event e;
                        std::memcpy(&e, &r, sizeof(e));
                        e.first_index = static_cast<int32_t>(n0) + r.first;
                        ctx->host_breakups.push_back(e);
                    }
                    if (used > 0) {
                        std::vector<double> sz(static_cast<size_t>(used));
                        KHIP_TRY(ctx, hipMemcpy(sz.data(), ctx->size_m + n0, sizeof(double) * used,
                                                hipMemcpyDeviceToHost));
                        ctx->host_size.insert(ctx->host_size.end(), sz.begin(), sz.end());
                        ctx->n = n0 + used;
                    }
                    ctx->n_launch = ctx->n;
                }
                if (ctx->timing) accumulate_timing_learned(ctx, len, sub, learned_windows(ctx->lcfg), screen != 0);
                ctx->cur = cur;
                ctx->t = t;
                ctx->force_check = force;
                CheckStats cs;
                ReverifyStats rs;
                KHIP_TRY(ctx, hipMemcpy(&cs, ctx->cstats, sizeof(cs), hipMemcpyDeviceToHost));
                KHIP_TRY(ctx, hipMemcpy(&rs, ctx->rstats, sizeof(rs), hipMemcpyDeviceToHost));
                khip_learned_report& r = ctx->lrep;
                if (checks > 0) {
                    r.learned = cs.count[net::CLASS_LEARNED];
                    r.plain = cs.count[net::CLASS_PLAIN];
                    r.base = cs.count[net::CLASS_BASE];
                }
                r.rejected = cs.rejected;
                r.dev_max_km = __builtin_bit_cast(double, cs.dev_max);
                r.checks += checks;
                r.events_verified += static_cast<int64_t>(rs.verified);
                r.events_dropped += static_cast<int64_t>(rs.dropped);
                r.events_unverified += static_cast<int64_t>(rs.kept_unverified);
                r.steps += len;
                if (follow) {
                    Bands lb;
                    KHIP_TRY(ctx, hipMemcpy(&lb, ctx->bands, sizeof(lb), hipMemcpyDeviceToHost));
                    r.band_km = lb.band;
                    r.band_v_km_s = lb.band_v;
                    unsigned long long dv[2] = {0, 0};
                    KHIP_TRY(ctx, hipMemcpy(dv, ctx->sample_dev, sizeof(dv), hipMemcpyDeviceToHost));
                    r.sample_dev_max_km = __builtin_bit_cast(double, dv[0]);
                    r.sample_dev_v_max_km_s = __builtin_bit_cast(double, dv[1]);
                    r.sample_objects = ctx->sample_n;
                }
                break;
            }
            ctx->tm.reruns += 1;
            ctx->tm.rerun_reasons |= ctx->last_rerun;
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->state[cur0], ctx->snap_state, sizeof(double) * 6 * c,
                                         hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->alive, ctx->snap_alive, c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->mode, ctx->snap_mode, c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->lists, ctx->snap_lists, sizeof(int) * 3 * c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->counts, ctx->snap_lists + 3 * c, sizeof(int) * 3,
                                         hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->cstats, &cs0, sizeof(cs0), hipMemcpyHostToDevice, s));
            if (follow) {
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->sample, ctx->snap_sample, sizeof(SampleObj) * ctx->sample_n,
                                             hipMemcpyDeviceToDevice, s));
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->sample_dev, dev0, sizeof(dev0), hipMemcpyHostToDevice, s));
            }
            if (bk) {
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->mass, ctx->snap_mass, sizeof(double) * c, hipMemcpyDeviceToDevice, s));
                KHIP_TRY(ctx, hipMemcpyAsync(ctx->bstate, ctx->snap_bstate, sizeof(bu::BreakupState),
                                             hipMemcpyDeviceToDevice, s));
            }
            if (retry_chunk) {
                if (bs.reserve_short) ctx->bu_reserve = std::max(2 * ctx->bu_reserve, bs.frag_used + 4096);
                if (bs.samples_short) {
                    const int want = std::max(2 * ctx->sample_cap, bs.nsamples + (1 << 16));
                    if (!dev_alloc(ctx, &ctx->samples, static_cast<size_t>(want))) return false;
                    ctx->sample_cap = want;
                }
                ctx->tm.rerun_reasons |= KHIP_RERUN_BREAKUPS;
                KHIP_TRY(ctx, hipStreamSynchronize(s));
                break;
            }
        }
        if (retry_chunk) {
            next_len = len;
            continue;
        }
        done += len;
    }
    ctx->step += nsteps;
    return true;
}

// Autotuner: times `steps` screened steps of the current configuration for
// each candidate launch / tiling (coordinate descent, one knob at a time),
// each trial from the same saved population and time, and keeps the
// fastest. The population, time, events and statistics are restored
// afterwards. Tilings never change results; cell_mult > 1 (allow_cell)
// changes the candidate counts in the statistics only.
bool autotune(khip_ctx* ctx, int64_t steps, bool allow_cell, double& base_ms, double& best_ms) {
    hipStream_t s = ctx->stream;
    const size_t c = static_cast<size_t>(ctx->cap);
    const bool learned = ctx->precision == KHIP_PRECISION_LEARNED;
    if (learned && !ensure_learned(ctx)) return false;
    TrainScratch sc;
    double* st;
    uint8_t *al, *md = nullptr;
    int* ls = nullptr;
    CheckStats* cs = nullptr;
    if (!sc.get(ctx, &st, 6 * c) || !sc.get(ctx, &al, c)) return false;
    if (learned && (!sc.get(ctx, &md, c) || !sc.get(ctx, &ls, 3 * c + 3) || !sc.get(ctx, &cs, 1))) return false;
    KHIP_TRY(ctx, hipMemcpyAsync(st, ctx->state[ctx->cur], sizeof(double) * 6 * c, hipMemcpyDeviceToDevice, s));
    KHIP_TRY(ctx, hipMemcpyAsync(al, ctx->alive, c, hipMemcpyDeviceToDevice, s));
    if (learned) {
        KHIP_TRY(ctx, hipMemcpyAsync(md, ctx->mode, c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ls, ctx->lists, sizeof(int) * 3 * c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ls + 3 * c, ctx->counts, sizeof(int) * 3, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(cs, ctx->cstats, sizeof(CheckStats), hipMemcpyDeviceToDevice, s));
    }
    const double t0 = ctx->t;
    const int64_t step0 = ctx->step;
    const int cur0 = ctx->cur;
    const bool force0 = ctx->force_check, timing0 = ctx->timing, graph0 = ctx->use_graph;
    const khip_learned_report lrep0 = ctx->lrep;
    // The followed sample, when one is running, goes back with the rest.
    const bool sreset0 = ctx->sample_reset;
    const int sn0 = ctx->sample_n;
    const double sample_t00 = ctx->sample_t0;
    SampleObj* sm = nullptr;
    unsigned long long* sd = nullptr;
    if (learned && !sreset0 && ctx->sample) {
        if (!sc.get(ctx, &sm, SAMPLE_MAX) || !sc.get(ctx, &sd, 2)) return false;
        KHIP_TRY(ctx, hipMemcpyAsync(sm, ctx->sample, sizeof(SampleObj) * SAMPLE_MAX, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(sd, ctx->sample_dev, 2 * sizeof(*sd), hipMemcpyDeviceToDevice, s));
    }
    const khip_timing tm0 = ctx->tm;
    std::vector<khip_event> ev0;
    std::vector<khip_stats> st0;
    ev0.swap(ctx->host_events);
    st0.swap(ctx->host_stats);
    ctx->timing = false;
    ctx->use_graph = false;
    auto restore = [&]() -> bool {
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->state[cur0], st, sizeof(double) * 6 * c, hipMemcpyDeviceToDevice, s));
        KHIP_TRY(ctx, hipMemcpyAsync(ctx->alive, al, c, hipMemcpyDeviceToDevice, s));
        if (learned) {
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->mode, md, c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->lists, ls, sizeof(int) * 3 * c, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->counts, ls + 3 * c, sizeof(int) * 3, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->cstats, cs, sizeof(CheckStats), hipMemcpyDeviceToDevice, s));
        }
        ctx->t = t0;
        ctx->step = step0;
        ctx->cur = cur0;
        ctx->force_check = force0;
        ctx->lrep = lrep0;
        ctx->sample_reset = sreset0;
        ctx->sample_n = sn0;
        ctx->sample_t0 = sample_t00;
        if (sm) {
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->sample, sm, sizeof(SampleObj) * SAMPLE_MAX, hipMemcpyDeviceToDevice, s));
            KHIP_TRY(ctx, hipMemcpyAsync(ctx->sample_dev, sd, 2 * sizeof(*sd), hipMemcpyDeviceToDevice, s));
        }
        ctx->host_events.clear();
        ctx->host_stats.clear();
        return true;
    };
    auto trial = [&](const TileCfg& tc, double& ms) -> bool {
        ctx->tile = tc;
        ctx->gen += 1;
        if (!restore()) return false;
        KHIP_TRY(ctx, hipStreamSynchronize(s));
        const auto w0 = std::chrono::steady_clock::now();
        if (!(learned ? run_steps_learned(ctx, steps, 1) : run_steps(ctx, steps, 1))) return false;
        KHIP_TRY(ctx, hipStreamSynchronize(s));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        return true;
    };
    TileCfg bestc = ctx->tile;
    double bt = 0.0;
    bool ok = trial(bestc, bt) && trial(bestc, bt);  // the first run grows buffers
    base_ms = bt;
    struct Knob {
        int TileCfg::*field;
        std::vector<int> values;
    };
    // The dense-cell threshold and the long-range cut act together (a large
    // threshold pays only with a long cut): they are searched as pairs.
    // The sparse-cell kernel first: partners on the lanes suits catalogues,
    // one own object per lane the crowded cells of fresh fragment clouds.
    for (int v : {0, 1}) {
        if (!ok || bestc.sparse == v) continue;
        TileCfg cand = bestc;
        cand.sparse = v;
        double t = 0.0;
        ok = trial(cand, t);
        if (ok && t < 0.98 * bt) {
            bestc = cand;
            bt = t;
        }
    }
    const std::pair<int, int> split[] = {{32, 64}, {64, 128}, {128, 192}, {128, 256}, {256, 256}, {256, 512}};
    for (const auto& dl : split) {
        if (!ok || (bestc.dense_min == dl.first && bestc.long_range == dl.second)) continue;
        TileCfg cand = bestc;
        cand.dense_min = dl.first;
        cand.long_range = dl.second;
        double t = 0.0;
        ok = trial(cand, t);
        if (ok && t < 0.98 * bt) {
            bestc = cand;
            bt = t;
        }
    }
    const Knob knobs[] = {{&TileCfg::seg, {128, 256, 512}},
                          {&TileCfg::dense_blocks, {512, 1024, 2048, 4096}},
                          {&TileCfg::exact_blocks, {256, 512, 1024, 2048}}};
    for (const Knob& k : knobs) {
        for (int v : k.values) {
            if (!ok || bestc.*(k.field) == v) continue;
            TileCfg cand = bestc;
            cand.*(k.field) = v;
            double t = 0.0;
            ok = trial(cand, t);
            if (ok && t < 0.98 * bt) {
                bestc = cand;
                bt = t;
            }
        }
    }
    if (allow_cell) {
        for (double m : {1.1, 1.25, 1.5}) {
            if (!ok) break;
            TileCfg cand = bestc;
            cand.cell_mult = m;
            double t = 0.0;
            ok = trial(cand, t);
            if (ok && t < 0.98 * bt) {
                bestc = cand;
                bt = t;
            }
        }
    }
    best_ms = bt;
    ctx->tile = bestc;
    ctx->gen += 1;
    const bool restored = restore();
    ctx->host_events.swap(ev0);
    ctx->host_stats.swap(st0);
    ctx->timing = timing0;
    ctx->use_graph = graph0;
    ctx->tm = tm0;
    return ok && restored;
}

bool upload_objects(khip_ctx* ctx, int64_t n, const double* states, const uint8_t* alive,
                    const double* size_m) {
    if (!ensure_capacity(ctx, n)) return false;
    ctx->n = n;
    ctx->n_launch = n;
    ctx->cur = 0;
    ctx->sample_reset = true;
    ctx->host_size.assign(static_cast<size_t>(n), 0.8);
    if (n == 0) return true;
    const size_t c = static_cast<size_t>(ctx->cap);
    std::vector<double> planes(6 * c, 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 6; ++k) planes[k * c + i] = states[6 * i + k];
    KHIP_TRY(ctx, hipMemcpy(ctx->state[0], planes.data(), sizeof(double) * 6 * c, hipMemcpyHostToDevice));
    KHIP_TRY(ctx, hipMemcpy(ctx->state[1], planes.data(), sizeof(double) * 6 * c, hipMemcpyHostToDevice));
    std::vector<uint8_t> a(c, 0);
    for (int64_t i = 0; i < n; ++i) a[i] = alive ? (alive[i] ? 1 : 0) : 1;
    KHIP_TRY(ctx, hipMemcpy(ctx->alive, a.data(), c, hipMemcpyHostToDevice));
    for (int64_t i = 0; i < n; ++i) ctx->host_size[i] = size_m ? size_m[i] : 0.8;
    std::vector<double> rad(c, 0.0);
    for (int64_t i = 0; i < n; ++i)
        rad[i] = 0.5 * ctx->host_size[i] * 1e-3 * ctx->cfg.radius_scale;
    KHIP_TRY(ctx, hipMemcpy(ctx->radius, rad.data(), sizeof(double) * c, hipMemcpyHostToDevice));
    std::vector<double> sz(c, 0.0);
    std::copy(ctx->host_size.begin(), ctx->host_size.end(), sz.begin());
    KHIP_TRY(ctx, hipMemcpy(ctx->size_m, sz.data(), sizeof(double) * c, hipMemcpyHostToDevice));
    KHIP_TRY(ctx, hipMemset(ctx->mass, 0, sizeof(double) * c));
    return true;
}

// Grows the capacity to at least `want` objects keeping the live population
// (states, alive flags, sizes, radii, masses, scenario ids).
bool grow_preserve(khip_ctx* ctx, int64_t want) {
    if (want <= ctx->cap) return true;
    const int64_t n = ctx->n;
    const size_t c0 = static_cast<size_t>(ctx->cap);
    std::vector<double> st(6 * static_cast<size_t>(n)), ms(static_cast<size_t>(n));
    std::vector<uint8_t> al(static_cast<size_t>(n));
    std::vector<int> sc;
    for (int k = 0; k < 6; ++k)
        KHIP_TRY(ctx, hipMemcpy(st.data() + k * n, ctx->state[ctx->cur] + k * c0, sizeof(double) * n,
                                hipMemcpyDeviceToHost));
    KHIP_TRY(ctx, hipMemcpy(al.data(), ctx->alive, n, hipMemcpyDeviceToHost));
    KHIP_TRY(ctx, hipMemcpy(ms.data(), ctx->mass, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (ctx->sbits) {
        sc.resize(static_cast<size_t>(n));
        KHIP_TRY(ctx, hipMemcpy(sc.data(), ctx->scen, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    if (!ensure_capacity(ctx, want)) return false;
    const size_t c = static_cast<size_t>(ctx->cap);
    std::vector<double> planes(6 * c, 0.0);
    for (int k = 0; k < 6; ++k) std::copy(st.begin() + k * n, st.begin() + (k + 1) * n, planes.begin() + k * c);
    for (int b = 0; b < 2; ++b)
        KHIP_TRY(ctx, hipMemcpy(ctx->state[b], planes.data(), sizeof(double) * 6 * c, hipMemcpyHostToDevice));
    al.resize(c, 0);
    KHIP_TRY(ctx, hipMemcpy(ctx->alive, al.data(), c, hipMemcpyHostToDevice));
    ms.resize(c, 0.0);
    KHIP_TRY(ctx, hipMemcpy(ctx->mass, ms.data(), sizeof(double) * c, hipMemcpyHostToDevice));
    std::vector<double> sz(c, 0.0), rad(c, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        sz[i] = ctx->host_size[i];
        rad[i] = 0.5 * ctx->host_size[i] * 1e-3 * ctx->cfg.radius_scale;
    }
    KHIP_TRY(ctx, hipMemcpy(ctx->size_m, sz.data(), sizeof(double) * c, hipMemcpyHostToDevice));
    KHIP_TRY(ctx, hipMemcpy(ctx->radius, rad.data(), sizeof(double) * c, hipMemcpyHostToDevice));
    if (ctx->sbits) {
        if (!dev_alloc(ctx, &ctx->scen, c)) return false;
        sc.resize(c, 0);
        KHIP_TRY(ctx, hipMemcpy(ctx->scen, sc.data(), sizeof(int) * c, hipMemcpyHostToDevice));
    }
    ctx->cur = 0;
    // The learned form's methods and lists move with the population
    // (ensure_learned), so the check pattern does not depend on growth.
    return true;
}

bool refresh_radius(khip_ctx* ctx) {
    if (ctx->n == 0) return true;
    std::vector<double> rad(static_cast<size_t>(ctx->n));
    for (int64_t i = 0; i < ctx->n; ++i)
        rad[i] = 0.5 * ctx->host_size[i] * 1e-3 * ctx->cfg.radius_scale;
    KHIP_TRY(ctx, hipMemcpy(ctx->radius, rad.data(), sizeof(double) * ctx->n, hipMemcpyHostToDevice));
    return true;
}

// ---------------------------------------------------------------------------
// Learned corrector: parameters, training, inference
// ---------------------------------------------------------------------------

bool net_alloc(khip_ctx* ctx) {
    if (ctx->net_p) return true;
    return dev_alloc(ctx, &ctx->net_p, net::NPARAM) && dev_alloc(ctx, &ctx->net_m1, net::NPARAM) &&
           dev_alloc(ctx, &ctx->net_m2, net::NPARAM) &&
           dev_alloc(ctx, &ctx->net_ffwd, static_cast<size_t>(net::NFRAG_FWD) * net::FRAG_HALVES) &&
           dev_alloc(ctx, &ctx->net_fbwd, static_cast<size_t>(net::NFRAG_BWD) * net::FRAG_HALVES);
}

void net_pack(khip_ctx* ctx) {
    const int total = (net::NFRAG_FWD + net::NFRAG_BWD) * net::FRAG_HALVES;
    net::k_net_pack<<<(total + 255) / 256, 256, 0, ctx->stream>>>(ctx->net_p, ctx->net_ffwd, ctx->net_fbwd);
}

bool net_upload(khip_ctx* ctx, const float* params) {
    if (!net_alloc(ctx)) return false;
    KHIP_TRY(ctx, hipMemcpyAsync(ctx->net_p, params, sizeof(float) * net::NPARAM, hipMemcpyHostToDevice,
                                 ctx->stream));
    KHIP_TRY(ctx, hipMemsetAsync(ctx->net_m1, 0, sizeof(float) * net::NPARAM, ctx->stream));
    KHIP_TRY(ctx, hipMemsetAsync(ctx->net_m2, 0, sizeof(float) * net::NPARAM, ctx->stream));
    net_pack(ctx);
    KHIP_TRY(ctx, hipGetLastError());
    KHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return true;
}

// PyTorch-style initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for
// weights and biases, from a seeded splitmix64 stream.
std::vector<float> net_init_params(uint64_t seed) {
    std::vector<float> p(net::NPARAM);
    uint64_t st = seed;
    auto uni = [&](float bound) {
        st += 1;
        const uint64_t z = net::mix64(st);
        const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
        return static_cast<float>((2.0 * u - 1.0) * bound);
    };
    struct L { int w, b, nin, nout; };
    const L layers[4] = {{net::OFF_W1, net::OFF_B1, net::NIN, net::NH},
                         {net::OFF_W2, net::OFF_B2, net::NH, net::NH},
                         {net::OFF_W3, net::OFF_B3, net::NH, net::NH},
                         {net::OFF_W4, net::OFF_B4, net::NH, net::NOUT}};
    for (const L& l : layers) {
        const float bound = 1.0f / std::sqrt(static_cast<float>(l.nin));
        for (int k = 0; k < l.nin * l.nout; ++k) p[l.w + k] = uni(bound);
        for (int k = 0; k < l.nout; ++k) p[l.b + k] = uni(bound);
    }
    return p;
}


bool net_alloc_tensors(khip_ctx* ctx, TrainScratch& sc, int B, net::TrainTensors& T) {
    const size_t b = static_cast<size_t>(B);
    if (!sc.get(ctx, &T.hx, 16 * b) || !sc.get(ctx, &T.h1, 128 * b) || !sc.get(ctx, &T.h2, 128 * b) ||
        !sc.get(ctx, &T.h3, 128 * b) || !sc.get(ctx, &T.g1, 128 * b) || !sc.get(ctx, &T.g2, 128 * b) ||
        !sc.get(ctx, &T.g3, 128 * b) || !sc.get(ctx, &T.g4, 16 * b) || !sc.get(ctx, &T.loss, b / 16))
        return false;
    KHIP_TRY(ctx, hipMemsetAsync(T.hx, 0, sizeof(uint16_t) * 16 * b, ctx->stream));
    return true;
}

constexpr int NET_KC = 1024;  // batch columns per weight-gradient chunk

// One forward + backward + weight-gradient pass over a batch of B columns
// (B a multiple of NET_KC); the chunk partials land in `part`.
void net_fwdbwd(khip_ctx* ctx, const float* X, const float* Y, const uint32_t* pool, uint32_t npool,
                uint64_t seed, int B, const net::TrainTensors& T, float* part) {
    hipStream_t s = ctx->stream;
    const int waves = B / 16;
    net::k_net_fwdbwd<<<(waves + 3) / 4, 256, 0, s>>>(X, Y, pool, npool, seed, B, ctx->net_ffwd, ctx->net_fbwd,
                                                      ctx->net_p, T);
    const int tasks = net::DW_TILES * (B / NET_KC);
    net::k_net_dw<<<(tasks + 3) / 4, 256, 0, s>>>(T, B, NET_KC, part);
}

bool net_train(khip_ctx* ctx, const double* orbits, int64_t n, const khip_train_config& cfg,
               khip_train_report& rep, float* curve) {
    hipStream_t s = ctx->stream;
    TrainScratch sc;
    const double tau = cfg.tau;
    const int per = cfg.samples_per_orbit, every = cfg.spacing_steps;
    const int steps = per * every;
    // Moon every tau / 12 over the sampling window.
    const int64_t nm = 12LL * steps + 1;
    std::vector<MoonSample> moon(static_cast<size_t>(nm));
    for (int64_t k = 0; k < nm; ++k) {
        const double t = ctx->t + static_cast<double>(k) * tau / 12.0;
        const Vec3 m = moon_position(ctx->epoch_jd + t / 86400.0);
        MoonSample& ms = moon[static_cast<size_t>(k)];
        ms.m[0] = m.x; ms.m[1] = m.y; ms.m[2] = m.z;
        const double mn = norm(m);
        const double f = kessler::MU_MOON / (mn * mn * mn);
        ms.ind[0] = m.x * f; ms.ind[1] = m.y * f; ms.ind[2] = m.z * f;
    }
    const size_t ns = static_cast<size_t>(n) * per;
    double* d_orb;
    MoonSample* d_moon;
    float *fx, *fy;
    uint8_t* valid;
    if (!sc.get(ctx, &d_orb, 6 * static_cast<size_t>(n)) || !sc.get(ctx, &d_moon, nm) ||
        !sc.get(ctx, &fx, ns * net::NIN) || !sc.get(ctx, &fy, ns * net::NOUT) || !sc.get(ctx, &valid, ns))
        return false;
    KHIP_TRY(ctx, hipMemcpyAsync(d_orb, orbits, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s));
    KHIP_TRY(ctx, hipMemcpyAsync(d_moon, moon.data(), sizeof(MoonSample) * nm, hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1, e2;
    KHIP_TRY(ctx, hipEventCreate(&e0));
    KHIP_TRY(ctx, hipEventCreate(&e1));
    KHIP_TRY(ctx, hipEventCreate(&e2));
    (void)hipEventRecord(e0, s);
    net::k_net_samples<<<grid_of(n), 256, 0, s>>>(d_orb, static_cast<int>(n), steps, every, d_moon, tau, fx, fy,
                                                  valid, per, cfg.e_max, cfg.a_max_km);
    (void)hipEventRecord(e1, s);
    KHIP_TRY(ctx, hipGetLastError());
    // Standardisation from the valid samples (host, FP64 sums).
    std::vector<float> hx(ns * net::NIN), hy(ns * net::NOUT);
    std::vector<uint8_t> hv(ns);
    KHIP_TRY(ctx, hipMemcpyAsync(hx.data(), fx, sizeof(float) * hx.size(), hipMemcpyDeviceToHost, s));
    KHIP_TRY(ctx, hipMemcpyAsync(hy.data(), fy, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, s));
    KHIP_TRY(ctx, hipMemcpyAsync(hv.data(), valid, ns, hipMemcpyDeviceToHost, s));
    KHIP_TRY(ctx, hipStreamSynchronize(s));
    float gen_ms = 0.0f;
    (void)hipEventElapsedTime(&gen_ms, e0, e1);
    std::vector<uint32_t> pool;
    pool.reserve(ns);
    double sx[net::NIN] = {}, sxx[net::NIN] = {}, sy[net::NOUT] = {}, syy[net::NOUT] = {};
    for (size_t k = 0; k < ns; ++k) {
        if (!hv[k]) continue;
        pool.push_back(static_cast<uint32_t>(k));
        for (int a = 0; a < net::NIN; ++a) {
            const double v = hx[k * net::NIN + a];
            sx[a] += v;
            sxx[a] += v * v;
        }
        for (int a = 0; a < net::NOUT; ++a) {
            const double v = hy[k * net::NOUT + a];
            sy[a] += v;
            syy[a] += v * v;
        }
    }
    if (pool.size() < 1024) {
        ctx->error = "too few valid training samples (orbits outside the learned domain?)";
        return false;
    }
    const double cnt = static_cast<double>(pool.size());
    net::NetNorm nn{};
    nn.tau = tau;
    for (int a = 0; a < net::NIN; ++a) {
        const double mu = sx[a] / cnt, var = std::max(0.0, sxx[a] / cnt - mu * mu);
        nn.fmu[a] = static_cast<float>(mu);
        nn.fsig[a] = var > 1e-24 ? static_cast<float>(std::sqrt(var)) : 1.0f;
    }
    double vb[2] = {0.0, 0.0};
    for (int a = 0; a < net::NOUT; ++a) {
        const double mu = sy[a] / cnt, var = std::max(0.0, syy[a] / cnt - mu * mu);
        nn.omu[a] = static_cast<float>(mu);
        vb[a / 3] += var / 3.0;
    }
    for (int a = 0; a < net::NOUT; ++a) {
        const double sd = std::sqrt(vb[a / 3]);
        nn.osig[a] = sd > 1e-30 ? static_cast<float>(sd) : 1.0f;
    }
    net::k_net_standardise<<<static_cast<unsigned>((ns + 255) / 256), 256, 0, s>>>(fx, fy, ns, nn);
    uint32_t* d_pool;
    if (!sc.get(ctx, &d_pool, pool.size())) return false;
    KHIP_TRY(ctx, hipMemcpyAsync(d_pool, pool.data(), sizeof(uint32_t) * pool.size(), hipMemcpyHostToDevice, s));
    // Fresh parameters.
    const std::vector<float> init = net_init_params(cfg.seed);
    if (!net_upload(ctx, init.data())) return false;
    const int B = cfg.batch;
    net::TrainTensors T;
    float *part, *losses;
    if (!net_alloc_tensors(ctx, sc, B, T) || !sc.get(ctx, &part, static_cast<size_t>(B / NET_KC) * net::NPARAM) ||
        !sc.get(ctx, &losses, cfg.iterations))
        return false;
    (void)hipEventRecord(e1, s);
    const double b1 = 0.9, b2 = 0.999;
    double pb1 = 1.0, pb2 = 1.0;
    const int warm = std::max(1, cfg.iterations / 20);
    for (int it = 0; it < cfg.iterations; ++it) {
        double lr;
        if (it < warm) {
            lr = cfg.lr_max * (it + 1) / warm;
        } else {
            const double q = static_cast<double>(it - warm) / std::max(1, cfg.iterations - warm);
            lr = cfg.lr_max * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.141592653589793 * q)));
        }
        pb1 *= b1;
        pb2 *= b2;
        net_fwdbwd(ctx, fx, fy, d_pool, static_cast<uint32_t>(pool.size()),
                   cfg.seed * 0x100000001b3ULL + static_cast<uint64_t>(it) * static_cast<uint64_t>(B), B, T, part);
        net::AdamArgs a{static_cast<float>(lr), static_cast<float>(b1), static_cast<float>(b2), 1e-8f,
                        static_cast<float>(1.0 - pb1), static_cast<float>(1.0 - pb2), B / NET_KC, 1};
        net::k_net_adam<<<(net::NPARAM + 255) / 256, 256, 0, s>>>(ctx->net_p, ctx->net_m1, ctx->net_m2, part,
                                                                  nullptr, a);
        net_pack(ctx);
        net::k_net_loss<<<1, 256, 0, s>>>(T.loss, B / 16, losses + it);
    }
    (void)hipEventRecord(e2, s);
    KHIP_TRY(ctx, hipGetLastError());
    std::vector<float> lh(static_cast<size_t>(cfg.iterations));
    KHIP_TRY(ctx, hipMemcpyAsync(lh.data(), losses, sizeof(float) * lh.size(), hipMemcpyDeviceToHost, s));
    KHIP_TRY(ctx, hipStreamSynchronize(s));
    float train_ms = 0.0f;
    (void)hipEventElapsedTime(&train_ms, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    // Per-batch sums of squared error -> mean per output.
    const double scale = 1.0 / (6.0 * B);
    const int win = std::min(100, cfg.iterations);
    double first = 0.0, last = 0.0;
    for (int k = 0; k < win; ++k) {
        first += lh[k] * scale / win;
        last += lh[cfg.iterations - 1 - k] * scale / win;
    }
    if (curve)
        for (int k = 0; k < cfg.iterations; ++k) curve[k] = static_cast<float>(lh[k] * scale);
    ctx->net_norm = nn;
    ctx->net_ready = true;
    ctx->gen += 1;
    rep.samples = static_cast<int64_t>(pool.size());
    rep.loss_first = first;
    rep.loss_last = last;
    rep.train_ms = train_ms;
    rep.gen_ms = gen_ms;
    return true;
}

bool net_infer(khip_ctx* ctx, int64_t n, const float* x, float* out) {
    TrainScratch sc;
    float *dx, *dout;
    if (!sc.get(ctx, &dx, static_cast<size_t>(n) * net::NIN) || !sc.get(ctx, &dout, static_cast<size_t>(n) * net::NOUT))
        return false;
    KHIP_TRY(ctx, hipMemcpyAsync(dx, x, sizeof(float) * n * net::NIN, hipMemcpyHostToDevice, ctx->stream));
    const int64_t waves = (n + 63) / 64;
    net::k_net_infer<<<static_cast<unsigned>((waves + 3) / 4), 256, 0, ctx->stream>>>(dx, static_cast<int>(n),
                                                                                     ctx->net_ffwd, ctx->net_p, dout);
    KHIP_TRY(ctx, hipGetLastError());
    KHIP_TRY(ctx, hipMemcpyAsync(out, dout, sizeof(float) * n * net::NOUT, hipMemcpyDeviceToHost, ctx->stream));
    KHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return true;
}

bool net_gradient(khip_ctx* ctx, int64_t n, const float* x, const float* y, float* grad, float* loss) {
    TrainScratch sc;
    const int B = static_cast<int>(n);
    float *dx, *dy, *part, *dgrad, *dloss;
    net::TrainTensors T;
    if (!sc.get(ctx, &dx, static_cast<size_t>(n) * net::NIN) || !sc.get(ctx, &dy, static_cast<size_t>(n) * net::NOUT) ||
        !sc.get(ctx, &part, static_cast<size_t>(B / NET_KC) * net::NPARAM) || !sc.get(ctx, &dgrad, net::NPARAM) ||
        !sc.get(ctx, &dloss, 1) || !net_alloc_tensors(ctx, sc, B, T))
        return false;
    hipStream_t s = ctx->stream;
    KHIP_TRY(ctx, hipMemcpyAsync(dx, x, sizeof(float) * n * net::NIN, hipMemcpyHostToDevice, s));
    KHIP_TRY(ctx, hipMemcpyAsync(dy, y, sizeof(float) * n * net::NOUT, hipMemcpyHostToDevice, s));
    net_fwdbwd(ctx, dx, dy, nullptr, 0, 0, B, T, part);
    net::AdamArgs a{0.f, 0.f, 0.f, 0.f, 1.f, 1.f, B / NET_KC, 0};
    net::k_net_adam<<<(net::NPARAM + 255) / 256, 256, 0, s>>>(ctx->net_p, ctx->net_m1, ctx->net_m2, part, dgrad, a);
    net::k_net_loss<<<1, 256, 0, s>>>(T.loss, B / 16, dloss);
    KHIP_TRY(ctx, hipGetLastError());
    KHIP_TRY(ctx, hipMemcpyAsync(grad, dgrad, sizeof(float) * net::NPARAM, hipMemcpyDeviceToHost, s));
    float l = 0.0f;
    KHIP_TRY(ctx, hipMemcpyAsync(&l, dloss, sizeof(float), hipMemcpyDeviceToHost, s));
    KHIP_TRY(ctx, hipStreamSynchronize(s));
    if (loss) *loss = l / static_cast<float>(6.0 * B);
    return true;
}

// ---------------------------------------------------------------------------
// Allocation-free screening (khip_screen_static)
// ---------------------------------------------------------------------------
//
// One screening of host-held states, as khip_screen does, for callers that
// cannot allocate device memory: every device buffer it uses (the staged
// inputs, the grid, the survivor queue, the events, the grid state) is a
// region of `scratch`, a file-scope device array of fixed size that each GPU
// holds from the library's load to the end of the process. Its address on a
// device comes from hipGetSymbolAddress in khip_create (the symbol's name is
// a host-side stand-in, not a device address). Inputs above the limits below
// are rejected. Calls on one GPU share its array: they must not overlap.
constexpr int64_t SCR_OBJECTS = 32768;     // objects per call
constexpr int64_t SCR_EVENTS = 1 << 17;    // events per call
constexpr int64_t SCR_SURVIVORS = 1 << 20; // pre-filter survivors per call
constexpr int64_t SCR_ITEMS = 1 << 16;     // dense-cell work items per call
constexpr size_t SCR_TILES = (SCR_OBJECTS + TBLOCK - 1) / TBLOCK;  // tiles of one element per thread at most
constexpr size_t SCR_BLOCKS = (SCR_OBJECTS + BLOCK - 1) / BLOCK;

enum ScrRegion {
    R_IN_BEFORE, R_IN_AFTER, R_IN_ALIVE, R_IN_SIZE, R_STATE0, R_STATE1, R_ALIVE, R_RADIUS, R_KEYS0, R_KEYS1,
    R_IDX0, R_IDX1, R_HIST, R_TOTALS, R_BCOUNT, R_CELL_KEY, R_CELL_BEGIN, R_CELL_COUNT, R_CELL_OF, R_RANGES,
    R_LREC, R_REC, R_STEP_PARTS, R_PAIR_PARTS, R_DENSE_PARTS, R_EXACT_PARTS, R_DENSE, R_ITEMS, R_SURV,
    R_EVENTS, R_STATS, R_GS, R_COUNT
};

constexpr size_t scr_bytes(int r) {
    constexpr size_t N = static_cast<size_t>(SCR_OBJECTS);
    switch (r) {
    case R_IN_BEFORE: case R_IN_AFTER: case R_STATE0: case R_STATE1: return 6 * N * sizeof(double);
    case R_IN_ALIVE: case R_ALIVE: return N;
    case R_IN_SIZE: case R_RADIUS: return N * sizeof(double);
    case R_KEYS0: case R_KEYS1: case R_CELL_KEY: return N * sizeof(uint64_t);
    case R_IDX0: case R_IDX1: case R_CELL_BEGIN: case R_CELL_COUNT: case R_CELL_OF: return N * sizeof(int);
    case R_HIST: return RADIX * SCR_TILES * sizeof(int);
    case R_TOTALS: return RADIX * sizeof(int);
    case R_BCOUNT: return SCR_TILES * sizeof(int);
    case R_RANGES: return NRANGES * N * sizeof(int2);
    case R_LREC: case R_REC: return 2 * N * sizeof(float4);
    case R_STEP_PARTS: return SCR_BLOCKS * sizeof(StepPartial);
    case R_PAIR_PARTS: return SCR_BLOCKS * sizeof(PairPartial);
    case R_DENSE_PARTS: return MAX_DENSE_BLOCKS * sizeof(PairPartial);
    case R_EXACT_PARTS: return MAX_EXACT_BLOCKS * sizeof(PairPartial);
    case R_DENSE: return N * sizeof(DenseCell);
    case R_ITEMS: return SCR_ITEMS * sizeof(PairItem);
    case R_SURV: return SCR_SURVIVORS * sizeof(uint2);
    case R_EVENTS: return SCR_EVENTS * sizeof(DevEvent);
    case R_STATS: return sizeof(DevStats);
    case R_GS: return sizeof(GridState);
    default: return 0;
    }
}

constexpr size_t scr_offset(int r) {
    size_t o = 0;
    for (int k = 0; k < r; ++k) o += (scr_bytes(k) + 255) & ~static_cast<size_t>(255);
    return o;
}

constexpr size_t SCR_TOTAL = scr_offset(R_COUNT);  // about 27 MB

static __device__ __align__(256) unsigned char scratch[SCR_TOTAL];

// Staged host inputs (array-of-structures states, alive bytes, sizes) into
// the structure-of-arrays planes, alive flags and collision radii the
// screening kernels read; radius as upload_objects computes it.
__global__ void __launch_bounds__(BLOCK) k_stage(const double* before, const double* after, const uint8_t* alive_in,
                                                 const double* size_in, int have_alive, int have_size, int n,
                                                 double* s0, double* s1, uint8_t* alive, double* radius,
                                                 double scale) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    constexpr size_t c = static_cast<size_t>(SCR_OBJECTS);
    for (int k = 0; k < 6; ++k) {
        s0[k * c + i] = before[6 * static_cast<size_t>(i) + k];
        s1[k * c + i] = after[6 * static_cast<size_t>(i) + k];
    }
    alive[i] = have_alive ? (alive_in[i] ? 1 : 0) : 1;
    const double size = have_size ? size_in[i] : 0.8;
    radius[i] = 0.5 * size * 1e-3 * scale;
}

// The context's grid pointers and sizes, pointed at the scratch regions for
// one allocation-free screening and swapped back afterwards, so the
// screening kernels run unchanged on the array's memory.
struct StaticView {
    double* state[2];
    int cur;
    uint8_t* alive;
    double* radius;
    uint64_t* keys[2];
    int* idx[2];
    int *hist, *totals, *bcount;
    uint64_t* cell_key;
    int *cell_begin, *cell_count, *cell_of;
    int2* ranges;
    float4 *lrec, *rec;
    int4* coltab;
    GridState* gs;
    StepPartial* step_parts;
    PairPartial *pair_parts, *dense_parts, *exact_parts;
    uint2* surv;
    int64_t surv_cap;
    DenseCell* dense;
    PairItem* items;
    int64_t item_cap;
    DevEvent* events;
    int64_t ev_cap;
    DevStats* stats;
    int64_t stats_cap, cap, ecap, n, n_launch;
    int passes, sbits;
    int* scen;
    bool timing;
};

void swap_view(khip_ctx* c, StaticView& v) {
    using std::swap;
    swap(c->state[0], v.state[0]); swap(c->state[1], v.state[1]); swap(c->cur, v.cur);
    swap(c->alive, v.alive); swap(c->radius, v.radius);
    swap(c->keys[0], v.keys[0]); swap(c->keys[1], v.keys[1]); swap(c->idx[0], v.idx[0]); swap(c->idx[1], v.idx[1]);
    swap(c->hist, v.hist); swap(c->totals, v.totals); swap(c->bcount, v.bcount);
    swap(c->cell_key, v.cell_key); swap(c->cell_begin, v.cell_begin); swap(c->cell_count, v.cell_count);
    swap(c->cell_of, v.cell_of); swap(c->ranges, v.ranges); swap(c->lrec, v.lrec); swap(c->rec, v.rec);
    swap(c->coltab, v.coltab); swap(c->gs, v.gs); swap(c->step_parts, v.step_parts);
    swap(c->pair_parts, v.pair_parts); swap(c->dense_parts, v.dense_parts); swap(c->exact_parts, v.exact_parts);
    swap(c->surv, v.surv); swap(c->surv_cap, v.surv_cap); swap(c->dense, v.dense); swap(c->items, v.items);
    swap(c->item_cap, v.item_cap); swap(c->events, v.events); swap(c->ev_cap, v.ev_cap);
    swap(c->stats, v.stats); swap(c->stats_cap, v.stats_cap); swap(c->cap, v.cap); swap(c->ecap, v.ecap);
    swap(c->n, v.n); swap(c->n_launch, v.n_launch); swap(c->passes, v.passes); swap(c->sbits, v.sbits);
    swap(c->scen, v.scen); swap(c->timing, v.timing);
}

// khip_screen_static's work. `code` receives the C-ABI status.
bool screen_static(khip_ctx* ctx, int64_t n, const double* before, const double* after, const uint8_t* alive_after,
                   const double* size_m, double t0, double dt, khip_event* events, int64_t cap, int64_t* nevents,
                   khip_stats* stats, int& code) {
    code = KHIP_ERR_HIP;
    hipStream_t s = ctx->stream;
    const double thr = ctx->cfg.threshold_km;
    if (n == 0) {
        if (nevents) *nevents = 0;
        if (stats) *stats = khip_stats{0, 0, 0, std::max(1.0, thr), 0, 0, 0, 0, 0};
        code = KHIP_OK;
        return true;
    }
    unsigned char* b = ctx->scratch;
    auto at = [&](int r) { return b + scr_offset(r); };
    KHIP_TRY(ctx, hipMemcpyAsync(at(R_IN_BEFORE), before, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s));
    KHIP_TRY(ctx, hipMemcpyAsync(at(R_IN_AFTER), after, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s));
    if (alive_after) KHIP_TRY(ctx, hipMemcpyAsync(at(R_IN_ALIVE), alive_after, n, hipMemcpyHostToDevice, s));
    if (size_m) KHIP_TRY(ctx, hipMemcpyAsync(at(R_IN_SIZE), size_m, sizeof(double) * n, hipMemcpyHostToDevice, s));
    const int ni = static_cast<int>(n);
    k_stage<<<grid_of(n), BLOCK, 0, s>>>(
        reinterpret_cast<const double*>(at(R_IN_BEFORE)), reinterpret_cast<const double*>(at(R_IN_AFTER)),
        at(R_IN_ALIVE), reinterpret_cast<const double*>(at(R_IN_SIZE)), alive_after ? 1 : 0, size_m ? 1 : 0, ni,
        reinterpret_cast<double*>(at(R_STATE0)), reinterpret_cast<double*>(at(R_STATE1)), at(R_ALIVE),
        reinterpret_cast<double*>(at(R_RADIUS)), ctx->cfg.radius_scale);
    StaticView v;
    v.state[0] = reinterpret_cast<double*>(at(R_STATE0));
    v.state[1] = reinterpret_cast<double*>(at(R_STATE1));
    v.cur = 0;
    v.alive = at(R_ALIVE);
    v.radius = reinterpret_cast<double*>(at(R_RADIUS));
    v.keys[0] = reinterpret_cast<uint64_t*>(at(R_KEYS0));
    v.keys[1] = reinterpret_cast<uint64_t*>(at(R_KEYS1));
    v.idx[0] = reinterpret_cast<int*>(at(R_IDX0));
    v.idx[1] = reinterpret_cast<int*>(at(R_IDX1));
    v.hist = reinterpret_cast<int*>(at(R_HIST));
    v.totals = reinterpret_cast<int*>(at(R_TOTALS));
    v.bcount = reinterpret_cast<int*>(at(R_BCOUNT));
    v.cell_key = reinterpret_cast<uint64_t*>(at(R_CELL_KEY));
    v.cell_begin = reinterpret_cast<int*>(at(R_CELL_BEGIN));
    v.cell_count = reinterpret_cast<int*>(at(R_CELL_COUNT));
    v.cell_of = reinterpret_cast<int*>(at(R_CELL_OF));
    v.ranges = reinterpret_cast<int2*>(at(R_RANGES));
    v.lrec = reinterpret_cast<float4*>(at(R_LREC));
    v.rec = reinterpret_cast<float4*>(at(R_REC));
    v.coltab = nullptr;  // the column table (256 MB) stays out: neighbour runs by binary search
    v.gs = reinterpret_cast<GridState*>(at(R_GS));
    v.step_parts = reinterpret_cast<StepPartial*>(at(R_STEP_PARTS));
    v.pair_parts = reinterpret_cast<PairPartial*>(at(R_PAIR_PARTS));
    v.dense_parts = reinterpret_cast<PairPartial*>(at(R_DENSE_PARTS));
    v.exact_parts = reinterpret_cast<PairPartial*>(at(R_EXACT_PARTS));
    v.surv = reinterpret_cast<uint2*>(at(R_SURV));
    v.surv_cap = SCR_SURVIVORS;
    v.dense = reinterpret_cast<DenseCell*>(at(R_DENSE));
    v.items = reinterpret_cast<PairItem*>(at(R_ITEMS));
    v.item_cap = SCR_ITEMS;
    v.events = reinterpret_cast<DevEvent*>(at(R_EVENTS));
    v.ev_cap = SCR_EVENTS;
    v.stats = reinterpret_cast<DevStats*>(at(R_STATS));
    v.stats_cap = 1;
    v.cap = SCR_OBJECTS;
    v.ecap = SCR_OBJECTS;
    v.n = n;
    v.n_launch = n;
    v.passes = ctx->static_passes;
    v.sbits = 0;
    v.scen = nullptr;
    v.timing = false;
    const TileCfg tile0 = ctx->tile;
    const TileCfg defaults{1.0, DENSE_MIN, LONG_RANGE, SEG, DENSE_BLOCKS, EXACT_BLOCKS, 0};
    GridState g{};
    bool ok = true;
    swap_view(ctx, v);
    ctx->tile = defaults;
    for (;;) {
        // Every field a screening starts from is cleared here: the array
        // keeps the previous call's contents.
        k_reset<<<1, 64, 0, s>>>(ctx->gs);
        k_zero_events<<<1, 64, 0, s>>>(ctx->gs);
        k_reduce<<<part_grid(n), BLOCK, 0, s>>>(ctx->planes(0), ctx->alive, ni, ctx->step_parts);
        ok = enqueue_screen(ctx, ctx->planes(0), ctx->planes(1), t0, 0, 0, dt, thr, 0);
        if (ok) ok = hipMemcpyAsync(&g, ctx->gs, sizeof(g), hipMemcpyDeviceToHost, s) == hipSuccess &&
                     hipStreamSynchronize(s) == hipSuccess;
        if (!ok || !g.passes_short) break;
        ctx->passes = std::min(MAX_PASSES, std::max(g.passes_max, ctx->passes + 1));  // more key bits: again
    }
    const int passes_used = ctx->passes;
    swap_view(ctx, v);
    ctx->tile = tile0;
    ctx->static_passes = passes_used;
    if (!ok) {
        if (ctx->error.empty()) ctx->error = "khip_screen_static: a HIP call failed";
        return false;
    }
    code = KHIP_ERR_STATE;
    if (g.surv_short || g.items_short || static_cast<int64_t>(g.events) > SCR_EVENTS) {
        ctx->error = "khip_screen_static: this screening exceeds the fixed buffers (" +
                     std::to_string(g.events) + " events, " + std::to_string(g.surv_max) +
                     " pre-filter survivors, " + std::to_string(g.items_max) + " work items); use khip_screen";
        return false;
    }
    const int64_t nev = static_cast<int64_t>(g.events);
    if (nevents) *nevents = nev;
    if (nev > cap || (nev > 0 && !events)) {
        code = KHIP_ERR_ARG;
        ctx->error = "khip_screen_static: the event buffer holds " + std::to_string(cap) + " events, the screening found " +
                     std::to_string(nev);
        return false;
    }
    code = KHIP_ERR_HIP;
    if (nev > 0)
        KHIP_TRY(ctx, hipMemcpy(events, at(R_EVENTS), sizeof(khip_event) * nev, hipMemcpyDeviceToHost));
    // (t, i, j) order (collision.cpp:739-747), in place in the caller's buffer.
    std::sort(events, events + nev, [](const khip_event& a, const khip_event& b2) {
        if (a.t != b2.t) return a.t < b2.t;
        if (a.i != b2.i) return a.i < b2.i;
        return a.j < b2.j;
    });
    if (stats) KHIP_TRY(ctx, hipMemcpy(stats, at(R_STATS), sizeof(khip_stats), hipMemcpyDeviceToHost));
    code = KHIP_OK;
    return true;
}

int status(khip_ctx* ctx, bool ok) { return ok ? KHIP_OK : KHIP_ERR_HIP; }

}  // namespace

// Device address of `scratch` per device, looked up once per device.
static bool scratch_address(int device, unsigned char** out) {
    static std::mutex lock;
    static std::vector<unsigned char*> addr;
    std::lock_guard<std::mutex> hold(lock);
    if (device >= static_cast<int>(addr.size())) addr.resize(static_cast<size_t>(device) + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
        addr[device] = static_cast<unsigned char*>(p);
    }
    *out = addr[device];
    return true;
}

extern "C" {

int khip_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int khip_create(int device, khip_ctx** out, char* msg, size_t msg_cap) {
    auto say = [&](const std::string& s) {
        if (msg && msg_cap > 0) std::snprintf(msg, msg_cap, "%s", s.c_str());
    };
    if (!out) return KHIP_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        say("no HIP device");
        return KHIP_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        say("device index out of range");
        return KHIP_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) {
        say("hipSetDevice failed");
        return KHIP_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        say("hipGetDeviceProperties failed");
        return KHIP_ERR_HIP;
    }
    khip_ctx* ctx = new khip_ctx;
    ctx->device = device;
    if (const char* f = std::getenv("KHIP_LEARNED_FUSED")) ctx->split_learned = std::atoi(f) == 0;
    if (const char* f = std::getenv("KHIP_WINDOW_STREAMS")) ctx->window_streams = std::atoi(f) != 0;
    if (const char* f = std::getenv("KHIP_TILE_ITEMS")) ctx->tile_items = std::min(3, std::max(0, std::atoi(f)));


    ctx->n_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s (%s, %d CUs, %.1f GB)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount, prop.totalGlobalMem / 1e9);
    ctx->name = buf;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&ctx->gs), sizeof(GridState)) == hipSuccess &&
              hipMemset(ctx->gs, 0, sizeof(GridState)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&ctx->coltab), sizeof(int4) << COL_BITS) == hipSuccess &&
              hipMemset(ctx->coltab, 0, sizeof(int4) << COL_BITS) == hipSuccess;
    if (ok) {
        k_reset<<<1, 64, 0, ctx->stream>>>(ctx->gs);
        k_zero_events<<<1, 64, 0, ctx->stream>>>(ctx->gs);
        ok = hipStreamSynchronize(ctx->stream) == hipSuccess;
    }
    // This device's copy of the allocation-free screening's array (the
    // symbol itself is a host-side stand-in): looked up once per device,
    // with the device current, and kept.
    if (ok) ok = scratch_address(device, &ctx->scratch);
    if (!ok) {
        say("initialisation failed on " + ctx->name);
        khip_destroy(ctx);
        return KHIP_ERR_HIP;
    }
    say(ctx->name);
    *out = ctx;
    return KHIP_OK;
}

void khip_destroy(khip_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    free_unit2(ctx);
    free_all(ctx);
    if (ctx->ev_pinned) (void)hipHostFree(ctx->ev_pinned);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* khip_last_error(const khip_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int khip_set_objects(khip_ctx* ctx, int64_t n, const double* states, const uint8_t* alive,
                     const double* size_m) {
    if (!ctx || n < 0 || (n > 0 && !states) || n >= (1LL << 31) - TILE) return KHIP_ERR_ARG;
    ctx->force_check = true;
    ctx->reset_checks = true;
    ctx->reset_breakups = true;
    ctx->host_breakups.clear();
    ctx->lrep = khip_learned_report{};
    ctx->sbits = 0;
    ctx->scen_off.clear();
    return status(ctx, upload_objects(ctx, n, states, alive, size_m));
}

int khip_set_scenarios(khip_ctx* ctx, int32_t nscen, const int64_t* offsets) {
    if (!ctx || nscen < 1 || nscen > (1 << 16) || !offsets || offsets[0] != 0 || offsets[nscen] != ctx->n)
        return KHIP_ERR_ARG;
    for (int32_t k = 0; k < nscen; ++k)
        if (offsets[k + 1] < offsets[k]) return KHIP_ERR_ARG;
    ctx->scen_off.assign(offsets, offsets + nscen + 1);
    ctx->gen += 1;
    if (nscen == 1) {
        ctx->sbits = 0;
        return KHIP_OK;
    }
    const size_t c = static_cast<size_t>(std::max<int64_t>(ctx->cap, 1));
    if (!dev_alloc(ctx, &ctx->scen, c)) return KHIP_ERR_HIP;
    std::vector<int> sid(c, 0);
    for (int32_t k = 0; k < nscen; ++k)
        for (int64_t i = offsets[k]; i < offsets[k + 1]; ++i) sid[static_cast<size_t>(i)] = k;
    if (hipMemcpy(ctx->scen, sid.data(), sizeof(int) * c, hipMemcpyHostToDevice) != hipSuccess) {
        ctx->error = "hipMemcpy of scenario ids failed";
        return KHIP_ERR_HIP;
    }
    int b = 0;
    while ((1LL << b) < nscen) ++b;
    ctx->sbits = b;
    return KHIP_OK;
}

int64_t khip_object_count(const khip_ctx* ctx) { return ctx ? ctx->n : -1; }

int khip_get_objects(khip_ctx* ctx, double* states, uint8_t* alive) {
    if (!ctx) return KHIP_ERR_ARG;
    if (ctx->n == 0) return KHIP_OK;
    const size_t c = static_cast<size_t>(ctx->cap);
    if (states) {
        std::vector<double> planes(6 * c);
        if (hipMemcpy(planes.data(), ctx->state[ctx->cur], sizeof(double) * 6 * c,
                      hipMemcpyDeviceToHost) != hipSuccess) {
            ctx->error = "hipMemcpy of states failed";
            return KHIP_ERR_HIP;
        }
        for (int64_t i = 0; i < ctx->n; ++i)
            for (int k = 0; k < 6; ++k) states[6 * i + k] = planes[k * c + i];
    }
    if (alive && hipMemcpy(alive, ctx->alive, ctx->n, hipMemcpyDeviceToHost) != hipSuccess) {
        ctx->error = "hipMemcpy of alive flags failed";
        return KHIP_ERR_HIP;
    }
    return KHIP_OK;
}

int khip_get_previous(khip_ctx* ctx, int64_t n, const int32_t* idx, double* states) {
    if (!ctx || n < 0 || (n > 0 && (!idx || !states))) return KHIP_ERR_ARG;
    const size_t c = static_cast<size_t>(ctx->cap);
    const double* prev = ctx->state[ctx->cur ^ 1];
    for (int64_t k = 0; k < n; ++k) {
        if (idx[k] < 0 || idx[k] >= ctx->n) return KHIP_ERR_ARG;
        for (int a = 0; a < 6; ++a)
            if (hipMemcpy(states + 6 * k + a, prev + a * c + idx[k], sizeof(double), hipMemcpyDeviceToHost) !=
                hipSuccess) {
                ctx->error = "hipMemcpy of previous states failed";
                return KHIP_ERR_HIP;
            }
    }
    return KHIP_OK;
}

int khip_set_config(khip_ctx* ctx, const khip_screen_config* cfg) {
    if (!ctx || !cfg || !(cfg->threshold_km >= 0) || !(cfg->radius_scale >= 0) ||
        !(cfg->min_encounter_speed_km_s >= 0))
        return KHIP_ERR_ARG;
    ctx->cfg = *cfg;
    ctx->gen += 1;
    return status(ctx, refresh_radius(ctx));
}

int khip_set_time(khip_ctx* ctx, double epoch_jd, double t, double dt) {
    if (!ctx || !(dt > 0)) return KHIP_ERR_ARG;
    ctx->epoch_jd = epoch_jd;
    ctx->t = t;
    ctx->dt = dt;
    ctx->step = 0;
    ctx->gen += 1;
    ctx->sample_reset = true;
    return KHIP_OK;
}

int khip_get_time(const khip_ctx* ctx, double* t, int64_t* step) {
    if (!ctx) return KHIP_ERR_ARG;
    if (t) *t = ctx->t;
    if (step) *step = ctx->step;
    return KHIP_OK;
}

int khip_run(khip_ctx* ctx, int64_t nsteps, int screen) {
    if (!ctx || nsteps < 0) return KHIP_ERR_ARG;
    if (ctx->precision == KHIP_PRECISION_LEARNED) return status(ctx, run_steps_learned(ctx, nsteps, screen));
    ctx->sample_reset = true;  // the learned form's sample does not follow other forms
    return status(ctx, run_steps(ctx, nsteps, screen));
}

int khip_screen(khip_ctx* ctx, int64_t n, const double* before, const double* after,
                const uint8_t* alive_after, const double* size_m, double t0, double dt,
                khip_event* events, int64_t cap, int64_t* nevents, khip_stats* stats) {
    if (!ctx || n < 0 || (n > 0 && (!before || !after)) || !(dt > 0)) return KHIP_ERR_ARG;
    const double dt_saved = ctx->dt;
    ctx->dt = dt;
    bool ok = upload_objects(ctx, n, before, alive_after, size_m);
    if (ok && n > 0) {
        // Second buffer <- end-of-step states.
        const size_t c = static_cast<size_t>(ctx->cap);
        std::vector<double> planes(6 * c, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 6; ++k) planes[k * c + i] = after[6 * i + k];
        ok = hipMemcpy(ctx->state[1], planes.data(), sizeof(double) * 6 * c, hipMemcpyHostToDevice) ==
             hipSuccess;
        if (!ok) ctx->error = "hipMemcpy of end-of-step states failed";
    }
    ok = ok && ensure_stats(ctx, 1) && ensure_events(ctx, std::max<int64_t>(ctx->ev_cap, kInitialEvents));
    while (ok) {
        hipStream_t s = ctx->stream;
        k_reset<<<1, 64, 0, s>>>(ctx->gs);
        k_zero_events<<<1, 64, 0, s>>>(ctx->gs);
        g_slot = 0;
        if (ctx->timing && !ensure_event_pool(ctx, 1)) { ok = false; break; }
        if (n > 0) {
            mark(ctx, 0);
            mark(ctx, 1);
            k_reduce<<<part_grid(n), BLOCK, 0, s>>>(ctx->planes(0), ctx->alive, static_cast<int>(n),
                                                  ctx->step_parts);
            ok = enqueue_screen(ctx, ctx->planes(0), ctx->planes(1), t0, 0, 0, dt, ctx->cfg.threshold_km, 0);
        } else {
            k_finish<<<1, TBLOCK, 0, s>>>(ctx->gs, ctx->stats, 1, ctx->pair_parts, 0, 0, ctx->dense_parts, 0,
                                         ctx->exact_parts, 0, 0);
        }
        if (!ok) break;
        std::vector<khip_event> saved_ev;
        std::vector<khip_stats> saved_st;
        saved_ev.swap(ctx->host_events);
        saved_st.swap(ctx->host_stats);
        bool overflow = false;
        ok = drain_chunk(ctx, 0, 1, 1, overflow);
        std::vector<khip_event> got_ev;
        std::vector<khip_stats> got_st;
        got_ev.swap(ctx->host_events);
        got_st.swap(ctx->host_stats);
        ctx->host_events.swap(saved_ev);
        ctx->host_stats.swap(saved_st);
        if (!ok || overflow) continue;
        if (nevents) *nevents = static_cast<int64_t>(got_ev.size());
        for (size_t k = 0; k < got_ev.size() && events && static_cast<int64_t>(k) < cap; ++k)
            events[k] = got_ev[k];
        if (stats && !got_st.empty()) *stats = got_st[0];
        break;
    }
    ctx->dt = dt_saved;
    return status(ctx, ok);
}

int khip_screen_static(khip_ctx* ctx, int64_t n, const double* before, const double* after,
                       const uint8_t* alive_after, const double* size_m, double t0, double dt,
                       khip_event* events, int64_t cap, int64_t* nevents, khip_stats* stats) {
    if (!ctx || n < 0 || (n > 0 && (!before || !after)) || !(dt > 0) || cap < 0) return KHIP_ERR_ARG;
    if (n > SCR_OBJECTS) {
        ctx->error = "khip_screen_static: " + std::to_string(n) + " objects, at most " + std::to_string(SCR_OBJECTS);
        return KHIP_ERR_ARG;
    }
    if (!ctx->scratch) {
        ctx->error = "khip_screen_static: the device array's address is unknown";
        return KHIP_ERR_STATE;
    }
    int code = KHIP_OK;
    screen_static(ctx, n, before, after, alive_after, size_m, t0, dt, events, cap, nevents, stats, code);
    return code;
}

int khip_take_events(khip_ctx* ctx, khip_event* out, int64_t cap, int64_t* n) {
    if (!ctx || !n) return KHIP_ERR_ARG;
    const int64_t have = static_cast<int64_t>(ctx->host_events.size());
    if (!out) {
        *n = have;
        return KHIP_OK;
    }
    const int64_t m = std::min(have, cap);
    std::copy(ctx->host_events.begin(), ctx->host_events.begin() + m, out);
    ctx->host_events.erase(ctx->host_events.begin(), ctx->host_events.begin() + m);
    *n = m;
    return KHIP_OK;
}

int khip_take_stats(khip_ctx* ctx, khip_stats* out, int64_t cap, int64_t* n) {
    if (!ctx || !n) return KHIP_ERR_ARG;
    const int64_t have = static_cast<int64_t>(ctx->host_stats.size());
    if (!out) {
        *n = have;
        return KHIP_OK;
    }
    const int64_t m = std::min(have, cap);
    std::copy(ctx->host_stats.begin(), ctx->host_stats.begin() + m, out);
    ctx->host_stats.erase(ctx->host_stats.begin(), ctx->host_stats.begin() + m);
    *n = m;
    return KHIP_OK;
}

int khip_set_timing(khip_ctx* ctx, int enabled) {
    if (!ctx) return KHIP_ERR_ARG;
    ctx->timing = enabled != 0;
    return KHIP_OK;
}

int khip_get_timing(const khip_ctx* ctx, khip_timing* out) {
    if (!ctx || !out) return KHIP_ERR_ARG;
    *out = ctx->tm;
    return KHIP_OK;
}

int khip_reset_timing(khip_ctx* ctx) {
    if (!ctx) return KHIP_ERR_ARG;
    ctx->tm = khip_timing{};
    return KHIP_OK;
}

int khip_set_precision(khip_ctx* ctx, int mode) {
    if (!ctx || (mode != KHIP_PRECISION_FP64 && mode != KHIP_PRECISION_MIXED && mode != KHIP_PRECISION_LEARNED))
        return KHIP_ERR_ARG;
    ctx->precision = mode;
    ctx->force_check = true;
    return KHIP_OK;
}

int khip_set_learned(khip_ctx* ctx, const khip_learned_config* cfg) {
    if (!ctx || !cfg || cfg->substeps < 1 || cfg->substeps > 12 || 12 % cfg->substeps != 0 ||
        cfg->check_every < 1 || !(cfg->tol_km >= 0) || !(cfg->margin_km >= 0) || !(cfg->e_max >= 0) ||
        !(cfg->a_max_km >= 0) || !(cfg->tol_a_km >= 0) || cfg->windows < 0 ||
        (cfg->windows > 0 && cfg->substeps % cfg->windows != 0) ||
        !(cfg->shell_km == 0 || (cfg->shell_km >= 1.0 && cfg->shell_km < 1e5)) ||
        !(cfg->margin_factor >= 0 && cfg->margin_factor < 1e6))
        return KHIP_ERR_ARG;
    ctx->lcfg = *cfg;
    ctx->force_check = true;
    ctx->sample_reset = true;
    ctx->gen += 1;
    return KHIP_OK;
}

int khip_set_breakup(khip_ctx* ctx, const khip_breakup_config* cfg) {
    if (!ctx || !cfg || cfg->reserve < 1 || !(cfg->lc_m > 0) || !(cfg->catastrophic_j_per_g > 0)) return KHIP_ERR_ARG;
    ctx->bcfg = *cfg;
    ctx->bu_reserve = cfg->reserve;
    ctx->gen += 1;
    return KHIP_OK;
}

int khip_set_masses(khip_ctx* ctx, int64_t n, const double* mass_kg) {
    if (!ctx || n != ctx->n || (n > 0 && !mass_kg)) return KHIP_ERR_ARG;
    if (n > 0 && hipMemcpy(ctx->mass, mass_kg, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
        ctx->error = "hipMemcpy of masses failed";
        return KHIP_ERR_HIP;
    }
    return KHIP_OK;
}

int khip_get_masses(khip_ctx* ctx, int64_t n, double* mass_kg) {
    if (!ctx || n != ctx->n || (n > 0 && !mass_kg)) return KHIP_ERR_ARG;
    if (n > 0 && hipMemcpy(mass_kg, ctx->mass, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) {
        ctx->error = "hipMemcpy of masses failed";
        return KHIP_ERR_HIP;
    }
    return KHIP_OK;
}

int khip_get_sizes(const khip_ctx* ctx, int64_t n, double* size_m) {
    if (!ctx || n != ctx->n || (n > 0 && !size_m)) return KHIP_ERR_ARG;
    std::copy(ctx->host_size.begin(), ctx->host_size.begin() + n, size_m);
    return KHIP_OK;
}

int khip_take_breakups(khip_ctx* ctx, khip_breakup_event* out, int64_t cap, int64_t* n) {
    static_assert(sizeof(khip_breakup_event) == sizeof(bu::LogRecord), "breakup record layout");
    if (!ctx || !n) return KHIP_ERR_ARG;
    const int64_t have = static_cast<int64_t>(ctx->host_breakups.size());
    if (!out) {
        *n = have;
        return KHIP_OK;
    }
    const int64_t m = std::min(have, cap);
    std::copy(ctx->host_breakups.begin(), ctx->host_breakups.begin() + m, out);
    ctx->host_breakups.erase(ctx->host_breakups.begin(), ctx->host_breakups.begin() + m);
    *n = m;
    return KHIP_OK;
}

int khip_set_tuning(khip_ctx* ctx, const khip_tuning* t) {
    static_assert(sizeof(khip_tuning) == sizeof(TileCfg), "tuning layout");
    if (!ctx || !t || !(t->cell_mult >= 1.0) || t->cell_mult > 4.0 || t->dense_min < 2 || t->long_range < 1 ||
        t->long_range > 16384 ||
        t->seg < 1 || t->dense_blocks < 1 || t->dense_blocks > MAX_DENSE_BLOCKS || t->exact_blocks < 1 ||
        t->exact_blocks > MAX_EXACT_BLOCKS || t->sparse_kernel < 0 || t->sparse_kernel > 1)
        return KHIP_ERR_ARG;
    std::memcpy(&ctx->tile, t, sizeof(TileCfg));
    ctx->gen += 1;
    return KHIP_OK;
}

int khip_get_tuning(const khip_ctx* ctx, khip_tuning* t) {
    if (!ctx || !t) return KHIP_ERR_ARG;
    std::memcpy(t, &ctx->tile, sizeof(TileCfg));
    return KHIP_OK;
}

int khip_autotune(khip_ctx* ctx, int64_t steps, int allow_cell, khip_tuning* best, double* base_ms,
                  double* best_ms) {
    if (!ctx || steps < 1 || steps > 4096) return KHIP_ERR_ARG;
    if (ctx->n == 0) {
        ctx->error = "autotune: no objects";
        return KHIP_ERR_STATE;
    }
    double b0 = 0.0, b1 = 0.0;
    const bool ok = autotune(ctx, steps, allow_cell != 0, b0, b1);
    if (best) std::memcpy(best, &ctx->tile, sizeof(TileCfg));
    if (base_ms) *base_ms = b0;
    if (best_ms) *best_ms = b1;
    return status(ctx, ok);
}

int khip_set_graph(khip_ctx* ctx, int enabled) {
    if (!ctx) return KHIP_ERR_ARG;
    ctx->use_graph = enabled != 0;
    return KHIP_OK;
}

int khip_get_graph_counts(const khip_ctx* ctx, int64_t* captures, int64_t* launches) {
    if (!ctx) return KHIP_ERR_ARG;
    if (captures) *captures = ctx->graph_captures;
    if (launches) *launches = ctx->graph_launches;
    return KHIP_OK;
}

int khip_get_learned_report(const khip_ctx* ctx, khip_learned_report* out) {
    if (!ctx || !out) return KHIP_ERR_ARG;
    *out = ctx->lrep;
    return KHIP_OK;
}

int khip_net_train(khip_ctx* ctx, int64_t n, const double* orbits, const khip_train_config* cfg,
                   khip_train_report* report, float* curve) {
    if (!ctx || n <= 0 || n >= (1LL << 31) || !orbits || !cfg || cfg->samples_per_orbit <= 0 ||
        cfg->spacing_steps <= 0 || cfg->iterations <= 0 || cfg->batch <= 0 || cfg->batch % NET_KC != 0 ||
        !(cfg->tau > 0) || !(cfg->lr_max > 0) ||
        static_cast<double>(n) * cfg->samples_per_orbit >= 4.0e9)
        return KHIP_ERR_ARG;
    khip_train_report r{};
    const bool ok = net_train(ctx, orbits, n, *cfg, r, curve);
    if (report) *report = r;
    return status(ctx, ok);
}

int khip_net_set(khip_ctx* ctx, const float* params, const khip_net_norm* norm) {
    if (!ctx || !params || !norm || !(norm->tau > 0)) return KHIP_ERR_ARG;
    static_assert(sizeof(khip_net_norm) == sizeof(net::NetNorm), "norm layout");
    if (!net_upload(ctx, params)) return KHIP_ERR_HIP;
    std::memcpy(&ctx->net_norm, norm, sizeof(net::NetNorm));
    ctx->net_ready = true;
    ctx->gen += 1;
    return KHIP_OK;
}

int khip_net_get(khip_ctx* ctx, float* params, khip_net_norm* norm) {
    if (!ctx) return KHIP_ERR_ARG;
    if (!ctx->net_ready) {
        ctx->error = "no corrector loaded or trained";
        return KHIP_ERR_STATE;
    }
    if (params && hipMemcpy(params, ctx->net_p, sizeof(float) * net::NPARAM, hipMemcpyDeviceToHost) != hipSuccess) {
        ctx->error = "hipMemcpy of corrector parameters failed";
        return KHIP_ERR_HIP;
    }
    if (norm) std::memcpy(norm, &ctx->net_norm, sizeof(net::NetNorm));
    return KHIP_OK;
}

int khip_net_infer(khip_ctx* ctx, int64_t n, const float* x, float* out) {
    if (!ctx || n <= 0 || n >= (1LL << 31) || !x || !out) return KHIP_ERR_ARG;
    if (!ctx->net_ready) {
        ctx->error = "no corrector loaded or trained";
        return KHIP_ERR_STATE;
    }
    return status(ctx, net_infer(ctx, n, x, out));
}

int khip_net_gradient(khip_ctx* ctx, int64_t n, const float* x, const float* y, float* grad, float* loss) {
    if (!ctx || n <= 0 || n % NET_KC != 0 || n > (1 << 24) || !x || !y || !grad) return KHIP_ERR_ARG;
    if (!ctx->net_ready) {
        ctx->error = "no corrector loaded or trained";
        return KHIP_ERR_STATE;
    }
    return status(ctx, net_gradient(ctx, n, x, y, grad, loss));
}

int khip_set_chunk(khip_ctx* ctx, int steps) {
    if (!ctx || steps <= 0) return KHIP_ERR_ARG;
    ctx->chunk = steps;
    return KHIP_OK;
}

int khip_fp64_selfcheck(khip_ctx* ctx, int64_t n, const double* a, const double* b, double* out) {
    if (!ctx || n <= 0 || n > (1 << 20) || !a || !b || !out) return KHIP_ERR_ARG;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    bool ok = hipMalloc(reinterpret_cast<void**>(&da), sizeof(double) * n) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&db), sizeof(double) * n) == hipSuccess &&
              hipMalloc(reinterpret_cast<void**>(&dout), sizeof(double) * 3 * n) == hipSuccess &&
              hipMemcpy(da, a, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        k_fp64_check<<<grid_of(n), BLOCK, 0, ctx->stream>>>(static_cast<int>(n), da, db, dout);
        ok = hipStreamSynchronize(ctx->stream) == hipSuccess &&
             hipMemcpy(out, dout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    if (!ok) ctx->error = "fp64 self-check failed";
    return ok ? KHIP_OK : KHIP_ERR_HIP;
}

}  // extern "C"
