// Device-side building blocks of the gfx950 engine: constants, the FP64
// dynamics, the Hermite closest-approach solve and wave64 helpers.
//
// All FP64 arithmetic keeps the reference's operation order (orbital.cpp:
// 91-118, collision.cpp:77-132) and the file is compiled with
// -ffp-contract=off: with IEEE-rounded division and square root on the
// device, every result is bit-identical to the host oracle.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../host/collision_filter.hpp"

namespace kessler {
namespace dev {

constexpr double MU_EARTH = 398600.4418;
constexpr double MU_MOON = 4902.800066;
constexpr double R_EARTH = 6378.137;
constexpr double J2 = 1.08262668e-3;
// Folded at compile time left to right, as the reference's expression is.
constexpr double K_J2 = 1.5 * J2 * MU_EARTH * R_EARTH * R_EARTH;

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WAVES_PER_BLOCK = BLOCK / WAVE;

// One lunar sample: position and the indirect (frame) acceleration
// rm * mu_M / |rm|^3, both computed on the host.
struct MoonSample {
    double m[3];
    double ind[3];
};

struct D3 {
    double x, y, z;
};

__device__ __forceinline__ D3 accel(D3 r, const MoonSample& mo) {
    const double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const double rn = sqrt(r2);
    const double c = -MU_EARTH / (r2 * rn);
    D3 a{r.x * c, r.y * c, r.z * c};
    const double k = K_J2 / (r2 * r2 * rn);
    const double z2 = 5.0 * r.z * r.z / r2;
    a.x = a.x + k * r.x * (z2 - 1.0);
    a.y = a.y + k * r.y * (z2 - 1.0);
    a.z = a.z + k * r.z * (z2 - 3.0);
    const double dx = mo.m[0] - r.x, dy = mo.m[1] - r.y, dz = mo.m[2] - r.z;
    const double dn = sqrt(dx * dx + dy * dy + dz * dz);
    const double s = MU_MOON / (dn * dn * dn);
    a.x = a.x + dx * s - mo.ind[0];
    a.y = a.y + dy * s - mo.ind[1];
    a.z = a.z + dz * s - mo.ind[2];
    return a;
}

// Mixed-precision acceleration (precision form): the central term -mu r/|r|^3
// in FP64 (in FP32 it alone would cost km over a day), J2 and the Moon's
// direct minus indirect term in FP32 with hardware reciprocal / rsqrt. The
// perturbations are 1e-3 (J2) and 1e-7 (Moon) of the central term, so their
// FP32 rounding is ~1e-10 of the acceleration.
struct F3 {
    float x, y, z;
};

// J2 + Moon (direct minus indirect) in FP32, from the FP64 position and |r|^2.
__device__ __forceinline__ F3 perturbation_fp32(D3 r, double r2, const MoonSample& mo) {
#pragma clang fp contract(fast)
    const float xf = static_cast<float>(r.x), yf = static_cast<float>(r.y), zf = static_cast<float>(r.z);
    const float r2f = static_cast<float>(r2);
    const float inv_r2 = __builtin_amdgcn_rcpf(r2f);
    const float inv_rn = __builtin_amdgcn_rsqf(r2f);
    const float k = static_cast<float>(K_J2) * inv_r2 * inv_r2 * inv_rn;
    const float z2 = 5.0f * zf * zf * inv_r2;
    const float dx = static_cast<float>(mo.m[0]) - xf;
    const float dy = static_cast<float>(mo.m[1]) - yf;
    const float dz = static_cast<float>(mo.m[2]) - zf;
    const float inv_d = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
    const float sm = static_cast<float>(MU_MOON) * inv_d * inv_d * inv_d;
    return F3{k * xf * (z2 - 1.0f) + (dx * sm - static_cast<float>(mo.ind[0])),
              k * yf * (z2 - 1.0f) + (dy * sm - static_cast<float>(mo.ind[1])),
              k * zf * (z2 - 3.0f) + (dz * sm - static_cast<float>(mo.ind[2]))};
}

__device__ __forceinline__ D3 accel_mixed(D3 r, const MoonSample& mo) {
    const double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const double rn = sqrt(r2);
    const double c = -MU_EARTH / (r2 * rn);
    const F3 p = perturbation_fp32(r, r2, mo);
    return D3{r.x * c + static_cast<double>(p.x), r.y * c + static_cast<double>(p.y),
              r.z * c + static_cast<double>(p.z)};
}

struct AccelFp64 {
    __device__ __forceinline__ D3 operator()(D3 r, const MoonSample& m) const { return accel(r, m); }
};
struct AccelMixed {
    __device__ __forceinline__ D3 operator()(D3 r, const MoonSample& m) const { return accel_mixed(r, m); }
};

// Classical RK4 (orbital.cpp:107-118) on one object; r, v updated in place.
// Acc is the acceleration model (AccelFp64: the reference's, bit for bit).
template <class Acc = AccelFp64>
__device__ __forceinline__ void rk4(D3& r, D3& v, double dt, const MoonSample& m0,
                                    const MoonSample& m1, const MoonSample& m2, Acc accel = Acc()) {
    const double h = dt / 2;
    const D3 k1v = accel(r, m0);
    const D3 k2r{v.x + k1v.x * h, v.y + k1v.y * h, v.z + k1v.z * h};
    const D3 k2v = accel(D3{r.x + v.x * h, r.y + v.y * h, r.z + v.z * h}, m1);
    const D3 k3r{v.x + k2v.x * h, v.y + k2v.y * h, v.z + k2v.z * h};
    const D3 k3v = accel(D3{r.x + k2r.x * h, r.y + k2r.y * h, r.z + k2r.z * h}, m1);
    const D3 k4r{v.x + k3v.x * dt, v.y + k3v.y * dt, v.z + k3v.z * dt};
    const D3 k4v = accel(D3{r.x + k3r.x * dt, r.y + k3r.y * dt, r.z + k3r.z * dt}, m2);
    const double w = dt / 6;
    const D3 nr{r.x + (v.x + k2r.x * 2 + k3r.x * 2 + k4r.x) * w,
                r.y + (v.y + k2r.y * 2 + k3r.y * 2 + k4r.y) * w,
                r.z + (v.z + k2r.z * 2 + k3r.z * 2 + k4r.z) * w};
    const D3 nv{v.x + (k1v.x + k2v.x * 2 + k3v.x * 2 + k4v.x) * w,
                v.y + (k1v.y + k2v.y * 2 + k3v.y * 2 + k4v.y) * w,
                v.z + (k1v.z + k2v.z * 2 + k3v.z * 2 + k4v.z) * w};
    r = nr;
    v = nv;
}

// Cubic Hermite of the relative state over a step (collision.cpp:77-90).
struct Hermite {
    D3 p0, m0, p1, m1;
    __device__ __forceinline__ D3 pos(double s) const {
        const double s2 = s * s, s3 = s2 * s;
        const double c0 = 2 * s3 - 3 * s2 + 1, c1 = s3 - 2 * s2 + s, c2 = -2 * s3 + 3 * s2,
                     c3 = s3 - s2;
        return D3{p0.x * c0 + m0.x * c1 + p1.x * c2 + m1.x * c3,
                  p0.y * c0 + m0.y * c1 + p1.y * c2 + m1.y * c3,
                  p0.z * c0 + m0.z * c1 + p1.z * c2 + m1.z * c3};
    }
    __device__ __forceinline__ D3 vel(double s) const {
        const double s2 = s * s;
        const double c0 = 6 * s2 - 6 * s, c1 = 3 * s2 - 4 * s + 1, c2 = -6 * s2 + 6 * s,
                     c3 = 3 * s2 - 2 * s;
        return D3{p0.x * c0 + m0.x * c1 + p1.x * c2 + m1.x * c3,
                  p0.y * c0 + m0.y * c1 + p1.y * c2 + m1.y * c3,
                  p0.z * c0 + m0.z * c1 + p1.z * c2 + m1.z * c3};
    }
};

// Minimum of |p(s)|^2 on [0, 1] by 60 golden-section iterations
// (collision.cpp:94-105).
__device__ __forceinline__ double golden_min(const Hermite& h) {
    const double g = 0.6180339887498949;
    double a = 0.0, b = 1.0;
    double c = b - g * (b - a), d = a + g * (b - a);
    D3 p = h.pos(c);
    double fc = p.x * p.x + p.y * p.y + p.z * p.z;
    p = h.pos(d);
    double fd = p.x * p.x + p.y * p.y + p.z * p.z;
    // Branch-free: the bracket and the new point are selected, then |p|^2 is
    // evaluated once (lanes of a wave that keep different sides do not run
    // both evaluations); the same operations on the same values as the
    // two-branch form, so the result is bit for bit the same.
    for (int it = 0; it < 60; ++it) {
        const bool left = fc < fd;
        if (left) b = d;
        else a = c;
        const double x = left ? b - g * (b - a) : a + g * (b - a);
        p = h.pos(x);
        const double fx = p.x * p.x + p.y * p.y + p.z * p.z;
        if (left) {
            d = c;
            fd = fc;
            c = x;
            fc = fx;
        } else {
            c = d;
            fc = fd;
            d = x;
            fd = fx;
        }
    }
    return 0.5 * (a + b);
}

// prefilter_keep (collision_filter.hpp) rearranged for throughput, still
// conservative with respect to the exact FP64 filter_pair:
//  * ts uses the hardware reciprocal (v_rcp_f32, about 1 ulp): an error e in
//    ts moves the computed distance by at most |v| |e| ~ 2^-22 |v| dt,
//    against the 1e-5 (|r| + |v| dt) relative margin;
//  * |r| and |v| enter only the gate and the margins, where the L1 norms
//    (>= the Euclidean ones) make the gate and every margin larger and the
//    receding tolerance larger: fewer rejections, never more;
//  * the distance test compares squares (monotone on non-negative values).
// Returns false only where prefilter_keep's own bound says filter_pair
// rejects. gdt3 = GRAVITY_GRADIENT * dt^3, base = threshold + 1e-3 +
// DR_ERR_KM + DV_ERR_KM_S * dt, folded by the caller.
__device__ __forceinline__ bool prefilter_keep_fast(float rx, float ry, float rz, float vx, float vy,
                                                    float vz, float dt, float gdt3, float base,
                                                    float vmin2, float lim_max2) {
    // FP32 only: fused multiply-adds round once instead of twice, which can
    // only shrink the error the margins cover.
#pragma clang fp contract(fast)
    const float dv2 = vx * vx + vy * vy + vz * vz;
    const float rv = rx * vx + ry * vy + rz * vz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    // Reciprocal-free first rejection. r2 dv2 - rv^2 = dv2 * (squared
    // distance of the line r + v t from the origin, over all t), a lower
    // bound of the distance over [0, dt]. The test
    //   fma(r2, K, -L) * dv2 > fma(rv, rv, 1e-30),  L = 1.00001 lim_max2,
    // K = 1 - 2e-6, holds only if that exact distance exceeds L: the three
    // dot products (relative error <= 3 ulp each, Cauchy-Schwarz for rv),
    // the fma and the two products err by less than 16 ulp of r2 dv2, which
    // K's 2e-6 covers, and the 1e-30 keeps every product above the
    // subnormal range. The d2 computed below (fma-rounded, >= (1 - 6 ulp)
    // of the exact distance at any t) then exceeds lim_max2, so this test
    // rejects only candidates that the d2 > lim_max2 test rejects too.
    if (fmaf(r2, 0.999998f, -1.00001f * lim_max2) * dv2 > fmaf(rv, rv, 1e-30f)) return false;
    // ts = -rv / dv2 clamped to [0, dt] in one v_max_f32 and one v_med3_f32:
    // dv2 is floored at 1e-30 (km/s)^2 so that rcp stays finite; below that
    // floor the relative motion over a step (< 1e-14 km) is far inside the
    // margins and any ts in [0, dt] gives the same distance.
    const float ts = __builtin_amdgcn_fmed3f(-rv * __builtin_amdgcn_rcpf(fmaxf(dv2, 1e-30f)), 0.0f, dt);
    const float cx = rx + vx * ts, cy = ry + vy * ts, cz = rz + vz * ts;
    const float d2 = cx * cx + cy * cy + cz * cz;
    // lim_max2 bounds lim^2 below from above for every candidate of the
    // screening (the caller's pair_lim_max2): most candidates leave here,
    // before the margins are computed.
    if (d2 > lim_max2) return false;
    const float dvn = fabsf(vx) + fabsf(vy) + fabsf(vz);
    const float rn = fabsf(rx) + fabsf(ry) + fabsf(rz);
    const float lim = base + gdt3 * dvn + 1e-5f * (rn + dvn * dt);
    if (d2 > lim * lim) return false;
    const float rv_tol = 1e-5f * rn * dvn + DR_ERR_KM * dvn + DV_ERR_KM_S * rn;
    return !(dv2 > 1.1f * vmin2 && rv > rv_tol &&
             dv2 > 1.01f * static_cast<float>(GRAVITY_GRADIENT) * r2);
}

// Monotone map of doubles onto uint64 (for atomic min / max).
__device__ __forceinline__ unsigned long long ordered_bits(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double from_ordered_bits(unsigned long long u) {
    const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffULL) : ~u;
    double x;
    __builtin_memcpy(&x, &b, sizeof(x));
    return x;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanes_below() {
    const unsigned l = __lane_id();
    return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, WAVE);
        v = o < v ? o : v;
    }
    return v;
}

}  // namespace dev
}  // namespace kessler
