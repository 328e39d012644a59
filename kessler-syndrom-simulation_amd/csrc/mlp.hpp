// Learned step corrector: a 10-128-128-128-6 perceptron evaluated and
// trained on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: BF16
// operands, FP32 accumulation), plus the features it reads and the
// on-device generation of its training set.
//
// What it predicts. One RK4 step of tau = 60 s misses the result of six
// RK4 steps of 10 s by a correction delta(r, v) that is, to leading order,
// the RK4 local truncation error. For the central force that error is
// rotation-equivariant: in the object's local orbital frame (R radial,
// S along-track, W cross-track) it depends only on |r|, the radial and the
// transverse speed, and it scales like r (omega tau)^5 in position and
// v_c (omega tau)^5 in velocity (omega = sqrt(mu / r^3), v_c = omega r).
// The network therefore reads ten frame-invariant inputs (the content of
// r, v, |r| and the lunar phase, expressed in that frame):
//   x0 = 10 omega tau          x1 = v.R / v_c          x2 = v.S / v_c - 1
//   x3, x4, x5 = z-axis in the (R, S, W) frame  (J2 orientation)
//   x6, x7, x8 = Moon direction in the (R, S, W) frame
//   x9 = (R_earth / |r|)^2     (J2 strength)
// each standardised by the training set's mean and spread, and returns
// the six frame components of delta divided by those scales, again
// standardised (mean removed, one spread for the position block and one
// for the velocity block).
//
// Matrix-core layout. The perceptron runs transposed, H^T = W X^T, so the
// objects are the MFMA N dimension and features the M / K dimensions.
// For a 16x16 output tile, lane (j + 16 g) holds rows 4g..4g+3 of column
// j; for a K = 32 operand it holds 8 values of row / column j. Reading the
// K index of the 8 values of lane group g in step ks as
//     kmap(ks, g, e) = 32 ks + 4 g + e          (e < 4)
//                    = 32 ks + 12 + 4 g + e     (e >= 4)
// makes the two output tiles 2ks and 2ks+1 of one layer, as they sit in
// the registers of a lane, exactly the B operand of K step ks of the next
// layer: activations never move between lanes. The weights are stored in
// that order ("fragments": 64 lanes x 8 BF16 per 16x32 tile), so every
// operand load is one 16-byte read per lane.
//
// Precision. Hidden layers take BF16 weights and activations. The input
// layer and the output layer split operands into BF16 high and low parts
// (a = a_hi + a_lo) and add hi*hi + hi*lo + lo*hi, which carries about 16
// mantissa bits: the inputs and the final correction keep FP32-like
// accuracy at the cost of one extra K step (input) and 3x the small output
// layer.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"

namespace kessler {
namespace net {

using dev::D3;
using dev::MoonSample;

constexpr int NIN = 10;
constexpr int NH = 128;
constexpr int NOUT = 6;

// FP32 master parameters, [out][in] row-major per layer.
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH * NIN;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NOUT * NH;
constexpr int NPARAM = OFF_B4 + NOUT;  // 35,206

// Fragment sets (each fragment: 64 lanes x 8 BF16 = 1 KiB).
//   forward:  F1[mt] (input layer, hi/lo extended), F2[mt][ks], F3[mt][ks],
//             F4hi[ks], F4lo[ks]
//   backward: T4[mt] (W4^T), T3[mt][ks] (W3^T), T2[mt][ks] (W2^T)
constexpr int FR_L1 = 0, FR_L2 = 8, FR_L3 = 40, FR_L4H = 72, FR_L4L = 76, NFRAG_FWD = 80;
constexpr int FR_T4 = 0, FR_T3 = 8, FR_T2 = 40, NFRAG_BWD = 72;
constexpr int FRAG_HALVES = 64 * 8;  // BF16 values per fragment

// Standardisation of inputs and outputs, and the step the net is trained for.
struct NetNorm {
    float fmu[NIN], fsig[NIN];
    float omu[NOUT], osig[NOUT];
    double tau;
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint32_t f32_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}
__host__ __device__ inline float bits_f32(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
// FP32 -> BF16 bits, round to nearest even (inputs are finite).
__host__ __device__ inline uint32_t bf16_rne(float f) {
    const uint32_t u = f32_bits(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__host__ __device__ inline float bf16_val(uint32_t h) { return bits_f32(h << 16); }
__host__ __device__ inline uint32_t bf16_lo(float f, uint32_t hi) { return bf16_rne(f - bf16_val(hi)); }

__host__ __device__ inline int kmap(int ks, int g, int e) { return 32 * ks + (e < 4 ? 4 * g + e : 12 + 4 * g + e); }

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ u32x4 pack8(uint32_t h0, uint32_t h1, uint32_t h2, uint32_t h3, uint32_t h4,
                                       uint32_t h5, uint32_t h6, uint32_t h7) {
    u32x4 r;
    r.x = h0 | (h1 << 16);
    r.y = h2 | (h3 << 16);
    r.z = h4 | (h5 << 16);
    r.w = h6 | (h7 << 16);
    return r;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two FP32 values rounded to BF16 (nearest even, as bf16_rne for finite
// values, FP32 subnormals kept) in one v_cvt_pk_bf16_f32: a in the low half.
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}
// Low BF16 parts of a and b, given w = cvt2(a, b): the rounding residues
// a - hi(a), b - hi(b) are exact in FP32.
__device__ __forceinline__ uint32_t cvt2_lo(float a, float b, uint32_t w) {
    return cvt2(a - bits_f32(w << 16), b - bits_f32(w & 0xffff0000u));
}

// B operand of K step ks from output tiles 2ks (a) and 2ks+1 (b) of the
// previous layer, rounded to BF16.
__device__ __forceinline__ u32x4 pack_tiles(f32x4 a, f32x4 b) {
    u32x4 r;
    r.x = cvt2(a.x, a.y);
    r.y = cvt2(a.z, a.w);
    r.z = cvt2(b.x, b.y);
    r.w = cvt2(b.z, b.w);
    return r;
}
// Low BF16 parts of the same values, given hi = pack_tiles(a, b).
__device__ __forceinline__ u32x4 pack_tiles_lo(f32x4 a, f32x4 b, u32x4 hi) {
    u32x4 r;
    r.x = cvt2_lo(a.x, a.y, hi.x);
    r.y = cvt2_lo(a.z, a.w, hi.y);
    r.z = cvt2_lo(b.x, b.y, hi.z);
    r.w = cvt2_lo(b.z, b.w, hi.w);
    return r;
}

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

__device__ __forceinline__ u32x4 frag_load(const uint16_t* frags, int f, int lane) {
    return *reinterpret_cast<const u32x4*>(frags + f * FRAG_HALVES + lane * 8);
}

// Hard swish, z clamp(z / 6 + 1 / 2, 0, 1), and its derivative: an FMA, a
// three-way median and a product per activation (SiLU needs an exp2 and a
// reciprocal, two quarter-rate instructions).
__device__ __forceinline__ float hswish(float z) {
    return z * __builtin_amdgcn_fmed3f(fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ f32x4 act4(f32x4 z) {
    f32x4 h;
    h.x = hswish(z.x);
    h.y = hswish(z.y);
    h.z = hswish(z.z);
    h.w = hswish(z.w);
    return h;
}
__device__ __forceinline__ float dact(float z) {
    return z < -3.0f ? 0.0f : (z > 3.0f ? 1.0f : fmaf(z, 1.0f / 3.0f, 0.5f));
}

__device__ __forceinline__ f32x4 bias4(const float* b, int mt, int g) {
    const float* p = b + 16 * mt + 4 * g;
    f32x4 r;
    r.x = p[0];
    r.y = p[1];
    r.z = p[2];
    r.w = p[3];
    return r;
}

// ---------------------------------------------------------------------------
// Features and the correction frame (FP64 geometry, FP32 network inputs)
// ---------------------------------------------------------------------------

struct Frame {
    D3 R, S, W;
    double scale_r, scale_v;  // r (omega tau)^5, v_c (omega tau)^5
};

__device__ __forceinline__ void features(D3 r, D3 v, const MoonSample& m, const NetNorm& nn, float x[NIN],
                                         Frame& fr) {
    const double rn = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    const D3 R{r.x / rn, r.y / rn, r.z / rn};
    D3 h{r.y * v.z - r.z * v.y, r.z * v.x - r.x * v.z, r.x * v.y - r.y * v.x};
    const double hn = sqrt(h.x * h.x + h.y * h.y + h.z * h.z);
    const D3 W{h.x / hn, h.y / hn, h.z / hn};
    const D3 S{W.y * R.z - W.z * R.y, W.z * R.x - W.x * R.z, W.x * R.y - W.y * R.x};
    const double vc = sqrt(dev::MU_EARTH / rn);
    const double w = vc / rn * nn.tau;
    const double w2 = w * w;
    const double w5 = w2 * w2 * w;
    const double mn = sqrt(m.m[0] * m.m[0] + m.m[1] * m.m[1] + m.m[2] * m.m[2]);
    const D3 M{m.m[0] / mn, m.m[1] / mn, m.m[2] / mn};
    const double re = dev::R_EARTH / rn;
    double f[NIN];
    f[0] = 10.0 * w;
    f[1] = (v.x * R.x + v.y * R.y + v.z * R.z) / vc;
    f[2] = (v.x * S.x + v.y * S.y + v.z * S.z) / vc - 1.0;
    f[3] = R.z;
    f[4] = S.z;
    f[5] = W.z;
    f[6] = M.x * R.x + M.y * R.y + M.z * R.z;
    f[7] = M.x * S.x + M.y * S.y + M.z * S.z;
    f[8] = M.x * W.x + M.y * W.y + M.z * W.z;
    f[9] = re * re;
    for (int k = 0; k < NIN; ++k)
        x[k] = static_cast<float>((f[k] - static_cast<double>(nn.fmu[k])) / static_cast<double>(nn.fsig[k]));
    fr.R = R;
    fr.S = S;
    fr.W = W;
    fr.scale_r = rn * w5;
    fr.scale_v = vc * w5;
}

// Raw (unstandardised) features and scaled targets of one training sample.
__device__ __forceinline__ void raw_features(D3 r, D3 v, const MoonSample& m, double tau, float f_out[NIN],
                                             Frame& fr) {
    NetNorm unit;
    for (int k = 0; k < NIN; ++k) {
        unit.fmu[k] = 0.0f;
        unit.fsig[k] = 1.0f;
    }
    unit.tau = tau;
    features(r, v, m, unit, f_out, fr);
}

// Applies the network output o (standardised) to the state.
__device__ __forceinline__ void apply_correction(D3& r, D3& v, const float o[NOUT], const NetNorm& nn,
                                                 const Frame& fr) {
    double c[NOUT];
    for (int k = 0; k < NOUT; ++k)
        c[k] = static_cast<double>(o[k]) * static_cast<double>(nn.osig[k]) + static_cast<double>(nn.omu[k]);
    const double pr = fr.scale_r, pv = fr.scale_v;
    r.x += (c[0] * fr.R.x + c[1] * fr.S.x + c[2] * fr.W.x) * pr;
    r.y += (c[0] * fr.R.y + c[1] * fr.S.y + c[2] * fr.W.y) * pr;
    r.z += (c[0] * fr.R.z + c[1] * fr.S.z + c[2] * fr.W.z) * pr;
    v.x += (c[3] * fr.R.x + c[4] * fr.S.x + c[5] * fr.W.x) * pv;
    v.y += (c[3] * fr.R.y + c[4] * fr.S.y + c[5] * fr.W.y) * pv;
    v.z += (c[3] * fr.R.z + c[4] * fr.S.z + c[5] * fr.W.z) * pv;
}

// Frame components of a difference (target construction).
__device__ __forceinline__ void to_frame(const Frame& fr, D3 dr, D3 dv, double out[NOUT]) {
    out[0] = (dr.x * fr.R.x + dr.y * fr.R.y + dr.z * fr.R.z) / fr.scale_r;
    out[1] = (dr.x * fr.S.x + dr.y * fr.S.y + dr.z * fr.S.z) / fr.scale_r;
    out[2] = (dr.x * fr.W.x + dr.y * fr.W.y + dr.z * fr.W.z) / fr.scale_r;
    out[3] = (dv.x * fr.R.x + dv.y * fr.R.y + dv.z * fr.R.z) / fr.scale_v;
    out[4] = (dv.x * fr.S.x + dv.y * fr.S.y + dv.z * fr.S.z) / fr.scale_v;
    out[5] = (dv.x * fr.W.x + dv.y * fr.W.y + dv.z * fr.W.z) / fr.scale_v;
}

// Input-layer B operand for lane group g from the standardised inputs x of
// the lane's column: the extended vector u = (x_hi, x_lo, x_hi, 0, 0)
// pairs with the weight row (W_hi, W_hi, W_lo, 0, 0): K rows 0-9 x_hi,
// 10-19 x_lo, 20-29 x_hi, 30-31 zero. It is built from packed words of the
// column's inputs (input_words): w[i] = the BF16 pair (x_2i, x_2i+1),
// w[5 + i] their low parts. With kmap(0, g, e) the lane group g holds K rows
// 4g..4g+3 and 16+4g..16+4g+3 of u, i.e. the pairs
//   g = 0: hi(0,1) hi(2,3) lo(6,7) lo(8,9)     g = 1: hi(4,5) hi(6,7) hi(0,1) hi(2,3)
//   g = 2: hi(8,9) lo(0,1) hi(4,5) hi(6,7)     g = 3: lo(2,3) lo(4,5) hi(8,9) 0
// selected by g (uniform per 16 lanes) instead of indexing x by a lane-
// dependent K row (compare-and-select chains over all ten inputs).
__device__ __forceinline__ void input_words(const float x[NIN], uint32_t w[10]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        w[i] = cvt2(x[2 * i], x[2 * i + 1]);
        w[5 + i] = cvt2_lo(x[2 * i], x[2 * i + 1], w[i]);
    }
}
__device__ __forceinline__ uint32_t sel4(int g, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return g == 0 ? a : (g == 1 ? b : (g == 2 ? c : d));
}
__device__ __forceinline__ u32x4 input_operand_words(const uint32_t w[10], int g) {
    u32x4 r;
    r.x = sel4(g, w[0], w[2], w[4], w[6]);
    r.y = sel4(g, w[1], w[3], w[5], w[7]);
    r.z = sel4(g, w[8], w[0], w[2], w[4]);
    r.w = sel4(g, w[9], w[1], w[3], 0u);
    return r;
}
__device__ __forceinline__ u32x4 input_operand(const float x[NIN], int g) {
    uint32_t w[10];
    input_words(x, w);
    return input_operand_words(w, g);
}

// ---------------------------------------------------------------------------
// Forward pass of NT tiles of 16 columns for one wave
// ---------------------------------------------------------------------------
//
// xin[t]: input operand of tile t (input_operand of column j of tile t).
// frags: forward fragments (LDS or global). bias: FP32 parameter vector.
// Returns out[t]: output tile (lane (j + 16 g) element e = output 4g + e).
template <int NT>
__device__ __forceinline__ void forward_tiles(const u32x4 xin[NT], const uint16_t* frags, const float* params,
                                              int lane, f32x4 out[NT]) {
    const int g = lane >> 4;
    u32x4 hb[4][NT];
    // Input layer.
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const u32x4 a0 = frag_load(frags, FR_L1 + 2 * p, lane);
        const u32x4 a1 = frag_load(frags, FR_L1 + 2 * p + 1, lane);
        const f32x4 b0 = bias4(params + OFF_B1, 2 * p, g), b1 = bias4(params + OFF_B1, 2 * p + 1, g);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const f32x4 z0 = mfma(a0, xin[t], b0);
            const f32x4 z1 = mfma(a1, xin[t], b1);
            hb[p][t] = pack_tiles(act4(z0), act4(z1));
        }
    }
    // Two hidden layers; the output layer (hi/lo split of weights and
    // activations) consumes the last one K step at a time.
    f32x4 bo;
    {
        float b[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) b[e] = 4 * g + e < NOUT ? params[OFF_B4 + 4 * g + e] : 0.0f;
        bo.x = b[0];
        bo.y = b[1];
        bo.z = b[2];
        bo.w = b[3];
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) out[t] = bo;
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int fr = layer == 0 ? FR_L2 : FR_L3;
        const float* bias = params + (layer == 0 ? OFF_B2 : OFF_B3);
        u32x4 nb[4][NT];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            f32x4 z0[NT], z1[NT];
            const f32x4 b0 = bias4(bias, 2 * p, g), b1 = bias4(bias, 2 * p + 1, g);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                z0[t] = b0;
                z1[t] = b1;
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const u32x4 a0 = frag_load(frags, fr + (2 * p) * 4 + ks, lane);
                const u32x4 a1 = frag_load(frags, fr + (2 * p + 1) * 4 + ks, lane);
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    z0[t] = mfma(a0, hb[ks][t], z0[t]);
                    z1[t] = mfma(a1, hb[ks][t], z1[t]);
                }
            }
            if (layer == 0) {
#pragma unroll
                for (int t = 0; t < NT; ++t) nb[p][t] = pack_tiles(act4(z0[t]), act4(z1[t]));
            } else {
                const u32x4 ah = frag_load(frags, FR_L4H + p, lane);
                const u32x4 al = frag_load(frags, FR_L4L + p, lane);
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const f32x4 h0 = act4(z0[t]), h1 = act4(z1[t]);
                    const u32x4 hi = pack_tiles(h0, h1), lo = pack_tiles_lo(h0, h1, hi);
                    out[t] = mfma(ah, hi, out[t]);
                    out[t] = mfma(ah, lo, out[t]);
                    out[t] = mfma(al, hi, out[t]);
                }
            }
        }
        if (layer == 0) {
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int t = 0; t < NT; ++t) hb[p][t] = nb[p][t];
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels: fragments, training step, weight gradients, Adam, samples
// ---------------------------------------------------------------------------

// Fragments from the FP32 master parameters. One thread per BF16 value.
__global__ void __launch_bounds__(256) k_net_pack(const float* __restrict__ P, uint16_t* __restrict__ fwd,
                                                  uint16_t* __restrict__ bwd) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    const int total = (NFRAG_FWD + NFRAG_BWD) * FRAG_HALVES;
    if (id >= total) return;
    const bool is_fwd = id < NFRAG_FWD * FRAG_HALVES;
    const int local = is_fwd ? id : id - NFRAG_FWD * FRAG_HALVES;
    const int f = local / FRAG_HALVES;
    const int lane = (local % FRAG_HALVES) / 8;
    const int e = local % 8;
    const int i = lane & 15, g = lane >> 4;
    uint32_t h = 0;
    if (is_fwd) {
        if (f < FR_L2) {  // input layer, extended K
            const int m = 16 * f + i;
            const int k = kmap(0, g, e);
            if (k < 30) {
                const float w = P[OFF_W1 + m * NIN + (k % 10)];
                const uint32_t hi = bf16_rne(w);
                h = k < 20 ? hi : bf16_lo(w, hi);
            }
        } else if (f < FR_L4H) {  // hidden layers
            const int base = f < FR_L3 ? OFF_W2 : OFF_W3;
            const int q = f - (f < FR_L3 ? FR_L2 : FR_L3);
            const int mt = q / 4, ks = q % 4;
            h = bf16_rne(P[base + (16 * mt + i) * NH + kmap(ks, g, e)]);
        } else {  // output layer, hi or lo
            const bool lo = f >= FR_L4L;
            const int ks = f - (lo ? FR_L4L : FR_L4H);
            if (i < NOUT) {
                const float w = P[OFF_W4 + i * NH + kmap(ks, g, e)];
                const uint32_t hi = bf16_rne(w);
                h = lo ? bf16_lo(w, hi) : hi;
            }
        }
        fwd[id] = static_cast<uint16_t>(h);
    } else {
        if (f < FR_T3) {  // W4^T: rows = hidden, K = outputs
            const int k = kmap(0, g, e);
            if (k < NOUT) h = bf16_rne(P[OFF_W4 + k * NH + 16 * f + i]);
        } else {
            const int base = f < FR_T2 ? OFF_W3 : OFF_W2;
            const int q = f - (f < FR_T2 ? FR_T3 : FR_T2);
            const int mt = q / 4, ks = q % 4;
            h = bf16_rne(P[base + kmap(ks, g, e) * NH + 16 * mt + i]);
        }
        bwd[local] = static_cast<uint16_t>(h);
    }
}

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Activation and gradient tensors of one training batch (BF16 bits,
// feature-major [rows][B]) for the weight-gradient products.
struct TrainTensors {
    uint16_t *hx, *h1, *h2, *h3;  // inputs of layers 1-4 (hx: 16 rows, x_hi)
    uint16_t *g1, *g2, *g3, *g4;  // dL/dz of layers 1-4 (g4: 16 rows)
    float* loss;                  // per-wave partial sums of squared error
};

__device__ __forceinline__ void store_tile(uint16_t* dst, int B, int mt, int g, int col, f32x4 v) {
    const int row = 16 * mt + 4 * g;
    dst[(row + 0) * B + col] = static_cast<uint16_t>(bf16_rne(v.x));
    dst[(row + 1) * B + col] = static_cast<uint16_t>(bf16_rne(v.y));
    dst[(row + 2) * B + col] = static_cast<uint16_t>(bf16_rne(v.z));
    dst[(row + 3) * B + col] = static_cast<uint16_t>(bf16_rne(v.w));
}

// Loss: mean over the batch of (1/6) sum_k (o_k - y_k)^2.
// Sample of column b: pool[mix64(seed + b) % npool] when pool != nullptr,
// else b itself (a fixed batch, for the gradient check).
__global__ void __launch_bounds__(256) k_net_fwdbwd(const float* __restrict__ X, const float* __restrict__ Y,
                                                    const uint32_t* __restrict__ pool, uint32_t npool,
                                                    uint64_t seed, int B, const uint16_t* __restrict__ ffwd,
                                                    const uint16_t* __restrict__ fbwd,
                                                    const float* __restrict__ P, TrainTensors T) {
    const int lane = static_cast<int>(__lane_id());
    const int wave = (blockIdx.x * 256 + threadIdx.x) / 64;
    const int b0 = wave * 16;
    if (b0 >= B) return;
    const int j = lane & 15, g = lane >> 4;
    const int col = b0 + j;
    uint32_t s = static_cast<uint32_t>(col);
    if (pool) s = pool[mix64(seed + static_cast<uint64_t>(col)) % npool];
    float x[NIN];
#pragma unroll
    for (int k = 0; k < NIN; ++k) x[k] = X[static_cast<size_t>(s) * NIN + k];
    const u32x4 xin = input_operand(x, g);
    // hx rows k = g, g + 4, g + 8 (x_hi), for the input layer's gradient.
    for (int k = g; k < NIN; k += 4) T.hx[k * B + col] = static_cast<uint16_t>(bf16_rne(x[k]));

    f32x4 z1[8], z2[8], z3[8];
    u32x4 hb1[4], hb2[4], hb3[4], hb3lo[4];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) z1[mt] = mfma(frag_load(ffwd, FR_L1 + mt, lane), xin, bias4(P + OFF_B1, mt, g));
#pragma unroll
    for (int p = 0; p < 4; ++p) hb1[p] = pack_tiles(act4(z1[2 * p]), act4(z1[2 * p + 1]));
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 z = bias4(P + OFF_B2, mt, g);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) z = mfma(frag_load(ffwd, FR_L2 + mt * 4 + ks, lane), hb1[ks], z);
        z2[mt] = z;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) hb2[p] = pack_tiles(act4(z2[2 * p]), act4(z2[2 * p + 1]));
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 z = bias4(P + OFF_B3, mt, g);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) z = mfma(frag_load(ffwd, FR_L3 + mt * 4 + ks, lane), hb2[ks], z);
        z3[mt] = z;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f32x4 h0 = act4(z3[2 * p]), h1 = act4(z3[2 * p + 1]);
        hb3[p] = pack_tiles(h0, h1);
        hb3lo[p] = pack_tiles_lo(h0, h1, hb3[p]);
    }
    f32x4 o;
    o.x = o.y = o.z = o.w = 0.0f;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const u32x4 ah = frag_load(ffwd, FR_L4H + ks, lane), al = frag_load(ffwd, FR_L4L + ks, lane);
        o = mfma(ah, hb3[ks], o);
        o = mfma(ah, hb3lo[ks], o);
        o = mfma(al, hb3[ks], o);
    }
    // Output gradient (rows 4g + e < 6).
    const float inv = 2.0f / (6.0f * static_cast<float>(B));
    float d[4] = {o.x, o.y, o.z, o.w};
    float sq = 0.0f;
    f32x4 g4;
    float gv[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int row = 4 * g + e;
        if (row < NOUT) {
            const float diff = d[e] + P[OFF_B4 + row] - Y[static_cast<size_t>(s) * NOUT + row];
            sq += diff * diff;
            gv[e] = diff * inv;
        } else {
            gv[e] = 0.0f;
        }
    }
    g4.x = gv[0];
    g4.y = gv[1];
    g4.z = gv[2];
    g4.w = gv[3];
    store_tile(T.g4, B, 0, g, col, g4);  // rows 0..15 (only < 6 nonzero)
    const float wsum = dev::wave_sum(sq);
    if (lane == 0) T.loss[wave] = wsum;
    // Backward through the output layer: dh3 = W4^T g4.
    const u32x4 gb4 = pack8(bf16_rne(g4.x), bf16_rne(g4.y), bf16_rne(g4.z), bf16_rne(g4.w), 0u, 0u, 0u, 0u);
    f32x4 gz[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 zero;
        zero.x = zero.y = zero.z = zero.w = 0.0f;
        const f32x4 dh = mfma(frag_load(fbwd, FR_T4 + mt, lane), gb4, zero);
        f32x4 r;
        r.x = dh.x * dact(z3[mt].x);
        r.y = dh.y * dact(z3[mt].y);
        r.z = dh.z * dact(z3[mt].z);
        r.w = dh.w * dact(z3[mt].w);
        gz[mt] = r;
    }
    // Layer-4 input activations and layer-3 gradients.
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        store_tile(T.h3, B, mt, g, col, act4(z3[mt]));
        store_tile(T.g3, B, mt, g, col, gz[mt]);
    }
    // dh2 = W3^T g3.
    u32x4 gb[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) gb[p] = pack_tiles(gz[2 * p], gz[2 * p + 1]);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 dh;
        dh.x = dh.y = dh.z = dh.w = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) dh = mfma(frag_load(fbwd, FR_T3 + mt * 4 + ks, lane), gb[ks], dh);
        gz[mt].x = dh.x * dact(z2[mt].x);
        gz[mt].y = dh.y * dact(z2[mt].y);
        gz[mt].z = dh.z * dact(z2[mt].z);
        gz[mt].w = dh.w * dact(z2[mt].w);
        store_tile(T.h2, B, mt, g, col, act4(z2[mt]));
        store_tile(T.g2, B, mt, g, col, gz[mt]);
    }
    // dh1 = W2^T g2.
#pragma unroll
    for (int p = 0; p < 4; ++p) gb[p] = pack_tiles(gz[2 * p], gz[2 * p + 1]);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        f32x4 dh;
        dh.x = dh.y = dh.z = dh.w = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) dh = mfma(frag_load(fbwd, FR_T2 + mt * 4 + ks, lane), gb[ks], dh);
        f32x4 r;
        r.x = dh.x * dact(z1[mt].x);
        r.y = dh.y * dact(z1[mt].y);
        r.z = dh.z * dact(z1[mt].z);
        r.w = dh.w * dact(z1[mt].w);
        store_tile(T.h1, B, mt, g, col, act4(z1[mt]));
        store_tile(T.g1, B, mt, g, col, r);
    }
}

// Weight and bias gradients: dW_l = G_l H_{l-1}^T over the batch, split
// along the batch into chunks of KC columns; one wave per (layer, 16x16
// tile, chunk). Layer 1 uses the x_hi rows (16, of which 10 are real),
// layer 4 the 16 gradient rows (6 real). Waves of the first column tile
// also sum their gradient rows (bias gradients). Partials go to
// part[chunk * NPARAM + parameter].
constexpr int DW_TILES = 8 + 64 + 64 + 8;

__global__ void __launch_bounds__(256) k_net_dw(TrainTensors T, int B, int KC, float* __restrict__ part) {
    const int lane = static_cast<int>(__lane_id());
    const int task = (blockIdx.x * 256 + threadIdx.x) / 64;
    const int nchunks = B / KC;
    if (task >= DW_TILES * nchunks) return;
    const int chunk = task / DW_TILES;
    int tile = task % DW_TILES;
    const uint16_t *A, *Bm;
    int layer, mt, nt;
    if (tile < 8) {
        layer = 1; mt = tile; nt = 0; A = T.g1; Bm = T.hx;
    } else if (tile < 72) {
        tile -= 8; layer = 2; mt = tile / 8; nt = tile % 8; A = T.g2; Bm = T.h1;
    } else if (tile < 136) {
        tile -= 72; layer = 3; mt = tile / 8; nt = tile % 8; A = T.g3; Bm = T.h2;
    } else {
        tile -= 136; layer = 4; mt = 0; nt = tile; A = T.g4; Bm = T.h3;
    }
    const int i = lane & 15, g = lane >> 4;
    const uint16_t* arow = A + static_cast<size_t>(16 * mt + i) * B + chunk * KC + 8 * g;
    const uint16_t* brow = Bm + static_cast<size_t>(16 * nt + i) * B + chunk * KC + 8 * g;
    f32x4 acc;
    acc.x = acc.y = acc.z = acc.w = 0.0f;
    float rowsum = 0.0f;
    for (int k = 0; k < KC; k += 32) {
        const u32x4 a = *reinterpret_cast<const u32x4*>(arow + k);
        const u32x4 b = *reinterpret_cast<const u32x4*>(brow + k);
        acc = mfma(a, b, acc);
        if (nt == 0) {
            rowsum += bf16_val(a.x & 0xffffu) + bf16_val(a.x >> 16) + bf16_val(a.y & 0xffffu) +
                      bf16_val(a.y >> 16) + bf16_val(a.z & 0xffffu) + bf16_val(a.z >> 16) +
                      bf16_val(a.w & 0xffffu) + bf16_val(a.w >> 16);
        }
    }
    float* out = part + static_cast<size_t>(chunk) * NPARAM;
    // D: lane (j + 16g) element e = (row 16mt + 4g + e, column 16nt + j).
    const int jcol = 16 * nt + i;
    const float v[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int row = 16 * mt + 4 * g + e;
        if (layer == 1) {
            if (jcol < NIN) out[OFF_W1 + row * NIN + jcol] = v[e];
        } else if (layer == 2) {
            out[OFF_W2 + row * NH + jcol] = v[e];
        } else if (layer == 3) {
            out[OFF_W3 + row * NH + jcol] = v[e];
        } else if (row < NOUT) {
            out[OFF_W4 + row * NH + jcol] = v[e];
        }
    }
    if (nt == 0) {
        rowsum += __shfl_xor(rowsum, 16, 64);
        rowsum += __shfl_xor(rowsum, 32, 64);
        if (g == 0) {
            const int row = 16 * mt + i;
            if (layer == 1) out[OFF_B1 + row] = rowsum;
            else if (layer == 2) out[OFF_B2 + row] = rowsum;
            else if (layer == 3) out[OFF_B3 + row] = rowsum;
            else if (row < NOUT) out[OFF_B4 + row] = rowsum;
        }
    }
}

struct AdamArgs {
    float lr, beta1, beta2, eps, bc1, bc2;  // bc = 1 - beta^t
    int nchunks;
    int apply;  // 0: only sum the gradient into grad_out
};

// Sums the chunk partials into the gradient and applies one Adam step to
// the FP32 master parameters.
__global__ void __launch_bounds__(256) k_net_adam(float* __restrict__ P, float* __restrict__ M1,
                                                  float* __restrict__ M2, const float* __restrict__ part,
                                                  float* __restrict__ grad_out, AdamArgs a) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NPARAM) return;
    float gsum = 0.0f;
    for (int c = 0; c < a.nchunks; ++c) gsum += part[static_cast<size_t>(c) * NPARAM + p];
    if (grad_out) grad_out[p] = gsum;
    if (!a.apply) return;
    const float m = a.beta1 * M1[p] + (1.0f - a.beta1) * gsum;
    const float v = a.beta2 * M2[p] + (1.0f - a.beta2) * gsum * gsum;
    M1[p] = m;
    M2[p] = v;
    const float mh = m / a.bc1, vh = v / a.bc2;
    P[p] = P[p] - a.lr * mh / (sqrtf(vh) + a.eps);
}

// Sum of the per-wave losses of one batch (one block).
__global__ void __launch_bounds__(256) k_net_loss(const float* __restrict__ wl, int nw, float* __restrict__ out) {
    __shared__ float s[4];
    float acc = 0.0f;
    for (int k = threadIdx.x; k < nw; k += 256) acc += wl[k];
    acc = dev::wave_sum(acc);
    if (__lane_id() == 0) s[threadIdx.x / 64] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *out = s[0] + s[1] + s[2] + s[3];
}

// Method class of an object from its osculating orbit:
//   0  corrected step: bound, eccentricity < e_max, semi-major axis < a_max,
//      perigee above R_earth + 120 km (the corrector's training domain);
//   2  base step alone: bound, eccentricity < e_max, semi-major axis >= a_max
//      (at MEO and GEO one RK4 step of 60 s is already within centimetres
//      of six steps of 10 s over a day);
//   1  plain sub-steps: everything else (eccentric, unbound, grazing).
constexpr int CLASS_LEARNED = 0, CLASS_PLAIN = 1, CLASS_BASE = 2;

__device__ __forceinline__ int domain_class(D3 r, D3 v, double e_max, double a_max) {
    const double rn = sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    const double v2 = v.x * v.x + v.y * v.y + v.z * v.z;
    const double en = 0.5 * v2 - dev::MU_EARTH / rn;
    if (!(en < 0.0)) return CLASS_PLAIN;
    const double hx = r.y * v.z - r.z * v.y, hy = r.z * v.x - r.x * v.z, hz = r.x * v.y - r.y * v.x;
    const double h2 = hx * hx + hy * hy + hz * hz;
    const double e2 = 1.0 + 2.0 * en * h2 / (dev::MU_EARTH * dev::MU_EARTH);
    if (!(e2 < e_max * e_max)) return CLASS_PLAIN;
    const double a = -dev::MU_EARTH / (2.0 * en);
    const double e = sqrt(e2 > 0.0 ? e2 : 0.0);
    if (a * (1.0 - e) < dev::R_EARTH + 120.0) return CLASS_PLAIN;
    return a < a_max ? CLASS_LEARNED : CLASS_BASE;
}

// Training samples: orbit o (one lane) is stepped by the 60 s base step
// (mixed-precision RK4) and, every `every` steps from step (o % every),
// emits (raw features, target) for states in the corrector's domain
// (domain_class == CLASS_LEARNED), where the target is
//   (six FP64 RK4 steps of tau/6) - (one mixed RK4 step of tau)
// in the frame components divided by the frame scales. moon5 holds the
// Moon every tau/12 from the start: steps p uses entries 12p .. 12p + 12.
__global__ void __launch_bounds__(256) k_net_samples(const double* __restrict__ orbits, int norbits, int steps,
                                                     int every, const MoonSample* __restrict__ moon5, double tau,
                                                     float* __restrict__ fx, float* __restrict__ fy,
                                                     uint8_t* __restrict__ valid, int per_orbit,
                                                     double e_max, double a_max) {
    const int o = blockIdx.x * 256 + threadIdx.x;
    if (o >= norbits) return;
    D3 r{orbits[6 * o], orbits[6 * o + 1], orbits[6 * o + 2]};
    D3 v{orbits[6 * o + 3], orbits[6 * o + 4], orbits[6 * o + 5]};
    int emitted = 0;
    const double h = tau / 6.0;
    for (int p = 0; p < steps && emitted < per_orbit; ++p) {
        const MoonSample* m = moon5 + 12 * p;
        if (p % every == o % every) {
            const size_t slot = static_cast<size_t>(o) * per_orbit + emitted;
            ++emitted;
            const bool ok = domain_class(r, v, e_max, a_max) == CLASS_LEARNED;
            valid[slot] = ok ? 1 : 0;
            if (ok) {
                D3 r6 = r, v6 = v;
                for (int k = 0; k < 6; ++k) dev::rk4(r6, v6, h, m[2 * k], m[2 * k + 1], m[2 * k + 2]);
                D3 r1 = r, v1 = v;
                dev::rk4(r1, v1, tau, m[0], m[6], m[12], dev::AccelMixed());
                Frame fr;
                float f[NIN];
                raw_features(r, v, m[0], tau, f, fr);
                double y[NOUT];
                to_frame(fr, D3{r6.x - r1.x, r6.y - r1.y, r6.z - r1.z}, D3{v6.x - v1.x, v6.y - v1.y, v6.z - v1.z}, y);
                for (int k = 0; k < NIN; ++k) fx[slot * NIN + k] = f[k];
                for (int k = 0; k < NOUT; ++k) fy[slot * NOUT + k] = static_cast<float>(y[k]);
            }
        }
        dev::rk4(r, v, tau, m[0], m[6], m[12], dev::AccelMixed());
        if (r.x * r.x + r.y * r.y + r.z * r.z < dev::R_EARTH * dev::R_EARTH) break;
    }
    for (; emitted < per_orbit; ++emitted) valid[static_cast<size_t>(o) * per_orbit + emitted] = 0;
}

// Standardises samples in place: x <- (x - fmu) / fsig, y <- (y - omu) / osig.
__global__ void __launch_bounds__(256) k_net_standardise(float* __restrict__ fx, float* __restrict__ fy, size_t n,
                                                         NetNorm nn) {
    const size_t s = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (s >= n) return;
    for (int k = 0; k < NIN; ++k) fx[s * NIN + k] = (fx[s * NIN + k] - nn.fmu[k]) / nn.fsig[k];
    for (int k = 0; k < NOUT; ++k) fy[s * NOUT + k] = (fy[s * NOUT + k] - nn.omu[k]) / nn.osig[k];
}

// Inference on given standardised inputs (tests): 64 columns per wave.
__global__ void __launch_bounds__(256) k_net_infer(const float* __restrict__ X, int n,
                                                   const uint16_t* __restrict__ ffwd, const float* __restrict__ P,
                                                   float* __restrict__ out) {
    const int lane = static_cast<int>(__lane_id());
    const int wave = (blockIdx.x * 256 + threadIdx.x) / 64;
    const int c0 = wave * 64;
    if (c0 >= n) return;
    const int j = lane & 15, g = lane >> 4;
    u32x4 xin[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int c = c0 + 16 * t + j;
        float x[NIN];
        for (int k = 0; k < NIN; ++k) x[k] = c < n ? X[static_cast<size_t>(c) * NIN + k] : 0.0f;
        xin[t] = input_operand(x, g);
    }
    f32x4 o[4];
    forward_tiles<4>(xin, ffwd, P, lane, o);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int c = c0 + 16 * t + j;
        const float v[4] = {o[t].x, o[t].y, o[t].z, o[t].w};
        for (int e = 0; e < 4; ++e) {
            const int row = 4 * g + e;
            if (c < n && row < NOUT) out[static_cast<size_t>(c) * NOUT + row] = v[e];
        }
    }
}

}  // namespace net
}  // namespace kessler
