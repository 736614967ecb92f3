// qk_device.hpp -- device-side building blocks of the fused path kernels
// (gfx950, wave64): counter-based RNG, per-precision math, Heston/Bates steps,
// payoff evaluation and wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qk_types.hpp"

namespace qk {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  The key is uniform across a launch,
// so the key schedule stays in scalar registers.  Each round is two 32x32->64
// products, written as one 64-bit product so that each becomes a single
// v_mad_u64_u32 giving both halves, and four XORs.
// ---------------------------------------------------------------------------
struct Words { uint32_t w0, w1, w2, w3; };

// Three-input XOR in one gfx950 v_bitop3_b32 (truth table 0x96); the compiler
// does not form it from two XORs by itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// PLAIN: rounds written with ordinary XORs, which the compiler keeps on the
// scalar unit when their operands are uniform.  With the counter (path low,
// path high, block, slot), a uniform path-high word and a uniform block,
// rounds 1 and 2 consist of products that are uniform or fixed per path
// (hoisted out of the step loop) and XORs of them: ordinary XORs there cost
// two vector instructions in all; bitop3 would force them onto the vector
// unit.
template <int PLAIN = 0>
__device__ __forceinline__ Words philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0, n2;
        if (r < PLAIN) {
            n0 = (uint32_t)(p1 >> 32) ^ (c1 ^ k0);
            n2 = (uint32_t)(p0 >> 32) ^ (c3 ^ k1);
        } else {
            n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
            n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        }
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Words{c0, c1, c2, c3};
}

// ---------------------------------------------------------------------------
// Per-precision math.  FP32 uses the CDNA transcendental unit directly
// (v_exp_f32 / v_log_f32 are base-2, v_sin_f32 / v_cos_f32 take revolutions).
// FP64 exp, log, sqrt, reciprocal and sin/cos(2 pi u) are written here for
// the arguments the kernels give them (finite; positive for log, sqrt and the
// reciprocal; u in [0, 1]), without the special-case handling of the library
// routines.  Measured against long-double references over 2e7 arguments:
// exp <= 0.9 ulp, log <= 1.5 ulp (99.97 % within 1 ulp), sqrt and reciprocal
// <= 0.5 ulp, sin/cos(2 pi u) <= 2.2 ulp where |value| > 0.01 and within
// 7.1e-17 everywhere.
// ---------------------------------------------------------------------------
template <typename Real> struct Math;

template <> struct Math<float> {
    __device__ static __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
    __device__ static __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
    __device__ static __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float fmax(float a, float b) { return __builtin_fmaxf(a, b); }
    __device__ static __forceinline__ float fmin(float a, float b) { return __builtin_fminf(a, b); }
    // Running max / min of non-negative values (spot prices) on their bit
    // patterns: the same result as fmax / fmin, without the NaN-quieting
    // operand fix-ups the compiler adds to a loop-carried fmax.
    __device__ static __forceinline__ float track_max(float a, float b) {
        return __uint_as_float(__builtin_elementwise_max(__float_as_uint(a), __float_as_uint(b)));
    }
    __device__ static __forceinline__ float track_min(float a, float b) {
        return __uint_as_float(__builtin_elementwise_min(__float_as_uint(a), __float_as_uint(b)));
    }
    // -2 ln u as log2(u) * (-2 ln 2): scaling by -2 commutes with rounding, so
    // this equals -2 * log(u) exactly, one multiply shorter.
    __device__ static __forceinline__ float neg2log(float u) {
        return __builtin_amdgcn_logf(u) * (-2.0f * 0.6931471805599453f);
    }
    // ln(x) / 2, equal to 0.5f * log(x) for the same reason.
    __device__ static __forceinline__ float half_log(float x) {
        return __builtin_amdgcn_logf(x) * (0.5f * 0.6931471805599453f);
    }
    // sin/cos of 2*pi*u for u in (0, 1).
    __device__ static __forceinline__ void sincos2pi(float u, float &s, float &c) {
        s = __builtin_amdgcn_sinf(u);
        c = __builtin_amdgcn_cosf(u);
    }
    // A Philox word rounded to FP32 (round to nearest even), scaled by 2^-32
    // and shifted by 2^-33: RN(RN(w) 2^-32 + 2^-33), in (0, 1].  Two
    // instructions (convert, fma); the spacing is that of FP32 itself, 2^-24
    // near 1 and finer towards 0 (smallest value 2^-33).
    __device__ static __forceinline__ float uniform(uint32_t w) {
        return __builtin_fmaf((float)w, 0x1.0p-32f, 0x1.0p-33f);
    }
};

// Polynomial coefficients, lowest order first; quantkos_amd/csrc/fit_fp64_poly.py
// derives them (Chebyshev near-minimax fits at 50 digits, rounded to double).
//   EXP_P: e^r on r in [-0.35, 0.35]
// sin/cos(2 pi u) from the table of the 64 angles 2 pi j/64 (each as a double
// and a double remainder) and the angle 2 pi r/64, |r| <= 1/2:
// SIN64_S(w) = sin(pi/32 sqrt w)/sqrt w, COS64_M(w) = (cos(pi/32 sqrt w) - 1)/w.
static __constant__ double SIN64_S[4] = {0x1.921fb54442d18p-4, -0x1.4abbce625bd02p-13,
                                        0x1.466bc66a361e0p-24, -0x1.32cd8b0fba547p-36};
static __constant__ double COS64_M[4] = {-0x1.3bd3cc9be45dep-8, 0x1.03c1f081b5a5ap-18,
                                        -0x1.55d3c7db89f7bp-30, 0x1.e1ee6b4c2f23fp-43};
struct SinCosEntry {
    double s, c, s_lo, c_lo;
};
static __constant__ SinCosEntry SINCOS_TAB[64] = {
    {0x0.0p+0, 0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
    {0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1, -0x1.e2718d26ed688p-60, -0x1.87df6378811c7p-55},
    {0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1, -0x1.26d19b9ff8d82p-57, 0x1.562172a361fd3p-56},
    {0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1, -0x1.5d28da2c4612dp-56, 0x1.760b1e2e3f81ep-55},
    {0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1, -0x1.72cedd3d5a610p-57, 0x1.457e610231ac2p-56},
    {0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1, 0x1.e0d891d3c6841p-58, -0x1.6e0b1757c8d07p-56},
    {0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1, 0x1.b25dd267f6600p-55, 0x1.9f630e8b6dac8p-60},
    {0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1, 0x1.8076a2cfdc6b3p-57, -0x1.2c5e12ed1336dp-55},
    {0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1, -0x1.bdd3413b26456p-55, -0x1.bdd3413b26456p-55},
    {0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1, -0x1.2c5e12ed1336dp-55, 0x1.8076a2cfdc6b3p-57},
    {0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1, 0x1.9f630e8b6dac8p-60, 0x1.b25dd267f6600p-55},
    {0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2, -0x1.6e0b1757c8d07p-56, 0x1.e0d891d3c6841p-58},
    {0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2, 0x1.457e610231ac2p-56, -0x1.72cedd3d5a610p-57},
    {0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2, 0x1.760b1e2e3f81ep-55, -0x1.5d28da2c4612dp-56},
    {0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3, 0x1.562172a361fd3p-56, -0x1.26d19b9ff8d82p-57},
    {0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4, -0x1.87df6378811c7p-55, -0x1.e2718d26ed688p-60},
    {0x1.0000000000000p+0, -0x1.8338bfdf4415ap-171, 0x0.0p+0, 0x1.d89cd9128a504p-226},
    {0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4, -0x1.87df6378811c7p-55, 0x1.e2718d26ed688p-60},
    {0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3, 0x1.562172a361fd3p-56, 0x1.26d19b9ff8d82p-57},
    {0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2, 0x1.760b1e2e3f81ep-55, 0x1.5d28da2c4612dp-56},
    {0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2, 0x1.457e610231ac2p-56, 0x1.72cedd3d5a610p-57},
    {0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2, -0x1.6e0b1757c8d07p-56, -0x1.e0d891d3c6841p-58},
    {0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1, 0x1.9f630e8b6dac8p-60, -0x1.b25dd267f6600p-55},
    {0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1, -0x1.2c5e12ed1336dp-55, -0x1.8076a2cfdc6b3p-57},
    {0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1, -0x1.bdd3413b26456p-55, 0x1.bdd3413b26456p-55},
    {0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1, 0x1.8076a2cfdc6b3p-57, 0x1.2c5e12ed1336dp-55},
    {0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1, 0x1.b25dd267f6600p-55, -0x1.9f630e8b6dac8p-60},
    {0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1, 0x1.e0d891d3c6841p-58, 0x1.6e0b1757c8d07p-56},
    {0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1, -0x1.72cedd3d5a610p-57, -0x1.457e610231ac2p-56},
    {0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1, -0x1.5d28da2c4612dp-56, -0x1.760b1e2e3f81ep-55},
    {0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1, -0x1.26d19b9ff8d82p-57, -0x1.562172a361fd3p-56},
    {0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1, -0x1.e2718d26ed688p-60, 0x1.87df6378811c7p-55},
    {-0x1.8338bfdf4415ap-170, -0x1.0000000000000p+0, 0x1.d89cd9128a504p-225, 0x0.0p+0},
    {-0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1, 0x1.e2718d26ed688p-60, 0x1.87df6378811c7p-55},
    {-0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1, 0x1.26d19b9ff8d82p-57, -0x1.562172a361fd3p-56},
    {-0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1, 0x1.5d28da2c4612dp-56, -0x1.760b1e2e3f81ep-55},
    {-0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1, 0x1.72cedd3d5a610p-57, -0x1.457e610231ac2p-56},
    {-0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1, -0x1.e0d891d3c6841p-58, 0x1.6e0b1757c8d07p-56},
    {-0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1, -0x1.b25dd267f6600p-55, -0x1.9f630e8b6dac8p-60},
    {-0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1, -0x1.8076a2cfdc6b3p-57, 0x1.2c5e12ed1336dp-55},
    {-0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1, 0x1.bdd3413b26456p-55, 0x1.bdd3413b26456p-55},
    {-0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1, 0x1.2c5e12ed1336dp-55, -0x1.8076a2cfdc6b3p-57},
    {-0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1, -0x1.9f630e8b6dac8p-60, -0x1.b25dd267f6600p-55},
    {-0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2, 0x1.6e0b1757c8d07p-56, -0x1.e0d891d3c6841p-58},
    {-0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2, -0x1.457e610231ac2p-56, 0x1.72cedd3d5a610p-57},
    {-0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2, -0x1.760b1e2e3f81ep-55, 0x1.5d28da2c4612dp-56},
    {-0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3, -0x1.562172a361fd3p-56, 0x1.26d19b9ff8d82p-57},
    {-0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4, 0x1.87df6378811c7p-55, 0x1.e2718d26ed688p-60},
    {-0x1.0000000000000p+0, 0x1.913547f3b9882p-168, 0x0.0p+0, -0x1.589d68b379ef1p-222},
    {-0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4, 0x1.87df6378811c7p-55, -0x1.e2718d26ed688p-60},
    {-0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3, -0x1.562172a361fd3p-56, -0x1.26d19b9ff8d82p-57},
    {-0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2, -0x1.760b1e2e3f81ep-55, -0x1.5d28da2c4612dp-56},
    {-0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2, -0x1.457e610231ac2p-56, -0x1.72cedd3d5a610p-57},
    {-0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2, 0x1.6e0b1757c8d07p-56, 0x1.e0d891d3c6841p-58},
    {-0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1, -0x1.9f630e8b6dac8p-60, 0x1.b25dd267f6600p-55},
    {-0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1, 0x1.2c5e12ed1336dp-55, 0x1.8076a2cfdc6b3p-57},
    {-0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1, 0x1.bdd3413b26456p-55, -0x1.bdd3413b26456p-55},
    {-0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1, -0x1.8076a2cfdc6b3p-57, -0x1.2c5e12ed1336dp-55},
    {-0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1, -0x1.b25dd267f6600p-55, 0x1.9f630e8b6dac8p-60},
    {-0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1, -0x1.e0d891d3c6841p-58, -0x1.6e0b1757c8d07p-56},
    {-0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1, 0x1.72cedd3d5a610p-57, 0x1.457e610231ac2p-56},
    {-0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1, 0x1.5d28da2c4612dp-56, 0x1.760b1e2e3f81ep-55},
    {-0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1, 0x1.26d19b9ff8d82p-57, 0x1.562172a361fd3p-56},
    {-0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1, 0x1.e2718d26ed688p-60, -0x1.87df6378811c7p-55},
};
static __constant__ double EXP_P[12] = {0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000012p-1,
                                        0x1.555555555555bp-3, 0x1.555555554e892p-5, 0x1.111111110ef9cp-7,
                                        0x1.6c16c189ae5d5p-10, 0x1.a01a01b24ec34p-13, 0x1.a0198d2d43afbp-16,
                                        0x1.71ddf2b060bb0p-19, 0x1.28b82404fa64fp-22, 0x1.af682e715258ap-26};
// log1p(r) = r + r^2 LOG1P_Q(r) on |r| <= 0.009, and the table of the
// table-driven log: for m in [sqrt(1/2), sqrt(2)) cut into 128 intervals by
// the high word of m, {-log(invc) rounded to double, invc, the float
// remainder of -log(invc)} with invc ~ 1 / (interval centre) of 7 significant
// bits (1 near m = 1).
static __constant__ double LOG1P_Q[6] = {-0x1.0000000000013p-1, 0x1.5555555555577p-2, -0x1.fffffff025cdcp-3,
                                         0x1.9999998b824e9p-3, -0x1.555d4be9e7898p-3, 0x1.24995d37a01f0p-3};
struct LogEntry {
    double lhi;
    float invc, llo;
};
static __constant__ LogEntry LOG_TAB[128] = {
    {-0x1.5d1bdbf5809cap-2, 0x1.6800000000000p+0f, -0x1.4236380000000p-56f},
    {-0x1.5d1bdbf5809cap-2, 0x1.6800000000000p+0f, -0x1.4236380000000p-56f},
    {-0x1.51aad872df82dp-2, 0x1.6400000000000p+0f, -0x1.3927ac0000000p-59f},
    {-0x1.51aad872df82dp-2, 0x1.6400000000000p+0f, -0x1.3927ac0000000p-59f},
    {-0x1.4618bc21c5ec2p-2, 0x1.6000000000000p+0f, -0x1.f42dec0000000p-56f},
    {-0x1.4618bc21c5ec2p-2, 0x1.6000000000000p+0f, -0x1.f42dec0000000p-56f},
    {-0x1.3a64c556945eap-2, 0x1.5c00000000000p+0f, 0x1.c686520000000p-57f},
    {-0x1.3a64c556945eap-2, 0x1.5c00000000000p+0f, 0x1.c686520000000p-57f},
    {-0x1.2e8e2bae11d31p-2, 0x1.5800000000000p+0f, 0x1.8f4cdc0000000p-56f},
    {-0x1.2e8e2bae11d31p-2, 0x1.5800000000000p+0f, 0x1.8f4cdc0000000p-56f},
    {-0x1.2e8e2bae11d31p-2, 0x1.5800000000000p+0f, 0x1.8f4cdc0000000p-56f},
    {-0x1.22941fbcf7966p-2, 0x1.5400000000000p+0f, 0x1.76f5ec0000000p-56f},
    {-0x1.22941fbcf7966p-2, 0x1.5400000000000p+0f, 0x1.76f5ec0000000p-56f},
    {-0x1.1675cababa60ep-2, 0x1.5000000000000p+0f, -0x1.ce63ea0000000p-61f},
    {-0x1.1675cababa60ep-2, 0x1.5000000000000p+0f, -0x1.ce63ea0000000p-61f},
    {-0x1.0a324e27390e3p-2, 0x1.4c00000000000p+0f, -0x1.7dcfde0000000p-56f},
    {-0x1.0a324e27390e3p-2, 0x1.4c00000000000p+0f, -0x1.7dcfde0000000p-56f},
    {-0x1.0a324e27390e3p-2, 0x1.4c00000000000p+0f, -0x1.7dcfde0000000p-56f},
    {-0x1.fb9186d5e3e2bp-3, 0x1.4800000000000p+0f, 0x1.caaae60000000p-57f},
    {-0x1.fb9186d5e3e2bp-3, 0x1.4800000000000p+0f, 0x1.caaae60000000p-57f},
    {-0x1.e27076e2af2e6p-3, 0x1.4400000000000p+0f, 0x1.6157800000000p-59f},
    {-0x1.e27076e2af2e6p-3, 0x1.4400000000000p+0f, 0x1.6157800000000p-59f},
    {-0x1.e27076e2af2e6p-3, 0x1.4400000000000p+0f, 0x1.6157800000000p-59f},
    {-0x1.c8ff7c79a9a22p-3, 0x1.4000000000000p+0f, 0x1.4f68a00000000p-57f},
    {-0x1.c8ff7c79a9a22p-3, 0x1.4000000000000p+0f, 0x1.4f68a00000000p-57f},
    {-0x1.af3c94e80bff3p-3, 0x1.3c00000000000p+0f, 0x1.398d000000000p-58f},
    {-0x1.af3c94e80bff3p-3, 0x1.3c00000000000p+0f, 0x1.398d000000000p-58f},
    {-0x1.af3c94e80bff3p-3, 0x1.3c00000000000p+0f, 0x1.398d000000000p-58f},
    {-0x1.9525a9cf456b4p-3, 0x1.3800000000000p+0f, -0x1.d904c20000000p-57f},
    {-0x1.9525a9cf456b4p-3, 0x1.3800000000000p+0f, -0x1.d904c20000000p-57f},
    {-0x1.7ab890210d909p-3, 0x1.3400000000000p+0f, -0x1.be36b20000000p-59f},
    {-0x1.7ab890210d909p-3, 0x1.3400000000000p+0f, -0x1.be36b20000000p-59f},
    {-0x1.7ab890210d909p-3, 0x1.3400000000000p+0f, -0x1.be36b20000000p-59f},
    {-0x1.5ff3070a793d4p-3, 0x1.3000000000000p+0f, 0x1.bc60f00000000p-58f},
    {-0x1.5ff3070a793d4p-3, 0x1.3000000000000p+0f, 0x1.bc60f00000000p-58f},
    {-0x1.5ff3070a793d4p-3, 0x1.3000000000000p+0f, 0x1.bc60f00000000p-58f},
    {-0x1.44d2b6ccb7d1ep-3, 0x1.2c00000000000p+0f, -0x1.9f4f660000000p-57f},
    {-0x1.44d2b6ccb7d1ep-3, 0x1.2c00000000000p+0f, -0x1.9f4f660000000p-57f},
    {-0x1.44d2b6ccb7d1ep-3, 0x1.2c00000000000p+0f, -0x1.9f4f660000000p-57f},
    {-0x1.29552f81ff523p-3, 0x1.2800000000000p+0f, -0x1.3017720000000p-57f},
    {-0x1.29552f81ff523p-3, 0x1.2800000000000p+0f, -0x1.3017720000000p-57f},
    {-0x1.29552f81ff523p-3, 0x1.2800000000000p+0f, -0x1.3017720000000p-57f},
    {-0x1.0d77e7cd08e59p-3, 0x1.2400000000000p+0f, -0x1.9a5dc60000000p-57f},
    {-0x1.0d77e7cd08e59p-3, 0x1.2400000000000p+0f, -0x1.9a5dc60000000p-57f},
    {-0x1.0d77e7cd08e59p-3, 0x1.2400000000000p+0f, -0x1.9a5dc60000000p-57f},
    {-0x1.e27076e2af2e6p-4, 0x1.2000000000000p+0f, 0x1.6157800000000p-60f},
    {-0x1.e27076e2af2e6p-4, 0x1.2000000000000p+0f, 0x1.6157800000000p-60f},
    {-0x1.e27076e2af2e6p-4, 0x1.2000000000000p+0f, 0x1.6157800000000p-60f},
    {-0x1.a926d3a4ad563p-4, 0x1.1c00000000000p+0f, -0x1.942f480000000p-58f},
    {-0x1.a926d3a4ad563p-4, 0x1.1c00000000000p+0f, -0x1.942f480000000p-58f},
    {-0x1.a926d3a4ad563p-4, 0x1.1c00000000000p+0f, -0x1.942f480000000p-58f},
    {-0x1.6f0d28ae56b4cp-4, 0x1.1800000000000p+0f, 0x1.906d9a0000000p-58f},
    {-0x1.6f0d28ae56b4cp-4, 0x1.1800000000000p+0f, 0x1.906d9a0000000p-58f},
    {-0x1.6f0d28ae56b4cp-4, 0x1.1800000000000p+0f, 0x1.906d9a0000000p-58f},
    {-0x1.6f0d28ae56b4cp-4, 0x1.1800000000000p+0f, 0x1.906d9a0000000p-58f},
    {-0x1.341d7961bd1d1p-4, 0x1.1400000000000p+0f, 0x1.b599f20000000p-58f},
    {-0x1.341d7961bd1d1p-4, 0x1.1400000000000p+0f, 0x1.b599f20000000p-58f},
    {-0x1.341d7961bd1d1p-4, 0x1.1400000000000p+0f, 0x1.b599f20000000p-58f},
    {-0x1.f0a30c01162a6p-5, 0x1.1000000000000p+0f, -0x1.85f3260000000p-59f},
    {-0x1.f0a30c01162a6p-5, 0x1.1000000000000p+0f, -0x1.85f3260000000p-59f},
    {-0x1.f0a30c01162a6p-5, 0x1.1000000000000p+0f, -0x1.85f3260000000p-59f},
    {-0x1.f0a30c01162a6p-5, 0x1.1000000000000p+0f, -0x1.85f3260000000p-59f},
    {-0x1.77458f632dcfcp-5, 0x1.0c00000000000p+0f, -0x1.18d3ca0000000p-59f},
    {-0x1.77458f632dcfcp-5, 0x1.0c00000000000p+0f, -0x1.18d3ca0000000p-59f},
    {-0x1.77458f632dcfcp-5, 0x1.0c00000000000p+0f, -0x1.18d3ca0000000p-59f},
    {-0x1.f829b0e783300p-6, 0x1.0800000000000p+0f, -0x1.33e3f00000000p-60f},
    {-0x1.f829b0e783300p-6, 0x1.0800000000000p+0f, -0x1.33e3f00000000p-60f},
    {-0x1.f829b0e783300p-6, 0x1.0800000000000p+0f, -0x1.33e3f00000000p-60f},
    {-0x1.f829b0e783300p-6, 0x1.0800000000000p+0f, -0x1.33e3f00000000p-60f},
    {-0x1.fc0a8b0fc03e4p-7, 0x1.0400000000000p+0f, 0x1.83092c0000000p-62f},
    {-0x1.fc0a8b0fc03e4p-7, 0x1.0400000000000p+0f, 0x1.83092c0000000p-62f},
    {-0x1.fc0a8b0fc03e4p-7, 0x1.0400000000000p+0f, 0x1.83092c0000000p-62f},
    {-0x1.fc0a8b0fc03e4p-7, 0x1.0400000000000p+0f, 0x1.83092c0000000p-62f},
    {0x0.0p+0, 0x1.0000000000000p+0f, 0x0.0p+0f},
    {0x0.0p+0, 0x1.0000000000000p+0f, 0x0.0p+0f},
    {0x0.0p+0, 0x1.0000000000000p+0f, 0x0.0p+0f},
    {0x1.0205658935847p-6, 0x1.f800000000000p-1f, 0x1.27c8e80000000p-60f},
    {0x1.0205658935847p-6, 0x1.f800000000000p-1f, 0x1.27c8e80000000p-60f},
    {0x1.8492528c8cabfp-6, 0x1.f400000000000p-1f, -0x1.d192d00000000p-60f},
    {0x1.0415d89e74444p-5, 0x1.f000000000000p-1f, 0x1.c05cf20000000p-59f},
    {0x1.466aed42de3eap-5, 0x1.ec00000000000p-1f, -0x1.cdd6f80000000p-59f},
    {0x1.894aa149fb343p-5, 0x1.e800000000000p-1f, 0x1.a8be980000000p-60f},
    {0x1.ccb73cdddb2ccp-5, 0x1.e400000000000p-1f, -0x1.e48fb00000000p-59f},
    {0x1.08598b59e3a07p-4, 0x1.e000000000000p-1f, -0x1.dd700a0000000p-58f},
    {0x1.2aa04a44717a5p-4, 0x1.dc00000000000p-1f, -0x1.d15d380000000p-58f},
    {0x1.4d3115d207eacp-4, 0x1.d800000000000p-1f, 0x1.769f420000000p-58f},
    {0x1.700d30aeac0e1p-4, 0x1.d400000000000p-1f, -0x1.7256620000000p-61f},
    {0x1.700d30aeac0e1p-4, 0x1.d400000000000p-1f, -0x1.7256620000000p-61f},
    {0x1.9335e5d594989p-4, 0x1.d000000000000p-1f, -0x1.478a860000000p-58f},
    {0x1.b6ac88dad5b1cp-4, 0x1.cc00000000000p-1f, -0x1.0057ee0000000p-59f},
    {0x1.da727638446a2p-4, 0x1.c800000000000p-1f, 0x1.401fa80000000p-58f},
    {0x1.fe89139dbd566p-4, 0x1.c400000000000p-1f, -0x1.ac9f420000000p-58f},
    {0x1.fe89139dbd566p-4, 0x1.c400000000000p-1f, -0x1.ac9f420000000p-58f},
    {0x1.1178e8227e47cp-3, 0x1.c000000000000p-1f, -0x1.0e63a60000000p-58f},
    {0x1.23d712a49c202p-3, 0x1.bc00000000000p-1f, -0x1.6e38160000000p-57f},
    {0x1.365fcb0159016p-3, 0x1.b800000000000p-1f, 0x1.7d411a0000000p-58f},
    {0x1.365fcb0159016p-3, 0x1.b800000000000p-1f, 0x1.7d411a0000000p-58f},
    {0x1.4913d8333b561p-3, 0x1.b400000000000p-1f, -0x1.0d56040000000p-58f},
    {0x1.5bf406b543db2p-3, 0x1.b000000000000p-1f, -0x1.1f5b440000000p-61f},
    {0x1.6f0128b756abcp-3, 0x1.ac00000000000p-1f, -0x1.8de59c0000000p-57f},
    {0x1.6f0128b756abcp-3, 0x1.ac00000000000p-1f, -0x1.8de59c0000000p-57f},
    {0x1.823c16551a3c2p-3, 0x1.a800000000000p-1f, -0x1.1232ce0000000p-57f},
    {0x1.95a5adcf7017fp-3, 0x1.a400000000000p-1f, 0x1.142c500000000p-58f},
    {0x1.95a5adcf7017fp-3, 0x1.a400000000000p-1f, 0x1.142c500000000p-58f},
    {0x1.a93ed3c8ad9e3p-3, 0x1.a000000000000p-1f, 0x1.bcafaa0000000p-57f},
    {0x1.bd087383bd8adp-3, 0x1.9c00000000000p-1f, 0x1.dd35600000000p-60f},
    {0x1.bd087383bd8adp-3, 0x1.9c00000000000p-1f, 0x1.dd35600000000p-60f},
    {0x1.d1037f2655e7bp-3, 0x1.9800000000000p-1f, 0x1.6062920000000p-57f},
    {0x1.e530effe71012p-3, 0x1.9400000000000p-1f, 0x1.2276040000000p-59f},
    {0x1.e530effe71012p-3, 0x1.9400000000000p-1f, 0x1.2276040000000p-59f},
    {0x1.f991c6cb3b379p-3, 0x1.9000000000000p-1f, 0x1.f665060000000p-57f},
    {0x1.f991c6cb3b379p-3, 0x1.9000000000000p-1f, 0x1.f665060000000p-57f},
    {0x1.07138604d5862p-2, 0x1.8c00000000000p-1f, 0x1.cdb16e0000000p-56f},
    {0x1.1178e8227e47cp-2, 0x1.8800000000000p-1f, -0x1.0e63a60000000p-57f},
    {0x1.1178e8227e47cp-2, 0x1.8800000000000p-1f, -0x1.0e63a60000000p-57f},
    {0x1.1bf99635a6b95p-2, 0x1.8400000000000p-1f, -0x1.12aeb80000000p-57f},
    {0x1.1bf99635a6b95p-2, 0x1.8400000000000p-1f, -0x1.12aeb80000000p-57f},
    {0x1.269621134db92p-2, 0x1.8000000000000p-1f, 0x1.e0efae0000000p-56f},
    {0x1.269621134db92p-2, 0x1.8000000000000p-1f, 0x1.e0efae0000000p-56f},
    {0x1.314f1e1d35ce4p-2, 0x1.7c00000000000p-1f, -0x1.3d69900000000p-56f},
    {0x1.3c25277333184p-2, 0x1.7800000000000p-1f, -0x1.2ad27e0000000p-56f},
    {0x1.3c25277333184p-2, 0x1.7800000000000p-1f, -0x1.2ad27e0000000p-56f},
    {0x1.4718dc271c41bp-2, 0x1.7400000000000p-1f, 0x1.8fb4c20000000p-60f},
    {0x1.4718dc271c41bp-2, 0x1.7400000000000p-1f, 0x1.8fb4c20000000p-60f},
    {0x1.522ae0738a3d8p-2, 0x1.7000000000000p-1f, -0x1.8f7e9c0000000p-57f},
    {0x1.522ae0738a3d8p-2, 0x1.7000000000000p-1f, -0x1.8f7e9c0000000p-57f},
    {0x1.5d5bddf595f30p-2, 0x1.6c00000000000p-1f, -0x1.6541140000000p-56f},
    {0x1.5d5bddf595f30p-2, 0x1.6c00000000000p-1f, -0x1.6541140000000p-56f},
};
// ln 2 = LN2_HI + LN2_LO; LN2_HI has 32 significant bits, so k * LN2_HI is exact.
__device__ constexpr double LN2_HI = 0x1.62e42ff000000p-1;
__device__ constexpr double LN2_LO = -0x1.718432a1b0e26p-35;

// Horner evaluation, lowest coefficient first.  The empty asm statement emits
// no instruction; its "+s" operand makes the compiler hold each coefficient in
// scalar registers at its use, so each step is one v_fma_f64 with a scalar
// operand (without it the coefficients are kept in vector registers and each
// step becomes a copy plus v_fmac_f64).
template <int N>
__device__ __forceinline__ double horner(const double (&c)[N], double x) {
    double p = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) {
        double ci = c[i];
        __asm__("" : "+s"(ci));
        p = __builtin_fma(p, x, ci);
    }
    return p;
}

template <> struct Math<double> {
    // e^x = 2^n e^r, n = rint(x log2 e), r = x - n LN2_HI - n LN2_LO (the first
    // product is exact, |r| <= ln2/2); e^r by EXP_P.  Overflows to inf above
    // ~709.8 and to 0 below ~-745 through the scaling.
    __device__ static __forceinline__ double exp(double x) {
        const double n = __builtin_rint(x * 0x1.71547652b82fep+0);
        double r = __builtin_fma(-n, LN2_HI, x);
        r = __builtin_fma(-n, LN2_LO, r);
        return __builtin_amdgcn_ldexp(horner(EXP_P, r), (int)n);
    }
    // Natural log of a positive normal finite x, table-driven (the scheme of
    // Tang, ACM TOMS 16 (1990), with an inexact reciprocal table): x = 2^k m,
    // m in [sqrt(1/2), sqrt(2)), k and m from the high word (subtracting the
    // high word of sqrt(1/2) puts k in its exponent field, and removing k
    // from the exponent of x leaves m); the next 7 bits pick a table entry;
    // r = m invc - 1 is exact in one fma (invc has 7 significant bits) and
    // log x = k ln2 - log(invc) + log1p(r), the two constant parts each split
    // into a high and a low term.
    __device__ static __forceinline__ double log(double x) {
        const uint64_t ix = __builtin_bit_cast(uint64_t, x);
        const uint32_t hx = (uint32_t)(ix >> 32);
        const int32_t tmp = (int32_t)(hx - 0x3fe6a09eu);
        const int k = tmp >> 20;
        const uint32_t hm = hx - ((uint32_t)tmp & 0xfff00000u);
        const double m = __builtin_bit_cast(double, ((uint64_t)hm << 32) | (uint32_t)ix);
        const LogEntry e = LOG_TAB[((uint32_t)tmp >> 13) & 127u];
        const double r = __builtin_fma(m, (double)e.invc, -1.0);
        const double lp = __builtin_fma(r * r, horner(LOG1P_Q, r), r);
        const double kd = (double)k;
        const double hi = __builtin_fma(kd, LN2_HI, e.lhi);
        const double lo = __builtin_fma(kd, LN2_LO, (double)e.llo);
        return hi + (lp + lo);
    }
    // sqrt of a finite x >= 0: one Goldschmidt step from v_rsq_f64 (g ~ sqrt x,
    // h ~ 1/(2 sqrt x)), then the Newton residual step g + h (x - g^2).  The max
    // with the smallest subnormal keeps rsq finite, so sqrt(0) = 0.
    __device__ static __forceinline__ double sqrt(double x) {
        const double y = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-1074));
        double g = x * y, h = 0.5 * y;
        const double e = __builtin_fma(-g, h, 0.5);
        g = __builtin_fma(g, e, g);
        h = __builtin_fma(h, e, h);
        return __builtin_fma(__builtin_fma(-g, g, x), h, g);
    }
    // 1/x for a finite normal x: v_rcp_f64 and two Newton steps.
    __device__ static __forceinline__ double rcp(double x) {
        double r = __builtin_amdgcn_rcp(x);
        r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
        return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    }
    // IEEE division (the LSM basis argument S/K, off the per-step path).
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
    __device__ static __forceinline__ double fmax(double a, double b) { return ::fmax(a, b); }
    __device__ static __forceinline__ double fmin(double a, double b) { return ::fmin(a, b); }
    __device__ static __forceinline__ double track_max(double a, double b) { return ::fmax(a, b); }
    __device__ static __forceinline__ double track_min(double a, double b) { return ::fmin(a, b); }
    __device__ static __forceinline__ double neg2log(double u) { return -2.0 * log(u); }
    __device__ static __forceinline__ double half_log(double x) { return 0.5 * log(x); }
    // sin/cos of 2 pi u for u in [0, 1]: 64u = j + r exactly (|r| <= 1/2), and
    // with A = 2 pi j/64 from the table and B = 2 pi r/64,
    // sin(A + B) = sin A + (sin A (cos B - 1) + cos A sin B) and
    // cos(A + B) = cos A + (cos A (cos B - 1) - sin A sin B), the table values
    // carried with their remainders.
    __device__ static __forceinline__ void sincos2pi(double u, double &s, double &c) {
        const double x = 64.0 * u;
        const double n = __builtin_rint(x);
        const double r = x - n;
        const double w = r * r;
        const double sb = r * horner(SIN64_S, w);
        const double cm1 = w * horner(COS64_M, w);
        const SinCosEntry e = SINCOS_TAB[(int)n & 63];
        s = e.s + (e.s_lo + __builtin_fma(e.s, cm1, e.c * sb));
        c = e.c + (e.c_lo + __builtin_fma(e.c, cm1, -(e.s * sb)));
    }
    // All 32 bits of a Philox word, centred: (w + 0.5) * 2^-32, exact in FP64
    // (0.5 is an inline operand and 2^-32 a 32-bit literal; 2^-33 as the
    // addend of an fma would take two moves per draw).
    __device__ static __forceinline__ double uniform(uint32_t w) {
        return ((double)w + 0.5) * 0x1.0p-32;
    }
};

// ---------------------------------------------------------------------------
// One Heston step on (S or log S, v) given the slot-0 words of (path, step).
// Formulas follow MSchemes.hpp: Euler :133-155, implicit Milstein :229-252,
// Andersen QE :363-441.  Returns the log-increment of S.
// ---------------------------------------------------------------------------
template <typename Real, int SCHEME>
__device__ __forceinline__ Real heston_step(const LaneC<Real> &c, Real z1, Real z2, uint32_t w2, Real &v) {
    using M = Math<Real>;
    const Real v_n = v;
    if constexpr (SCHEME == SCHEME_EULER || SCHEME == SCHEME_MILSTEIN) {
        const Real zs = c.rho * z1 + c.rho_c * z2;
        const Real sv = M::sqrt(v_n);
        const Real expo = c.rqdt - v_n * c.half_dt + (sv * c.sqrt_dt) * zs;
        Real vn;
        if constexpr (SCHEME == SCHEME_EULER) {
            vn = v_n * c.omkdt + c.ktdt + (c.ssdt * sv) * z1;
        } else {
            vn = c.iden * (v_n + c.ktdt + (c.ssdt * sv) * z1 + c.qs2dt * (z1 * z1 - Real(1)));
        }
        v = M::fmax(vn, Real(0));
        return expo;
    } else {
        // Andersen QE with martingale correction.  The exponential branch draws
        // its uniform from the step's third word (a2; same law as Phi(Z_V)).
        // FP32: the branches as written; the transcendental unit makes a
        // divergent wavefront's second branch cheap.  FP64: both branches share
        // their (software) transcendentals through lane selects.
        Real vn, logM;
        if constexpr (sizeof(Real) == 4) {
            const Real m = v_n * c.ekdt + c.tmf;
            const Real s2 = v_n * c.s2f1 + c.s2f2;
            // psi = s^2 / m^2 enters only through comparisons and the ratios
            // 2/psi and (psi -+ 1)/(psi + 1), formed from s^2 and m^2 without
            // dividing by m.
            if (m <= c.eps) {
                vn = Real(0);
                logM = Real(0);
            } else {
                const Real m2 = m * m;
                if (s2 <= c.eps * m2) {                          // psi <= eps
                    vn = m;
                    logM = Real(0);
                } else if (s2 <= Real(1.5) * m2) {               // psi <= psi_c
                    const Real two_ip = (m2 + m2) * M::rcp(s2);  // 2 / psi
                    const Real b2 = two_ip - Real(1) + M::sqrt(two_ip * (two_ip - Real(1)));
                    const Real a = M::div(m, Real(1) + b2);
                    const Real bz = M::sqrt(b2) + z1;
                    vn = a * bz * bz;
                    // A clamped to 1/(2a) - eps, applied to the product A a
                    const Real Aa = M::fmin(c.A * a, Real(0.5) - c.eps * a);
                    const Real tAa = Real(1) - Real(2) * Aa;
                    // log M = A b^2 a / (1 - 2 A a) - log(1 - 2 A a) / 2
                    logM = M::div(Aa * b2, tAa) - M::half_log(tAa);
                } else {
                    const Real rq = M::rcp(s2 + m2);
                    const Real p = (s2 - m2) * rq;               // (psi - 1) / (psi + 1)
                    const Real beta = (m + m) * rq;              // (1 - p) / m
                    const Real U = M::uniform(w2);
                    if (U <= p) {
                        vn = Real(0);
                    } else {
                        const Real safe_U = M::track_min(U, Real(1) - c.eps);   // both positive
                        vn = M::div(M::log(M::div(Real(1) - p, Real(1) - safe_U)), beta);
                    }
                    // A clamped to beta - eps: beta - A becomes max(beta - A, eps)
                    logM = M::log(p + M::div(beta * (Real(1) - p), M::fmax(beta - c.A, c.eps)));
                }
            }
        } else {
            // Both branches are evaluated by every lane with lane-selected
            // arguments, so a wavefront whose lanes split between the branches
            // (nearly all of them when v often nears zero) issues each
            // transcendental once instead of twice:
            //   quadratic (psi <= 1.5):  x = 2/psi = 2m^2/s^2, b^2 = x - 1 + sqrt(x(x-1)),
            //     a = m/(1+b^2), v' = a (b + Z_V)^2,
            //     log M = A b^2 a/(1-2Aa) - log(1-2Aa)/2  (A a clamped to 1/2 - eps a)
            //   exponential:  p = (psi-1)/(psi+1) = (s^2-m^2)/(s^2+m^2),
            //     beta = (1-p)/m = 2m/(s^2+m^2),
            //     v' = 0 if U <= p else log((1-p)/(1-U)) / beta  (= max(0, log(..)/beta)),
            //     M = p + beta (1-p) / max(beta - A, eps)          (A clamped to beta - eps)
            //   psi <= eps: v' = m, M = 1;  m <= eps: v' = 0, M = 1.
            // Three reciprocals per lane: s^2 or s^2+m^2; 1+b^2 or (1-U) beta,
            // which gives both 1/(1-U) = beta r2 and 1/beta = (1-U) r2; 1-2Aa
            // or max(beta - A, eps).  psi itself is never formed.
            const Real m = v_n * c.ekdt + c.tmf;
            const Real s2 = v_n * c.s2f1 + c.s2f2;
            const bool mpos = m > c.eps;
            const Real ms = mpos ? m : Real(1);
            const Real m2 = ms * ms;
            const bool quad = s2 <= Real(1.5) * m2;
            const Real r1 = M::rcp(quad ? s2 : s2 + m2);
            const Real p = (s2 - m2) * r1;                           // exponential
            const Real omp = Real(1) - p;
            const Real beta = (ms + ms) * r1;
            const Real x = (m2 + m2) * r1;                           // quadratic
            const Real b2 = x - Real(1) + M::sqrt(quad ? x * (x - Real(1)) : Real(0));
            const Real U = M::uniform(w2);
            const Real omU = Real(1) - M::fmin(U, Real(1) - c.eps);
            const Real r2 = M::rcp(quad ? Real(1) + b2 : omU * beta);
            const Real a = ms * r2;
            const Real bz = M::sqrt(quad ? b2 : Real(0)) + z1;
            const Real Aa = M::fmin(c.A * a, Real(0.5) - c.eps * a);
            const Real tAa = Real(1) - Real(2) * Aa;
            const Real r3 = M::rcp(quad ? tAa : M::fmax(beta - c.A, c.eps));
            const Real L1 = M::log(quad ? tAa : omp * (beta * r2));
            Real L2 = Real(0);
            const bool expo_lane = mpos && !quad;
            if (__any(expo_lane)) L2 = M::log(expo_lane ? p + beta * omp * r3 : Real(1));
            const Real vn_quad = a * bz * bz;
            // U <= p: the ratio (1-p)/(1-U) is at most 1, L1 <= 0, and the max gives v' = 0
            const Real vn_exp = M::fmax(L1 * (omU * r2), Real(0));
            const Real logM_quad = Aa * b2 * r3 - Real(0.5) * L1;
            const bool degenerate = s2 <= c.eps * m2;                 // psi <= eps
            const Real vn_sel = !mpos ? Real(0) : degenerate ? m : quad ? vn_quad : vn_exp;
            const Real logM_sel = (!mpos || degenerate) ? Real(0) : quad ? logM_quad : L2;
            vn = vn_sel;
            logM = logM_sel;
        }
        v = vn;
        // K0 + K1 v = -log M - (K3/2) v  (MSchemes.hpp:434-439)
        return c.rqdt - logM + c.cv * v_n + c.K2 * vn + M::sqrt(c.K3 * v_n + c.K4 * vn) * z2;
    }
}

// Extra Bates uniform e_j of (path, step): word (j mod 4) of block slot 1 + j/4.
template <typename Real>
__device__ __forceinline__ Real bates_extra(uint32_t p0, uint32_t p1, uint32_t t, uint32_t j,
                                            uint32_t k0, uint32_t k1) {
    const Words w = philox(p0, p1, t, 1u + (j >> 2), k0, k1);
    const uint32_t x = (j & 3u) == 0 ? w.w0 : (j & 3u) == 1 ? w.w1 : (j & 3u) == 2 ? w.w2 : w.w3;
    return Math<Real>::uniform(x);
}

// Compound-Poisson log-normal jump of one step (MSchemes.hpp:493-527, Knuth's
// product of uniforms).  Returns the log jump (0 when no jump).
template <typename Real>
__device__ __noinline__ Real bates_jump_slow(const LaneC<Real> &c, Real u3, uint32_t p0, uint32_t p1,
                                             uint32_t t, uint32_t k0, uint32_t k1) {
    using M = Math<Real>;
    int N = 0;
    Real p = u3;
    uint32_t j = 0;
    while (p >= c.eml_dt && N < 64) {
        ++N;
        p *= bates_extra<Real>(p0, p1, t, j++, k0, k1);
    }
    const Real e0 = bates_extra<Real>(p0, p1, t, j, k0, k1);
    const Real e1 = bates_extra<Real>(p0, p1, t, j + 1, k0, k1);
    Real s, co;
    M::sincos2pi(e1, s, co);
    const Real z = M::sqrt(M::neg2log(e0)) * co;
    const Real n = Real(N);
    return n * c.mu_j + M::sqrt(n) * c.sigma_j * z;
}

// Vanilla payoff (Payoff.hpp:41-72).
template <typename Real>
__device__ __forceinline__ Real vanilla(int right, Real S, Real K) {
    return right == QK_CALL ? (S > K ? S - K : Real(0)) : (K > S ? K - S : Real(0));
}

// evaluate_final_payoff (Payoff.hpp:157-242).  sum excludes S0; max/min start at S0.
template <typename Real>
__device__ __forceinline__ Real payoff_eval(const PayD<Real> &p, Real S, double sum, Real mx, Real mn,
                                            int n_steps) {
    const Real K = p.K, B = p.B;
    const bool call = p.right == QK_CALL;
    switch (p.type) {
    case QK_EUROPEAN: return vanilla<Real>(p.right, S, K);
    case QK_ASIAN: return vanilla<Real>(p.right, Real(sum / (double)n_steps), K);
    case QK_BARRIER_UP_OUT: return mx < B ? vanilla<Real>(p.right, S, K) : Real(0);
    case QK_BARRIER_UP_IN: return mx >= B ? vanilla<Real>(p.right, S, K) : Real(0);
    case QK_BARRIER_DOWN_OUT: return mn > B ? vanilla<Real>(p.right, S, K) : Real(0);
    case QK_BARRIER_DOWN_IN: return mn <= B ? vanilla<Real>(p.right, S, K) : Real(0);
    case QK_LOOKBACK_FIXED: return call ? vanilla<Real>(QK_CALL, mx, K) : vanilla<Real>(QK_PUT, mn, K);
    case QK_LOOKBACK_FLOATING: return call ? vanilla<Real>(QK_CALL, S, mn) : vanilla<Real>(QK_PUT, S, mx);
    case QK_BINARY_CASH: return call ? (S > K ? Real(1) : Real(0)) : (S < K ? Real(1) : Real(0));
    case QK_BINARY_ASSET: return call ? (S > K ? S : Real(0)) : (S < K ? S : Real(0));
    default: return Real(0);
    }
}

// ---------------------------------------------------------------------------
// Raw buffer access.  A gfx9 buffer resource (base address, no stride,
// num_records in bytes, dword 3 = 0x00020000) is uniform, so the address is
// a 32-bit byte offset in one vector register plus the instruction's offset
// field; a load at an offset >= num_records returns 0 and a store there is
// dropped, so reads past the end of a range need no branch or clamp.
// ---------------------------------------------------------------------------
typedef uint32_t u32x2_t __attribute__((__vector_size__(8)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename Real> __device__ __forceinline__ Real buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, off, 0, 0);
}

// ---------------------------------------------------------------------------
// Wavefront (64-lane) reductions.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Cross-lane exchanges of a 64-lane wavefront without LDS traffic (every lane
// of the wavefront must be active).  DPP moves inside rows of 16 lanes:
// QP_XOR1 / QP_XOR2 take lane l ^ 1 / l ^ 2 (quad permutations), HALF_MIRROR
// lane 7 - l and MIRROR lane 15 - l within each group of 8 / 16.  Across rows,
// gfx950's permlane swaps exchange whole halves of two registers: after
// swap32(a, b), a + b holds a's sum over lanes l, l + 32 in lanes 0..31 and
// b's in lanes 32..63; swap16 does the same for row pairs (0,1) and (2,3).
// ---------------------------------------------------------------------------
enum : int { DPP_QP_XOR1 = 0xB1, DPP_QP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140 };

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <bool ROWS16>
__device__ __forceinline__ void swap_halves_f64(double &a, double &b) {
    const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
    const unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32), blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
    unsigned nlo_a, nlo_b, nhi_a, nhi_b;
    if constexpr (ROWS16) {
        const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
        nlo_a = lo[0]; nlo_b = lo[1]; nhi_a = hi[0]; nhi_b = hi[1];
    } else {
        const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
        nlo_a = lo[0]; nlo_b = lo[1]; nhi_a = hi[0]; nhi_b = hi[1];
    }
    a = __longlong_as_double((long long)(((uint64_t)nhi_a << 32) | nlo_a));
    b = __longlong_as_double((long long)(((uint64_t)nhi_b << 32) | nlo_b));
}

// x + x of the partner lane l ^ 16 (ROWS16) or l ^ 32; the same bits in both.
template <bool ROWS16>
__device__ __forceinline__ double add_across(double x) {
    double a = x, b = x;
    swap_halves_f64<ROWS16>(a, b);
    return a + b;
}

// Sum over the wavefront, the same bits in every lane (each step adds two
// lanes' values, which both lanes of the pair compute in the same order).
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_f64<DPP_QP_XOR1>(x);
    x += dpp_f64<DPP_QP_XOR2>(x);
    x += dpp_f64<DPP_HALF_MIRROR>(x);
    x += dpp_f64<DPP_MIRROR>(x);
    x = add_across<true>(x);
    return add_across<false>(x);
}

// Chan et al. pairwise merge of (n, mean, M2) moments.
__device__ __forceinline__ void chan_merge(double &n_a, double &mean_a, double &m2_a, double n_b,
                                           double mean_b, double m2_b) {
    if (n_b <= 0.0) return;
    if (n_a <= 0.0) {
        n_a = n_b; mean_a = mean_b; m2_a = m2_b;
        return;
    }
    const double n = n_a + n_b;
    const double delta = mean_b - mean_a;
    mean_a = mean_a + delta * (n_b / n);
    m2_a = m2_a + m2_b + delta * delta * (n_a * n_b / n);
    n_a = n;
}

}  // namespace qk
