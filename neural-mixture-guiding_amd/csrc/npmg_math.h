// npmg_math.h -- deterministic single-precision elementary functions.
//
// The guiding path needs bit-identical results on the host and on gfx950 for
// two things: the lobe selected by a sample (a running subtraction of float
// weights lambda_i / sum(lambda), where lambda comes from exp() and, with the
// cosine product, from sin/cos/acos/atan2), and the hash-grid indices. The
// vendor libraries (glibc libm on the host, ocml on the device) do not promise
// identical rounding, so this header provides one implementation of each
// function, written only with IEEE-754 operations that both targets round
// correctly: +, -, *, /, sqrt and fused multiply-add. Every translation unit
// that includes it must be compiled with -ffp-contract=off so that no extra
// fusions are introduced by the compiler.
//
// Methods (all from the mathematics, no vendor constants):
//   exp    Cody & Waite additive range reduction x = k ln2 + r, |r| <= ln2/2,
//          degree-7 Taylor polynomial of e^r, scaling by 2^k via the exponent.
//   log    x = m 2^e with m in [sqrt(2)/2, sqrt(2)); log m = 2 atanh(s),
//          s = (m-1)/(m+1), odd series of atanh to s^11.
//   log1p  Goldberg's correction log1p(x) = log(u) * x / (u - 1), u = 1 + x.
//   expm1  Taylor polynomial to degree 11 on |x| <= 3/4, e^x - 1 beyond.
//   sin/cos  Cody & Waite reduction by pi/2 (three-part constant), Taylor
//          polynomials on [-pi/4, pi/4] (sin to r^11, cos to r^12).
//   atan   reduction |x|>1 -> pi/2 - atan(1/x) and, above tan(pi/12),
//          atan(x) = pi/6 + atan((sqrt(3) x - 1) / (x + sqrt(3))); odd Taylor
//          series to t^15 on |t| <= tan(pi/12).
//   atan2, acos  from atan and sqrt.
// Accuracy is a few ulp; what matters for parity is that host and device run
// the same operation sequence.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define NPMG_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define NPMG_HD static inline
#endif

namespace npmg {
namespace fm {

NPMG_HD float bits_to_float(uint32_t u) { return __builtin_bit_cast(float, u); }
NPMG_HD uint32_t float_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

NPMG_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
NPMG_HD float floor_(float x) { return __builtin_floorf(x); }
NPMG_HD float sqrt_(float x) { return __builtin_sqrtf(x); }
NPMG_HD float abs_(float x) { return __builtin_fabsf(x); }
NPMG_HD float min_(float a, float b) { return a < b ? a : b; }
NPMG_HD float max_(float a, float b) { return a > b ? a : b; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPiO2 = 1.57079632679489661923f;
constexpr float kPiO6 = 0.52359877559829887308f;
constexpr float kLn2Hi = 0.693145751953125f;          // 0x3f317200, 15 significant bits
constexpr float kLn2Lo = 1.428606765330187045e-06f;   // ln2 - kLn2Hi
constexpr float kLog2e = 1.44269504088896340736f;

// x rounded to the nearest IEEE binary16 value (ties to even) and returned
// as a float: what storing a float into a __half and reading it back gives,
// including the binary16 subnormals (quantum 2^-24) and overflow to +-inf at
// and above 65520.
NPMG_HD float round_half(float x) {
    const uint32_t u = float_to_bits(x);
    const uint32_t sign = u & 0x80000000u, a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return x;                                // inf, nan
    if (a >= 0x477ff000u) return bits_to_float(sign | 0x7f800000u);  // |x| >= 65520
    if (a < 0x38800000u) {                                         // |x| < 2^-14: multiples of 2^-24
        const float q = bits_to_float(a) * 16777216.0f;             // exact: |q| < 2^10
        const float r = (q + 12582912.0f) - 12582912.0f;            // round to an integer, ties to even
        return bits_to_float(sign | float_to_bits(r * 5.9604644775390625e-8f));
    }
    const uint32_t b = (a + 0xfffu + ((a >> 13) & 1u)) & 0xffffe000u;  // 10 fraction bits, ties to even
    return bits_to_float(sign | b);
}

NPMG_HD float inf_() { return bits_to_float(0x7f800000u); }
NPMG_HD float nan_() { return bits_to_float(0x7fc00000u); }
NPMG_HD bool isnan_(float x) { return x != x; }
NPMG_HD bool isfinite_(float x) { return (float_to_bits(x) & 0x7f800000u) != 0x7f800000u; }

// 2^k for integer k in [-149, 127], exact.
NPMG_HD float exp2i(int k) {
    if (k >= -126) return bits_to_float((uint32_t)(k + 127) << 23);
    return bits_to_float(1u << (k + 149));
}

NPMG_HD float exp(float x) {
    if (isnan_(x)) return x;
    if (x > 88.7228394f) return inf_();
    if (x < -103.972084f) return 0.0f;
    const float kf = floor_(fma_(x, kLog2e, 0.5f));
    const int k = (int)kf;
    float r = fma_(kf, -kLn2Hi, x);
    r = fma_(kf, -kLn2Lo, r);
    float p = 1.0f / 5040.0f;
    p = fma_(p, r, 1.0f / 720.0f);
    p = fma_(p, r, 1.0f / 120.0f);
    p = fma_(p, r, 1.0f / 24.0f);
    p = fma_(p, r, 1.0f / 6.0f);
    p = fma_(p, r, 0.5f);
    p = fma_(p, r, 1.0f);
    p = fma_(p, r, 1.0f);
    // Split the scaling so that neither factor over- or underflows.
    if (k > 127) return p * 2.0f * exp2i(k - 1);
    if (k < -126) return (p * exp2i(k + 64)) * exp2i(-64);
    return p * exp2i(k);
}

NPMG_HD float log(float x) {
    if (isnan_(x) || x < 0.0f) return nan_();
    if (x == 0.0f) return -inf_();
    if (!isfinite_(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        e = -23;
    }
    const uint32_t ix = float_to_bits(x);
    e += (int)((ix >> 23) & 0xffu) - 127;
    float m = bits_to_float((ix & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float q = 1.0f / 11.0f;
    q = fma_(q, s2, 1.0f / 9.0f);
    q = fma_(q, s2, 1.0f / 7.0f);
    q = fma_(q, s2, 1.0f / 5.0f);
    q = fma_(q, s2, 1.0f / 3.0f);
    const float two_s = 2.0f * s;
    const float lm = fma_(two_s * s2, q, two_s);
    const float ef = (float)e;
    return fma_(ef, kLn2Hi, fma_(ef, kLn2Lo, lm));
}

NPMG_HD float log1p(float x) {
    const float u = 1.0f + x;
    if (u == 1.0f) return x;
    if (!isfinite_(u)) return log(u);
    return log(u) * (x / (u - 1.0f));
}

NPMG_HD float expm1(float x) {
    if (isnan_(x)) return x;
    if (x > 88.7228394f) return inf_();
    if (abs_(x) <= 0.75f) {
        // x (1 + x/2 + x^2/6 + ... + x^10/11!), Taylor to degree 11
        float p = 1.0f / 39916800.0f;
        p = fma_(p, x, 1.0f / 3628800.0f);
        p = fma_(p, x, 1.0f / 362880.0f);
        p = fma_(p, x, 1.0f / 40320.0f);
        p = fma_(p, x, 1.0f / 5040.0f);
        p = fma_(p, x, 1.0f / 720.0f);
        p = fma_(p, x, 1.0f / 120.0f);
        p = fma_(p, x, 1.0f / 24.0f);
        p = fma_(p, x, 1.0f / 6.0f);
        p = fma_(p, x, 0.5f);
        return fma_(p * x, x, x);
    }
    return exp(x) - 1.0f;
}

NPMG_HD void sincos(float x, float* s_out, float* c_out) {
    // pi/2 split into three floats: hi + md + lo.
    const float kPio2Hi = 1.57079625129699707031f;   // 0x3fc90fda
    const float kPio2Md = 7.549790126404332e-08f;    // float(pi/2 - hi)
    const float kPio2Lo = -1.7151245100058819e-15f;  // float(pi/2 - hi - md)
    const float qf = floor_(fma_(x, 0.636619772367581343076f, 0.5f));
    float r = fma_(qf, -kPio2Hi, x);
    r = fma_(qf, -kPio2Md, r);
    r = fma_(qf, -kPio2Lo, r);
    const float r2 = r * r;
    float ps = -1.0f / 39916800.0f;
    ps = fma_(ps, r2, 1.0f / 362880.0f);
    ps = fma_(ps, r2, -1.0f / 5040.0f);
    ps = fma_(ps, r2, 1.0f / 120.0f);
    ps = fma_(ps, r2, -1.0f / 6.0f);
    const float s = fma_(r * r2, ps, r);
    float pc = 1.0f / 479001600.0f;
    pc = fma_(pc, r2, -1.0f / 3628800.0f);
    pc = fma_(pc, r2, 1.0f / 40320.0f);
    pc = fma_(pc, r2, -1.0f / 720.0f);
    pc = fma_(pc, r2, 1.0f / 24.0f);
    pc = fma_(pc, r2, -0.5f);
    const float c = fma_(r2, pc, 1.0f);
    const int q = ((int)qf) & 3;
    if (q == 0) { *s_out = s; *c_out = c; }
    else if (q == 1) { *s_out = c; *c_out = -s; }
    else if (q == 2) { *s_out = -s; *c_out = -c; }
    else { *s_out = -c; *c_out = s; }
}

NPMG_HD float sin(float x) { float s, c; sincos(x, &s, &c); return s; }
NPMG_HD float cos(float x) { float s, c; sincos(x, &s, &c); return c; }

// atan on [0, inf), result in [0, pi/2].
NPMG_HD float atan_pos(float a) {
    const bool invert = a > 1.0f;
    float t = invert ? 1.0f / a : a;
    float off = 0.0f;
    if (t > 0.267949192431122706f) {  // tan(pi/12)
        const float kSqrt3 = 1.73205080756887729353f;
        t = fma_(t, kSqrt3, -1.0f) / (t + kSqrt3);
        off = kPiO6;
    }
    const float t2 = t * t;
    float p = -1.0f / 15.0f;
    p = fma_(p, t2, 1.0f / 13.0f);
    p = fma_(p, t2, -1.0f / 11.0f);
    p = fma_(p, t2, 1.0f / 9.0f);
    p = fma_(p, t2, -1.0f / 7.0f);
    p = fma_(p, t2, 1.0f / 5.0f);
    p = fma_(p, t2, -1.0f / 3.0f);
    float res = off + fma_(t * t2, p, t);
    if (invert) res = kPiO2 - res;
    return res;
}

NPMG_HD float atan(float x) {
    if (isnan_(x)) return x;
    const float r = atan_pos(abs_(x));
    return x < 0.0f ? -r : r;
}

NPMG_HD float atan2(float y, float x) {
    if (isnan_(x) || isnan_(y)) return nan_();
    const float ax = abs_(x), ay = abs_(y);
    float a;
    if (ax == 0.0f && ay == 0.0f) {
        a = 0.0f;
    } else if (ay <= ax) {
        a = atan_pos(ay / ax);
    } else {
        a = kPiO2 - atan_pos(ax / ay);
    }
    const bool x_neg = (float_to_bits(x) >> 31) != 0u;
    const bool y_neg = (float_to_bits(y) >> 31) != 0u;
    if (x_neg) a = kPi - a;
    return y_neg ? -a : a;
}

NPMG_HD float acos(float x) {
    // acos(x) = atan2(sqrt((1 - x)(1 + x)), x); NaN outside [-1, 1].
    return atan2(sqrt_((1.0f - x) * (1.0f + x)), x);
}

NPMG_HD float logistic(float x) { return 1.0f / (1.0f + exp(-x)); }

}  // namespace fm
}  // namespace npmg
