// device.h -- device building blocks of the guiding path: pcg32, hash-grid
// corner indexing, OneBlob/SH encodings, and the vMF mixture (map, cosine
// product, pdf, sample, adjoint).
//
// Floating-point expressions are written in a fixed order and the library is
// built with -ffp-contract=off; explicit fma marks the accumulations that the
// reference's nvcc build contracts. Together with npmg_math.h this makes the
// FP32 path reproduce the CPU oracle bit for bit (indices, lobe selection,
// encoded features, network outputs and mixture values).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "config.h"
#include "kernels.h"
#include "npmg_math.h"

namespace npmg {
namespace dev {

namespace fm = ::npmg::fm;

// Constants of krrmath (src/backend/src/core/math/include/krrmath/constants.h:15-25).
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647693f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kEps = 1e-5f;  // M_EPSILON

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float a, float b, float c) { return {a, b, c}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return fm::sqrt_(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, length(a)); }

// Math policies of the mixture stage. ExactMath is the library's own
// deterministic FP32 functions (npmg_math.h), bit-identical to the oracle's
// FP32 instantiation. FastMath ("numerics": "fast") uses the hardware
// transcendentals (v_exp_f32 / v_log_f32 base 2, v_sin_f32 / v_cos_f32 in
// revolutions, v_rcp_f32, v_sqrt_f32: about 1 ulp each) for exp, log,
// sin/cos, division and sqrt; expm1 is a Taylor polynomial below |x| = 1/2
// (relative error below 3e-7) and the hardware exponential minus 1 beyond;
// log1p stays exact.
struct ExactMath {
    static __device__ __forceinline__ float exp(float x) { return fm::exp(x); }
    static __device__ __forceinline__ float log(float x) { return fm::log(x); }
    static __device__ __forceinline__ float expm1(float x) { return fm::expm1(x); }
    static __device__ __forceinline__ float log1p(float x) { return fm::log1p(x); }
    static __device__ __forceinline__ void sincos(float x, float* s, float* c) { fm::sincos(x, s, c); }
    static __device__ __forceinline__ float sqrt(float x) { return fm::sqrt_(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a / b; }
};
struct FastMath {
    static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
    static __device__ __forceinline__ float expm1(float x) {
        float p = __builtin_fmaf(x, 1.0f / 5040.0f, 1.0f / 720.0f);
        p = __builtin_fmaf(p, x, 1.0f / 120.0f);
        p = __builtin_fmaf(p, x, 1.0f / 24.0f);
        p = __builtin_fmaf(p, x, 1.0f / 6.0f);
        p = __builtin_fmaf(p, x, 0.5f);
        p = __builtin_fmaf(p, x, 1.0f);
        const float e = __builtin_amdgcn_exp2f(x * 1.44269504088896341f) - 1.0f;
        return __builtin_fabsf(x) < 0.5f ? p * x : e;
    }
    static __device__ __forceinline__ float log1p(float x) { return fm::log1p(x); }
    static __device__ __forceinline__ void sincos(float x, float* s, float* c) {
        const float r = x * 0.159154943091895336f;  // revolutions
        *s = __builtin_amdgcn_sinf(r);
        *c = __builtin_amdgcn_cosf(r);
    }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};
template <class MT>
__device__ __forceinline__ V3 normalize_m(V3 a) {
    if constexpr (__is_same(MT, ExactMath)) return normalize(a);
    else return mul(a, __builtin_amdgcn_rsqf(dot(a, a)));
}
// std::min / std::max semantics (second argument wins only when strictly smaller / larger).
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float clampv(float v, float lo, float hi) { return smin(smax(v, lo), hi); }

// math_utils.h:191-201
template <class MT = ExactMath>
__device__ __forceinline__ V3 sph_to_cart(float theta, float phi) {
    float st, ct, sp, cp;
    MT::sincos(theta, &st, &ct);
    MT::sincos(phi, &sp, &cp);
    return {st * cp, st * sp, ct};
}
// math_utils.h:176-183
__device__ __forceinline__ void cart_to_sph(V3 v, float* theta, float* phi) {
    const V3 vn = normalize(v);
    *theta = fm::acos(vn.z);
    float p = fm::atan2(vn.y, vn.x);
    if (p < 0) p += kTwoPi;
    *phi = p;
}
// math_utils.h:141-151
template <class MT = ExactMath>
__device__ __forceinline__ V3 get_perpendicular(V3 u) {
    const float ax = fm::abs_(u.x), ay = fm::abs_(u.y), az = fm::abs_(u.z);
    const uint32_t uyx = (ax - ay) < 0 ? 1u : 0u;
    const uint32_t uzx = (ax - az) < 0 ? 1u : 0u;
    const uint32_t uzy = (ay - az) < 0 ? 1u : 0u;
    const uint32_t xm = uyx & uzx;
    const uint32_t ym = (1u ^ xm) & uzy;
    const uint32_t zm = 1u ^ (xm | ym);
    return normalize_m<MT>(cross(u, mk((float)xm, (float)ym, (float)zm)));
}
// raytracing.h:35-56: Frame(n).toWorld(v)
template <class MT = ExactMath>
__device__ __forceinline__ V3 frame_to_world(V3 n, V3 v) {
    const V3 t = get_perpendicular<MT>(n);
    const V3 b = normalize_m<MT>(cross(n, t));
    return add(add(mul(t, v.x), mul(b, v.y)), mul(n, v.z));
}
// sampling.h:32-37
template <class MT = ExactMath>
__device__ __forceinline__ V3 uniform_sample_sphere(float u0, float u1) {
    const float z = 1.0f - 2.0f * u0;
    const float r = MT::sqrt(smax(0.0f, 1.0f - z * z));
    float s, c;
    MT::sincos(kTwoPi * u1, &s, &c);
    return {r * c, r * s, z};
}

// ---------------------------------------------------------------------------
// PCG32 (O'Neill 2014) with a per-query stream, for synthetic records.
struct Pcg32 {
    uint64_t state, inc;
    __device__ __forceinline__ Pcg32(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    __device__ __forceinline__ uint32_t next_uint() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    __device__ __forceinline__ float next_float() {
        return fm::bits_to_float((next_uint() >> 9) | 0x3f800000u) - 1.0f;
    }
};

// ---------------------------------------------------------------------------
// Hash grid (tiny-cuda-nn encodings/grid.h:111-187, common_device.h:470-496).
__device__ __forceinline__ uint32_t grid_index(const DevCfg& c, uint32_t hsize, uint32_t res, uint32_t p0, uint32_t p1,
                                               uint32_t p2) {
    uint32_t stride = 1, index = 0;
    const uint32_t p[3] = {p0, p1, p2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (stride <= hsize) {
            index += p[d] * stride;
            stride *= res;
        }
    }
    if (c.grid_type == GRID_HASH && hsize < stride) {
        uint32_t f0 = 1u, f1 = 2654435761u, f2 = 805459861u;
        if (c.hash_type == HASH_PRIME) f0 = 1958374283u;
        else if (c.hash_type == HASH_REVERSED_PRIME) { f0 = 2165219737u; f1 = 1434869437u; f2 = 2097192037u; }
        index = (p0 * f0) ^ (p1 * f1) ^ (p2 * f2);
    }
    // index % hsize without a 32-bit division: hashed levels have power-of-two
    // sizes; a dense index is below 2 hsize (cells reach resolution + 1 on the
    // cube's faces), so one conditional subtraction is exact.
    if ((hsize & (hsize - 1u)) == 0u) return index & (hsize - 1u);
    if (index < hsize) return index;
    if (index - hsize < hsize) return index - hsize;
    return index % hsize;
}

// Fractional position and cell of one coordinate; the reference's nvcc build
// contracts input * scale + 0.5 into one fma.
__device__ __forceinline__ void pos_fract(float x, float scale, int interp, float* frac, uint32_t* cell) {
    const float p = fm::fma_(x, scale, 0.5f);
    const float fl = fm::floor_(p);
    *cell = (uint32_t)(int)fl;
    float f = p - fl;
    if (interp == INTERP_SMOOTHSTEP) f = f * f * (3.0f - 2.0f * f);
    *frac = f;
}

// The 8 corners of one level: index relative to the level and weight.
// Weights are products over dims in order x, y, z of (1 - frac) or frac.
__device__ __forceinline__ int level_corners(const DevCfg& c, int l, float x0, float x1, float x2, uint32_t idx[8],
                                             float w[8]) {
    const uint32_t hsize = c.offsets[l + 1] - c.offsets[l];
    const float scale = c.scales[l];
    const uint32_t res = c.res[l];
    float f0, f1, f2;
    uint32_t c0, c1, c2;
    pos_fract(x0, scale, c.interp, &f0, &c0);
    pos_fract(x1, scale, c.interp, &f1, &c1);
    pos_fract(x2, scale, c.interp, &f2, &c2);
    if (c.interp == INTERP_NEAREST) {
        idx[0] = grid_index(c, hsize, res, c0, c1, c2);
        w[0] = 1.0f;
        return 1;
    }
    // grid_index of the 8 corners with its decisions taken once per level
    // (uniform): the stride walk (dims that fit the level), dense or hashed,
    // and the hash products formed incrementally, (p + 1) m = p m + m in
    // uint32 arithmetic, so every index is bit-identical to grid_index's.
    uint32_t stride = 1u, s1 = 0u, s2 = 0u;
    if (stride <= hsize) stride *= res;
    if (stride <= hsize) {
        s1 = stride;
        stride *= res;
    }
    if (s1 != 0u && stride <= hsize) {
        s2 = stride;
        stride *= res;
    }
    const bool pow2 = (hsize & (hsize - 1u)) == 0u;
    if (c.grid_type == GRID_HASH && hsize < stride) {
        uint32_t m0 = 1u, m1 = 2654435761u, m2 = 805459861u;
        if (c.hash_type == HASH_PRIME) m0 = 1958374283u;
        else if (c.hash_type == HASH_REVERSED_PRIME) {
            m0 = 2165219737u;
            m1 = 1434869437u;
            m2 = 2097192037u;
        }
        const uint32_t hx0 = c0 * m0, hx1 = hx0 + m0, hy0 = c1 * m1, hy1 = hy0 + m1, hz0 = c2 * m2, hz1 = hz0 + m2;
        const uint32_t yz[4] = {hy0 ^ hz0, hy1 ^ hz0, hy0 ^ hz1, hy1 ^ hz1};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t hv = ((k & 1) ? hx1 : hx0) ^ yz[k >> 1];
            idx[k] = pow2 ? (hv & (hsize - 1u)) : (hv % hsize);
        }
    } else {
        const uint32_t base = c0 + c1 * s1 + c2 * s2;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t i = base + ((k & 1) ? 1u : 0u) + ((k & 2) ? s1 : 0u) + ((k & 4) ? s2 : 0u);
            uint32_t r;
            if (pow2) r = i & (hsize - 1u);
            else if (i < hsize) r = i;
            else if (i - hsize < hsize) r = i - hsize;
            else r = i % hsize;
            idx[k] = r;
        }
    }
    const float a0 = 1.0f - f0, a1 = 1.0f - f1, a2 = 1.0f - f2;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float wk = 1.0f;
        wk *= (k & 1) ? f0 : a0;
        wk *= (k & 2) ? f1 : a1;
        wk *= (k & 4) ? f2 : a2;
        w[k] = wk;
    }
    return 8;
}

// The F features of one table entry. An entry is F * sizeof(T) bytes at a
// multiple of its size from the table start, so when that size is 4, 8 or 16
// bytes and the table start is aligned to it (`vec`), one vector load fetches
// the entry: one gather per corner instead of one per feature.
template <class T, int F>
__device__ __forceinline__ void load_entry(const T* __restrict__ tl, uint32_t idx, bool vec, T (&v)[F]) {
    constexpr int B = F * (int)sizeof(T);
    if constexpr (B == 4 || B == 8 || B == 16) {
        using U = std::conditional_t<B == 4, uint32_t, std::conditional_t<B == 8, uint2, uint4>>;
        if (vec) {
            const U u = *reinterpret_cast<const U*>(tl + (size_t)idx * F);
            __builtin_memcpy(&v[0], &u, B);
            return;
        }
    }
#pragma unroll
    for (int f = 0; f < F; ++f) v[f] = tl[(size_t)idx * F + f];
}

// Whether load_entry may use one vector load on this table (a uniform test).
template <class T, int F>
__device__ __forceinline__ bool entry_aligned(const T* table) {
    return (reinterpret_cast<uintptr_t>(table) & (uintptr_t)(F * sizeof(T) - 1)) == 0;
}

// The entries of the 8 corners of one level (corner k at idx[k]; bit 0 of k
// is the x offset). Corners 2m and 2m + 1 differ only in x, and their entries
// usually share an aligned 16-byte group: dense rows put x + 1 at idx + 1, and
// the coherent-prime hash (x * 1 ^ ...) keeps x + 1 within the group unless
// the increment carries past its low bits. So when an entry is 4 bytes and
// the table start is 16-byte aligned (`g16`; level offsets are multiples
// of 8 entries), one 16-byte load of corner 2m's group also yields corner
// 2m + 1 for most lanes; the others load it on their own. Each gather costs a
// cache-line lookup, so this saves about a third of them (same values, same
// order of accumulation afterwards). Only for 4-byte entries (16-bit tables,
// F = 2): with FP32 entries the gathers of the FP32 paths, which read every
// level of a 52 MB table per query, are bound by the bytes they move, and a
// 16-byte group doubles those (measured slower).
template <class T, int F>
__device__ __forceinline__ void gather_corners(const T* __restrict__ tl, const uint32_t (&idx)[8], bool g16, bool vec,
                                               T (&v)[8][F]) {
    constexpr int B = F * (int)sizeof(T);
    if constexpr (B == 4) {
        if (g16) {
            constexpr uint32_t G = 4u;  // entries per group
            using E = uint32_t;
            const E* te = reinterpret_cast<const E*>(tl);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t i0 = idx[2 * m], i1 = idx[2 * m + 1];
                const uint4 p = *reinterpret_cast<const uint4*>(te + (i0 & ~(G - 1u)));
                auto pick = [&](uint32_t i) {
                    const uint32_t lo = (i & 1u) ? p.y : p.x, hi = (i & 1u) ? p.w : p.z;
                    return (i & 2u) ? hi : lo;
                };
                const E e0 = pick(i0);
                E e1;
                if ((i0 ^ i1) < G) e1 = pick(i1);
                else e1 = te[i1];
                __builtin_memcpy(&v[2 * m][0], &e0, B);
                __builtin_memcpy(&v[2 * m + 1][0], &e1, B);
            }
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) load_entry<T, F>(tl, idx[k], vec, v[k]);
}

// Whether gather_corners may use 16-byte group loads on this table (uniform).
template <class T>
__device__ __forceinline__ bool group_aligned(const T* table) {
    return (reinterpret_cast<uintptr_t>(table) & (uintptr_t)15) == 0;
}

// tiny-cuda-nn common_device.h:533-548. The kernel's CDF is monotonic and
// reaches its clamps; beyond |u| >= 1.0625 the clamped result is exactly 0 or
// 1, so those evaluations are skipped (bit-identical).
__device__ __forceinline__ float quartic_poly(float u) {
    const float u2 = u * u;
    const float u4 = u2 * u2;
    const float v = 15.0f / 16.0f * u * (1.0f - 2.0f / 3.0f * u2 + 1.0f / 5.0f * u4) + 0.5f;
    return smax(0.0f, smin(1.0f, v));
}
__device__ __forceinline__ float quartic_cdf(float x, float inv_radius) {
    const float u = x * inv_radius;
    if (u <= -1.0625f) return 0.0f;
    if (u >= 1.0625f) return 1.0f;
    return quartic_poly(u);
}

// One OneBlob bin edge (oneblob.h:114-123): the sum of the three wrapped
// quartic CDFs cdf(d) + cdf(d - 1) + cdf(d + 1), d = edge - x, in that order.
// The three arguments are n_bins apart in u, so for n_bins >= 4 at most one
// of them lies inside the support |u| < 1.0625: the polynomial is evaluated
// once, at that argument, and the other two take their exact clamp values
// (bit-identical to three quartic_cdf calls).
__device__ __forceinline__ float oneblob_edge(float d, float fnb) {
    if (fnb < 4.0f)
        return quartic_cdf(d, fnb) + quartic_cdf(d - 1.0f, fnb) + quartic_cdf(d + 1.0f, fnb);
    const float u1 = d * fnb, u2 = (d - 1.0f) * fnb, u3 = (d + 1.0f) * fnb;
    const bool in1 = u1 > -1.0625f && u1 < 1.0625f;
    const bool in2 = u2 > -1.0625f && u2 < 1.0625f;
    const bool in3 = u3 > -1.0625f && u3 < 1.0625f;
    const float p = quartic_poly(in1 ? u1 : (in2 ? u2 : u3));
    const float c1 = in1 ? p : (u1 <= -1.0625f ? 0.0f : 1.0f);
    const float c2 = in2 ? p : (u2 <= -1.0625f ? 0.0f : 1.0f);
    const float c3 = in3 ? p : (u3 <= -1.0625f ? 0.0f : 1.0f);
    return c1 + c2 + c3;
}

// q_in[d] for a runtime d without indexing private memory (keeps q_in in
// VGPRs). Integer masks rather than `d == k ? q_in[k] : v`, which the
// compiler folds back into an indexed scratch load.
__device__ __forceinline__ float pick(const float* q_in, int d) {
    uint32_t v = 0u;
#pragma unroll
    for (int k = 0; k < 16; ++k) v |= fm::float_to_bits(q_in[k]) & (0u - (uint32_t)(d == k));
    return fm::bits_to_float(v);
}

// Feeds every encoded feature of one query, in order, to `feed(value)`:
// grid levels (tiny-cuda-nn kernel_grid), OneBlob bins (kernel_one_blob_soa,
// oneblob.h:98-127), SH (spherical_harmonics.h:45-63), identity, then the
// 1.0 padding of the composite to a multiple of 16.
template <class T, class Feed>
__device__ __forceinline__ void encode_query(const DevCfg& c, const T* __restrict__ table, const float* q_in,
                                             Feed&& feed) {
    for (int e = 0; e < c.n_enc; ++e) {
        const int b = c.enc_begin[e];
        const int type = c.enc_type[e];
        if (type == ENC_GRID) {
            const float x0 = pick(q_in, b), x1 = pick(q_in, b + 1), x2 = pick(q_in, b + 2);
            for (int l = 0; l < c.n_levels; ++l) {
                uint32_t idx[8];
                float w[8];
                const int nc = level_corners(c, l, x0, x1, x2, idx, w);
                const T* tl = table + (size_t)c.offsets[l] * c.F;
                if (c.F == 2) {
                    float a0 = 0.0f, a1 = 0.0f;
                    if (nc == 1) {
                        a0 = (float)tl[idx[0] * 2u];
                        a1 = (float)tl[idx[0] * 2u + 1u];
                    } else {
                        T v[8][2];
                        gather_corners<T, 2>(tl, idx, group_aligned(table), entry_aligned<T, 2>(table), v);
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            a0 = fm::fma_(w[k], (float)v[k][0], a0);
                            a1 = fm::fma_(w[k], (float)v[k][1], a1);
                        }
                    }
                    feed(a0);
                    feed(a1);
                } else {
                    for (int f = 0; f < c.F; ++f) {
                        float a = 0.0f;
                        if (nc == 1) {
                            a = (float)tl[idx[0] * (uint32_t)c.F + f];
                        } else {
#pragma unroll
                            for (int k = 0; k < 8; ++k) a = fm::fma_(w[k], (float)tl[idx[k] * (uint32_t)c.F + f], a);
                        }
                        feed(a);
                    }
                }
            }
        } else if (type == ENC_ONEBLOB) {
            const int nb = c.enc_param[e];
            const float fnb = (float)nb;
            const float inv_nb = 1.0f / fnb;  // n_bins is a power of two: k * inv_nb == k / n_bins exactly
            for (int d = 0; d < c.enc_ndims[e]; ++d) {
                const float xv = pick(q_in, b + d);
                float left = oneblob_edge(-xv, fnb);
                for (int k = 0; k < nb; ++k) {
                    const float right = oneblob_edge((float)(k + 1) * inv_nb - xv, fnb);
                    feed(right - left);
                    left = right;
                }
            }
        } else if (type == ENC_SH) {
            const float sx = pick(q_in, b) * 2.0f - 1.0f, sy = pick(q_in, b + 1) * 2.0f - 1.0f,
                        sz = pick(q_in, b + 2) * 2.0f - 1.0f;
            feed(0.28209479177387814f);
            if (c.enc_param[e] > 1) {
                feed(-0.48860251190291987f * sy);
                feed(0.48860251190291987f * sz);
                feed(-0.48860251190291987f * sx);
            }
        } else {
            for (int d = 0; d < c.enc_ndims[e]; ++d) feed(pick(q_in, b + d));
        }
    }
    for (int j = c.enc_width; j < c.enc_padded; ++j) feed(1.0f);
}

// ---------------------------------------------------------------------------
// vMF mixture (src/render/guided/distribution.h, src/backend/src/util/vmf.h).
template <int K>
struct Mixture {
    float lambda[K], kappa[K], norm[K], lognorm[K], weight[K];
    V3 mu[K];
    float total;
};

// The vMF normaliser with the e^{-kappa} of the exponent folded in,
// C(k) = k e^{-k} / (4 pi sinh k) = k / (2 pi (-expm1(-2 k))), exactly 1/4pi at
// k = 0; the pdf is C(k) exp(k (mu.w - 1)), evaluated as exp(k (mu.w - 1) + log C).
template <class MT = ExactMath>
__device__ __forceinline__ float stable_norm(float k) {
    if (!(k > 1e-30f)) return kInv4Pi;
    return MT::div(k * kInv2Pi, -MT::expm1(-2.0f * k));
}
// Reference normaliser (vmf.h:22-27 and distribution.h:40-42) with its 1/4pi threshold.
__device__ __forceinline__ float ref_norm(float k, float thresh) {
    return k < thresh ? kInv4Pi : k * kInv2Pi / (1.0f - fm::exp(-2.0f * k));
}

// network_to_params (train.h:199-210) + constructor (distribution.h:85-98).
// x points at the first raw output of the query; `stride` is the distance
// between consecutive outputs. Mapping "softmax" reads (theta, phi, log
// kappa, logit) per lobe.
template <int K, class MT = ExactMath, class Raw>
__device__ __forceinline__ void build_mixture(const DevCfg& c, Raw&& x, Mixture<K>& m) {
    m.total = 0.0f;
    float theta[K], phi[K];
    if (c.mapping == MAP_REFERENCE) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            m.lambda[i] = MT::exp(clampv(x(4 * i + 0), -10.0f, 15.0f));
            m.kappa[i] = MT::exp(clampv(x(4 * i + 1), -10.0f, 15.0f));
            theta[i] = kPi * MT::div(1.0f, 1.0f + MT::exp(-x(4 * i + 2)));
            phi[i] = kTwoPi * MT::div(1.0f, 1.0f + MT::exp(-x(4 * i + 3)));
            m.total += m.lambda[i];
        }
    } else {
        float mx = x(3);
#pragma unroll
        for (int i = 1; i < K; ++i) mx = smax(mx, x(4 * i + 3));
#pragma unroll
        for (int i = 0; i < K; ++i) {
            theta[i] = kPi * MT::div(1.0f, 1.0f + MT::exp(-x(4 * i + 0)));
            phi[i] = kTwoPi * MT::div(1.0f, 1.0f + MT::exp(-x(4 * i + 1)));
            m.kappa[i] = clampv(MT::exp(x(4 * i + 2)), 1e-2f, 1e3f);
            m.lambda[i] = MT::exp(x(4 * i + 3) - mx);
            m.total += m.lambda[i];
        }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
        m.weight[i] = MT::div(m.lambda[i], m.total);
        m.mu[i] = sph_to_cart<MT>(theta[i], phi[i]);
        m.norm[i] = stable_norm<MT>(m.kappa[i]);
        m.lognorm[i] = MT::log(m.norm[i]);
    }
}

// applyCosineLobe (distribution.h:184-195) with VMFKernel::product (:39-64).
// Reference numerics take each new mean back through spherical coordinates
// like the reference; stable numerics keep the normalised cartesian mean.
template <int K, class MT = ExactMath>
__device__ __forceinline__ void apply_cosine(const DevCfg& c, V3 nrm, Mixture<K>& m, bool want_angles,
                                             float (&theta_out)[K], float (&phi_out)[K]) {
    const bool ref = c.numerics == NUM_REFERENCE;
    const float kb = c.cos_kappa;
    V3 mu_b;
    if constexpr (__is_same(MT, ExactMath)) {
        // the reference's round trip of the normal through (theta, phi)
        float nt, np;
        cart_to_sph(nrm, &nt, &np);
        mu_b = sph_to_cart<MT>(nt, np);
    } else {
        mu_b = normalize_m<MT>(nrm);
    }
    const float norm_b = ref ? ref_norm(kb, 1e-3f) : stable_norm<MT>(kb);
    m.total = 0.0f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const V3 mu_a = m.mu[i];
        const float ka = m.kappa[i];
        V3 nm = add(mul(mu_a, ka), mul(mu_b, kb));
        float nk = MT::sqrt(dot(nm, nm));
        if (nk >= 1e-3f) {
            if constexpr (__is_same(MT, ExactMath)) nm = divs(nm, nk);
            else nm = mul(nm, MT::div(1.0f, nk));
        } else {
            nk = 0.0f;
            nm = mu_a;
        }
        V3 mu_new = nm;
        if (ref || want_angles) {
            float th, ph;
            cart_to_sph(nm, &th, &ph);
            if (ref) mu_new = sph_to_cart(th, ph);
            theta_out[i] = th;
            phi_out[i] = ph;
        }
        const float e = MT::exp(ka * (dot(nm, mu_a) - 1.0f) + kb * (dot(nm, mu_b) - 1.0f));
        const float na = ref ? ref_norm(ka, 1e-3f) : m.norm[i];
        const float cn = stable_norm<MT>(nk);
        const float nn = ref ? ref_norm(nk, 1e-3f) : cn;
        const float s = MT::div(na * norm_b, nn) * e;
        m.lambda[i] = s * m.lambda[i] * 1.0f;
        m.kappa[i] = nk;
        m.mu[i] = mu_new;
        m.norm[i] = cn;
        m.lognorm[i] = MT::log(cn);
        m.total += m.lambda[i];
    }
    m.total = smax(m.total, kEps);
#pragma unroll
    for (int i = 0; i < K; ++i) m.weight[i] = MT::div(m.lambda[i], m.total);
}

// One lobe's pdf at unit direction w (vmf.h:22-27, or the log-space form).
template <int K, class MT = ExactMath>
__device__ __forceinline__ float lobe_pdf(const DevCfg& c, const Mixture<K>& m, int i, V3 w) {
    const float cth = dot(w, m.mu[i]);
    const float d = smin(0.0f, cth - 1.0f);
    if (c.numerics == NUM_REFERENCE) {
        const float k = m.kappa[i];
        if (k < kEps) return kInv4Pi;
        return fm::exp(k * d) * k / (kTwoPi * (1.0f - fm::exp(-2.0f * k)));
    }
    return MT::exp(fm::fma_(m.kappa[i], d, m.lognorm[i]));
}

// distribution.h:113-119
template <int K, class MT = ExactMath>
__device__ __forceinline__ float mixture_pdf(const DevCfg& c, const Mixture<K>& m, V3 w) {
    float p = 0.0f;
#pragma unroll
    for (int i = 0; i < K; ++i) p = fm::fma_(m.weight[i], lobe_pdf<K, MT>(c, m, i, w), p);
    return p;
}

// distribution.h:132-140 (lobe choice by running subtraction, fallback lobe 0)
// and vmf.h:48-71 (inverse-CDF sample about +z rotated into Frame(mu)).
template <int K, class MT = ExactMath>
__device__ __forceinline__ V3 mixture_sample(const DevCfg& c, const Mixture<K>& m, float u0, float u1, float u2,
                                             uint32_t* lobe) {
    int sel = -1;
    float u = u0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (sel < 0) {
            if (u < m.weight[i]) sel = i;
            else u -= m.weight[i];
        }
    }
    if (sel < 0) sel = 0;
    *lobe = (uint32_t)sel;
    // Bitwise selects (not `if (i == sel)`, which the compiler folds back into
    // an indexed load that forces the lobe arrays out of registers).
    uint32_t kb = fm::float_to_bits(m.kappa[0]), xb = fm::float_to_bits(m.mu[0].x),
             yb = fm::float_to_bits(m.mu[0].y), zb = fm::float_to_bits(m.mu[0].z);
#pragma unroll
    for (int i = 1; i < K; ++i) {
        const uint32_t mask = 0u - (uint32_t)(i == sel);
        kb = (kb & ~mask) | (fm::float_to_bits(m.kappa[i]) & mask);
        xb = (xb & ~mask) | (fm::float_to_bits(m.mu[i].x) & mask);
        yb = (yb & ~mask) | (fm::float_to_bits(m.mu[i].y) & mask);
        zb = (zb & ~mask) | (fm::float_to_bits(m.mu[i].z) & mask);
    }
    const float kap = fm::bits_to_float(kb);
    const V3 mu = mk(fm::bits_to_float(xb), fm::bits_to_float(yb), fm::bits_to_float(zb));
    const bool ref = c.numerics == NUM_REFERENCE;
    V3 loc;
    if (ref ? kap < kEps : !(kap > 1e-30f)) {
        loc = uniform_sample_sphere<MT>(u1, u2);
    } else {
        float ct;
        if (ref) ct = 1.0f + fm::log1p(-u1 + fm::exp(-2.0f * kap) * u1) / kap;
        else ct = 1.0f + MT::div(fm::log1p(u1 * fm::expm1(-2.0f * kap)), kap);
        const float st = MT::sqrt(smax(0.0f, 1.0f - ct * ct));
        float sp, cp;
        MT::sincos(kTwoPi * u2, &sp, &cp);
        loc = {cp * st, sp * st, ct};
    }
    return frame_to_world<MT>(mu, loc);
}

// The consumer side of one query (integrator.cpp:197-201, 256-260, 297-317):
// the mixture from raw outputs x, the cosine product when io.normal is set,
// one sample with its pdf and the pdfs of io.n_eval directions, and the lobe
// parameters when io.lobes_out is set (LOBES: the consumer kernel only).
template <int K, class MT, bool LOBES, class Raw>
__device__ __forceinline__ void vmf_query(const DevCfg& c, Raw&& x, const QueryIO& io, uint32_t q, uint32_t n) {
    Mixture<K> m;
    build_mixture<K, MT>(c, x, m);
    float th[K], ph[K];
    if (io.normal) {
        const V3 nrm = mk(io.normal[q], io.normal[(size_t)n + q], io.normal[2 * (size_t)n + q]);
        apply_cosine<K, MT>(c, nrm, m, LOBES && io.lobes_out != nullptr, th, ph);
    } else if (LOBES && io.lobes_out) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const bool r = c.mapping == MAP_REFERENCE;
            th[i] = kPi * MT::div(1.0f, 1.0f + MT::exp(-(r ? x(4 * i + 2) : x(4 * i))));
            ph[i] = kTwoPi * MT::div(1.0f, 1.0f + MT::exp(-(r ? x(4 * i + 3) : x(4 * i + 1))));
        }
    }
    if (LOBES && io.lobes_out) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            io.lobes_out[(size_t)(5 * i + 0) * n + q] = m.lambda[i];
            io.lobes_out[(size_t)(5 * i + 1) * n + q] = m.kappa[i];
            io.lobes_out[(size_t)(5 * i + 2) * n + q] = th[i];
            io.lobes_out[(size_t)(5 * i + 3) * n + q] = ph[i];
            io.lobes_out[(size_t)(5 * i + 4) * n + q] = m.weight[i];
        }
    }
    if (c.sel && io.out_sel) io.out_sel[q] = MT::div(1.0f, 1.0f + MT::exp(-x(4 * K)));  // extractSelectionProbability
    if (io.u) {
        uint32_t lobe;
        const V3 d = mixture_sample<K, MT>(c, m, io.u[q], io.u[(size_t)n + q], io.u[2 * (size_t)n + q], &lobe);
        io.out_dir[q] = d.x;
        io.out_dir[(size_t)n + q] = d.y;
        io.out_dir[2 * (size_t)n + q] = d.z;
        io.out_pdf[q] = mixture_pdf<K, MT>(c, m, d);
        if (io.out_lobe) io.out_lobe[q] = lobe;
    }
    for (int e = 0; e < io.n_eval; ++e) {
        const float* ed = io.eval_dirs + (size_t)e * 3 * n;
        const V3 w = mk(ed[q], ed[(size_t)n + q], ed[2 * (size_t)n + q]);
        io.out_eval_pdf[(size_t)e * n + q] = mixture_pdf<K, MT>(c, m, w);
    }
}

// vmf_query with the math policy of the configuration's numerics.
template <int K, bool LOBES, class Raw>
__device__ __forceinline__ void vmf_query_cfg(const DevCfg& c, Raw&& x, const QueryIO& io, uint32_t q, uint32_t n) {
    if (c.numerics == NUM_FAST) vmf_query<K, FastMath, LOBES>(c, x, io, q, n);
    else vmf_query<K, ExactMath, LOBES>(c, x, io, q, n);
}

// Derivative of 1/k - coth(k): reference polynomial below 1 (distribution.h:
// 162-168) or the exact value (series below 0.5).
template <class MT = ExactMath>
__device__ __forceinline__ float inv_k_minus_coth(float k, bool ref) {
    if (ref) {
        if (k < 1.0f) return 0.000926f + -0.344883f * k + 0.030147f * (k * k);
        return 1.0f / k - (1.0f + fm::exp(-2.0f * k)) / (1.0f - fm::exp(-2.0f * k));
    }
    if (k < 0.5f) {
        const float k2 = k * k;
        return -k * (1.0f / 3.0f - k2 * (1.0f / 45.0f - k2 * (2.0f / 945.0f - k2 * (1.0f / 4725.0f))));
    }
    const float e = MT::exp(-2.0f * k);
    return MT::div(1.0f, k) - MT::div(1.0f + e, 1.0f - e);
}

// KL loss -f/q log(p + eps) and its adjoint with respect to the raw outputs
// (train.cu:69-124 with gradients_probability, distribution.h:143-181).
// dp/dlambda_s is evaluated as (v_s - p) / Lambda, the closed form of the
// reference's O(K^2) expression. `ls_over_n` is loss_scale / batch size;
// clamp_hi the upper clamp of the scaled gradient (0: none).
// With a learned selection probability s = logistic(x[4K]) (c.sel), the loss
// and gradients of compute_dL_doutput_with_selection (train.cu:126-172,
// distribution.h:216-231): P = s (p + eps) + (1 - s) bsdf_pdf, loss
// -f/P log P, dL/dx = -f/P^2 dP/dx with dP/dlobe = s dp/dlobe and the
// selection slot 0.01 (p - bsdf_pdf) s (1 - s); no NaN zeroing or clamp.
//
// MT: the math policy of the configuration's numerics (FastMath with "fast",
// whose formulas are the stable ones on hardware transcendentals).
// c.half_adjoint reproduces the reference's FP16 intermediates: the raw
// outputs are read as the network's __half outputs (train.cu:80-81), the four
// per-lobe dp/d(param) are stored in __half by gradients_probability
// (distribution.h:143-181, its output is precision_t), and each dL/dx is
// stored in __half before the NaN test and the clamp (train.cu:95-123), so a
// value beyond the FP16 range becomes infinite and is zeroed, not clamped.
template <int K, class MT = ExactMath, class Raw0, class Out>
__device__ __forceinline__ float mixture_adjoint(const DevCfg& c, Raw0&& x0, float th_t, float ph_t, float r, float g,
                                                 float b, float wipdf, float bsdf_pdf, float ls_over_n,
                                                 float clamp_hi, Out&& out) {
    const bool ref = c.numerics == NUM_REFERENCE;
    const bool h16 = c.half_adjoint != 0;
    auto x = [&](int j) { return h16 ? fm::round_half(x0(j)) : x0(j); };
    auto hq = [&](float v) { return h16 ? fm::round_half(v) : v; };
    Mixture<K> m;
    build_mixture<K, MT>(c, x, m);
    const V3 wi = sph_to_cart<MT>(th_t, ph_t);
    const float Li = (r + g + b) / 3.0f;
    const float wiPdf = wipdf + kEps;
    float vmf[K];
    float p = 0.0f;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        vmf[s] = lobe_pdf<K, MT>(c, m, s, wi);
        p = fm::fma_(m.weight[s], vmf[s], p);
    }
    const float guide = p + kEps;
    float prefix, loss, sel_s = 1.0f, sel_grad = 0.0f;
    if (c.sel) {
        const float s = MT::div(1.0f, 1.0f + MT::exp(-x(4 * K)));
        const float P = s * guide + (1.0f - s) * bsdf_pdf;
        prefix = -Li / P / P * ls_over_n;
        loss = -Li / P * MT::log(P);
        sel_s = s;
        sel_grad = hq(0.01f * prefix * hq(p - bsdf_pdf) * (s * (1.0f - s)));
    } else {
        prefix = -Li / wiPdf / guide * ls_over_n;
        loss = -Li / wiPdf * MT::log(guide);
    }
    // dp/d(param) of a lobe, times s with a learned selection (stored again:
    // distribution.h:224-226)
    auto sc = [&](float v) { return c.sel ? hq(v * sel_s) : v; };
    auto finish = [&](int j, float v) {
        v = hq(v);
        if (!c.sel) {
            if (!fm::isfinite_(v)) v = 0.0f;
            if (clamp_hi > 0.0f) v = smin(clamp_hi, v);
        }
        out(j, v);
    };
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const float kap = m.kappa[s];
        const V3 mu = m.mu[s];
        // mu = (sin t cos p, sin t sin p, cos t): recover the derivative
        // directions from mu and sin t (no second trigonometric evaluation).
        const float st = MT::sqrt(smax(0.0f, mu.x * mu.x + mu.y * mu.y));
        const float wv = m.weight[s] * vmf[s];
        const float dl = hq(MT::div(vmf[s] - p, m.total));
        const float dk = hq(wv * (dot(wi, mu) + inv_k_minus_coth<MT>(kap, ref)));
        // d mu / d theta = (cos t cos p, cos t sin p, -sin t); d mu / d phi = sin t (-sin p, cos p, 0).
        float cp = 1.0f, sp = 0.0f;
        if (st > 0.0f) { cp = MT::div(mu.x, st); sp = MT::div(mu.y, st); }
        const float dth = hq(wv * kap * (mu.z * (wi.x * cp + wi.y * sp) - wi.z * st));
        const float dph = hq(wv * kap * (wi.y * cp - wi.x * sp) * st);
        if (c.mapping == MAP_REFERENCE) {
            const float x0 = x(4 * s + 0), x1 = x(4 * s + 1), x2 = x(4 * s + 2), x3 = x(4 * s + 3);
            const float e0 = fm::exp(clampv(x0, -10.0f, 15.0f));
            const float e1 = fm::exp(clampv(x1, -10.0f, 15.0f));
            const float d0 = (!ref && (x0 < -10.0f || x0 > 15.0f)) ? 0.0f : e0;
            const float d1 = (!ref && (x1 < -10.0f || x1 > 15.0f)) ? 0.0f : e1;
            const float s2 = 1.0f / (1.0f + fm::exp(-x2));
            const float s3 = 1.0f / (1.0f + fm::exp(-x3));
            finish(4 * s + 0, prefix * sc(dl) * d0);
            finish(4 * s + 1, prefix * sc(dk) * d1);
            finish(4 * s + 2, prefix * sc(dth) * kPi * (s2 * (1.0f - s2)));
            finish(4 * s + 3, prefix * sc(dph) * kTwoPi * (s3 * (1.0f - s3)));
        } else {
            const float x0 = x(4 * s + 0), x1 = x(4 * s + 1), x2 = x(4 * s + 2);
            const float s0 = MT::div(1.0f, 1.0f + MT::exp(-x0));
            const float s1 = MT::div(1.0f, 1.0f + MT::exp(-x1));
            const float ek = MT::exp(x2);
            const float dkx = (ek > 1e-2f && ek < 1e3f) ? ek : 0.0f;
            finish(4 * s + 0, prefix * sc(dth) * kPi * (s0 * (1.0f - s0)));
            finish(4 * s + 1, prefix * sc(dph) * kTwoPi * (s1 * (1.0f - s1)));
            finish(4 * s + 2, prefix * sc(dk) * dkx);
            // lambda_k = exp(logit_k - max): d lambda_k / d logit_s = delta_ks lambda_s.
            finish(4 * s + 3, prefix * sc(dl) * m.lambda[s]);
        }
    }
    if (c.sel) out(4 * K, sel_grad);
    for (int j = 4 * K + c.sel; j < c.n_out_pad; ++j) out(j, 0.0f);
    return loss;
}

}  // namespace dev
}  // namespace npmg

namespace npmg {
namespace dev {

// ---------------------------------------------------------------------------
// Chunked encoding: the 8 features [8 ch, 8 ch + 8) of one query, bit-identical
// to the values encode_query produces (same expressions, same order), so the
// MFMA kernels see exactly the FP32 path's features before rounding.

// Sum of the three wrapped quartic CDFs at bin boundary b (oneblob.h:114-123).
__device__ __forceinline__ float oneblob_cdf(float b, float x, float fnb) { return oneblob_edge(b - x, fnb); }

// Bins k0 .. k0 + 7 of one OneBlob dimension for the fast numerics. With
// d = edge - x in [-1, 1], exactly one of d - 1, d, d + 1 lies within half a
// period of 0, namely d - rint(d); the other two copies lie at least 8 kernel
// radii away and contribute exactly 0 or 1, rint(d) + 1 in total. So the edge
// value is P(nb (d - rint(d))) + rint(d) + 1 and a bin, the difference of two
// edges, is dP + d(rint): the same quantity as the three-copy sum, without its
// cancellation against the integer parts. P is the quartic CDF with the
// polynomial in FMA form, clamped by evaluating at |u| <= 1.0625.
__device__ __forceinline__ void oneblob_bins_fast(float xv, int k0, float fnb, float inv_nb, float* v) {
    float P[9], R[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float d = (float)(k0 + i) * inv_nb - xv;
        const float r = __builtin_rintf(d);
        const float u = __builtin_amdgcn_fmed3f((d - r) * fnb, -1.0625f, 1.0625f);
        const float u2 = u * u;
        const float poly = __builtin_fmaf(u2, __builtin_fmaf(u2, 0.2f * 0.9375f, -2.0f / 3.0f * 0.9375f), 0.9375f);
        P[i] = __builtin_amdgcn_fmed3f(__builtin_fmaf(u, poly, 0.5f), 0.0f, 1.0f);
        R[i] = r;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = (P[i + 1] - P[i]) + (R[i + 1] - R[i]);
}

// Features of one grid level (F values) accumulated over the corners in order.
template <class T, int F>
__device__ __forceinline__ void grid_level_features(const DevCfg& c, const T* __restrict__ table, int l, float x0,
                                                    float x1, float x2, float* out) {
    uint32_t idx[8];
    float w[8];
    const int nc = level_corners(c, l, x0, x1, x2, idx, w);
    const T* tl = table + (size_t)c.offsets[l] * F;
    const bool vec = entry_aligned<T, F>(table);
    if (nc == 1) {
        T v0[F];
        load_entry<T, F>(tl, idx[0], vec, v0);
#pragma unroll
        for (int f = 0; f < F; ++f) out[f] = (float)v0[f];
        return;
    }
    T v[8][F];
    gather_corners<T, F>(tl, idx, group_aligned(table), vec, v);
#pragma unroll
    for (int f = 0; f < F; ++f) {
        float a = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) a = fm::fma_(w[k], (float)v[k][f], a);
        out[f] = a;
    }
}

// One feature, computed on its own (irregular chunks only).
template <class T>
__device__ __forceinline__ float single_feature(const DevCfg& c, const T* __restrict__ table, const float* q_in, int f) {
    if (f >= c.enc_width) return 1.0f;
    int e = 0;
    for (int i = 1; i < c.n_enc; ++i)
        if (f >= c.enc_offset[i]) e = i;
    const int r = f - c.enc_offset[e];
    const int b = c.enc_begin[e];
    const int type = c.enc_type[e];
    if (type == ENC_GRID) {
        const int l = r / c.F, ff = r % c.F;
        uint32_t idx[8];
        float w[8];
        const int nc = level_corners(c, l, pick(q_in, b), pick(q_in, b + 1), pick(q_in, b + 2), idx, w);
        const T* tl = table + (size_t)c.offsets[l] * c.F;
        if (nc == 1) return (float)tl[idx[0] * (uint32_t)c.F + ff];
        float a = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) a = fm::fma_(w[k], (float)tl[idx[k] * (uint32_t)c.F + ff], a);
        return a;
    }
    if (type == ENC_ONEBLOB) {
        const int nb = c.enc_param[e];
        const float fnb = (float)nb;
        const int d = r / nb, k = r % nb;
        const float xv = pick(q_in, b + d);
        const float left = oneblob_cdf((float)k * (1.0f / fnb), xv, fnb);
        const float right = oneblob_cdf((float)(k + 1) * (1.0f / fnb), xv, fnb);
        return right - left;
    }
    if (type == ENC_SH) {
        const float sx = pick(q_in, b) * 2.0f - 1.0f, sy = pick(q_in, b + 1) * 2.0f - 1.0f,
                    sz = pick(q_in, b + 2) * 2.0f - 1.0f;
        if (r == 0) return 0.28209479177387814f;
        if (r == 1) return -0.48860251190291987f * sy;
        if (r == 2) return 0.48860251190291987f * sz;
        return -0.48860251190291987f * sx;
    }
    return pick(q_in, b + r);
}

// The 8 features of chunk ch of one query into v[0..7]. When genc is given,
// the grid features come from the level-major encoding [L F][n] (already in
// the operand type) instead of being gathered here.
template <class T>
__device__ __forceinline__ void encode_chunk(const DevCfg& c, const T* __restrict__ table, const float* q_in, int ch,
                                             float v[8], const T* __restrict__ genc = nullptr, uint32_t q = 0,
                                             uint32_t n = 0, bool fast = false) {
    const int f0 = 8 * ch;
    if (genc) {
        const int g0 = c.enc_offset[c.grid_enc], gw = c.n_levels * c.F;
        if (f0 >= g0 && f0 + 8 <= g0 + gw) {
            const T* src = genc + (size_t)(f0 - g0) * n + q;
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (float)src[(size_t)i * n];
            return;
        }
    }
    int e = 0;
    for (int i = 1; i < c.n_enc; ++i)
        if (f0 >= c.enc_offset[i]) e = i;
    const int r0 = f0 - c.enc_offset[e];
    const int b = c.enc_begin[e];
    const int type = c.enc_type[e];
    const bool inside = f0 + 8 <= c.enc_width && r0 + 8 <= (e + 1 < c.n_enc ? c.enc_offset[e + 1] : c.enc_width) - c.enc_offset[e];
    if (f0 >= c.enc_width) {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = 1.0f;
        return;
    }
    if (inside && type == ENC_GRID && c.F == 2) {
        const float x0 = pick(q_in, b), x1 = pick(q_in, b + 1), x2 = pick(q_in, b + 2);
        const int l0 = r0 / 2;
#pragma unroll
        for (int i = 0; i < 4; ++i) grid_level_features<T, 2>(c, table, l0 + i, x0, x1, x2, v + 2 * i);
        return;
    }
    if (inside && type == ENC_GRID && c.F == 4) {
        const float x0 = pick(q_in, b), x1 = pick(q_in, b + 1), x2 = pick(q_in, b + 2);
        const int l0 = r0 / 4;
#pragma unroll
        for (int i = 0; i < 2; ++i) grid_level_features<T, 4>(c, table, l0 + i, x0, x1, x2, v + 4 * i);
        return;
    }
    if (inside && type == ENC_ONEBLOB && c.enc_param[e] >= 8) {
        const int nb = c.enc_param[e];
        const float fnb = (float)nb;
        const int d = r0 / nb, k0 = r0 % nb;
        const float xv = pick(q_in, b + d);
        const float inv_nb = 1.0f / fnb;  // power of two: exact
        if (fast) {
            oneblob_bins_fast(xv, k0, fnb, inv_nb, v);
            return;
        }
        float left = oneblob_cdf((float)k0 * inv_nb, xv, fnb);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float right = oneblob_cdf((float)(k0 + i + 1) * inv_nb, xv, fnb);
            v[i] = right - left;
            left = right;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = single_feature<T>(c, table, q_in, f0 + i);
}

}  // namespace dev
}  // namespace npmg
