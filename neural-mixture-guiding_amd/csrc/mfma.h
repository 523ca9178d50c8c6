// mfma.h -- fragment layouts of the gfx950 32x32x16 MFMA and the network
// pieces built on them (shared by the inference and training kernels).
//
// v_mfma_f32_32x32x16_{bf16,f16}: D[32][32] += A[32][16] * B[16][32], one
// wave. Lane l = 32 h + j holds
//   A: row j, k = 8h .. 8h+7        (8 values)
//   B: column j, k = 8h .. 8h+7     (8 values)
//   D: column j, rows 8a + 4h + b for a, b in 0..3 (16 floats, register 4a + b)
//
// Network layers use A = weights (rows = output neurons), B = activations
// (columns = queries). A wave holds 64 queries as two 32-column tiles c = 0, 1.
//
// Layer chaining without data movement: the accumulator registers 8m .. 8m+7
// (m = 0, 1) of output tile t are exactly the B fragment of K-block
// kb = 2t + m of the next layer if the next layer's weight columns are
// permuted so that logical input k = 16 kb + 8 h + e reads hidden neuron
//   p(kb, h, e) = 32 (kb >> 1) + 8 (2 (kb & 1) + (e >> 2)) + 4 h + (e & 3).
// The packing kernel applies this permutation to every layer after the first.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace npmg {
namespace mf {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;

// Split-BF16 operand ("precision": "bf16x3"): every value x is carried as
// hi = bf16(x) and lo = bf16(x - hi), about 16 significant bits, and each
// product of two operands is formed from three BF16 products, hi*hi + hi*lo
// + lo*hi (the lo*lo term is below the FP32 accumulator's rounding). A
// fragment is two planes: the hi fragment, then the lo fragment.
struct Bf16x3 {};
struct bf16x8x2 {
    bf16x8 hi, lo;
};

// FP8 E4M3 operand byte (inference only). v_mfma_f32_32x32x16_fp8_fp8 has the
// fragment layout of the 16-bit types above with 8 one-byte values per lane.
struct Fp8 {
    uint8_t bits;
};
constexpr float kFp8Max = 448.0f;  // largest finite E4M3 value

// Hidden neuron read by logical input k = 16 kb + 8 h + e of a chained layer.
__host__ __device__ __forceinline__ int chained_column(int kb, int h, int e) {
    return 32 * (kb >> 1) + 8 * (2 * (kb & 1) + (e >> 2)) + 4 * h + (e & 3);
}

template <class T> struct Frag;
template <> struct Frag<__bf16> {
    using vec = bf16x8;
    __device__ static __forceinline__ f32x16 mma(vec a, vec b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};
template <> struct Frag<_Float16> {
    using vec = f16x8;
    __device__ static __forceinline__ f32x16 mma(vec a, vec b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <> struct Frag<Bf16x3> {
    using vec = bf16x8x2;
    __device__ static __forceinline__ f32x16 mma(const vec& a, const vec& b, f32x16 c) {
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.lo, b.hi, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi, b.lo, c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.hi, b.hi, c, 0, 0, 0);
    }
};

// Storage of an operand type: element type and planes (hi, lo) per value, and
// the type of the grid table and level-major features that feed it.
template <class O> struct Op {
    using elem = O;
    using table = O;
    static constexpr int planes = 1;
};
template <> struct Op<Bf16x3> {
    using elem = __bf16;
    using table = _Float16;
    static constexpr int planes = 2;
};
template <> struct Op<Fp8> {
    using elem = uint8_t;
    using table = __bf16;
    static constexpr int planes = 1;
};

// Element e of plane p of a fragment value, as stored.
template <class O>
__device__ __forceinline__ typename Op<O>::elem get(const typename Frag<O>::vec& v, int p, int e) {
    if constexpr (Op<O>::planes == 2) return p ? v.lo[e] : v.hi[e];
    else return v[e];
}
template <class O>
__device__ __forceinline__ void set(typename Frag<O>::vec& v, int p, int e, typename Op<O>::elem x) {
    if constexpr (Op<O>::planes == 2) {
        if (p) v.lo[e] = x;
        else v.hi[e] = x;
    } else {
        v[e] = x;
    }
}
// Value of element e as a float (hi + lo for the split type).
template <class O>
__device__ __forceinline__ float value(const typename Frag<O>::vec& v, int e) {
    if constexpr (Op<O>::planes == 2) return (float)v.hi[e] + (float)v.lo[e];
    else return (float)v[e];
}
// 8 consecutive 16-bit elements of each plane (16-byte loads; p0 + plane_stride
// is the lo plane of the split type).
template <class O>
__device__ __forceinline__ typename Frag<O>::vec load_rows16(const typename Op<O>::elem* p0, size_t plane_stride) {
    typename Frag<O>::vec v;
    if constexpr (Op<O>::planes == 2) {
        v.hi = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p0));
        v.lo = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p0 + plane_stride));
    } else {
        v = __builtin_bit_cast(typename Frag<O>::vec, *reinterpret_cast<const uint4*>(p0));
    }
    return v;
}

template <> struct Frag<Fp8> {
    using vec = long;  // 8 E4M3 values, element e in byte e
    __device__ static __forceinline__ f32x16 mma(vec a, vec b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a, b, c, 0, 0, 0);
    }
};

// 8 floats -> one B/A fragment (round to nearest even; FP8 inputs must lie
// within +-kFp8Max, the callers clamp).
template <class T>
__device__ __forceinline__ typename Frag<T>::vec pack8(const float* v) {
    if constexpr (Op<T>::planes == 2) {
        bf16x8x2 r;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            r.hi[e] = (__bf16)v[e];
            r.lo[e] = (__bf16)(v[e] - (float)r.hi[e]);
        }
        return r;
    } else if constexpr (sizeof(typename Frag<T>::vec) == 8) {
        int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
        lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
        int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0, false);
        hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
        return (long)(((unsigned long)(uint32_t)hi << 32) | (unsigned long)(uint32_t)lo);
    } else {
        typename Frag<T>::vec r;
#pragma unroll
        for (int e = 0; e < 8; ++e) r[e] = (T)v[e];
        return r;
    }
}

// One float -> E4M3 byte (round to nearest even; |x| <= kFp8Max).
__device__ __forceinline__ uint8_t to_fp8(float x) {
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xff);
}

// Fragment of the weights, fragments stored lane-major so a wave reads one
// contiguous block (16 bytes per lane, or 8 for FP8): conflict-free LDS reads.
template <class T>
__device__ __forceinline__ typename Frag<T>::vec load_frag(const T* base, int frag, int lane) {
    if constexpr (Op<T>::planes == 2) {
        const uint4* p = reinterpret_cast<const uint4*>(base);
        return bf16x8x2{__builtin_bit_cast(bf16x8, p[(2 * frag) * 64 + lane]),
                        __builtin_bit_cast(bf16x8, p[(2 * frag + 1) * 64 + lane])};
    } else if constexpr (sizeof(typename Frag<T>::vec) == 8) {
        const uint2 raw = reinterpret_cast<const uint2*>(base)[frag * 64 + lane];
        return __builtin_bit_cast(long, raw);
    } else {
        const uint4 raw = reinterpret_cast<const uint4*>(base)[frag * 64 + lane];
        return __builtin_bit_cast(typename Frag<T>::vec, raw);
    }
}

// Two 16-bit operands in one register, `lo` in bits 0..15 (the element order of
// a fragment), each rounded to nearest even like mf::pack8.
template <class T>
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    const T a = (T)lo, b = (T)hi;
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// Bins 8h .. 8h+7 of a 16-bin OneBlob encoding of x, as one B fragment (fast
// numerics). The quartic kernel of radius one bin (tiny-cuda-nn
// common_device.h:533-548), centred at s = 16 x = kx + t with t in [0, 1),
// puts all its mass into bins kx - 1, kx and kx + 1, wrapped mod 16 like
// kernel_one_blob_soa (oneblob.h:114-123): 1/2 - S(t), S(t) + S(1 - t) and
// 1/2 - S(1 - t), with S(u) = 15/16 (u - 2u^3/3 + u^5/5) the odd part of the
// kernel's CDF. Bins b, b + 1 are the 32-bit window at 16-bit slot
// e = (b - kx + 2) mod 16 of the slot sequence (0, m[kx-1], m[kx], m[kx+1], 0, ...).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The 8 16-bit slots of bins 8h .. 8h+7 taken from the slot sequence `slots`.
__device__ __forceinline__ u32x4 oneblob16_place(uint64_t slots, int kx, int h) {
    const int base = 8 * h + 2 - kx;
    u32x4 d;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t e = (uint32_t)(base + 2 * j) & 15u;
        d[j] = e < 4u ? (uint32_t)(slots >> (16u * e)) : 0u;
    }
    return d;
}

template <class T>
__device__ __forceinline__ typename mf::Frag<T>::vec oneblob16_frag(float x, int h) {
    const float s = x * 16.0f;
    const float fl = __builtin_floorf(s);
    const float t = s - fl;
    const int kx = (int)fl;
    auto odd = [](float u) {
        const float u2 = u * u;
        return u * __builtin_fmaf(u2, __builtin_fmaf(u2, 0.1875f, -0.625f), 0.9375f);
    };
    const float st = odd(t), sr = odd(1.0f - t);
    const float mp = 0.5f - st, mc = st + sr, mn = 0.5f - sr;
    if constexpr (Op<T>::planes == 2) {
        const __bf16 hp = (__bf16)mp, hc = (__bf16)mc, hn = (__bf16)mn;
        const uint64_t sh = ((uint64_t)pack2<__bf16>(mc, mn) << 32) | (uint64_t)pack2<__bf16>(0.0f, mp);
        const uint64_t sl = ((uint64_t)pack2<__bf16>(mc - (float)hc, mn - (float)hn) << 32) |
                            (uint64_t)pack2<__bf16>(0.0f, mp - (float)hp);
        return bf16x8x2{__builtin_bit_cast(bf16x8, oneblob16_place(sh, kx, h)),
                        __builtin_bit_cast(bf16x8, oneblob16_place(sl, kx, h))};
    } else {
        const uint64_t slots = ((uint64_t)pack2<T>(mc, mn) << 32) | (uint64_t)pack2<T>(0.0f, mp);
        return __builtin_bit_cast(typename mf::Frag<T>::vec, oneblob16_place(slots, kx, h));
    }
}

}  // namespace mf
}  // namespace npmg
