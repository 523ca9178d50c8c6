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
    if constexpr (sizeof(typename Frag<T>::vec) == 8) {
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
    if constexpr (sizeof(typename Frag<T>::vec) == 8) {
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
    const uint64_t slots = ((uint64_t)pack2<T>(st + sr, 0.5f - sr) << 32) | (uint64_t)pack2<T>(0.0f, 0.5f - st);
    const int base = 8 * h + 2 - kx;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 d;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t e = (uint32_t)(base + 2 * j) & 15u;
        d[j] = e < 4u ? (uint32_t)(slots >> (16u * e)) : 0u;
    }
    return __builtin_bit_cast(typename mf::Frag<T>::vec, d);
}

}  // namespace mf
}  // namespace npmg
