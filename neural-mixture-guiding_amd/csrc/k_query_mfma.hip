// k_query_mfma.hip -- host dispatch of the fused MFMA query kernel (BF16,
// FP16 or FP8 operands, FP32 accumulation; the kernel is in query_mfma.h),
// the level-major grid encode, and the packing kernels that produce its
// weight fragments and half-precision grid tables.
//
// The grid features come from k_grid_encode (level-major, operand type, [L F][n])
// when the grid rows form whole 8-feature chunks; the host pipelines it with the
// fused kernel over chunks of the batch on two streams, since the encode is
// bound by L2 gather bandwidth and the fused kernel by VALU issue.
#include <algorithm>

#include "asg.h"
#include "device.h"
#include "kernels.h"
#include "mfma.h"

namespace npmg {

// Fused-kernel launchers for one shape, defined in query_mfma.h and
// instantiated in k_query_mfma_inst.hip.
template <class T, class O, int W, int K, int P, bool FAST>
hipError_t launch_mfma_t(const DevCfg& c, const void* wfrag, const void* table, const void* genc, const QueryIO& io,
                         uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu, int bpc, const float* qscale,
                         hipStream_t s);
template <class T, class O, int W, int K, int P, bool FAST>
hipError_t launch_mfma_queues_t(const DevCfg& c, const void* wfrag, const void* table, const QueueSet& qs, int n_cu,
                                int blocks_per_cu, const float* qscale, hipStream_t s);

namespace {

using namespace dev;

// Level-major grid encoding: grid (ceil(n / 256), n_levels). Blocks run in
// linear order, so the whole GPU works on one level at a time and that
// level's table slice (at most 2^19 x F entries) stays resident in each XCD's
// L2 while it is gathered; the features go out in the operand type, [L F][n].
// l0: the first level of the launch (the levels below it are gathered from LDS,
// k_grid_encode_lds).
template <class T>
__global__ void __launch_bounds__(256) k_grid_encode(const DevCfg c, const T* __restrict__ table,
                                                     const float* __restrict__ in, T* __restrict__ genc, uint32_t n,
                                                     uint32_t q_begin, uint32_t q_end, int l0) {
    const uint32_t q = q_begin + blockIdx.x * 256u + threadIdx.x;
    const int l = blockIdx.y + l0;
    if (q >= q_end) return;
    const int b = c.enc_begin[c.grid_enc];
    const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
    if (c.F == 2) {
        float f[2];
        grid_level_features<T, 2>(c, table, l, x0, x1, x2, f);
        genc[(size_t)(2 * l) * n + q] = (T)f[0];
        genc[(size_t)(2 * l + 1) * n + q] = (T)f[1];
    } else if (c.F == 4) {
        float f[4];
        grid_level_features<T, 4>(c, table, l, x0, x1, x2, f);
#pragma unroll
        for (int i = 0; i < 4; ++i) genc[(size_t)(4 * l + i) * n + q] = (T)f[i];
    } else {
        uint32_t idx[8];
        float w[8];
        const int nc = level_corners(c, l, x0, x1, x2, idx, w);
        const T* tl = table + (size_t)c.offsets[l] * c.F;
        for (int ff = 0; ff < c.F; ++ff) {
            float a = 0.0f;
            if (nc == 1) a = (float)tl[idx[0] * (uint32_t)c.F + ff];
            else
                for (int k = 0; k < 8; ++k) a = fm::fma_(w[k], (float)tl[idx[k] * (uint32_t)c.F + ff], a);
            genc[(size_t)(c.F * l + ff) * n + q] = (T)a;
        }
    }
}

// The coarse levels whose table fits in LDS (F = 2: one 32-bit word per
// entry, at most kLdsLevelBytes): grid (workgroups, levels); a workgroup of
// 1024 threads stages its level's table in LDS once and gathers the corners
// of many queries from there (grid-stride over the queries), instead of four
// L2 requests per query and level; two such workgroups fill a CU. The
// interpolation is grid_level_features' (same operations, same order), so the
// features are the same bits.
constexpr uint32_t kLdsLevelBytes = 64u * 1024u;
constexpr uint32_t kLdsEncodeThreads = 1024u;

// Level l's table (F = 2 16-bit features: one word per entry) into LDS.
template <class T>
__device__ __forceinline__ void lds_stage_level(const DevCfg& c, const T* __restrict__ table, int l, uint32_t* lt) {
    static_assert(sizeof(T) == 2, "16-bit features");
    const uint32_t entries = c.offsets[l + 1] - c.offsets[l];
    const uint32_t* src = reinterpret_cast<const uint32_t*>(table + (size_t)c.offsets[l] * 2u);
    for (uint32_t i = threadIdx.x; i < entries; i += kLdsEncodeThreads) lt[i] = src[i];
    __syncthreads();
}

// Features of query q at level l from the staged table: grid_level_features'
// corners, weights and FMA order.
template <class T>
__device__ __forceinline__ void lds_level_query(const DevCfg& c, const uint32_t* lt, int l, const float* __restrict__ in,
                                                int b, uint32_t n, uint32_t q, T* __restrict__ genc) {
    auto entry = [&](uint32_t i, float& f0, float& f1) {
        const uint32_t wv = lt[i];
        T v[2];
        __builtin_memcpy(v, &wv, 4);
        f0 = (float)v[0];
        f1 = (float)v[1];
    };
    const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
    uint32_t idx[8];
    float w[8];
    const int nc = level_corners(c, l, x0, x1, x2, idx, w);
    float a0 = 0.0f, a1 = 0.0f;
    if (nc == 1) {
        entry(idx[0], a0, a1);
    } else {
        float v0[8], v1[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) entry(idx[k], v0[k], v1[k]);
#pragma unroll
        for (int k = 0; k < 8; ++k) a0 = fm::fma_(w[k], v0[k], a0);
#pragma unroll
        for (int k = 0; k < 8; ++k) a1 = fm::fma_(w[k], v1[k], a1);
    }
    genc[(size_t)(2 * l) * n + q] = (T)a0;
    genc[(size_t)(2 * l + 1) * n + q] = (T)a1;
}
template <class T>
__global__ void __launch_bounds__(1024) k_grid_encode_lds(const DevCfg c, const T* __restrict__ table,
                                                          const float* __restrict__ in, T* __restrict__ genc,
                                                          uint32_t n, uint32_t q_begin, uint32_t q_end) {
    extern __shared__ uint32_t lt[];
    const int l = blockIdx.y;
    lds_stage_level(c, table, l, lt);
    const int b = c.enc_begin[c.grid_enc];
    for (uint32_t q = q_begin + blockIdx.x * kLdsEncodeThreads + threadIdx.x; q < q_end;
         q += gridDim.x * kLdsEncodeThreads)
        lds_level_query<T>(c, lt, l, in, b, n, q, genc);
}

// The coarse levels of several queues (k_grid_encode_lds over each queue's
// live queries; the counts are read on the device).
template <class T>
__global__ void __launch_bounds__(1024) k_grid_encode_lds_queues(const DevCfg c, const T* __restrict__ table,
                                                                 const QueueSet qs) {
    extern __shared__ uint32_t lt[];
    const int l = blockIdx.y;
    lds_stage_level(c, table, l, lt);
    const int b = c.enc_begin[c.grid_enc];
    for (uint32_t qi = 0; qi < qs.n_queues; ++qi) {
        const uint32_t n = qs.stride[qi];
        const uint32_t cnt = min(qs.counts[qi], n);
        for (uint32_t q = blockIdx.x * kLdsEncodeThreads + threadIdx.x; q < cnt; q += gridDim.x * kLdsEncodeThreads)
            lds_level_query<T>(c, lt, l, qs.io[qi].in, b, n, q, static_cast<T*>(qs.genc[qi]));
    }
}

// Level-major grid encoding of several queues in one persistent launch:
// workgroups pull (level, 256-query tile) items in level-major order, a batch
// of qs.batch consecutive items per atomic (one counter serves every
// workgroup; one atomic per item serialised on it and took most of the time).

template <class T>
__global__ void __launch_bounds__(256) k_grid_encode_queues(const DevCfg c, const T* __restrict__ table,
                                                            const QueueSet qs, int l0) {
    __shared__ uint32_t item;
    uint32_t first[kMaxQueues + 1];
    queue_tile_prefix(qs, 256u, first);
    const uint32_t per_level = first[kMaxQueues];
    const uint32_t total = per_level * (uint32_t)(c.n_levels - l0);  // levels l0 .. n_levels - 1
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(qs.counter, qs.batch);
        __syncthreads();
        const uint32_t w0 = item;
        __syncthreads();
        if (w0 >= total) break;
        const uint32_t w1 = min(total, w0 + qs.batch);
        for (uint32_t w = w0; w < w1; ++w) {
            const int l = l0 + (int)(w / per_level);
            const uint32_t t = w % per_level;
            int qi = 0;
            uint32_t base = 0;
#pragma unroll
            for (int i = 1; i < kMaxQueues; ++i)
                if (t >= first[i] && first[i] < first[i + 1]) {
                    qi = i;
                    base = first[i];
                }
            const uint32_t n = qs.stride[qi];
            const uint32_t q = (t - base) * 256u + threadIdx.x;
            if (q >= min(qs.counts[qi], n)) continue;
            const float* in = qs.io[qi].in;
            T* genc = static_cast<T*>(qs.genc[qi]);
            const int b = c.enc_begin[c.grid_enc];
            const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q],
                        x2 = in[(size_t)(b + 2) * n + q];
            if (c.F == 2) {
                float f[2];
                grid_level_features<T, 2>(c, table, l, x0, x1, x2, f);
                genc[(size_t)(2 * l) * n + q] = (T)f[0];
                genc[(size_t)(2 * l + 1) * n + q] = (T)f[1];
            } else {
                float f[4];
                grid_level_features<T, 4>(c, table, l, x0, x1, x2, f);
#pragma unroll
                for (int i = 0; i < 4; ++i) genc[(size_t)(4 * l + i) * n + q] = (T)f[i];
            }
        }
    }
}

// Weight fragments from full-precision row-major matrices: fragment
// (layer, t, kb), lane (h, i), element e holds W_l[32 t + i][col] with
// col = 16 kb + 8 h + e on the first layer and the chained permutation after.
// Split-BF16 operands (mf::Bf16x3): fragment f is stored as the hi fragment
// (2 f) and the lo fragment (2 f + 1) of BF16 elements.
template <class T>
__global__ void __launch_bounds__(256) k_pack_frags(const DevCfg c, const float* __restrict__ w,
                                                    typename mf::Op<T>::elem* __restrict__ frag) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    uint32_t base = 0;
    for (int l = 0; l < c.n_layers; ++l) {
        const uint32_t iw = (uint32_t)c.layer_in[l], ow = (uint32_t)c.layer_out[l];
        const uint32_t cnt = iw * ow;
        if (gid < base + cnt) {
            const uint32_t r = gid - base;
            const uint32_t f = r / 512u, within = r % 512u;
            const uint32_t lane = within / 8u, e = within % 8u;
            const uint32_t kbs = iw / 16u;
            const uint32_t t = f / kbs, kb = f % kbs;
            const uint32_t h = lane / 32u, i = lane % 32u;
            const uint32_t col = l == 0 ? 16u * kb + 8u * h + e : (uint32_t)mf::chained_column((int)kb, (int)h, (int)e);
            const float x = w[c.layer_off[l] + (32u * t + i) * iw + col];
            if constexpr (mf::Op<T>::planes == 2) {
                const size_t g0 = (size_t)(base / 512u + f) * 1024u + within;
                const __bf16 hi = (__bf16)x;
                frag[g0] = hi;
                frag[g0 + 512u] = (__bf16)(x - (float)hi);
            } else {
                frag[gid] = (T)x;
            }
            return;
        }
        base += cnt;
    }
}

// 2^j for the largest j with m * 2^j <= kFp8Max (m = f 2^k, f in [0.5, 1)):
// the power-of-two scale that maps a maximum magnitude m into E4M3 range.
__device__ __forceinline__ float fp8_pow2_scale(float m) {
    if (!(m > 0.0f) || !__builtin_isfinite(m)) return 1.0f;
    int k;
    const float f = frexpf(m, &k);
    int j = (f <= 0.875f ? 9 : 8) - k;  // kFp8Max = 0.875 * 2^9
    j = j < -100 ? -100 : (j > 100 ? 100 : j);
    return ldexpf(1.0f, j);
}

// FP8 weight fragments, one power-of-two scale per layer: block l reduces
// max |W_l|, sets wscale[l] = fp8_pow2_scale(max) and packs W_l * wscale[l]
// in the fragment order of k_pack_frags.
__global__ void __launch_bounds__(256) k_pack_frags_fp8(const DevCfg c, const float* __restrict__ w,
                                                        uint8_t* __restrict__ frag, float* __restrict__ wscale) {
    __shared__ float red[4];
    const int l = blockIdx.x;
    uint32_t base = 0;
    for (int i = 0; i < l; ++i) base += (uint32_t)(c.layer_in[i] * c.layer_out[i]);
    const uint32_t iw = (uint32_t)c.layer_in[l], ow = (uint32_t)c.layer_out[l], cnt = iw * ow;
    const float* wl = w + c.layer_off[l];
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < cnt; i += 256u) m = fmaxf(m, fabsf(wl[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    const float sw = fp8_pow2_scale(m);
    if (threadIdx.x == 0) wscale[l] = sw;
    const uint32_t kbs = iw / 16u;
    for (uint32_t r = threadIdx.x; r < cnt; r += 256u) {
        const uint32_t f = r / 512u, within = r % 512u;
        const uint32_t lane = within / 8u, e = within % 8u;
        const uint32_t t = f / kbs, kb = f % kbs;
        const uint32_t h = lane / 32u, i = lane % 32u;
        const uint32_t col = l == 0 ? 16u * kb + 8u * h + e : (uint32_t)mf::chained_column((int)kb, (int)h, (int)e);
        const float v = wl[(32u * t + i) * iw + col] * sw;
        frag[base + r] = mf::to_fp8(__builtin_amdgcn_fmed3f(v, -mf::kFp8Max, mf::kFp8Max));
    }
}

// Per-layer activation scales from the calibrated input maxima (one binade of
// headroom: amax maps into [kFp8Max / 4, kFp8Max / 2]) and the multipliers of
// the fused kernel: qscale[0] = sx_0, qscale[l] = sx_l / (sw_{l-1} sx_{l-1}),
// qscale[L] = 1 / (sw_{L-1} sx_{L-1}). All powers of two, so exact.
__global__ void k_fp8_scales(const DevCfg c, const float* __restrict__ wscale, const float* __restrict__ amax,
                             float* __restrict__ ascale, float* __restrict__ qscale) {
    if (threadIdx.x != 0) return;
    float prev = 1.0f;
    for (int l = 0; l < c.n_layers; ++l) {
        const float sx = fp8_pow2_scale(2.0f * amax[l]);
        ascale[l] = sx;
        qscale[l] = l == 0 ? sx : sx / prev;
        prev = wscale[l] * sx;
    }
    qscale[c.n_layers] = 1.0f / prev;
}

// max |x| over count values into *out (non-negative floats order like their bits).
__global__ void __launch_bounds__(256) k_amax(const float* __restrict__ x, size_t count, float* __restrict__ out) {
    __shared__ float red[4];
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < count; i += (size_t)gridDim.x * 256u)
        m = fmaxf(m, fabsf(x[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        atomicMax(reinterpret_cast<unsigned int*>(out), __float_as_uint(m));
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_to_half(const float* __restrict__ src, T* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = (T)src[i];
}

template <class T, class O>
hipError_t launch_mfma_queues_type(const DevCfg& c, const void* wfrag, const void* table, const QueueSet& qs,
                                   int n_cu, int bpc, const float* qscale, hipStream_t s) {
#define NPMG_QQ(WW, KK, PP)                                \
    if (c.width == WW && c.K == KK && c.lobe_params == PP) \
        return c.numerics == NUM_FAST                                                                 \
                   ? launch_mfma_queues_t<T, O, WW, KK, PP, true>(c, wfrag, table, qs, n_cu, bpc, qscale, s)     \
                   : launch_mfma_queues_t<T, O, WW, KK, PP, false>(c, wfrag, table, qs, n_cu, bpc, qscale, s);
    NPMG_QQ(64, 7, 4)
    NPMG_QQ(64, 8, 4)
    NPMG_QQ(64, 16, 4)
    NPMG_QQ(64, 16, 6)
    if constexpr (mf::Op<O>::planes == 1) {  // split weights of a 128-wide network do not fit in LDS
        NPMG_QQ(128, 7, 4)
        NPMG_QQ(128, 8, 4)
        NPMG_QQ(128, 16, 4)
        NPMG_QQ(128, 16, 6)
    }
#undef NPMG_QQ
    return hipErrorInvalidValue;
}

template <class T, class O>
hipError_t launch_mfma_type(const DevCfg& c, const void* wfrag, const void* table, const void* genc,
                            const QueryIO& io, uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu, int bpc,
                            const float* qscale, hipStream_t s) {
#define NPMG_QM(WW, KK, PP)                                      \
    if (c.width == WW && c.K == KK && c.lobe_params == PP)       \
        return c.numerics == NUM_FAST                                                                            \
                   ? launch_mfma_t<T, O, WW, KK, PP, true>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, bpc,   \
                                                          qscale, s)                                             \
                   : launch_mfma_t<T, O, WW, KK, PP, false>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, bpc,  \
                                                           qscale, s);
    NPMG_QM(64, 7, 4)
    NPMG_QM(64, 8, 4)
    NPMG_QM(64, 16, 4)
    NPMG_QM(64, 16, 6)
    if constexpr (mf::Op<O>::planes == 1) {  // split weights of a 128-wide network do not fit in LDS
        NPMG_QM(128, 7, 4)
        NPMG_QM(128, 8, 4)
        NPMG_QM(128, 16, 4)
        NPMG_QM(128, 16, 6)
    }
#undef NPMG_QM
    return hipErrorInvalidValue;
}

}  // namespace

// The number of leading levels a batch of `batch` queries gathers from LDS:
// those whose F = 2 tables fit in kLdsLevelBytes, when the batch gives each
// of 256 workgroups per level at least 4096 queries to repay the staging
// (82 KB per CU for the full configuration's levels 0 and 1; at 2^18 queries
// the staging costs what it saves).
constexpr uint32_t kLdsMinBatch = 1u << 20;
int grid_encode_lds_levels(const DevCfg& c, const void* table, uint32_t batch) {
    int n_lds = 0;
    if (batch < kLdsMinBatch || c.F != 2 || (reinterpret_cast<uintptr_t>(table) & 3u) != 0u) return 0;
    while (n_lds < c.n_levels && (c.offsets[n_lds + 1] - c.offsets[n_lds]) * 4u <= kLdsLevelBytes) ++n_lds;
    return n_lds;
}

// lds_reserve: unused dynamic LDS per workgroup, which caps the encode's
// workgroups per CU so that fused-kernel workgroups running beside it (the
// pipelined query path) still find room; the encode loses nothing down to
// about 5 workgroups per CU. n_lds: the leading levels gathered from LDS
// (grid_encode_lds_levels). levels: ENC_ALL, or only those n_lds levels
// (ENC_LDS_LEVELS) or only the others (ENC_GATHER_LEVELS), so that a
// pipelined caller can encode the coarse levels of the whole batch once (a
// 1024-thread, 64 KB workgroup does not fit beside the fused kernel).
hipError_t launch_grid_encode_range(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                                    uint32_t n, uint32_t q_begin, uint32_t q_end, hipStream_t s, uint32_t lds_reserve,
                                    int n_lds, int levels) {
    if (q_end <= q_begin) return hipSuccess;
    if (precision != PREC_BF16 && precision != PREC_FP16 && precision != PREC_BF16X3) return hipErrorInvalidValue;
    if (n_lds < 0 || n_lds > c.n_levels || (n_lds > 0 && (c.F != 2 || (reinterpret_cast<uintptr_t>(table) & 3u))))
        return hipErrorInvalidValue;
    const bool bf = precision == PREC_BF16;
    uint32_t lds_bytes = 0;
    for (int l = 0; l < n_lds; ++l) {
        const uint32_t b = (c.offsets[l + 1] - c.offsets[l]) * 4u;
        if (b > kLdsLevelBytes) return hipErrorInvalidValue;
        lds_bytes = std::max(lds_bytes, b);
    }
    const uint32_t tiles = (q_end - q_begin + 255) / 256;
    if (n_lds > 0 && levels != ENC_GATHER_LEVELS) {
        // one workgroup per CU and level at most, each over >= 1 tile of kLdsEncodeThreads queries
        const uint32_t wgs = std::min((q_end - q_begin + kLdsEncodeThreads - 1) / kLdsEncodeThreads, 256u);
        const dim3 grid_l(wgs, (uint32_t)n_lds);
        const uint32_t lds = std::max(lds_bytes, lds_reserve);
        if (bf)
            hipLaunchKernelGGL((k_grid_encode_lds<__bf16>), grid_l, dim3(kLdsEncodeThreads), lds, s, c,
                               (const __bf16*)table, in, (__bf16*)genc, n, q_begin, q_end);
        else
            hipLaunchKernelGGL((k_grid_encode_lds<_Float16>), grid_l, dim3(kLdsEncodeThreads), lds, s, c,
                               (const _Float16*)table, in, (_Float16*)genc, n, q_begin, q_end);
    }
    if (n_lds < c.n_levels && levels != ENC_LDS_LEVELS) {
        const dim3 grid(tiles, (uint32_t)(c.n_levels - n_lds));
        if (bf)
            hipLaunchKernelGGL((k_grid_encode<__bf16>), grid, dim3(256), lds_reserve, s, c, (const __bf16*)table, in,
                               (__bf16*)genc, n, q_begin, q_end, n_lds);
        else
            hipLaunchKernelGGL((k_grid_encode<_Float16>), grid, dim3(256), lds_reserve, s, c, (const _Float16*)table,
                               in, (_Float16*)genc, n, q_begin, q_end, n_lds);
    }
    return hipGetLastError();
}

hipError_t launch_grid_encode(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                              uint32_t n, hipStream_t s) {
    return launch_grid_encode_range(c, precision, table, in, genc, n, 0u, n, s, 0u,
                                    grid_encode_lds_levels(c, table, n), ENC_ALL);
}

bool mfma_split_grid(const DevCfg& c) {
    return c.grid_enc >= 0 && c.enc_offset[c.grid_enc] % 8 == 0 && (c.n_levels * c.F) % 8 == 0;
}

// The fused kernel over queries [q_begin, q_end); with genc, the grid features
// of those queries must already be there (launch_grid_encode_range).
hipError_t launch_query_mfma(const DevCfg& c, int precision, const void* wfrag, const void* table, const void* genc,
                             const QueryIO& io, uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu,
                             int blocks_per_cu, const float* fp8_scale, hipStream_t s) {
    if (q_end <= q_begin) return hipSuccess;
    if (q_end > n || c.enc_padded % 16 != 0) return hipErrorInvalidValue;
    if (genc && !mfma_split_grid(c)) return hipErrorInvalidValue;
    if (fp8_scale) {
        if (precision != PREC_BF16) return hipErrorInvalidValue;
        return launch_mfma_type<__bf16, mf::Fp8>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, blocks_per_cu,
                                                 fp8_scale, s);
    }
    if (precision == PREC_BF16)
        return launch_mfma_type<__bf16, __bf16>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, blocks_per_cu,
                                                nullptr, s);
    if (precision == PREC_FP16)
        return launch_mfma_type<_Float16, _Float16>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu,
                                                    blocks_per_cu, nullptr, s);
    if (precision == PREC_BF16X3)
        return launch_mfma_type<_Float16, mf::Bf16x3>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu,
                                                      blocks_per_cu, nullptr, s);
    return hipErrorInvalidValue;
}

// Persistent launches over several queues (qs.counter must be zero).
hipError_t launch_query_mfma_queues(const DevCfg& c, int precision, const void* wfrag, const void* table,
                                    const QueueSet& qs, int n_cu, int blocks_per_cu, const float* fp8_scale,
                                    hipStream_t s) {
    if (qs.n_queues == 0) return hipSuccess;
    if (qs.n_queues > (uint32_t)kMaxQueues || c.enc_padded % 16 != 0) return hipErrorInvalidValue;
    if (fp8_scale) {
        if (precision != PREC_BF16) return hipErrorInvalidValue;
        return launch_mfma_queues_type<__bf16, mf::Fp8>(c, wfrag, table, qs, n_cu, blocks_per_cu, fp8_scale, s);
    }
    if (precision == PREC_BF16)
        return launch_mfma_queues_type<__bf16, __bf16>(c, wfrag, table, qs, n_cu, blocks_per_cu, nullptr, s);
    if (precision == PREC_FP16)
        return launch_mfma_queues_type<_Float16, _Float16>(c, wfrag, table, qs, n_cu, blocks_per_cu, nullptr, s);
    if (precision == PREC_BF16X3)
        return launch_mfma_queues_type<_Float16, mf::Bf16x3>(c, wfrag, table, qs, n_cu, blocks_per_cu, nullptr, s);
    return hipErrorInvalidValue;
}

// capacity: the queues' total capacity; when it reaches the LDS batch size the
// coarse levels go to k_grid_encode_lds_queues first.
hipError_t launch_grid_encode_queues(const DevCfg& c, int precision, const void* table, const QueueSet& qs, int n_cu,
                                     uint64_t capacity, hipStream_t s) {
    if (qs.n_queues == 0 || c.grid_enc < 0) return hipSuccess;
    if (c.F != 2 && c.F != 4) return hipErrorInvalidValue;
    if (precision != PREC_BF16 && precision != PREC_FP16 && precision != PREC_BF16X3) return hipErrorInvalidValue;
    const bool bf = precision == PREC_BF16;
    const int n_lds = grid_encode_lds_levels(c, table, (uint32_t)std::min<uint64_t>(capacity, 0xffffffffu));
    if (n_lds > 0) {
        uint32_t lds = 0;
        for (int l = 0; l < n_lds; ++l) lds = std::max(lds, (c.offsets[l + 1] - c.offsets[l]) * 4u);
        const dim3 grid_l(256u, (uint32_t)n_lds);
        if (bf)
            hipLaunchKernelGGL((k_grid_encode_lds_queues<__bf16>), grid_l, dim3(kLdsEncodeThreads), lds, s, c,
                               (const __bf16*)table, qs);
        else
            hipLaunchKernelGGL((k_grid_encode_lds_queues<_Float16>), grid_l, dim3(kLdsEncodeThreads), lds, s, c,
                               (const _Float16*)table, qs);
    }
    const dim3 grid((uint32_t)n_cu * 8u);
    if (bf)
        hipLaunchKernelGGL((k_grid_encode_queues<__bf16>), grid, dim3(256), 0, s, c, (const __bf16*)table, qs, n_lds);
    else
        hipLaunchKernelGGL((k_grid_encode_queues<_Float16>), grid, dim3(256), 0, s, c, (const _Float16*)table, qs,
                           n_lds);
    return hipGetLastError();
}

// FP8 fragments of w with per-layer weight scales, then the kernel multipliers
// from the activation maxima amax [n_layers] (fp8 holds wscale, ascale,
// qscale: [n_layers], [n_layers], [n_layers + 1]).
hipError_t launch_pack_frags_fp8(const DevCfg& c, const float* w, void* frag, const float* amax, float* fp8,
                                 hipStream_t s) {
    float* wscale = fp8;
    float* ascale = fp8 + c.n_layers;
    float* qscale = fp8 + 2 * c.n_layers;
    hipLaunchKernelGGL(k_pack_frags_fp8, dim3((uint32_t)c.n_layers), dim3(256), 0, s, c, w, (uint8_t*)frag, wscale);
    hipLaunchKernelGGL(k_fp8_scales, dim3(1), dim3(64), 0, s, c, wscale, amax, ascale, qscale);
    return hipGetLastError();
}

// Calibration maxima: amax[0] = max |encoded input|, amax[l] = max |input of
// layer l| from the FP32 forward's stored activations of m queries.
hipError_t launch_fp8_amax(const DevCfg& c, const float* enc, const float* acts, uint32_t m, float* amax,
                           hipStream_t s) {
    hipError_t e = launch_zero_u32(reinterpret_cast<uint32_t*>(amax), (uint32_t)c.n_layers, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_amax, dim3(256), dim3(256), 0, s, enc, (size_t)c.enc_padded * m, amax);
    for (int l = 1; l < c.n_layers; ++l)
        hipLaunchKernelGGL(k_amax, dim3(256), dim3(256), 0, s, acts + (size_t)(l - 1) * c.width * m,
                           (size_t)c.width * m, amax + l);
    return hipGetLastError();
}

hipError_t launch_pack_frags(const DevCfg& c, int precision, const float* w, void* frag, hipStream_t s) {
    const uint32_t n = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const dim3 grid((n + 255) / 256);
    if (precision == PREC_BF16) hipLaunchKernelGGL((k_pack_frags<__bf16>), grid, dim3(256), 0, s, c, w, (__bf16*)frag);
    else if (precision == PREC_FP16) hipLaunchKernelGGL((k_pack_frags<_Float16>), grid, dim3(256), 0, s, c, w, (_Float16*)frag);
    else if (precision == PREC_BF16X3)
        hipLaunchKernelGGL((k_pack_frags<mf::Bf16x3>), grid, dim3(256), 0, s, c, w, (__bf16*)frag);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_to_half(int precision, const float* src, void* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (precision == PREC_BF16) hipLaunchKernelGGL((k_to_half<__bf16>), grid, dim3(256), 0, s, src, (__bf16*)dst, n);
    else if (precision == PREC_FP16 || precision == PREC_BF16X3)
        hipLaunchKernelGGL((k_to_half<_Float16>), grid, dim3(256), 0, s, src, (_Float16*)dst, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace npmg
