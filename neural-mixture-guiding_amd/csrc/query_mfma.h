// query_mfma.h -- the fused MFMA query kernel (included by
// k_query_mfma_inst.hip only, which instantiates it per operand type and lobe
// family so the instantiations compile in parallel).
//
// One wave handles 64 queries. Lane l = 32 h + j encodes, for queries j and
// 32 + j of the tile, the feature chunks 2 kb + h (8 features each) of every
// K-block kb of the first layer, and hands them to v_mfma_f32_32x32x16 as B
// fragments straight from registers. Weights live in LDS in
// fragment order (1 KB per 32x16 fragment, conflict-free reads). Hidden
// layers chain register to register (see mfma.h). The output tile is split
// between lanes l and l ^ 32 and exchanged with one shuffle, after which
// lane l owns every raw output of query l and runs the FP32 mixture code
// (normaliser, softmax, log-sum-exp and sampling stay FP32).
#pragma once

#include <type_traits>

#include "asg.h"
#include "device.h"
#include "kernels.h"
#include "mfma.h"

namespace npmg {
namespace {

using namespace dev;


// Layer 0 of the "grid, then OneBlob(16)" layout (DevCfg::ob_nkb > 0) for a
// 16-bit operand type: the grid K-blocks straight from the level-major
// features, then one K-block of mf::oneblob16_frag bins per OneBlob input dim.
// G is the type of the level-major features, T the operand type: the same
// 16-bit type, or the split BF16 operand built from FP16 features (exact: an
// FP16 value has 11 significant bits, hi and lo hold 16).
template <class G, class T, int TH>
__device__ __forceinline__ void layer0_oneblob16(const DevCfg& c, const T* sw, const G* __restrict__ genc,
                                                 const QueryIO& io, uint32_t n, uint32_t qa, uint32_t qb, int lane,
                                                 mf::f32x16 (&acc)[TH][2]) {
    using V = typename mf::Frag<T>::vec;
    const int h = lane >> 5, KB0 = c.enc_padded / 16;
    for (int kb = 0; kb < c.ob_kb0; ++kb) {
        V ba, bb;
        const G* src = genc + (size_t)(16 * kb + 8 * h) * n;
        if constexpr (std::is_same_v<G, T>) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                ba[i] = src[(size_t)i * n + qa];
                bb[i] = src[(size_t)i * n + qb];
            }
        } else {
            float va[8], vb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                va[i] = (float)src[(size_t)i * n + qa];
                vb[i] = (float)src[(size_t)i * n + qb];
            }
            ba = mf::pack8<T>(va);
            bb = mf::pack8<T>(vb);
        }
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            const V a = mf::load_frag<T>(sw, t * KB0 + kb, lane);
            acc[t][0] = mf::Frag<T>::mma(a, ba, acc[t][0]);
            acc[t][1] = mf::Frag<T>::mma(a, bb, acc[t][1]);
        }
    }
    for (int d = 0; d < c.ob_nkb; ++d) {
        const float* row = io.in + (size_t)(c.ob_in + d) * n;
        const V ba = mf::oneblob16_frag<T>(row[qa], h), bb = mf::oneblob16_frag<T>(row[qb], h);
        const int kb = c.ob_kb0 + d;
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            const V a = mf::load_frag<T>(sw, t * KB0 + kb, lane);
            acc[t][0] = mf::Frag<T>::mma(a, ba, acc[t][0]);
            acc[t][1] = mf::Frag<T>::mma(a, bb, acc[t][1]);
        }
    }
}

// The same 8 OneBlob(16) bins as float values (bin 8h + i is m[kx - 1],
// m[kx] or m[kx + 1] when (8h + i - kx + 1) mod 16 is 0, 1 or 2, else 0), for
// the FP8 operands, which scale and round them like any other feature.
__device__ __forceinline__ void oneblob16_bins(float x, int h, float (&v)[8]) {
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
    const int base = 8 * h + 1 - kx;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t r = (uint32_t)(base + i) & 15u;
        v[i] = r == 0u ? mp : (r == 1u ? mc : (r == 2u ? mn : 0.0f));
    }
}

// Layer 0 of the "grid, then OneBlob(16)" layout with FP8 E4M3 operands:
// features scaled by qscale0 and clamped to the E4M3 range, as in the general
// path; the grid K-blocks come from the level-major BF16 features.
template <class T, int TH>
__device__ __forceinline__ void layer0_oneblob16_fp8(const DevCfg& c, const mf::Fp8* sw, const T* __restrict__ genc,
                                                     const QueryIO& io, uint32_t n, uint32_t qa, uint32_t qb, int lane,
                                                     float qscale0, mf::f32x16 (&acc)[TH][2]) {
    using V = typename mf::Frag<mf::Fp8>::vec;
    const int h = lane >> 5, KB0 = c.enc_padded / 16;
    auto put = [&](float (&va)[8], float (&vb)[8], int kb) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            va[e] = __builtin_amdgcn_fmed3f(va[e] * qscale0, -mf::kFp8Max, mf::kFp8Max);
            vb[e] = __builtin_amdgcn_fmed3f(vb[e] * qscale0, -mf::kFp8Max, mf::kFp8Max);
        }
        const V ba = mf::pack8<mf::Fp8>(va), bb = mf::pack8<mf::Fp8>(vb);
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            const V a = mf::load_frag<mf::Fp8>(sw, t * KB0 + kb, lane);
            acc[t][0] = mf::Frag<mf::Fp8>::mma(a, ba, acc[t][0]);
            acc[t][1] = mf::Frag<mf::Fp8>::mma(a, bb, acc[t][1]);
        }
    };
    for (int kb = 0; kb < c.ob_kb0; ++kb) {
        float va[8], vb[8];
        const T* src = genc + (size_t)(16 * kb + 8 * h) * n;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            va[i] = (float)src[(size_t)i * n + qa];
            vb[i] = (float)src[(size_t)i * n + qb];
        }
        put(va, vb, kb);
    }
    for (int d = 0; d < c.ob_nkb; ++d) {
        const float* row = io.in + (size_t)(c.ob_in + d) * n;
        float va[8], vb[8];
        oneblob16_bins(row[qa], h, va);
        oneblob16_bins(row[qb], h, vb);
        put(va, vb, c.ob_kb0 + d);
    }
}

// Queries [q_begin, q_end) of SoA arrays with row stride n.
// One 64-query tile [q0, min(q0 + 64, q_end)) of a query stream with row stride n:
// encode, MLP on MFMA (weights in LDS at sw), mixture, sample and pdfs.
//
// O is the MFMA operand type: T itself, or mf::Fp8 for FP8 inference, where
// qscale holds the per-layer power-of-two multipliers: [0] input features ->
// E4M3, [l] accumulators of layer l - 1 -> E4M3 inputs of layer l, [n_layers]
// output accumulators -> raw outputs.
template <class T, class O, int W, int K, int P, bool FAST>
__device__ __forceinline__ void query_tile_mfma(const DevCfg& c, const O* sw, const T* __restrict__ table,
                                                const T* __restrict__ genc, const QueryIO& io, uint32_t n,
                                                uint32_t q0, uint32_t q_end, int lane,
                                                const float* __restrict__ qscale) {
    using V = typename mf::Frag<O>::vec;
    constexpr bool F8 = std::is_same_v<O, mf::Fp8>;
    constexpr int NOUT = (P * K + 15) / 16 * 16;
    static_assert(NOUT % 32 == 0, "whole 32-row output tiles");
    constexpr int TO = NOUT / 32;  // output tiles
    constexpr int TH = W / 32;     // hidden tiles
    constexpr int KBH = W / 16;    // K-blocks of a hidden layer's input
    const int h = lane >> 5, j = lane & 31;
    const int KB0 = c.enc_padded / 16;
    const uint32_t qa = min(q0 + (uint32_t)j, q_end - 1u), qb = min(q0 + 32u + (uint32_t)j, q_end - 1u);
    constexpr bool fast = FAST;  // fast OneBlob bins and mixture ("numerics": "fast")
    mf::f32x16 acc[TH][2];
#pragma unroll
    for (int t = 0; t < TH; ++t) acc[t][0] = acc[t][1] = mf::f32x16{};
    bool layer0_done = false;
    if constexpr (FAST) {
        if (c.ob_nkb > 0 && genc) {
            if constexpr (F8) layer0_oneblob16_fp8<T, TH>(c, sw, genc, io, n, qa, qb, lane, qscale[0], acc);
            else layer0_oneblob16<T, O, TH>(c, sw, genc, io, n, qa, qb, lane, acc);
            layer0_done = true;
        }
    }
    if (!layer0_done) {
        float ina[16], inb[16];
#pragma unroll
        for (int d = 0; d < 16; ++d) {
            ina[d] = d < c.n_in ? io.in[(size_t)d * n + qa] : 0.0f;
            inb[d] = d < c.n_in ? io.in[(size_t)d * n + qb] : 0.0f;
        }
        // Layer 0: encode chunk by chunk straight into B fragments.
        for (int kb = 0; kb < KB0; ++kb) {
            float va[8], vb[8];
            encode_chunk<T>(c, table, ina, 2 * kb + h, va, genc, qa, n, fast);
            encode_chunk<T>(c, table, inb, 2 * kb + h, vb, genc, qb, n, fast);
            if constexpr (F8) {
                const float s0 = qscale[0];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    va[e] = __builtin_amdgcn_fmed3f(va[e] * s0, -mf::kFp8Max, mf::kFp8Max);
                    vb[e] = __builtin_amdgcn_fmed3f(vb[e] * s0, -mf::kFp8Max, mf::kFp8Max);
                }
            }
            const V ba = mf::pack8<O>(va), bb = mf::pack8<O>(vb);
#pragma unroll
            for (int t = 0; t < TH; ++t) {
                const V a = mf::load_frag<O>(sw, t * KB0 + kb, lane);
                acc[t][0] = mf::Frag<O>::mma(a, ba, acc[t][0]);
                acc[t][1] = mf::Frag<O>::mma(a, bb, acc[t][1]);
            }
        }
    }
    int fb = TH * KB0;
    // Hidden layers: ReLU, round, and feed the accumulators back as B fragments.
    for (int l = 1; l < c.n_layers - 1; ++l) {
        const float m = F8 ? qscale[l] : 1.0f;
        V bx[2][KBH];
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                float v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float a = acc[kb >> 1][cc][8 * (kb & 1) + e];
                    if constexpr (F8) v[e] = __builtin_amdgcn_fmed3f(a * m, 0.0f, mf::kFp8Max);
                    else v[e] = a > 0.0f ? a : 0.0f;
                }
                bx[cc][kb] = mf::pack8<O>(v);
            }
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            acc[t][0] = acc[t][1] = mf::f32x16{};
#pragma unroll
            for (int kb = 0; kb < KBH; ++kb) {
                const V a = mf::load_frag<O>(sw, fb + t * KBH + kb, lane);
                acc[t][0] = mf::Frag<O>::mma(a, bx[0][kb], acc[t][0]);
                acc[t][1] = mf::Frag<O>::mma(a, bx[1][kb], acc[t][1]);
            }
        }
        fb += TH * KBH;
    }
    // Output layer.
    const float m_last = F8 ? qscale[c.n_layers - 1] : 1.0f;
    V bx[2][KBH];
#pragma unroll
    for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float a = acc[kb >> 1][cc][8 * (kb & 1) + e];
                if constexpr (F8) v[e] = __builtin_amdgcn_fmed3f(a * m_last, 0.0f, mf::kFp8Max);
                else v[e] = a > 0.0f ? a : 0.0f;
            }
            bx[cc][kb] = mf::pack8<O>(v);
        }
    const float m_out = F8 ? qscale[c.n_layers] : 1.0f;
    float out[NOUT];
#pragma unroll
    for (int t = 0; t < TO; ++t) {
        mf::f32x16 o0 = {}, o1 = {};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
            const V a = mf::load_frag<O>(sw, fb + t * KBH + kb, lane);
            o0 = mf::Frag<O>::mma(a, bx[0][kb], o0);
            o1 = mf::Frag<O>::mma(a, bx[1][kb], o1);
        }
        if constexpr (F8) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                o0[r] *= m_out;
                o1[r] *= m_out;
            }
        }
        // Lane (h, j) keeps the tile of its own query (c = h) and swaps the other with lane l ^ 32.
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float mine = h ? o1[r] : o0[r];
            const float recv = __shfl_xor(h ? o0[r] : o1[r], 32);
            const int a = r >> 2, b = r & 3;
            out[32 * t + 8 * a + b] = h ? recv : mine;
            out[32 * t + 8 * a + 4 + b] = h ? mine : recv;
        }
    }
    const uint32_t q = q0 + (uint32_t)lane;
    if (q < q_end) {
        if (io.raw_out) {
#pragma unroll
            for (int o = 0; o < NOUT; ++o) io.raw_out[(size_t)o * n + q] = out[o];
        }
        if constexpr (P == 6) {
            if (io.u || io.n_eval > 0)
                asg_query_m<K, std::conditional_t<FAST, FastMath, ExactMath>>([&](int jj) { return out[jj]; }, io, q,
                                                                               n);
        } else if (io.u || io.n_eval > 0 || io.out_sel) {
            vmf_query<K, std::conditional_t<FAST, FastMath, ExactMath>, false>(c, [&](int jj) { return out[jj]; }, io,
                                                                                q, n);
        }
    }
}

// Waves per workgroup: the weights of a 128-wide network fill most of a CU's
// LDS (120 KB in BF16), so one workgroup per CU is all that fits; 8 waves
// then put two waves on each SIMD to hide the MFMA and memory latencies. The
// two weight planes of the split BF16 operands (72 KB at width 64) get 8 waves
// as well, so that one workgroup per CU leaves LDS for the grid encode that
// runs beside it in the pipelined query path.
template <int W, class O>
constexpr int query_waves() { return W >= 128 || mf::Op<O>::planes == 2 ? 8 : 4; }

// Queries [q_begin, q_end) of SoA arrays with row stride n.
template <class T, class O, int W, int K, int P, bool FAST>
__global__ void __launch_bounds__((64 * query_waves<W, O>())) k_query_mfma(const DevCfg c, const O* __restrict__ wfrag,
                                                    const T* __restrict__ table, const T* __restrict__ genc,
                                                    const QueryIO io, uint32_t n, uint32_t n_frag, uint32_t q_begin,
                                                    uint32_t q_end, const float* __restrict__ qscale) {
    extern __shared__ uint4 smem_frag[];
    constexpr uint32_t u4_per_frag = 64u * sizeof(typename mf::Frag<O>::vec) / 16u;
    constexpr uint32_t NW = query_waves<W, O>();
    for (uint32_t i = threadIdx.x; i < n_frag * u4_per_frag; i += 64u * NW)
        smem_frag[i] = reinterpret_cast<const uint4*>(wfrag)[i];
    __syncthreads();
    const O* sw = reinterpret_cast<const O*>(smem_frag);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t n_tiles = (q_end - q_begin + 63u) / 64u;
    for (uint32_t tile = blockIdx.x * NW + (uint32_t)wave; tile < n_tiles; tile += gridDim.x * NW)
        query_tile_mfma<T, O, W, K, P, FAST>(c, sw, table, genc, io, n, q_begin + tile * 64u, q_end, lane, qscale);
}

// Several queues in one persistent launch: each wave pulls 64-query tiles
// from a device counter over the concatenated tiles of the queues (counts on
// the device) until all are taken, qs.batch consecutive tiles per atomic
// (every wave shares the one counter, whose same-address atomics serialise;
// the autotuner picks the batch).

template <class T, class O, int W, int K, int P, bool FAST>
__global__ void __launch_bounds__((64 * query_waves<W, O>())) k_query_mfma_queues(const DevCfg c, const O* __restrict__ wfrag,
                                                           const T* __restrict__ table, const QueueSet qs,
                                                           uint32_t n_frag, const float* __restrict__ qscale) {
    extern __shared__ uint4 smem_frag[];
    constexpr uint32_t u4_per_frag = 64u * sizeof(typename mf::Frag<O>::vec) / 16u;
    for (uint32_t i = threadIdx.x; i < n_frag * u4_per_frag; i += 64u * query_waves<W, O>())
        smem_frag[i] = reinterpret_cast<const uint4*>(wfrag)[i];
    __syncthreads();
    const O* sw = reinterpret_cast<const O*>(smem_frag);
    const int lane = threadIdx.x & 63;
    uint32_t first[kMaxQueues + 1];
    queue_tile_prefix(qs, 64u, first);
    const uint32_t total = first[kMaxQueues];
    for (;;) {
        uint32_t w0 = 0;
        if (lane == 0) w0 = atomicAdd(qs.counter, qs.batch);
        w0 = __shfl(w0, 0);  // wave-uniform first tile; every wave leaves once the items run out
        if (w0 >= total) break;
        const uint32_t w1 = min(total, w0 + qs.batch);
        for (uint32_t w = w0; w < w1; ++w) {
            int qi = 0;
            uint32_t base = 0;
#pragma unroll
            for (int i = 1; i < kMaxQueues; ++i)
                if (w >= first[i] && first[i] < first[i + 1]) {
                    qi = i;
                    base = first[i];
                }
            query_tile_mfma<T, O, W, K, P, FAST>(c, sw, table, static_cast<const T*>(qs.genc[qi]), qs.io[qi],
                                                 qs.stride[qi], (w - base) * 64u, min(qs.counts[qi], qs.stride[qi]),
                                                 lane, qscale);
        }
    }
}

}  // namespace

template <class T, class O, int W, int K, int P, bool FAST>
hipError_t launch_mfma_t(const DevCfg& c, const void* wfrag, const void* table, const void* genc, const QueryIO& io,
                         uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu, int bpc, const float* qscale,
                         hipStream_t s) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const uint32_t n_frag = n_mlp / 512u;
    const size_t lds = (size_t)n_frag * 64u * sizeof(typename mf::Frag<O>::vec);
    auto kern = k_query_mfma<T, O, W, K, P, FAST>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * query_waves<W, O>(), lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (bpc > 0 && bpc < per_cu) per_cu = bpc;
    const uint32_t tiles = (q_end - q_begin + 63u) / 64u;
    constexpr uint32_t NW = query_waves<W, O>();
    const uint32_t want = (tiles + NW - 1u) / NW;
    const uint32_t grid = want < (uint32_t)(n_cu * per_cu) ? want : (uint32_t)(n_cu * per_cu);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * NW), lds, s, c, static_cast<const O*>(wfrag),
                       static_cast<const T*>(table), static_cast<const T*>(genc), io, n, n_frag, q_begin, q_end, qscale);
    return hipGetLastError();
}

template <class T, class O, int W, int K, int P, bool FAST>
hipError_t launch_mfma_queues_t(const DevCfg& c, const void* wfrag, const void* table, const QueueSet& qs, int n_cu,
                                int blocks_per_cu, const float* qscale, hipStream_t s) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const uint32_t n_frag = n_mlp / 512u;
    const size_t lds = (size_t)n_frag * 64u * sizeof(typename mf::Frag<O>::vec);
    auto kern = k_query_mfma_queues<T, O, W, K, P, FAST>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * query_waves<W, O>(), lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < per_cu) per_cu = blocks_per_cu;
    hipLaunchKernelGGL(kern, dim3((uint32_t)(n_cu * per_cu)), dim3(64 * query_waves<W, O>()), lds, s, c,
                       static_cast<const O*>(wfrag),
                       static_cast<const T*>(table), qs, n_frag, qscale);
    return hipGetLastError();
}

}  // namespace npmg
