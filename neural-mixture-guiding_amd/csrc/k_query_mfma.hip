// k_query_mfma.hip -- fused query kernel with the MLP on MFMA (BF16 or FP16
// operands, FP32 accumulation), plus the packing kernels that produce its
// weight fragments and half-precision grid tables.
//
// The grid features come from k_grid_encode (level-major, operand type, [L F][n])
// when the grid rows form whole 8-feature chunks; the host pipelines it with the
// fused kernel over chunks of the batch on two streams, since the encode is
// bound by L2 gather bandwidth and the fused kernel by VALU issue.
// One wave handles 64 queries. Lane l = 32 h + j encodes, for queries j and
// 32 + j of the tile, the feature chunks 2 kb + h (8 features each) of every
// K-block kb of the first layer, and hands them to v_mfma_f32_32x32x16 as B
// fragments straight from registers. Weights live in LDS in
// fragment order (1 KB per 32x16 fragment, conflict-free reads). Hidden
// layers chain register to register (see mfma.h). The output tile is split
// between lanes l and l ^ 32 and exchanged with one shuffle, after which
// lane l owns every raw output of query l and runs the FP32 mixture code
// (normaliser, softmax, log-sum-exp and sampling stay FP32).
#include "asg.h"
#include "device.h"
#include "kernels.h"
#include "mfma.h"

namespace npmg {
namespace {

using namespace dev;

// Level-major grid encoding: grid (ceil(n / 256), n_levels). Blocks run in
// linear order, so the whole GPU works on one level at a time and that
// level's table slice (at most 2^19 x F entries) stays resident in each XCD's
// L2 while it is gathered; the features go out in the operand type, [L F][n].
template <class T>
__global__ void __launch_bounds__(256) k_grid_encode(const DevCfg c, const T* __restrict__ table,
                                                     const float* __restrict__ in, T* __restrict__ genc, uint32_t n,
                                                     uint32_t q_begin, uint32_t q_end) {
    const uint32_t q = q_begin + blockIdx.x * 256u + threadIdx.x;
    const int l = blockIdx.y;
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

// Queries [q_begin, q_end) of SoA arrays with row stride n.
// One 64-query tile [q0, min(q0 + 64, q_end)) of a query stream with row stride n:
// encode, MLP on MFMA (weights in LDS at sw), mixture, sample and pdfs.
template <class T, int W, int K, int P>
__device__ __forceinline__ void query_tile_mfma(const DevCfg& c, const T* sw, const T* __restrict__ table,
                                                const T* __restrict__ genc, const QueryIO& io, uint32_t n,
                                                uint32_t q0, uint32_t q_end, int lane) {
    using V = typename mf::Frag<T>::vec;
    constexpr int NOUT = (P * K + 15) / 16 * 16;
    static_assert(NOUT % 32 == 0, "whole 32-row output tiles");
    constexpr int TO = NOUT / 32;  // output tiles
    constexpr int TH = W / 32;     // hidden tiles
    constexpr int KBH = W / 16;    // K-blocks of a hidden layer's input
    const int h = lane >> 5, j = lane & 31;
    const int KB0 = c.enc_padded / 16;
    const uint32_t qa = min(q0 + (uint32_t)j, q_end - 1u), qb = min(q0 + 32u + (uint32_t)j, q_end - 1u);
    float ina[16], inb[16];
#pragma unroll
    for (int d = 0; d < 16; ++d) {
        ina[d] = d < c.n_in ? io.in[(size_t)d * n + qa] : 0.0f;
        inb[d] = d < c.n_in ? io.in[(size_t)d * n + qb] : 0.0f;
    }
    mf::f32x16 acc[TH][2];
#pragma unroll
    for (int t = 0; t < TH; ++t) acc[t][0] = acc[t][1] = mf::f32x16{};
    // Layer 0: encode chunk by chunk straight into B fragments.
    for (int kb = 0; kb < KB0; ++kb) {
        float va[8], vb[8];
        encode_chunk<T>(c, table, ina, 2 * kb + h, va, genc, qa, n);
        encode_chunk<T>(c, table, inb, 2 * kb + h, vb, genc, qb, n);
        const V ba = mf::pack8<T>(va), bb = mf::pack8<T>(vb);
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            const V a = mf::load_frag<T>(sw, t * KB0 + kb, lane);
            acc[t][0] = mf::Frag<T>::mma(a, ba, acc[t][0]);
            acc[t][1] = mf::Frag<T>::mma(a, bb, acc[t][1]);
        }
    }
    int fb = TH * KB0;
    // Hidden layers: ReLU, round, and feed the accumulators back as B fragments.
    for (int l = 1; l < c.n_layers - 1; ++l) {
        V bx[2][KBH];
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                float v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float a = acc[kb >> 1][cc][8 * (kb & 1) + e];
                    v[e] = a > 0.0f ? a : 0.0f;
                }
                bx[cc][kb] = mf::pack8<T>(v);
            }
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            acc[t][0] = acc[t][1] = mf::f32x16{};
#pragma unroll
            for (int kb = 0; kb < KBH; ++kb) {
                const V a = mf::load_frag<T>(sw, fb + t * KBH + kb, lane);
                acc[t][0] = mf::Frag<T>::mma(a, bx[0][kb], acc[t][0]);
                acc[t][1] = mf::Frag<T>::mma(a, bx[1][kb], acc[t][1]);
            }
        }
        fb += TH * KBH;
    }
    // Output layer.
    V bx[2][KBH];
#pragma unroll
    for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float a = acc[kb >> 1][cc][8 * (kb & 1) + e];
                v[e] = a > 0.0f ? a : 0.0f;
            }
            bx[cc][kb] = mf::pack8<T>(v);
        }
    float out[NOUT];
#pragma unroll
    for (int t = 0; t < TO; ++t) {
        mf::f32x16 o0 = {}, o1 = {};
#pragma unroll
        for (int kb = 0; kb < KBH; ++kb) {
            const V a = mf::load_frag<T>(sw, fb + t * KBH + kb, lane);
            o0 = mf::Frag<T>::mma(a, bx[0][kb], o0);
            o1 = mf::Frag<T>::mma(a, bx[1][kb], o1);
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
            if (io.u || io.n_eval > 0) asg_query<K>([&](int jj) { return out[jj]; }, io, q, n);
        } else if (io.u || io.n_eval > 0) {
            Mixture<K> m;
            build_mixture<K>(c, [&](int jj) { return out[jj]; }, m);
            float th[K], ph[K];
            if (io.normal) {
                const V3 nrm = mk(io.normal[q], io.normal[(size_t)n + q], io.normal[2 * (size_t)n + q]);
                apply_cosine<K>(c, nrm, m, false, th, ph);
            }
            if (io.u) {
                uint32_t lobe;
                const V3 d = mixture_sample<K>(c, m, io.u[q], io.u[(size_t)n + q], io.u[2 * (size_t)n + q], &lobe);
                io.out_dir[q] = d.x;
                io.out_dir[(size_t)n + q] = d.y;
                io.out_dir[2 * (size_t)n + q] = d.z;
                io.out_pdf[q] = mixture_pdf<K>(c, m, d);
                if (io.out_lobe) io.out_lobe[q] = lobe;
            }
            for (int e = 0; e < io.n_eval; ++e) {
                const float* ed = io.eval_dirs + (size_t)e * 3 * n;
                io.out_eval_pdf[(size_t)e * n + q] =
                    mixture_pdf<K>(c, m, mk(ed[q], ed[(size_t)n + q], ed[2 * (size_t)n + q]));
            }
        }
    }
}

// Queries [q_begin, q_end) of SoA arrays with row stride n.
template <class T, int W, int K, int P>
__global__ void __launch_bounds__(256) k_query_mfma(const DevCfg c, const T* __restrict__ wfrag,
                                                    const T* __restrict__ table, const T* __restrict__ genc,
                                                    const QueryIO io, uint32_t n, uint32_t n_frag, uint32_t q_begin,
                                                    uint32_t q_end) {
    extern __shared__ uint4 smem_frag[];
    for (uint32_t i = threadIdx.x; i < n_frag * 64u; i += 256u) smem_frag[i] = reinterpret_cast<const uint4*>(wfrag)[i];
    __syncthreads();
    const T* sw = reinterpret_cast<const T*>(smem_frag);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t n_tiles = (q_end - q_begin + 63u) / 64u;
    for (uint32_t tile = blockIdx.x * 4u + (uint32_t)wave; tile < n_tiles; tile += gridDim.x * 4u)
        query_tile_mfma<T, W, K, P>(c, sw, table, genc, io, n, q_begin + tile * 64u, q_end, lane);
}

// Several queues in one persistent launch: each wave pulls 64-query tiles
// from a device counter over the concatenated tiles of the queues (counts on
// the device) until all are taken.
template <class T, int W, int K, int P>
__global__ void __launch_bounds__(256) k_query_mfma_queues(const DevCfg c, const T* __restrict__ wfrag,
                                                           const T* __restrict__ table, const QueueSet qs,
                                                           uint32_t n_frag) {
    extern __shared__ uint4 smem_frag[];
    for (uint32_t i = threadIdx.x; i < n_frag * 64u; i += 256u) smem_frag[i] = reinterpret_cast<const uint4*>(wfrag)[i];
    __syncthreads();
    const T* sw = reinterpret_cast<const T*>(smem_frag);
    const int lane = threadIdx.x & 63;
    uint32_t first[kMaxQueues + 1];
    queue_tile_prefix(qs, 64u, first);
    const uint32_t total = first[kMaxQueues];
    for (;;) {
        uint32_t w = 0;
        if (lane == 0) w = atomicAdd(qs.counter, 1u);
        w = __shfl(w, 0);  // wave-uniform work item; every wave leaves once the items run out
        if (w >= total) break;
        int qi = 0;
        uint32_t base = 0;
#pragma unroll
        for (int i = 1; i < kMaxQueues; ++i)
            if (w >= first[i] && first[i] < first[i + 1]) {
                qi = i;
                base = first[i];
            }
        query_tile_mfma<T, W, K, P>(c, sw, table, static_cast<const T*>(qs.genc[qi]), qs.io[qi], qs.stride[qi],
                                    (w - base) * 64u, min(qs.counts[qi], qs.stride[qi]), lane);
    }
}

// Level-major grid encoding of several queues in one persistent launch:
// workgroups pull (level, 256-query tile) items in level-major order.
template <class T>
__global__ void __launch_bounds__(256) k_grid_encode_queues(const DevCfg c, const T* __restrict__ table,
                                                            const QueueSet qs) {
    __shared__ uint32_t item;
    uint32_t first[kMaxQueues + 1];
    queue_tile_prefix(qs, 256u, first);
    const uint32_t per_level = first[kMaxQueues];
    const uint32_t total = per_level * (uint32_t)c.n_levels;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(qs.counter, 1u);
        __syncthreads();
        const uint32_t w = item;
        __syncthreads();
        if (w >= total) break;
        const int l = (int)(w / per_level);
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
        const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
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

// Weight fragments from full-precision row-major matrices: fragment
// (layer, t, kb), lane (h, i), element e holds W_l[32 t + i][col] with
// col = 16 kb + 8 h + e on the first layer and the chained permutation after.
template <class T>
__global__ void __launch_bounds__(256) k_pack_frags(const DevCfg c, const float* __restrict__ w, T* __restrict__ frag) {
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
            frag[gid] = (T)w[c.layer_off[l] + (32u * t + i) * iw + col];
            return;
        }
        base += cnt;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_to_half(const float* __restrict__ src, T* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) dst[i] = (T)src[i];
}

template <class T, int W, int K, int P>
hipError_t launch_mfma_t(const DevCfg& c, const void* wfrag, const void* table, const void* genc, const QueryIO& io,
                         uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu, int bpc, hipStream_t s) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const uint32_t n_frag = n_mlp / 512u;
    const size_t lds = (size_t)n_frag * 1024u;
    auto kern = k_query_mfma<T, W, K, P>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (bpc > 0 && bpc < per_cu) per_cu = bpc;
    const uint32_t tiles = (q_end - q_begin + 63u) / 64u;
    const uint32_t want = (tiles + 3u) / 4u;
    const uint32_t grid = want < (uint32_t)(n_cu * per_cu) ? want : (uint32_t)(n_cu * per_cu);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, c, static_cast<const T*>(wfrag),
                       static_cast<const T*>(table), static_cast<const T*>(genc), io, n, n_frag, q_begin, q_end);
    return hipGetLastError();
}

template <class T, int W, int K, int P>
hipError_t launch_mfma_queues_t(const DevCfg& c, const void* wfrag, const void* table, const QueueSet& qs, int n_cu,
                                int blocks_per_cu, hipStream_t s) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const uint32_t n_frag = n_mlp / 512u;
    const size_t lds = (size_t)n_frag * 1024u;
    auto kern = k_query_mfma_queues<T, W, K, P>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < per_cu) per_cu = blocks_per_cu;
    hipLaunchKernelGGL(kern, dim3((uint32_t)(n_cu * per_cu)), dim3(256), lds, s, c, static_cast<const T*>(wfrag),
                       static_cast<const T*>(table), qs, n_frag);
    return hipGetLastError();
}

template <class T>
hipError_t launch_mfma_queues_type(const DevCfg& c, const void* wfrag, const void* table, const QueueSet& qs,
                                   int n_cu, int bpc, hipStream_t s) {
#define NPMG_QQ(WW, KK, PP)                                \
    if (c.width == WW && c.K == KK && c.lobe_params == PP) \
        return launch_mfma_queues_t<T, WW, KK, PP>(c, wfrag, table, qs, n_cu, bpc, s);
    NPMG_QQ(64, 8, 4)
    NPMG_QQ(64, 16, 4)
    NPMG_QQ(128, 8, 4)
    NPMG_QQ(128, 16, 4)
    NPMG_QQ(64, 16, 6)
    NPMG_QQ(128, 16, 6)
#undef NPMG_QQ
    return hipErrorInvalidValue;
}

template <class T>
hipError_t launch_mfma_type(const DevCfg& c, const void* wfrag, const void* table, const void* genc,
                            const QueryIO& io, uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu, int bpc,
                            hipStream_t s) {
#define NPMG_QM(WW, KK, PP)                                      \
    if (c.width == WW && c.K == KK && c.lobe_params == PP)       \
        return launch_mfma_t<T, WW, KK, PP>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, bpc, s);
    NPMG_QM(64, 8, 4)
    NPMG_QM(64, 16, 4)
    NPMG_QM(128, 8, 4)
    NPMG_QM(128, 16, 4)
    NPMG_QM(64, 16, 6)
    NPMG_QM(128, 16, 6)
#undef NPMG_QM
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_grid_encode_range(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                                    uint32_t n, uint32_t q_begin, uint32_t q_end, hipStream_t s) {
    if (q_end <= q_begin) return hipSuccess;
    const dim3 grid((q_end - q_begin + 255) / 256, c.n_levels);
    if (precision == PREC_BF16)
        hipLaunchKernelGGL((k_grid_encode<__bf16>), grid, dim3(256), 0, s, c, (const __bf16*)table, in, (__bf16*)genc,
                           n, q_begin, q_end);
    else if (precision == PREC_FP16)
        hipLaunchKernelGGL((k_grid_encode<_Float16>), grid, dim3(256), 0, s, c, (const _Float16*)table, in,
                           (_Float16*)genc, n, q_begin, q_end);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_grid_encode(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                              uint32_t n, hipStream_t s) {
    return launch_grid_encode_range(c, precision, table, in, genc, n, 0u, n, s);
}

bool mfma_split_grid(const DevCfg& c) {
    return c.grid_enc >= 0 && c.enc_offset[c.grid_enc] % 8 == 0 && (c.n_levels * c.F) % 8 == 0;
}

// The fused kernel over queries [q_begin, q_end); with genc, the grid features
// of those queries must already be there (launch_grid_encode_range).
hipError_t launch_query_mfma(const DevCfg& c, int precision, const void* wfrag, const void* table, const void* genc,
                             const QueryIO& io, uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu,
                             int blocks_per_cu, hipStream_t s) {
    if (q_end <= q_begin) return hipSuccess;
    if (q_end > n || c.enc_padded % 16 != 0) return hipErrorInvalidValue;
    if (genc && !mfma_split_grid(c)) return hipErrorInvalidValue;
    if (precision == PREC_BF16)
        return launch_mfma_type<__bf16>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, blocks_per_cu, s);
    if (precision == PREC_FP16)
        return launch_mfma_type<_Float16>(c, wfrag, table, genc, io, n, q_begin, q_end, n_cu, blocks_per_cu, s);
    return hipErrorInvalidValue;
}

// Persistent launches over several queues (qs.counter must be zero).
hipError_t launch_query_mfma_queues(const DevCfg& c, int precision, const void* wfrag, const void* table,
                                    const QueueSet& qs, int n_cu, int blocks_per_cu, hipStream_t s) {
    if (qs.n_queues == 0) return hipSuccess;
    if (qs.n_queues > (uint32_t)kMaxQueues || c.enc_padded % 16 != 0) return hipErrorInvalidValue;
    if (precision == PREC_BF16) return launch_mfma_queues_type<__bf16>(c, wfrag, table, qs, n_cu, blocks_per_cu, s);
    if (precision == PREC_FP16) return launch_mfma_queues_type<_Float16>(c, wfrag, table, qs, n_cu, blocks_per_cu, s);
    return hipErrorInvalidValue;
}

hipError_t launch_grid_encode_queues(const DevCfg& c, int precision, const void* table, const QueueSet& qs, int n_cu,
                                     hipStream_t s) {
    if (qs.n_queues == 0 || c.grid_enc < 0) return hipSuccess;
    if (c.F != 2 && c.F != 4) return hipErrorInvalidValue;
    const dim3 grid((uint32_t)n_cu * 8u);
    if (precision == PREC_BF16)
        hipLaunchKernelGGL((k_grid_encode_queues<__bf16>), grid, dim3(256), 0, s, c, (const __bf16*)table, qs);
    else if (precision == PREC_FP16)
        hipLaunchKernelGGL((k_grid_encode_queues<_Float16>), grid, dim3(256), 0, s, c, (const _Float16*)table, qs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_pack_frags(const DevCfg& c, int precision, const float* w, void* frag, hipStream_t s) {
    const uint32_t n = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const dim3 grid((n + 255) / 256);
    if (precision == PREC_BF16) hipLaunchKernelGGL((k_pack_frags<__bf16>), grid, dim3(256), 0, s, c, w, (__bf16*)frag);
    else if (precision == PREC_FP16) hipLaunchKernelGGL((k_pack_frags<_Float16>), grid, dim3(256), 0, s, c, w, (_Float16*)frag);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_to_half(int precision, const float* src, void* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (precision == PREC_BF16) hipLaunchKernelGGL((k_to_half<__bf16>), grid, dim3(256), 0, s, src, (__bf16*)dst, n);
    else if (precision == PREC_FP16) hipLaunchKernelGGL((k_to_half<_Float16>), grid, dim3(256), 0, s, src, (_Float16*)dst, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace npmg
