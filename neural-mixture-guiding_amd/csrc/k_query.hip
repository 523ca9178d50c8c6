// k_query.hip -- FP32 fused query kernel, consumer-side mixture kernel,
// encoding and grid-corner kernels.
//
// FP32 query kernel layout: one thread per query, 256 threads per block,
// blocks persistent over tiles of 256 queries. The MLP weights, transposed per
// layer to [in][out], are staged once per block into LDS and read as float4
// broadcasts; each thread keeps its current layer's accumulators in VGPRs and
// its previous layer's activations in a private LDS column ([width][256],
// lane-consecutive, so conflict-free). Encoded features are streamed straight
// into the first layer as they are produced, so the encoding never touches
// HBM. Accumulation order over inputs is ascending, as in the oracle.
#include "asg.h"
#include "device.h"
#include "kernels.h"

namespace npmg {
namespace {

using namespace dev;

template <int W>
constexpr int block_size() { return W == 64 ? 256 : 128; }

__host__ __device__ __forceinline__ uint32_t n_mlp_of(const DevCfg& c) {
    const int L = c.n_layers - 1;
    return c.layer_off[L] + (uint32_t)(c.layer_in[L] * c.layer_out[L]);
}

// acc[o] = fma(w[o], v, acc[o]) for o in [0, N), w a 16-byte aligned row.
template <int N>
__device__ __forceinline__ void axpy_row(const float* __restrict__ w, float v, float* acc) {
    const float4* w4 = reinterpret_cast<const float4*>(w);
#pragma unroll
    for (int o4 = 0; o4 < N / 4; ++o4) {
        const float4 ww = w4[o4];
        acc[4 * o4 + 0] = fm::fma_(ww.x, v, acc[4 * o4 + 0]);
        acc[4 * o4 + 1] = fm::fma_(ww.y, v, acc[4 * o4 + 1]);
        acc[4 * o4 + 2] = fm::fma_(ww.z, v, acc[4 * o4 + 2]);
        acc[4 * o4 + 3] = fm::fma_(ww.w, v, acc[4 * o4 + 3]);
    }
}

// Network forward for one query up to the raw outputs. sw: transposed weights
// (LDS or global), act: this thread's activation column (stride BS).
// If store_acts, the post-ReLU hidden activations go to acts_out[(l*W+o)*n+q]
// and the encoded features to enc_out[j*n+q].
template <int W, int NOUT, int BS>
__device__ __forceinline__ void mlp_forward(const DevCfg& c, const float* __restrict__ sw,
                                            const float* __restrict__ table, const float* qin, float* act,
                                            float* out, float* enc_out, float* acts_out, uint32_t q, uint32_t n) {
    float acc[W];
#pragma unroll
    for (int o = 0; o < W; ++o) acc[o] = 0.0f;
    int j = 0;
    encode_query(c, table, qin, [&](float v) {
        if (enc_out) enc_out[(size_t)j * n + q] = v;
        axpy_row<W>(sw + (size_t)j * W, v, acc);
        ++j;
    });
    for (int l = 1; l < c.n_layers; ++l) {
#pragma unroll
        for (int o = 0; o < W; ++o) {
            const float a = acc[o] > 0.0f ? acc[o] : 0.0f;
            act[o * BS] = a;
            if (acts_out) acts_out[((size_t)(l - 1) * W + o) * n + q] = a;
        }
        const float* wl = sw + c.layer_off[l];
        if (l < c.n_layers - 1) {
#pragma unroll
            for (int o = 0; o < W; ++o) acc[o] = 0.0f;
            for (int i = 0; i < W; ++i) axpy_row<W>(wl + (size_t)i * W, act[i * BS], acc);
        } else {
#pragma unroll
            for (int o = 0; o < NOUT; ++o) out[o] = 0.0f;
            for (int i = 0; i < W; ++i) axpy_row<NOUT>(wl + (size_t)i * NOUT, act[i * BS], out);
        }
    }
}

// Mixture work of one query from its raw outputs (P raw outputs per lobe:
// 4 vMF, 6 anisotropic).
template <int K, int P, int NOUT>
__device__ __forceinline__ void mixture_query(const DevCfg& c, const float* out, const QueryIO& io, uint32_t q,
                                              uint32_t n) {
    if constexpr (P == 6) asg_query<K>(c, [&](int j) { return out[j]; }, io, q, n);
    else vmf_query_cfg<K, true>(c, [&](int j) { return out[j]; }, io, q, n);
}

// Weights to LDS when they fit; returns the weights' base and this thread's activation column.
template <int BS, bool LDS_W>
__device__ __forceinline__ const float* stage_weights(const DevCfg& c, const float* __restrict__ wt, float* smem,
                                                      float** act) {
    const uint32_t n_mlp = n_mlp_of(c);
    const float* sw = wt;
    float* act_base = smem;
    if (LDS_W) {
        for (uint32_t i = threadIdx.x; i < n_mlp; i += BS) smem[i] = wt[i];
        __syncthreads();
        sw = smem;
        act_base = smem + ((n_mlp + 3u) & ~3u);
    }
    *act = act_base + threadIdx.x;
    return sw;
}

// One query q of a stream with row stride n.
template <int W, int K, int P, int BS>
__device__ __forceinline__ void query_one_fp32(const DevCfg& c, const float* sw, const float* __restrict__ table,
                                               float* act, const QueryIO& io, uint32_t n, uint32_t q) {
    constexpr int NOUT = (P * K + 15) / 16 * 16;
    float qin[16];
#pragma unroll
    for (int d = 0; d < 16; ++d) qin[d] = d < c.n_in ? io.in[(size_t)d * n + q] : 0.0f;
    float out[NOUT];
    mlp_forward<W, NOUT, BS>(c, sw, table, qin, act, out, nullptr, nullptr, q, n);
    if (io.raw_out) {
#pragma unroll
        for (int o = 0; o < NOUT; ++o) io.raw_out[(size_t)o * n + q] = out[o];
    }
    if (io.u || io.n_eval > 0 || io.out_sel) mixture_query<K, P, NOUT>(c, out, io, q, n);
}

template <int W, int K, int P, bool LDS_W>
__global__ void __launch_bounds__(block_size<W>()) k_query_fp32(const DevCfg c, const float* __restrict__ wt,
                                                                const float* __restrict__ table, const QueryIO io,
                                                                uint32_t n) {
    constexpr int BS = block_size<W>();
    extern __shared__ float smem[];
    float* act;
    const float* sw = stage_weights<BS, LDS_W>(c, wt, smem, &act);
    for (uint32_t tile = blockIdx.x; (size_t)tile * BS < n; tile += gridDim.x) {
        const uint32_t q = tile * BS + threadIdx.x;
        if (q < n) query_one_fp32<W, K, P, BS>(c, sw, table, act, io, n, q);
    }
}

// Several queues in one persistent launch: workgroups pull tiles of BS
// queries from a device counter (see k_query_mfma_queues).
template <int W, int K, int P, bool LDS_W>
__global__ void __launch_bounds__(block_size<W>()) k_query_fp32_queues(const DevCfg c, const float* __restrict__ wt,
                                                                       const float* __restrict__ table,
                                                                       const QueueSet qs) {
    constexpr int BS = block_size<W>();
    extern __shared__ float smem[];
    __shared__ uint32_t item;
    float* act;
    const float* sw = stage_weights<BS, LDS_W>(c, wt, smem, &act);
    uint32_t first[kMaxQueues + 1];
    queue_tile_prefix(qs, (uint32_t)BS, first);
    const uint32_t total = first[kMaxQueues];
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(qs.counter, 1u);
        __syncthreads();
        const uint32_t w = item;
        __syncthreads();
        if (w >= total) break;
        int qi = 0;
        uint32_t base = 0;
#pragma unroll
        for (int i = 1; i < kMaxQueues; ++i)
            if (w >= first[i] && first[i] < first[i + 1]) {
                qi = i;
                base = first[i];
            }
        const uint32_t q = (w - base) * BS + threadIdx.x;
        if (q < min(qs.counts[qi], qs.stride[qi]))
            query_one_fp32<W, K, P, BS>(c, sw, table, act, qs.io[qi], qs.stride[qi], q);
    }
}

template <int K, int P>
__global__ void __launch_bounds__(256) k_mixture(const DevCfg c, const float* __restrict__ raw, const QueryIO io,
                                                 uint32_t n) {
    constexpr int NOUT = (P * K + 15) / 16 * 16;
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n) return;
    float out[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; ++o) out[o] = o < c.n_out ? raw[(size_t)o * n + q] : 0.0f;
    mixture_query<K, P, NOUT>(c, out, io, q, n);
}

__global__ void __launch_bounds__(256) k_encode(const DevCfg c, const float* __restrict__ table,
                                                const float* __restrict__ in, float* __restrict__ enc, uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n) return;
    float qin[16];
#pragma unroll
    for (int d = 0; d < 16; ++d) qin[d] = d < c.n_in ? in[(size_t)d * n + q] : 0.0f;
    int j = 0;
    encode_query(c, table, qin, [&](float v) { enc[(size_t)(j++) * n + q] = v; });
}

__global__ void __launch_bounds__(256) k_grid_corners(const DevCfg c, const float* __restrict__ in,
                                                      uint32_t* __restrict__ idx, float* __restrict__ w, uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n) return;
    const int b = c.enc_begin[c.grid_enc];
    const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
    for (int l = 0; l < c.n_levels; ++l) {
        uint32_t ix[8] = {};
        float wk[8] = {};
        level_corners(c, l, x0, x1, x2, ix, wk);
        for (int k = 0; k < 8; ++k) {
            idx[((size_t)l * 8 + k) * n + q] = ix[k];
            w[((size_t)l * 8 + k) * n + q] = wk[k];
        }
    }
}

template <int W, int K, int P, bool LDS_W>
hipError_t launch_query_t(const DevCfg& c, const float* wt, const float* table, const QueryIO& io, uint32_t n,
                          int n_cu, hipStream_t s) {
    constexpr int BS = block_size<W>();
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const size_t lds = (LDS_W ? ((n_mlp + 3u) & ~3u) * sizeof(float) : 0) + (size_t)W * BS * sizeof(float);
    hipError_t e = hipFuncSetAttribute((const void*)k_query_fp32<W, K, P, LDS_W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t tiles = (n + BS - 1) / BS;
    const uint32_t grid = tiles < (uint32_t)n_cu ? tiles : (uint32_t)n_cu;
    hipLaunchKernelGGL((k_query_fp32<W, K, P, LDS_W>), dim3(grid), dim3(BS), lds, s, c, wt, table, io, n);
    return hipGetLastError();
}

}  // namespace

bool fp32_weights_fit_lds(const DevCfg& c) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    const int bs = c.width == 64 ? 256 : 128;
    return ((size_t)n_mlp + 4) * 4 + (size_t)c.width * bs * 4 <= 160 * 1024;
}

hipError_t launch_query_fp32(const DevCfg& c, const float* wt, const float* table, const QueryIO& io, uint32_t n,
                             int n_cu, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const bool lds = fp32_weights_fit_lds(c);
#define NPMG_Q(WW, KK, PP)                                                           \
    if (c.width == WW && c.K == KK && c.lobe_params == PP)                           \
        return lds ? launch_query_t<WW, KK, PP, true>(c, wt, table, io, n, n_cu, s) \
                   : launch_query_t<WW, KK, PP, false>(c, wt, table, io, n, n_cu, s);
    NPMG_Q(64, 7, 4)
    NPMG_Q(64, 8, 4)
    NPMG_Q(64, 16, 4)
    NPMG_Q(128, 7, 4)
    NPMG_Q(128, 8, 4)
    NPMG_Q(128, 16, 4)
    NPMG_Q(64, 16, 6)
    NPMG_Q(128, 16, 6)
#undef NPMG_Q
    return hipErrorInvalidValue;
}

template <int W, int K, int P, bool LDS_W>
hipError_t launch_queues_t(const DevCfg& c, const float* wt, const float* table, const QueueSet& qs, int n_cu,
                           hipStream_t s) {
    constexpr int BS = block_size<W>();
    const uint32_t n_mlp = n_mlp_of(c);
    const size_t lds = (LDS_W ? ((n_mlp + 3u) & ~3u) * sizeof(float) : 0) + (size_t)W * BS * sizeof(float);
    auto kern = k_query_fp32_queues<W, K, P, LDS_W>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((uint32_t)n_cu), dim3(BS), lds, s, c, wt, table, qs);
    return hipGetLastError();
}

// Persistent FP32 launch over several queues (qs.counter must be zero).
hipError_t launch_query_fp32_queues(const DevCfg& c, const float* wt, const float* table, const QueueSet& qs,
                                    int n_cu, hipStream_t s) {
    if (qs.n_queues == 0) return hipSuccess;
    if (qs.n_queues > (uint32_t)kMaxQueues) return hipErrorInvalidValue;
    const bool lds = fp32_weights_fit_lds(c);
#define NPMG_QF(WW, KK, PP)                                                  \
    if (c.width == WW && c.K == KK && c.lobe_params == PP)                   \
        return lds ? launch_queues_t<WW, KK, PP, true>(c, wt, table, qs, n_cu, s) \
                   : launch_queues_t<WW, KK, PP, false>(c, wt, table, qs, n_cu, s);
    NPMG_QF(64, 7, 4)
    NPMG_QF(64, 8, 4)
    NPMG_QF(64, 16, 4)
    NPMG_QF(128, 7, 4)
    NPMG_QF(128, 8, 4)
    NPMG_QF(128, 16, 4)
    NPMG_QF(64, 16, 6)
    NPMG_QF(128, 16, 6)
#undef NPMG_QF
    return hipErrorInvalidValue;
}

hipError_t launch_mixture(const DevCfg& c, const float* raw, const QueryIO& io, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256);
    if (c.lobe_params == 6) hipLaunchKernelGGL((k_mixture<16, 6>), grid, dim3(256), 0, s, c, raw, io, n);
    else if (c.K == 7) hipLaunchKernelGGL((k_mixture<7, 4>), grid, dim3(256), 0, s, c, raw, io, n);
    else if (c.K == 8) hipLaunchKernelGGL((k_mixture<8, 4>), grid, dim3(256), 0, s, c, raw, io, n);
    else if (c.K == 16) hipLaunchKernelGGL((k_mixture<16, 4>), grid, dim3(256), 0, s, c, raw, io, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_encode(const DevCfg& c, const float* table, const float* in, float* enc, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_encode, dim3((n + 255) / 256), dim3(256), 0, s, c, table, in, enc, n);
    return hipGetLastError();
}

hipError_t launch_grid_corners(const DevCfg& c, const float* in, uint32_t* idx, float* w, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_corners, dim3((n + 255) / 256), dim3(256), 0, s, c, in, idx, w, n);
    return hipGetLastError();
}

}  // namespace npmg
