// k_train.hip -- FP32 training kernels: forward with stored activations and
// fused mixture adjoint, MLP/grid backward, weight-gradient GEMM with a
// deterministic split-K reduction, Adam + EMA, weight transposition and the
// loss mean.
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

// Forward for training: encoded features and post-ReLU activations are kept
// in HBM for the weight gradients; the loss and dL/draw of the mixture-KL
// adjoint are computed in the same thread (train.cu:69-124).
template <int W, int K, int P, bool LDS_W>
__global__ void __launch_bounds__(block_size<W>()) k_train_fwd(const DevCfg c, const float* __restrict__ wt,
                                                               const float* __restrict__ table,
                                                               const float* __restrict__ in,
                                                               const float* __restrict__ tgt_dir,
                                                               const float* __restrict__ tgt_rgb,
                                                               const float* __restrict__ tgt_pdf,
                                                               const float* __restrict__ tgt_bsdf, const LossScale ls,
                                                               const TrainWS ws, uint32_t n) {
    constexpr int BS = block_size<W>();
    constexpr int NOUT = (P * K + 15) / 16 * 16;
    extern __shared__ float smem[];
    const uint32_t n_mlp = n_mlp_of(c);
    const float* sw = wt;
    float* act_base = smem;
    if (LDS_W) {
        for (uint32_t i = threadIdx.x; i < n_mlp; i += BS) smem[i] = wt[i];
        __syncthreads();
        sw = smem;
        act_base = smem + ((n_mlp + 3u) & ~3u);
    }
    float* act = act_base + threadIdx.x;
    for (uint32_t tile = blockIdx.x; (size_t)tile * BS < n; tile += gridDim.x) {
        const uint32_t q = tile * BS + threadIdx.x;
        if (q >= n) continue;
        float qin[16];
#pragma unroll
        for (int d = 0; d < 16; ++d) qin[d] = d < c.n_in ? in[(size_t)d * n + q] : 0.0f;
        float acc[W];
#pragma unroll
        for (int o = 0; o < W; ++o) acc[o] = 0.0f;
        int j = 0;
        encode_query(c, table, qin, [&](float v) {
            ws.enc[(size_t)j * n + q] = v;
            axpy_row<W>(sw + (size_t)j * W, v, acc);
            ++j;
        });
        float out[NOUT];
        for (int l = 1; l < c.n_layers; ++l) {
#pragma unroll
            for (int o = 0; o < W; ++o) {
                const float a = acc[o] > 0.0f ? acc[o] : 0.0f;
                act[o * BS] = a;
                ws.acts[((size_t)(l - 1) * W + o) * n + q] = a;
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
        if (tgt_dir) {
            const float loss = lobe_adjoint<K, P>(
                c, [&](int jj) { return out[jj]; }, tgt_dir[q], tgt_dir[(size_t)n + q], tgt_rgb[q],
                tgt_rgb[(size_t)n + q], tgt_rgb[2 * (size_t)n + q], tgt_pdf[q], tgt_bsdf ? tgt_bsdf[q] : 0.0f,
                ls.over_n(), ls.clamp(c), [&](int jj, float v) { ws.dout[(size_t)jj * n + q] = v; });
            ws.loss[q] = loss;
        }
    }
}

// Standalone adjoint from raw outputs (harness split pass).
template <int K, int P>
__global__ void __launch_bounds__(256) k_adjoint(const DevCfg c, const float* __restrict__ raw,
                                                 const float* __restrict__ tgt_dir, const float* __restrict__ tgt_rgb,
                                                 const float* __restrict__ tgt_pdf, const float* __restrict__ tgt_bsdf,
                                                 const LossScale ls, float* __restrict__ dout,
                                                 float* __restrict__ loss, uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n) return;
    loss[q] = lobe_adjoint<K, P>(
        c, [&](int j) { return raw[(size_t)j * n + q]; }, tgt_dir[q], tgt_dir[(size_t)n + q], tgt_rgb[q],
        tgt_rgb[(size_t)n + q], tgt_rgb[2 * (size_t)n + q], tgt_pdf[q], tgt_bsdf ? tgt_bsdf[q] : 0.0f, ls.over_n(),
        ls.clamp(c), [&](int j, float v) { dout[(size_t)j * n + q] = v; });
}

// Backward: g_{l-1} = (W_l^T g_l) masked by ReLU, from the output layer down,
// then dL/denc = W_0^T g_0; the grid rows of dL/denc go to dgrid [L F][n] for
// the scatter kernels. w: row-major [out][in] weights of every layer (the
// parameter vector), staged in LDS when they fit.
template <int W, int NOUT, bool LDS_W>
__global__ void __launch_bounds__(block_size<W>()) k_train_bwd(const DevCfg c, const float* __restrict__ w,
                                                               const TrainWS ws, float* __restrict__ denc,
                                                               float* __restrict__ dgrid, uint32_t n) {
    constexpr int BS = block_size<W>();
    extern __shared__ float smem[];
    const uint32_t n_mlp = n_mlp_of(c);
    const float* sw = w;
    float* g_base = smem;
    if (LDS_W) {
        for (uint32_t i = threadIdx.x; i < n_mlp; i += BS) smem[i] = w[i];
        __syncthreads();
        sw = smem;
        g_base = smem + ((n_mlp + 3u) & ~3u);
    }
    float* gcol = g_base + threadIdx.x;  // [max(W, NOUT)][BS]
    const int H = c.n_layers - 1;
    for (uint32_t tile = blockIdx.x; (size_t)tile * BS < n; tile += gridDim.x) {
        const uint32_t q = tile * BS + threadIdx.x;
        if (q >= n) continue;
        for (int o = 0; o < NOUT; ++o) gcol[o * BS] = ws.dout[(size_t)o * n + q];
        int ow = NOUT;
        for (int l = H; l >= 1; --l) {
            float acc[W];
#pragma unroll
            for (int i = 0; i < W; ++i) acc[i] = 0.0f;
            const float* wl = sw + c.layer_off[l];
            for (int o = 0; o < ow; ++o) axpy_row<W>(wl + (size_t)o * W, gcol[o * BS], acc);
            const float* a = ws.acts + (size_t)(l - 1) * W * n + q;
            float* gout = ws.g + (size_t)(l - 1) * W * n + q;
#pragma unroll
            for (int i = 0; i < W; ++i) {
                const float v = a[(size_t)i * n] > 0.0f ? acc[i] : 0.0f;
                gout[(size_t)i * n] = v;
                gcol[i * BS] = v;
            }
            ow = W;
        }
        // Layer 0: dL/denc for the grid features (kept for the level-major
        // scatter) and for every feature when denc is requested.
        const int nin = c.layer_in[0];
        const int ge = c.grid_enc;
        const int gbeg = ge >= 0 ? c.enc_offset[ge] : 0;
        const int gend = ge >= 0 ? gbeg + c.n_levels * c.F : 0;
        const int jb = denc ? 0 : gbeg, je = denc ? nin : gend;
        for (int j0 = jb & ~15; j0 < je; j0 += 16) {
            float acc[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
            // W_0 is [W][nin] row-major: row o, columns j0..j0+15 (nin is a multiple of 16).
            for (int o = 0; o < W; ++o) axpy_row<16>(sw + (size_t)o * nin + j0, gcol[o * BS], acc);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int jj = j0 + i;
                if (jj < jb || jj >= je) continue;
                if (denc) denc[(size_t)jj * n + q] = acc[i];
                if (dgrid && jj >= gbeg && jj < gend) dgrid[(size_t)(jj - gbeg) * n + q] = acc[i];
            }
        }
    }
}

// Weight gradients: one block per (64 x 64 tile of one layer's dW, split of
// the batch). Partial sums go to partials[split][param]; k_reduce_splits adds
// the splits in order, so the result does not depend on scheduling.
struct WgradTiles {
    int n;
    int layer[64], o0[64], i0[64];
};

__global__ void __launch_bounds__(256) k_wgrad(const DevCfg c, const WgradTiles tiles, const TrainWS ws,
                                               float* __restrict__ partials, uint32_t n, uint32_t q_per_split) {
    __shared__ float sg[32][65];
    __shared__ float sx[32][65];
    const int t = blockIdx.x;
    const int l = tiles.layer[t], o0 = tiles.o0[t], i0 = tiles.i0[t];
    const int H = c.n_layers - 1;
    const int iw = c.layer_in[l], ow = c.layer_out[l];
    const float* G = l == H ? ws.dout : ws.g + (size_t)l * c.width * n;        // [ow][n]
    const float* X = l == 0 ? ws.enc : ws.acts + (size_t)(l - 1) * c.width * n;  // [iw][n]
    const uint32_t qb = blockIdx.y * q_per_split;
    const uint32_t qe = min(n, qb + q_per_split);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // 16 x 16 threads, 4 x 4 outputs each
    float acc[4][4] = {};
    for (uint32_t q0 = qb; q0 < qe; q0 += 32) {
        // Load 64 rows x 32 samples of G and X (coalesced along the batch).
        for (int r = threadIdx.x; r < 64 * 32; r += 256) {
            const int row = r >> 5, col = r & 31;
            const uint32_t qq = q0 + col;
            const bool okq = qq < qe;
            sg[col][row] = (okq && o0 + row < ow) ? G[(size_t)(o0 + row) * n + qq] : 0.0f;
            sx[col][row] = (okq && i0 + row < iw) ? X[(size_t)(i0 + row) * n + qq] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < 32; ++k) {
            float gv[4], xv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) { gv[a] = sg[k][ty * 4 + a]; xv[a] = sx[k][tx * 4 + a]; }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fm::fma_(gv[a], xv[b], acc[a][b]);
        }
        __syncthreads();
    }
    const uint32_t n_mlp = n_mlp_of(c);
    float* P = partials + (size_t)blockIdx.y * n_mlp + c.layer_off[l];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int o = o0 + ty * 4 + a, i = i0 + tx * 4 + b;
            if (o < ow && i < iw) P[(size_t)o * iw + i] = acc[a][b];
        }
}

// out[i] = sum over the splits k of partials[k][i], in a fixed order: a
// workgroup takes 32 columns; thread (g, col) adds the rows of group g
// (8 contiguous ranges, in row order), then the 8 group sums are added in
// group order. Coalesced 128-byte row segments and 8 loads in flight per
// column instead of one thread walking every split.
__global__ void __launch_bounds__(256) k_reduce_splits(const float* __restrict__ partials, uint32_t n_splits,
                                                       uint32_t len, float* __restrict__ out,
                                                       uint32_t* __restrict__ overflow) {
    __shared__ float part[8][32];
    const uint32_t col = threadIdx.x & 31u, g = threadIdx.x >> 5;
    const uint32_t i = blockIdx.x * 32u + col;
    const uint32_t per = (n_splits + 7u) / 8u;
    const uint32_t k0 = min(n_splits, g * per), k1 = min(n_splits, k0 + per);
    float s = 0.0f;
    if (i < len)
        for (uint32_t k = k0; k < k1; ++k) s += partials[(size_t)k * len + i];
    part[g][col] = s;
    __syncthreads();
    if (g == 0) {
        float t = 0.0f;
#pragma unroll
        for (int r = 0; r < 8; ++r) t += part[r][col];
        if (i < len) out[i] = t;
        const bool finite = !(i < len) || __builtin_isfinite(t);
        if (overflow && !__all(finite ? 1 : 0) && threadIdx.x == 0) atomicOr(overflow, 1u);
    }
}

// Adam's bias correction at step t (optimizers/adam.h:88-90):
// sqrt(1 - beta2^t) / (1 - beta1^t), beta^t as exp(t log beta) in FP32.
__device__ __forceinline__ float adam_bias(float beta1, float beta2, uint32_t t) {
    const float b2t = fm::exp((float)t * fm::log(beta2)), b1t = fm::exp((float)t * fm::log(beta1));
    return fm::sqrt_(1 - b2t) / (1 - b1t);
}

// The bias correction of every step below len, tabulated once per context: it
// depends on the step count alone, and the two exponentials and the logarithm
// per parameter were most of the optimiser's arithmetic.
__global__ void __launch_bounds__(256) k_adam_bias_table(float beta1, float beta2, float* __restrict__ out,
                                                         uint32_t len) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t < len) out[t] = adam_bias(beta1, beta2, t);
}

// tiny-cuda-nn optimizers/adam.h:47-119 followed by the EMA of
// optimizers/ema.h:62-76 (full-precision EMA).
__global__ void __launch_bounds__(256) k_adam(const AdamArgs a, float* __restrict__ master,
                                              const float* __restrict__ grad, float* __restrict__ m,
                                              float* __restrict__ v, uint32_t* __restrict__ steps,
                                              float* __restrict__ ema) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= a.n) return;
    if (a.check_overflow && a.st->overflow) return;
    float g = grad[i] / (a.scale_from_state ? a.st->scale : a.loss_scale);
    const bool skip = i >= a.n_matrix && g == 0.0f;
    float w = master[i];
    if (!skip) {
        if (i < a.n_matrix) g += a.l2_reg * w;
        const float g2 = g * g;
        const float m1 = a.beta1 * m[i] + (1 - a.beta1) * g;
        const float v1 = a.beta2 * v[i] + (1 - a.beta2) * g2;
        m[i] = m1;
        v[i] = v1;
        const uint32_t t = ++steps[i];
        const float lr_t = a.lr * (t < a.adam_bias_len ? a.adam_bias[t] : adam_bias(a.beta1, a.beta2, t));
        const float eff = smin(smax(lr_t / (fm::sqrt_(v1) + a.eps), 0.0f), 3.402823466e38f);
        w = w - eff * m1;
        master[i] = w;
    }
    float e = w;
    if (a.ema) {
        const uint32_t t = a.st->step + 1u, last = a.ema_pow_len - 1u;
        const float deb_old = 1 - a.ema_pow[t - 1u < last ? t - 1u : last];
        const float deb_new = 1.0f / (1 - a.ema_pow[t < last ? t : last]);
        if (a.ema_half)
            e = fm::round_half((ema[i] * a.ema_decay * deb_old + fm::round_half(w) * (1 - a.ema_decay)) * deb_new);
        else
            e = (ema[i] * a.ema_decay * deb_old + w * (1 - a.ema_decay)) * deb_new;
        ema[i] = e;
    }
    if (a.half_precision != PREC_FP32 && i >= a.n_matrix) {
        const size_t j = i - a.n_matrix;
        if (a.half_precision == PREC_BF16) {
            static_cast<__bf16*>(a.half_train)[j] = (__bf16)w;
            static_cast<__bf16*>(a.half_infer)[j] = (__bf16)e;
        } else {
            static_cast<_Float16*>(a.half_train)[j] = (_Float16)w;
            static_cast<_Float16*>(a.half_infer)[j] = (_Float16)e;
        }
    }
}

__global__ void k_scale_update(ScaleState* st, bool dynamic, uint32_t window, float max_scale) {
    if (threadIdx.x != 0) return;
    if (dynamic && st->overflow) {
        st->scale = st->scale * 0.5f < 1.0f ? 1.0f : st->scale * 0.5f;
        st->good = 0;
    } else {
        st->step += 1u;
        if (dynamic && ++st->good >= window) {
            st->scale = st->scale * 2.0f > max_scale ? max_scale : st->scale * 2.0f;
            st->good = 0;
        }
    }
    st->overflow = 0;
}

__global__ void __launch_bounds__(256) k_transpose_mlp(const DevCfg c, const float* __restrict__ w,
                                                       float* __restrict__ wt) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n_mlp_of(c)) return;
    int l = 0;
    while (l + 1 < c.n_layers && i >= c.layer_off[l + 1]) ++l;
    const uint32_t r = i - c.layer_off[l];
    const uint32_t iw = (uint32_t)c.layer_in[l], ow = (uint32_t)c.layer_out[l];
    const uint32_t o = r / iw, ii = r % iw;
    wt[c.layer_off[l] + ii * ow + o] = w[i];
}

// Sum of x[first, last) in a fixed order: per-thread strided sums, then a tree.
__device__ __forceinline__ float block_sum(const float* __restrict__ x, uint32_t first, uint32_t last, float* red) {
    float s = 0.0f;
    for (uint32_t i = first + threadIdx.x; i < last; i += 256u) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// Stage 1: 256 blocks, block b sums its contiguous slice into partials[b].
__global__ void __launch_bounds__(256) k_mean_partial(const float* __restrict__ x, uint32_t n,
                                                      float* __restrict__ partials) {
    __shared__ float red[256];
    const uint32_t per = (n + 255u) / 256u;
    const uint32_t lo = blockIdx.x * per < n ? blockIdx.x * per : n, hi = lo + per < n ? lo + per : n;
    const float s = block_sum(x, lo, hi, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Stage 2: one block, the 256 partials -> the mean.
__global__ void __launch_bounds__(256) k_mean_final(const float* __restrict__ partials, uint32_t n,
                                                    float* __restrict__ out) {
    __shared__ float red[256];
    const float s = block_sum(partials, 0u, 256u, red);
    if (threadIdx.x == 0) out[0] = n ? s / (float)n : 0.0f;
}

template <int W, int K, int P, bool LDS_W>
hipError_t launch_fwd_t(const DevCfg& c, const float* wt, const float* table, const float* in, const float* td,
                        const float* tr, const float* tp, const float* tb, LossScale ls, const TrainWS& ws, uint32_t n,
                        int n_cu, hipStream_t s) {
    constexpr int BS = block_size<W>();
    const uint32_t n_mlp = n_mlp_of(c);
    const size_t lds = (LDS_W ? ((n_mlp + 3u) & ~3u) * sizeof(float) : 0) + (size_t)W * BS * sizeof(float);
    hipError_t e = hipFuncSetAttribute((const void*)k_train_fwd<W, K, P, LDS_W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t tiles = (n + BS - 1) / BS;
    const uint32_t grid = tiles < (uint32_t)n_cu ? tiles : (uint32_t)n_cu;
    hipLaunchKernelGGL((k_train_fwd<W, K, P, LDS_W>), dim3(grid), dim3(BS), lds, s, c, wt, table, in, td, tr, tp, tb,
                       ls, ws, n);
    return hipGetLastError();
}

template <int W, int NOUT, bool LDS_W>
hipError_t launch_bwd_t(const DevCfg& c, const float* w, const TrainWS& ws, float* denc, float* dgrid, uint32_t n,
                        int n_cu, hipStream_t s) {
    constexpr int BS = block_size<W>();
    constexpr int GR = W > NOUT ? W : NOUT;
    const uint32_t n_mlp = n_mlp_of(c);
    const size_t lds = (LDS_W ? ((n_mlp + 3u) & ~3u) * sizeof(float) : 0) + (size_t)GR * BS * sizeof(float);
    hipError_t e = hipFuncSetAttribute((const void*)k_train_bwd<W, NOUT, LDS_W>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t tiles = (n + BS - 1) / BS;
    const uint32_t grid = tiles < (uint32_t)n_cu ? tiles : (uint32_t)n_cu;
    hipLaunchKernelGGL((k_train_bwd<W, NOUT, LDS_W>), dim3(grid), dim3(BS), lds, s, c, w, ws, denc, dgrid, n);
    return hipGetLastError();
}

}  // namespace

bool fp32_weights_fit_lds(const DevCfg& c);

hipError_t launch_train_forward_fp32(const DevCfg& c, const float* wt, const float* table, const float* in,
                                     const float* td, const float* tr, const float* tp, const float* tb, LossScale ls,
                                     const TrainWS& ws, uint32_t n, int n_cu, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const bool lds = fp32_weights_fit_lds(c);
#define NPMG_F(WW, KK, PP)                                                                       \
    if (c.width == WW && c.K == KK && c.lobe_params == PP)                                       \
        return lds ? launch_fwd_t<WW, KK, PP, WW == 64>(c, wt, table, in, td, tr, tp, tb, ls, ws, n, n_cu, s) \
                   : launch_fwd_t<WW, KK, PP, false>(c, wt, table, in, td, tr, tp, tb, ls, ws, n, n_cu, s);
    NPMG_F(64, 7, 4)
    NPMG_F(64, 8, 4)
    NPMG_F(64, 16, 4)
    NPMG_F(128, 7, 4)
    NPMG_F(128, 8, 4)
    NPMG_F(128, 16, 4)
    NPMG_F(64, 16, 6)
    NPMG_F(128, 16, 6)
#undef NPMG_F
    return hipErrorInvalidValue;
}

hipError_t launch_train_backward_fp32(const DevCfg& c, const float* w, const TrainWS& ws, float* denc, float* dgrid,
                                      uint32_t n, int n_cu, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const bool lds = fp32_weights_fit_lds(c);
    if (c.width == 64 && c.n_out_pad == 32) return lds ? launch_bwd_t<64, 32, true>(c, w, ws, denc, dgrid, n, n_cu, s)
                                                       : launch_bwd_t<64, 32, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    if (c.width == 64 && c.n_out_pad == 64) return lds ? launch_bwd_t<64, 64, true>(c, w, ws, denc, dgrid, n, n_cu, s)
                                                       : launch_bwd_t<64, 64, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    if (c.width == 128 && c.n_out_pad == 32) return launch_bwd_t<128, 32, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    if (c.width == 128 && c.n_out_pad == 64) return launch_bwd_t<128, 64, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    if (c.width == 64 && c.n_out_pad == 96) return lds ? launch_bwd_t<64, 96, true>(c, w, ws, denc, dgrid, n, n_cu, s)
                                                       : launch_bwd_t<64, 96, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    if (c.width == 128 && c.n_out_pad == 96) return launch_bwd_t<128, 96, false>(c, w, ws, denc, dgrid, n, n_cu, s);
    return hipErrorInvalidValue;
}

size_t wgrad_partials_needed(const DevCfg& c, uint32_t n, uint32_t* n_splits, uint32_t* q_per_split) {
    uint32_t splits = (n + 4095) / 4096;
    if (splits > 64) splits = 64;
    if (splits < 1) splits = 1;
    *n_splits = splits;
    *q_per_split = ((n + splits - 1) / splits + 31) / 32 * 32;
    return (size_t)splits * n_mlp_of(c);
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s);

hipError_t launch_wgrad_fp32(const DevCfg& c, const TrainWS& ws, float* partials, size_t partials_cap, float* dw,
                             uint32_t n, uint32_t* overflow, hipStream_t s) {
    if (n == 0) return launch_fill(dw, 0.0f, n_mlp_of(c), s);
    WgradTiles t{};
    for (int l = 0; l < c.n_layers; ++l)
        for (int o0 = 0; o0 < c.layer_out[l]; o0 += 64)
            for (int i0 = 0; i0 < c.layer_in[l]; i0 += 64) {
                if (t.n >= 64) return hipErrorInvalidValue;
                t.layer[t.n] = l; t.o0[t.n] = o0; t.i0[t.n] = i0; ++t.n;
            }
    uint32_t splits, qps;
    const size_t need = wgrad_partials_needed(c, n, &splits, &qps);
    if (need > partials_cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_wgrad, dim3(t.n, splits), dim3(256), 0, s, c, t, ws, partials, n, qps);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t len = n_mlp_of(c);
    hipLaunchKernelGGL(k_reduce_splits, dim3((len + 31) / 32), dim3(256), 0, s, partials, splits, len, dw, overflow);
    return hipGetLastError();
}

hipError_t launch_reduce_splits(const float* partials, uint32_t n_splits, uint32_t len, float* out, hipStream_t s,
                                uint32_t* overflow) {
    hipLaunchKernelGGL(k_reduce_splits, dim3((len + 31) / 32), dim3(256), 0, s, partials, n_splits, len, out,
                       overflow);
    return hipGetLastError();
}

hipError_t launch_scale_update(ScaleState* st, bool dynamic, uint32_t window, float max_scale, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_update, dim3(1), dim3(64), 0, s, st, dynamic, window, max_scale);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, float* master, const float* grad, float* m, float* v, uint32_t* steps,
                       float* ema, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a, master, grad, m, v, steps, ema);
    return hipGetLastError();
}

hipError_t launch_adam_bias_table(float beta1, float beta2, float* out, uint32_t len, hipStream_t s) {
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adam_bias_table, dim3((len + 255u) / 256u), dim3(256), 0, s, beta1, beta2, out, len);
    return hipGetLastError();
}

hipError_t launch_transpose_mlp(const DevCfg& c, const float* w, float* wt, hipStream_t s) {
    const uint32_t n = n_mlp_of(c);
    hipLaunchKernelGGL(k_transpose_mlp, dim3((n + 255) / 256), dim3(256), 0, s, c, w, wt);
    return hipGetLastError();
}

hipError_t launch_mean(const float* x, uint32_t n, float* partials, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mean_partial, dim3(256), dim3(256), 0, s, x, n, partials);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(256), 0, s, partials, n, out);
    return hipGetLastError();
}

hipError_t launch_adjoint(const DevCfg& c, const float* raw, const float* td, const float* tr, const float* tp,
                          const float* tb, LossScale ls, float* dout, float* loss, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256);
    if (c.lobe_params == 6)
        hipLaunchKernelGGL((k_adjoint<16, 6>), grid, dim3(256), 0, s, c, raw, td, tr, tp, tb, ls, dout, loss, n);
    else if (c.K == 7)
        hipLaunchKernelGGL((k_adjoint<7, 4>), grid, dim3(256), 0, s, c, raw, td, tr, tp, tb, ls, dout, loss, n);
    else if (c.K == 8)
        hipLaunchKernelGGL((k_adjoint<8, 4>), grid, dim3(256), 0, s, c, raw, td, tr, tp, tb, ls, dout, loss, n);
    else if (c.K == 16)
        hipLaunchKernelGGL((k_adjoint<16, 4>), grid, dim3(256), 0, s, c, raw, td, tr, tp, tb, ls, dout, loss, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace npmg
