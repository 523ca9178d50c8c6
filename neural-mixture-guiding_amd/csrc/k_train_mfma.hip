// k_train_mfma.hip -- training step with the MLP on MFMA (BF16/FP16 operands,
// FP32 accumulation): forward and backward chains, weight gradients, and the
// packing of the transposed weight fragments used by the backward chain.
//
// One wave per 64 samples in both chains (fragment layouts as in mfma.h):
//   forward   layer 0 from the level-major grid features and OneBlob chunks;
//             hidden layers chained register to register; the ReLU sign of
//             every hidden unit is kept as one bit per accumulator register;
//             the raw outputs go to HBM in FP32 for the mixture-KL adjoint
//             (train.cu:69-124), which runs lane-per-sample in FP32;
//   backward  dX = W^T dY on MFMA with transposed weight fragments whose K
//             columns use the same chaining permutation, so each layer's
//             gradient feeds the next backward step from registers; ReLU'
//             comes from the stored sign bits (fully_fused_mlp.cu:150-259).
// Stored for the weight gradients (operand type, [dim][n]): layer inputs and
// output gradients of every layer (the layer-0 inputs only outside the "grid,
// then OneBlob(16)" layout with the fast numerics, whose inputs the weight
// gradient rebuilds); dL/d(grid features) [L F][n] in FP32 for the grid
// scatter.
#include <type_traits>

#include "device.h"
#include "kernels.h"
#include "mfma.h"

namespace npmg {

namespace {

using namespace dev;

// (a split value is finite when its hi part is: lo = x - hi)
template <class T>
__device__ __forceinline__ bool finite_v(const typename mf::Frag<T>::vec& v) {
    bool ok = true;
#pragma unroll
    for (int e = 0; e < 8; ++e) ok = ok && __builtin_isfinite((float)mf::get<T>(v, 0, e));
    return ok;
}

// Forward chain; mask word (tile, l, t, lane) bit 16 cc + r = ReLU active for
// accumulator register r of half cc of column tile t.
// T is the operand type; G = mf::Op<T>::table the type of the grid table and
// features; the stored operands are mf::Op<T>::planes planes of elements E
// (hi then lo for the split type) at a stride of one whole buffer.
template <class T, int W, int NOUT, int H>
__global__ void __launch_bounds__(256) k_train_fwd_mfma(const DevCfg c, const T* __restrict__ wfrag,
                                                        const typename mf::Op<T>::table* __restrict__ table,
                                                        const typename mf::Op<T>::table* __restrict__ genc,
                                                        const float* __restrict__ in, const TrainMfmaWS ws, uint32_t n,
                                                        uint32_t n_frag) {
    using V = typename mf::Frag<T>::vec;
    using E = typename mf::Op<T>::elem;
    using G = typename mf::Op<T>::table;
    constexpr int NP = mf::Op<T>::planes;
    constexpr int TO = NOUT / 32;
    constexpr int TH = W / 32;
    constexpr int KBH = W / 16;
    static_assert(H >= 1 && NOUT % 32 == 0, "shape");
    extern __shared__ uint4 smem_frag[];
    for (uint32_t i = threadIdx.x; i < n_frag * 64u * NP; i += 256u)
        smem_frag[i] = reinterpret_cast<const uint4*>(wfrag)[i];
    __syncthreads();
    const T* sw = reinterpret_cast<const T*>(smem_frag);
    E* x0 = static_cast<E*>(ws.x0);
    E* xh = static_cast<E*>(ws.xh);
    const size_t px0 = (size_t)c.enc_padded * n, pxh = (size_t)H * W * n;  // plane strides
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    const int KB0 = c.enc_padded / 16;
    const uint32_t n_tiles = (n + 63u) / 64u;
    for (uint32_t tile = blockIdx.x * 4u + (uint32_t)wave; tile < n_tiles; tile += gridDim.x * 4u) {
        const uint32_t q0 = tile * 64u;
        const uint32_t qra = q0 + (uint32_t)j, qrb = q0 + 32u + (uint32_t)j;
        const bool oka = qra < n, okb = qrb < n;
        const uint32_t qa = oka ? qra : n - 1u, qb = okb ? qrb : n - 1u;
        // With the fast numerics and the "grid, then OneBlob(16)" layout the
        // fragments come as in the query kernel: grid K-blocks from the
        // level-major features as they are, OneBlob K-blocks from the three
        // non-zero bins (mf::oneblob16_frag); the inputs are read per K-block.
        const bool ob16 = c.numerics == NUM_FAST && c.ob_nkb > 0 && genc;
        float ina[16], inb[16];
#pragma unroll
        for (int d = 0; d < 16; ++d) {
            ina[d] = d < c.n_in && !ob16 ? in[(size_t)d * n + qa] : 0.0f;
            inb[d] = d < c.n_in && !ob16 ? in[(size_t)d * n + qb] : 0.0f;
        }
        mf::f32x16 acc[TH][2];
#pragma unroll
        for (int t = 0; t < TH; ++t) acc[t][0] = acc[t][1] = mf::f32x16{};
        for (int kb = 0; kb < KB0; ++kb) {
            V ba, bb;
            if (ob16 && kb < c.ob_kb0) {
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
            } else if (ob16 && kb < c.ob_kb0 + c.ob_nkb) {
                const float* row = in + (size_t)(c.ob_in + kb - c.ob_kb0) * n;
                ba = mf::oneblob16_frag<T>(row[qa], h);
                bb = mf::oneblob16_frag<T>(row[qb], h);
            } else {
                float va[8], vb[8];
                encode_chunk<G>(c, table, ina, 2 * kb + h, va, genc, qa, n);
                encode_chunk<G>(c, table, inb, 2 * kb + h, vb, genc, qb, n);
                ba = mf::pack8<T>(va);
                bb = mf::pack8<T>(vb);
            }
            // The grid and OneBlob K-blocks are not stored: the weight gradient
            // rebuilds them, bit for bit, from the level-major features and the
            // input records (k_wgrad_mfma).
            const int f0 = 8 * (2 * kb + h);
            if (!ob16) {
#pragma unroll
                for (int p = 0; p < NP; ++p)
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        if (oka) x0[p * px0 + (size_t)(f0 + e) * n + qra] = mf::get<T>(ba, p, e);
                        if (okb) x0[p * px0 + (size_t)(f0 + e) * n + qrb] = mf::get<T>(bb, p, e);
                    }
            }
#pragma unroll
            for (int t = 0; t < TH; ++t) {
                const V a = mf::load_frag<T>(sw, t * KB0 + kb, lane);
                acc[t][0] = mf::Frag<T>::mma(a, ba, acc[t][0]);
                acc[t][1] = mf::Frag<T>::mma(a, bb, acc[t][1]);
            }
        }
        int fb = TH * KB0;
        uint32_t* mask_tile = ws.mask + (size_t)tile * H * TH * 64u + (uint32_t)lane;
#pragma unroll
        for (int l = 1; l <= H; ++l) {
            // ReLU of layer l-1's outputs -> B fragments of layer l; keep the sign bits and the inputs.
            V bx[2][KBH];
            uint32_t mask[TH] = {};
#pragma unroll
            for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    float v[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int r = 8 * (kb & 1) + e;
                        const float a = acc[kb >> 1][cc][r];
                        const bool on = a > 0.0f;
                        v[e] = on ? a : 0.0f;
                        mask[kb >> 1] |= (on ? 1u : 0u) << (16 * cc + r);
                    }
                    bx[cc][kb] = mf::pack8<T>(v);
                    const uint32_t qr = cc ? qrb : qra;
                    if (cc ? okb : oka) {
                        E* dst = xh + ((size_t)(l - 1) * W) * n + qr;
#pragma unroll
                        for (int p = 0; p < NP; ++p)
#pragma unroll
                            for (int e = 0; e < 8; ++e)
                                dst[p * pxh + (size_t)mf::chained_column(kb, h, e) * n] = mf::get<T>(bx[cc][kb], p, e);
                    }
                }
#pragma unroll
            for (int t = 0; t < TH; ++t) mask_tile[((l - 1) * TH + t) * 64] = mask[t];
            if (l < H) {
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
            } else {
#pragma unroll
                for (int t = 0; t < TO; ++t) {
                    mf::f32x16 o0 = {}, o1 = {};
#pragma unroll
                    for (int kb = 0; kb < KBH; ++kb) {
                        const V a = mf::load_frag<T>(sw, fb + t * KBH + kb, lane);
                        o0 = mf::Frag<T>::mma(a, bx[0][kb], o0);
                        o1 = mf::Frag<T>::mma(a, bx[1][kb], o1);
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const size_t row = (size_t)(32 * t + 8 * (r >> 2) + 4 * h + (r & 3)) * n;
                        if (oka) ws.raw[row + qra] = o0[r];
                        if (okb) ws.raw[row + qrb] = o1[r];
                    }
                }
            }
        }
    }
}

// Backward chain from dL/draw (FP32 [NOUT][n]): stores dL/draw in the operand
// type for the output layer's weight gradient, the pre-activation gradients of
// every hidden layer, and the grid rows of dL/denc (all of dL/denc into denc
// when it is given). Non-finite operands raise *overflow (dynamic loss scaling).
template <class T, int W, int NOUT, int H>
__global__ void __launch_bounds__(256) k_train_bwd_mfma(const DevCfg c, const T* __restrict__ bfrag,
                                                        const float* __restrict__ dout, const TrainMfmaWS ws,
                                                        float* __restrict__ denc, uint32_t n, uint32_t n_bfrag,
                                                        uint32_t* __restrict__ overflow) {
    using V = typename mf::Frag<T>::vec;
    using E = typename mf::Op<T>::elem;
    constexpr int NP = mf::Op<T>::planes;
    constexpr int TH = W / 32;
    constexpr int KBH = W / 16;
    constexpr int KBO = NOUT / 16;
    extern __shared__ uint4 smem_frag[];
    for (uint32_t i = threadIdx.x; i < n_bfrag * 64u * NP; i += 256u)
        smem_frag[i] = reinterpret_cast<const uint4*>(bfrag)[i];
    __syncthreads();
    const T* sb = reinterpret_cast<const T*>(smem_frag);
    E* dyh = static_cast<E*>(ws.dyh);
    E* dyo = static_cast<E*>(ws.dyo);
    const size_t pdh = (size_t)H * W * n, pdo = (size_t)NOUT * n;  // plane strides
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    // Fragment bases of the transposed layers: layer 0 has ceil(enc / 32) x (W / 16)
    // fragments, every hidden layer (W / 32) x (W / 16), the output layer (W / 32) x (NOUT / 16).
    const int tb0 = (c.enc_padded + 31) / 32 * KBH;
    const uint32_t n_tiles = (n + 63u) / 64u;
    bool finite = true;
    for (uint32_t tile = blockIdx.x * 4u + (uint32_t)wave; tile < n_tiles; tile += gridDim.x * 4u) {
        const uint32_t q0 = tile * 64u;
        const uint32_t qra = q0 + (uint32_t)j, qrb = q0 + 32u + (uint32_t)j;
        const bool oka = qra < n, okb = qrb < n;
        const uint32_t qa = oka ? qra : n - 1u, qb = okb ? qrb : n - 1u;
        // dL/draw -> B fragments (K = output rows 16 kb + 8 h + e, N = samples)
        V bd[2][KBO];
#pragma unroll
        for (int kb = 0; kb < KBO; ++kb)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                float v[8];
                const uint32_t q = cc ? qb : qa;
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (cc ? okb : oka) ? dout[(size_t)(16 * kb + 8 * h + e) * n + q] : 0.0f;
                bd[cc][kb] = mf::pack8<T>(v);
                finite = finite && finite_v<T>(bd[cc][kb]);
                if (cc ? okb : oka) {
#pragma unroll
                    for (int p = 0; p < NP; ++p)
#pragma unroll
                        for (int e = 0; e < 8; ++e)
                            dyo[p * pdo + (size_t)(16 * kb + 8 * h + e) * n + q] = mf::get<T>(bd[cc][kb], p, e);
                }
            }
        mf::f32x16 acc[TH][2];
        const int bo = tb0 + (H - 1) * TH * KBH;
#pragma unroll
        for (int t = 0; t < TH; ++t) {
            acc[t][0] = acc[t][1] = mf::f32x16{};
#pragma unroll
            for (int kb = 0; kb < KBO; ++kb) {
                const V a = mf::load_frag<T>(sb, bo + t * KBO + kb, lane);
                acc[t][0] = mf::Frag<T>::mma(a, bd[0][kb], acc[t][0]);
                acc[t][1] = mf::Frag<T>::mma(a, bd[1][kb], acc[t][1]);
            }
        }
        const uint32_t* mask_tile = ws.mask + (size_t)tile * H * TH * 64u + (uint32_t)lane;
#pragma unroll
        for (int l = H; l >= 1; --l) {
            // ReLU' of hidden layer l-1, store its gradient, chain into layer l-1.
            uint32_t mask[TH];
#pragma unroll
            for (int t = 0; t < TH; ++t) mask[t] = mask_tile[((l - 1) * TH + t) * 64];
            V bg[2][KBH];
#pragma unroll
            for (int kb = 0; kb < KBH; ++kb)
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    float v[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int r = 8 * (kb & 1) + e;
                        const bool on = (mask[kb >> 1] >> (16 * cc + r)) & 1u;
                        v[e] = on ? acc[kb >> 1][cc][r] : 0.0f;
                    }
                    bg[cc][kb] = mf::pack8<T>(v);
                    finite = finite && finite_v<T>(bg[cc][kb]);
                    if (cc ? okb : oka) {
                        E* dst = dyh + ((size_t)(l - 1) * W) * n + (cc ? qrb : qra);
#pragma unroll
                        for (int p = 0; p < NP; ++p)
#pragma unroll
                            for (int e = 0; e < 8; ++e)
                                dst[p * pdh + (size_t)mf::chained_column(kb, h, e) * n] = mf::get<T>(bg[cc][kb], p, e);
                    }
                }
            if (l >= 2) {
                const int bl = tb0 + (l - 2) * TH * KBH;
#pragma unroll
                for (int t = 0; t < TH; ++t) {
                    acc[t][0] = acc[t][1] = mf::f32x16{};
#pragma unroll
                    for (int kb = 0; kb < KBH; ++kb) {
                        const V a = mf::load_frag<T>(sb, bl + t * KBH + kb, lane);
                        acc[t][0] = mf::Frag<T>::mma(a, bg[0][kb], acc[t][0]);
                        acc[t][1] = mf::Frag<T>::mma(a, bg[1][kb], acc[t][1]);
                    }
                }
            } else {
                // Layer 0: the grid rows of dL/denc (the scatter input), or every row for denc.
                const int g0 = c.grid_enc >= 0 ? c.enc_offset[c.grid_enc] : 0;
                const int gw = c.grid_enc >= 0 ? c.n_levels * c.F : 0;
                const int t_lo = denc ? 0 : g0 / 32, t_hi = denc ? c.enc_padded : g0 + gw;
                for (int t = t_lo; t * 32 < t_hi; ++t) {
                    mf::f32x16 d0 = {}, d1 = {};
#pragma unroll
                    for (int kb = 0; kb < KBH; ++kb) {
                        const V a = mf::load_frag<T>(sb, t * KBH + kb, lane);
                        d0 = mf::Frag<T>::mma(a, bg[0][kb], d0);
                        d1 = mf::Frag<T>::mma(a, bg[1][kb], d1);
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int er = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3), row = er - g0;
                        if (denc && er < c.enc_padded) {
                            if (oka) denc[(size_t)er * n + qra] = d0[r];
                            if (okb) denc[(size_t)er * n + qrb] = d1[r];
                        }
                        if (row < 0 || row >= gw) continue;
                        if (oka) ws.dgrid[(size_t)row * n + qra] = d0[r];
                        if (okb) ws.dgrid[(size_t)row * n + qrb] = d1[r];
                    }
                }
            }
        }
    }
    // The tile loop is uniform per wave, so every lane of the wave reaches the vote.
    if (overflow && !__all(finite ? 1 : 0) && lane == 0) atomicOr(overflow, 1u);
}

// Transposed weight fragments for the backward chain: fragment (layer, t, kb)
// over input rows 32 t .. 32 t + 31 and output K-block kb; lane (h, i),
// element e holds W_l[o][32 t + i] with o = 16 kb + 8 h + e on the output
// layer and the chained permutation on hidden layers.
// Split-BF16 operands: hi fragment 2 f, lo fragment 2 f + 1, as k_pack_frags.
template <class T>
__global__ void __launch_bounds__(256) k_pack_bfrags(const DevCfg c, const float* __restrict__ w,
                                                     typename mf::Op<T>::elem* __restrict__ frag) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    uint32_t base = 0;
    const int H = c.n_layers - 1;
    for (int l = 0; l <= H; ++l) {
        const uint32_t iw = (uint32_t)c.layer_in[l], ow = (uint32_t)c.layer_out[l];
        const uint32_t cnt = (iw + 31u) / 32u * 32u * ow;  // rows padded to whole 32-row tiles
        if (gid < base + cnt) {
            const uint32_t r = gid - base;
            const uint32_t f = r / 512u, within = r % 512u;
            const uint32_t lane = within / 8u, e = within % 8u;
            const uint32_t kbs = ow / 16u;
            const uint32_t t = f / kbs, kb = f % kbs;
            const uint32_t hh = lane / 32u, i = lane % 32u;
            const uint32_t o = l == H ? 16u * kb + 8u * hh + e : (uint32_t)mf::chained_column((int)kb, (int)hh, (int)e);
            const uint32_t row = 32u * t + i;
            const float x = row < iw ? w[c.layer_off[l] + o * iw + row] : 0.0f;
            if constexpr (mf::Op<T>::planes == 2) {
                const size_t g0 = (size_t)(gid / 512u) * 1024u + gid % 512u;
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

__host__ uint32_t n_mlp_frags(const DevCfg& c) {
    const int L = c.n_layers - 1;
    return (c.layer_off[L] + (uint32_t)(c.layer_in[L] * c.layer_out[L])) / 512u;
}

// Persistent grid: as many 4-wave blocks as fit on the chip at once, capped by the work.
__host__ uint32_t persistent_grid(const void* kern, size_t lds, uint32_t n, int n_cu) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const uint32_t want = ((n + 63u) / 64u + 3u) / 4u;
    const uint32_t cap = (uint32_t)n_cu * (uint32_t)per_cu;
    return want < cap ? want : cap;
}

__host__ uint32_t n_bfrag_of(const DevCfg& c) {
    uint32_t f = 0;
    for (int l = 0; l < c.n_layers; ++l) f += (uint32_t)((c.layer_in[l] + 31) / 32) * (uint32_t)(c.layer_out[l] / 16);
    return f;
}

template <class T, int W, int NOUT, int H>
hipError_t launch_fwd_t(const DevCfg& c, const void* wfrag, const void* table, const void* genc, const float* in,
                        const TrainMfmaWS& ws, uint32_t n, int n_cu, hipStream_t s) {
    using G = typename mf::Op<T>::table;
    const uint32_t n_frag = n_mlp_frags(c);
    const size_t lds = (size_t)n_frag * 1024u * mf::Op<T>::planes;
    auto kern = k_train_fwd_mfma<T, W, NOUT, H>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t grid = persistent_grid((const void*)kern, lds, n, n_cu);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, c, static_cast<const T*>(wfrag),
                       static_cast<const G*>(table), static_cast<const G*>(genc), in, ws, n, n_frag);
    return hipGetLastError();
}

template <class T, int W, int NOUT, int H>
hipError_t launch_bwd_t(const DevCfg& c, const void* bfrag, const float* dout, const TrainMfmaWS& ws, float* denc,
                        uint32_t n, uint32_t* overflow, int n_cu, hipStream_t s) {
    const uint32_t n_bfrag = n_bfrag_of(c);
    const size_t lds = (size_t)n_bfrag * 1024u * mf::Op<T>::planes;
    auto kern = k_train_bwd_mfma<T, W, NOUT, H>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const uint32_t grid = persistent_grid((const void*)kern, lds, n, n_cu);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, s, c, static_cast<const T*>(bfrag), dout, ws, denc, n,
                       n_bfrag, overflow);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_grid_encode(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                              uint32_t n, hipStream_t s);
hipError_t launch_reduce_splits(const float* partials, uint32_t n_splits, uint32_t len, float* out, hipStream_t s,
                                uint32_t* overflow);
bool mfma_split_grid(const DevCfg& c);

// Shapes with compiled MFMA training kernels: width 64 with 2-4 hidden layers or
// width 128 with 2-3, and 32 or 64 raw outputs; 96 raw outputs (the
// anisotropic mixture) with 3 hidden layers.
bool mfma_train_supported(const DevCfg& c) {
    const bool shape = c.n_out_pad == 96 ? ((c.width == 64 || c.width == 128) && c.n_hidden == 3)
                                         : (c.width == 64 && c.n_hidden >= 2 && c.n_hidden <= 4) ||
                                               (c.width == 128 && c.n_hidden >= 2 && c.n_hidden <= 3);
    return shape && (c.n_out_pad == 32 || c.n_out_pad == 64 || c.n_out_pad == 96) && c.enc_padded % 16 == 0;
}

#define NPMG_TRAIN_SHAPES(CASE, TT) \
    CASE(TT, 64, 32, 2)             \
    CASE(TT, 64, 32, 3)             \
    CASE(TT, 64, 32, 4)             \
    CASE(TT, 64, 64, 2)             \
    CASE(TT, 64, 64, 3)             \
    CASE(TT, 64, 64, 4)             \
    CASE(TT, 128, 32, 2)            \
    CASE(TT, 128, 32, 3)            \
    CASE(TT, 128, 64, 2)            \
    CASE(TT, 128, 64, 3)            \
    CASE(TT, 64, 96, 3)             \
    CASE(TT, 128, 96, 3)
// split-BF16 operands: width 64 (two weight planes of a 128-wide network exceed the LDS)
#define NPMG_TRAIN_SHAPES_NARROW(CASE, TT) \
    CASE(TT, 64, 32, 2)                    \
    CASE(TT, 64, 32, 3)                    \
    CASE(TT, 64, 32, 4)                    \
    CASE(TT, 64, 64, 2)                    \
    CASE(TT, 64, 64, 3)                    \
    CASE(TT, 64, 64, 4)                    \
    CASE(TT, 64, 96, 3)

hipError_t launch_train_forward_mfma(const DevCfg& c, int precision, const void* wfrag, const void* table, void* genc,
                                     const float* in, const TrainMfmaWS& ws, uint32_t n, int n_cu, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!mfma_train_supported(c)) return hipErrorInvalidValue;
    if (mfma_split_grid(c)) {
        if (!genc) return hipErrorInvalidValue;
        hipError_t e = launch_grid_encode(c, precision, table, in, genc, n, s);
        if (e != hipSuccess) return e;
    } else {
        genc = nullptr;
    }
#define NPMG_FWD_CASE(TT, WW, NO, HH)                              \
    if (c.width == WW && c.n_out_pad == NO && c.n_hidden == HH) \
        return launch_fwd_t<TT, WW, NO, HH>(c, wfrag, table, genc, in, ws, n, n_cu, s);
    if (precision == PREC_BF16) { NPMG_TRAIN_SHAPES(NPMG_FWD_CASE, __bf16) }
    else if (precision == PREC_FP16) { NPMG_TRAIN_SHAPES(NPMG_FWD_CASE, _Float16) }
    else if (precision == PREC_BF16X3) { NPMG_TRAIN_SHAPES_NARROW(NPMG_FWD_CASE, mf::Bf16x3) }
#undef NPMG_FWD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_train_backward_mfma(const DevCfg& c, int precision, const void* bfrag, const float* dout,
                                      const TrainMfmaWS& ws, float* denc, uint32_t n, uint32_t* overflow, int n_cu,
                                      hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!mfma_train_supported(c)) return hipErrorInvalidValue;
#define NPMG_BWD_CASE(TT, WW, NO, HH)                              \
    if (c.width == WW && c.n_out_pad == NO && c.n_hidden == HH) \
        return launch_bwd_t<TT, WW, NO, HH>(c, bfrag, dout, ws, denc, n, overflow, n_cu, s);
    if (precision == PREC_BF16) { NPMG_TRAIN_SHAPES(NPMG_BWD_CASE, __bf16) }
    else if (precision == PREC_FP16) { NPMG_TRAIN_SHAPES(NPMG_BWD_CASE, _Float16) }
    else if (precision == PREC_BF16X3) { NPMG_TRAIN_SHAPES_NARROW(NPMG_BWD_CASE, mf::Bf16x3) }
#undef NPMG_BWD_CASE
    return hipErrorInvalidValue;
}
#undef NPMG_TRAIN_SHAPES
#undef NPMG_TRAIN_SHAPES_NARROW

size_t train_mask_words(const DevCfg& c, uint32_t n) {
    return (size_t)((n + 63u) / 64u) * (size_t)c.n_hidden * (size_t)(c.width / 32) * 64u;
}

size_t bfrag_elements(const DevCfg& c) { return (size_t)n_bfrag_of(c) * 512u; }

hipError_t launch_pack_bfrags(const DevCfg& c, int precision, const float* w, void* frag, hipStream_t s) {
    const uint32_t n = n_bfrag_of(c) * 512u;
    const dim3 grid((n + 255) / 256);
    if (precision == PREC_BF16) hipLaunchKernelGGL((k_pack_bfrags<__bf16>), grid, dim3(256), 0, s, c, w, (__bf16*)frag);
    else if (precision == PREC_FP16) hipLaunchKernelGGL((k_pack_bfrags<_Float16>), grid, dim3(256), 0, s, c, w, (_Float16*)frag);
    else if (precision == PREC_BF16X3)
        hipLaunchKernelGGL((k_pack_bfrags<mf::Bf16x3>), grid, dim3(256), 0, s, c, w, (__bf16*)frag);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace npmg
