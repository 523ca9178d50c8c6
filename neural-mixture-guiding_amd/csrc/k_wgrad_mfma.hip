// k_wgrad_mfma.hip -- weight gradients of the MFMA training step: dW_l = dY_l
// X_l^T over the batch (tiny-cuda-nn's weight-gradient GEMMs after the fully
// fused backward, fully_fused_mlp.cu:806-836 via cutlass_matmul.h:440-482),
// one partial row per wave, summed in a fixed order by k_reduce_splits.
// The operands are the layer inputs and output gradients that the forward and
// backward chains stored ([dim][n], operand type, hi and lo planes for the
// split type), except the layer-0 inputs of the "grid, then OneBlob(16)"
// layout with the fast numerics: those are rebuilt here from the level-major
// grid features and the input records with the forward's own operations (so
// the same bits), instead of a stored copy of 2 or 4 B per feature and sample
// written by the forward and read back (measured: 0.08 ms less per 2^18-sample
// split-BF16 step, although each OneBlob bin is computed per row here).
#include "device.h"
#include "kernels.h"
#include "mfma.h"

namespace npmg {

namespace {

// Bin `bin` of the OneBlob(16) encoding of x with the fast numerics: the
// value mf::oneblob16_frag puts there (the same operations, so the same bits).
__device__ __forceinline__ float oneblob16_bin(float x, int bin) {
    const float s = x * 16.0f;
    const float fl = __builtin_floorf(s);
    const float t = s - fl;
    const int kx = (int)fl;
    auto odd = [](float u) {
        const float u2 = u * u;
        return u * __builtin_fmaf(u2, __builtin_fmaf(u2, 0.1875f, -0.625f), 0.9375f);
    };
    const float st = odd(t), sr = odd(1.0f - t);
    const uint32_t r = (uint32_t)(bin + 1 - kx) & 15u;
    return r == 0u ? 0.5f - st : (r == 1u ? st + sr : (r == 2u ? 0.5f - sr : 0.0f));
}

// Weight gradients on MFMA: dW_l[o][i] = sum_q dY_l[o][q] X_l[i][q] with the
// batch as the K dimension. One launch per layer; block = one range of
// q_per_row samples (a multiple of 64), whose partial row of dW_l a
// deterministic reduction over the ranges sums afterwards. The layer's 32 x 32 output tiles are cut into
// groups of one input tile and up to two output tiles; the 4 waves of the
// block take the groups in turn over the same samples, so an operand row
// loaded by several waves is fetched once and served to the others by the
// CU's L1, and each wave holds few rows (occupancy 2).
// The sum over the batch does not depend on which samples share a K-step, so
// a block of 64 samples is read with lane (h, j) loading the 32 consecutive
// samples 32 h .. 32 h + 31 of its row (four 16-byte loads per plane) and
// K-step s taking samples 8 s .. 8 s + 7 of that run on both operands: the two
// lanes of a row read one whole 128-byte line in four back-to-back loads.
// (Reading 16 samples per K-step, 32 bytes of a line at a time, keeps every
// line of every row open across four K-steps; with hundreds of lines per wave
// and every wave of an XCD doing so, the L2 evicted the lines in between and
// fetched each about four times.)
// genc (with in): the layer-0 inputs are rebuilt from the level-major grid
// features (K-blocks below ob_kb0) and the OneBlob(16) bins of the input
// records, instead of read from ws.x0.
template <class T>
__global__ void __launch_bounds__(256) k_wgrad_mfma(const DevCfg c, const TrainMfmaWS ws,
                                                       const typename mf::Op<T>::table* __restrict__ genc,
                                                       const float* __restrict__ in, float* __restrict__ partials,
                                                       uint32_t n, uint32_t q_per_row, uint32_t n_mlp, int layer) {
    using V = typename mf::Frag<T>::vec;
    using E = typename mf::Op<T>::elem;
    using GT = typename mf::Op<T>::table;
    constexpr int NP = mf::Op<T>::planes;
    const int l = layer;
    const int H = c.n_layers - 1;
    const int iw = c.layer_in[l], ow = c.layer_out[l];
    const E* X = l == 0 ? static_cast<const E*>(ws.x0) : static_cast<const E*>(ws.xh) + (size_t)(l - 1) * c.width * n;
    const E* G = l == H ? static_cast<const E*>(ws.dyo) : static_cast<const E*>(ws.dyh) + (size_t)l * c.width * n;
    // plane strides of the split type: the buffers' sizes
    const size_t px = (l == 0 ? (size_t)c.enc_padded : (size_t)H * c.width) * n;
    const size_t pg = (l == H ? (size_t)c.n_out_pad : (size_t)H * c.width) * n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    const uint32_t row = blockIdx.x;
    const uint32_t qb = row * q_per_row;
    const uint32_t qe = min(n, qb + q_per_row);
    const int n_to = ow / 32, n_ti = (iw + 31) / 32;
    // groups: (input tile, pair of output tiles), or single output tiles when
    // pairs would leave waves without work
    // (split operands: one output tile per group, so that a wave's operands
    // of a whole block fit at two waves per SIMD)
    constexpr int TPG = NP == 2 ? 1 : 2;
    const int tpg = TPG == 2 && n_ti * ((n_to + 1) / 2) >= 4 ? 2 : 1;
    const int n_grp = n_ti * ((n_to + tpg - 1) / tpg);
    float* P = partials + (size_t)row * n_mlp + c.layer_off[l];
    // 16-byte fragment loads need 16-byte aligned rows: n a multiple of 8.
    const bool vec = (n & 7u) == 0u;
    // K-step s of the block at q: samples q + 32 h + 8 s + e, e = 0..7.
    auto load = [&](const E* rowp, size_t stride, uint32_t q, int s, bool full) {
        V v;
        const uint32_t q0 = q + 32u * (uint32_t)h + 8u * (uint32_t)s;
        if (full) {
            v = mf::load_rows16<T>(rowp + q0, stride);
        } else {
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t qq = q0 + (uint32_t)e;
                    mf::set<T>(v, p, e, qq < qe ? rowp[p * stride + qq] : (E)0.0f);
                }
        }
        return v;
    };
    const bool rebuild = l == 0 && genc;
    // Layer-0 input row i rebuilt: grid features (FP16 or BF16) or a OneBlob bin.
    auto load_x0 = [&](int i, uint32_t q, int s, bool full) {
        float v[8];
        const uint32_t q0 = q + 32u * (uint32_t)h + 8u * (uint32_t)s;
        const int kb = i >> 4;
        if (kb < c.ob_kb0) {
            const GT* src = genc + (size_t)i * n + q0;
            if (full) {
                const uint4 raw = *reinterpret_cast<const uint4*>(src);
                GT g8[8];
                __builtin_memcpy(g8, &raw, 16);
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = (float)g8[e];
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) v[e] = q0 + (uint32_t)e < qe ? (float)src[e] : 0.0f;
            }
        } else {
            const float* src = in + (size_t)(c.ob_in + kb - c.ob_kb0) * n + q0;
            float x[8];
            if (full) {
                const float4 a = *reinterpret_cast<const float4*>(src), b = *reinterpret_cast<const float4*>(src + 4);
                x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
                x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = q0 + (uint32_t)e < qe ? src[e] : 0.0f;
            }
            const int bin = i & 15;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = oneblob16_bin(x[e], bin);
        }
        return mf::pack8<T>(v);
    };
    for (int grp = wave; grp < n_grp; grp += 4) {
        const int per_ti = (n_to + tpg - 1) / tpg;
        const int ti = grp / per_ti, to0 = (grp % per_ti) * tpg;
        const int nto = n_to - to0 < tpg ? n_to - to0 : tpg;
        mf::f32x16 acc[TPG];
#pragma unroll
        for (int a = 0; a < TPG; ++a) acc[a] = mf::f32x16{};
        const E* ga[TPG];
#pragma unroll
        for (int a = 0; a < TPG; ++a) ga[a] = G + (size_t)(32 * (to0 + (a < nto ? a : 0)) + j) * n;
        // rows beyond iw (a partial tile) read row iw - 1 and are not stored
        const int xi = 32 * ti + j < iw ? 32 * ti + j : iw - 1;
        const E* xb = X + (size_t)xi * n;
        // Where the group's input rows come from, the same for the whole wave:
        // the stored operands (0), or for rebuilt layer-0 inputs a tile of grid
        // features (1), of OneBlob bins (2) or of both (3, lane by lane). Each
        // case issues all of a block's loads before any use (a per-lane choice
        // between them made the compiler wait for each load in turn).
        const int kind = __builtin_amdgcn_readfirstlane(
            !rebuild ? 0 : ((32 * ti + 31) >> 4) < c.ob_kb0 ? 1 : ((32 * ti) >> 4) >= c.ob_kb0 ? 2 : 3);
        auto load_a = [&](uint32_t q, bool full, V (&av)[TPG][4]) {
#pragma unroll
            for (int a = 0; a < TPG; ++a)
#pragma unroll
                for (int s = 0; s < 4; ++s) av[a][s] = load(ga[a], pg, q, s, full);
        };
        auto block = [&](uint32_t q, bool full) {
            V av[TPG][4], bv[4];
            if (kind == 0) {
#pragma unroll
                for (int s = 0; s < 4; ++s) bv[s] = load(xb, px, q, s, full);
                load_a(q, full, av);
            } else if (kind == 1) {
                // grid features of row xi (16-bit), converted after all loads are out
                GT g[4][8];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const uint32_t q0 = q + 32u * (uint32_t)h + 8u * (uint32_t)s;
                    const GT* src = genc + (size_t)xi * n + q0;
                    if (full) {
                        const uint4 raw = *reinterpret_cast<const uint4*>(src);
                        __builtin_memcpy(g[s], &raw, 16);
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e) g[s][e] = q0 + (uint32_t)e < qe ? src[e] : (GT)0.0f;
                    }
                }
                load_a(q, full, av);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    float v[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = (float)g[s][e];
                    bv[s] = mf::pack8<T>(v);
                }
            } else if (kind == 2) {
                // OneBlob input of row xi's dimension (FP32), binned after all loads are out
                const float* row = in + (size_t)(c.ob_in + (xi >> 4) - c.ob_kb0) * n;
                float x[4][8];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const uint32_t q0 = q + 32u * (uint32_t)h + 8u * (uint32_t)s;
                    if (full) {
                        const float4 a = *reinterpret_cast<const float4*>(row + q0);
                        const float4 b = *reinterpret_cast<const float4*>(row + q0 + 4);
                        x[s][0] = a.x; x[s][1] = a.y; x[s][2] = a.z; x[s][3] = a.w;
                        x[s][4] = b.x; x[s][5] = b.y; x[s][6] = b.z; x[s][7] = b.w;
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e) x[s][e] = q0 + (uint32_t)e < qe ? row[q0 + e] : 0.0f;
                    }
                }
                load_a(q, full, av);
                const int bin = xi & 15;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    float v[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = oneblob16_bin(x[s][e], bin);
                    bv[s] = mf::pack8<T>(v);
                }
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) bv[s] = load_x0(xi, q, s, full);
                load_a(q, full, av);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int a = 0; a < TPG; ++a)
                    if (a < nto) acc[a] = mf::Frag<T>::mma(av[a][s], bv[s], acc[a]);
        };
        uint32_t q = qb;
        if (vec)
            for (; q + 64u <= qe; q += 64u) block(q, true);
        for (; q < qe; q += 64u) block(q, false);
        const int i = 32 * ti + j;
        if (i < iw) {
#pragma unroll
            for (int a = 0; a < TPG; ++a) {
                if (a >= nto) break;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int o = 32 * (to0 + a) + 8 * (r >> 2) + 4 * h + (r & 3);
                    P[(size_t)o * iw + i] = acc[a][r];
                }
            }
        }
    }
}

// Partial rows of the weight gradient: one workgroup per row and layer, about
// 512 samples each in whole 64-sample blocks (at most 1024 rows; the rows are
// summed by k_reduce_splits).
void wgrad_split(uint32_t n, uint32_t* rows, uint32_t* q_per_row) {
    uint32_t r = (n + 511u) / 512u;
    if (r > 1024u) r = 1024u;
    if (r < 1u) r = 1u;
    *rows = r;
    *q_per_row = ((n + r - 1u) / r + 63u) / 64u * 64u;
}

}  // namespace

hipError_t launch_reduce_splits(const float* partials, uint32_t n_splits, uint32_t len, float* out, hipStream_t s,
                                uint32_t* overflow);

hipError_t launch_wgrad_mfma(const DevCfg& c, int precision, const TrainMfmaWS& ws, const void* genc, const float* in,
                             float* partials, size_t partials_cap, float* dw, uint32_t n, uint32_t* overflow,
                             hipStream_t s) {
    // the forward left the layer-0 inputs unstored exactly when it built them
    // from the level-major features and OneBlob(16) bins
    const bool rebuild = c.numerics == NUM_FAST && c.ob_nkb > 0;
    if (rebuild && (!genc || !in)) return hipErrorInvalidValue;
    if (!rebuild) genc = nullptr;
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    uint32_t rows, qpw;
    wgrad_split(n, &rows, &qpw);
    if ((size_t)rows * n_mlp > partials_cap) return hipErrorInvalidValue;
    // one launch per layer: each streams one layer's operands (a single launch
    // over all layers interleaved their workgroups and took 0.27 ms against
    // 0.22 ms for the four launches one after another, split BF16 at 2^18)
    const dim3 grid(rows);
    for (int l = 0; l < c.n_layers; ++l) {
        if (precision == PREC_BF16)
            hipLaunchKernelGGL((k_wgrad_mfma<__bf16>), grid, dim3(256), 0, s, c, ws, static_cast<const __bf16*>(genc),
                               in, partials, n, qpw, n_mlp, l);
        else if (precision == PREC_FP16)
            hipLaunchKernelGGL((k_wgrad_mfma<_Float16>), grid, dim3(256), 0, s, c, ws,
                               static_cast<const _Float16*>(genc), in, partials, n, qpw, n_mlp, l);
        else if (precision == PREC_BF16X3)
            hipLaunchKernelGGL((k_wgrad_mfma<mf::Bf16x3>), grid, dim3(256), 0, s, c, ws,
                               static_cast<const _Float16*>(genc), in, partials, n, qpw, n_mlp, l);
        else return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_splits(partials, rows, n_mlp, dw, s, overflow);
}

size_t wgrad_mfma_partials_needed(const DevCfg& c, uint32_t n) {
    const uint32_t n_mlp = c.layer_off[c.n_layers - 1] + (uint32_t)(c.layer_in[c.n_layers - 1] * c.layer_out[c.n_layers - 1]);
    uint32_t rows, qpw;
    wgrad_split(n, &rows, &qpw);
    return (size_t)rows * n_mlp;
}

}  // namespace npmg
