// kernels.h -- launchers of the HIP kernels (host-callable, return hipError_t).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "config.h"

namespace npmg {

// Per-query inputs and outputs of the fused query path (all SoA, device).
struct QueryIO {
    const float* in;         // [n_in][n]
    const float* normal;     // [3][n] or null (cosine product off)
    const float* u;          // [3][n] or null (no sampling)
    const float* eval_dirs;  // [n_eval][3][n]
    int n_eval;
    float* out_dir;          // [3][n]
    float* out_pdf;          // [n]
    float* out_eval_pdf;     // [n_eval][n]
    uint32_t* out_lobe;      // [n] or null
    float* raw_out;          // [n_out_pad][n] or null
    float* lobes_out;        // [K][5][n] or null (consumer kernel only)
    float* out_sel;          // [n] learned selection probability, or null
};

// Several query streams (the per-bounce queues of a wavefront frame) served by
// one persistent launch: the query counts live on the device, everything
// else is per queue.
constexpr int kMaxQueues = 16;
struct QueueSet {
    QueryIO io[kMaxQueues];
    uint32_t stride[kMaxQueues];  // row stride (capacity) of each queue's arrays
    void* genc[kMaxQueues];       // level-major grid features of each queue (MFMA path)
    const uint32_t* counts;       // [n_queues] queries per queue (device)
    uint32_t* counter;            // work counter (device), zero before the launch
    uint32_t n_queues;
    uint32_t batch;               // consecutive work items a workgroup or wave claims per atomic (>= 1)
};

// first[i] = tiles of `tile` queries in queues 0 .. i-1; first[n_queues] = total.
__device__ __forceinline__ void queue_tile_prefix(const QueueSet& qs, uint32_t tile, uint32_t* first) {
    first[0] = 0u;
    for (uint32_t i = 0; i < (uint32_t)kMaxQueues; ++i) {
        const uint32_t t = i < qs.n_queues ? (min(qs.counts[i], qs.stride[i]) + tile - 1u) / tile : 0u;
        first[i + 1] = first[i] + t;
    }
}

// Training-side activations kept for the backward pass (SoA, device).
struct TrainWS {
    float* enc;   // [enc_padded][n]
    float* acts;  // [n_hidden][width][n] post-ReLU
    float* g;     // [n_hidden][width][n] dL/d(pre-activation) of each hidden layer
    float* dout;  // [n_out_pad][n]
    float* loss;  // [n]
    float* dgrid; // [n_levels F][n] dL/d(grid features)
};

// Loss scale seen by the adjoint: a device scalar (dynamic loss scaling) or a
// host value, and 1 / batch size. The reference clamps the scaled gradient at
// loss scale 128 (train.cu:116-122); the threshold here is grad_clamp * scale /
// clamp_scale, so the clamp acts on the same unscaled values at any scale.
struct LossScale {
    const float* dev;  // device scalar or null
    float scale;       // used when dev is null
    float inv_n;
    __device__ float value() const { return dev ? dev[0] : scale; }
    __device__ float over_n() const { return value() * inv_n; }
    __device__ float clamp(const DevCfg& c) const {
        return c.grad_clamp > 0.0f ? c.grad_clamp * (value() / c.clamp_scale) : 0.0f;
    }
};

// Device-resident optimiser/loss-scale state, so the train step needs no host
// round trip (and can be captured in a graph).
struct ScaleState {
    float scale;        // current loss scale
    uint32_t good;      // finite steps since the scale last changed
    uint32_t overflow;  // raised by the backward kernels when a gradient is not finite
    uint32_t step;      // optimiser steps applied (EMA debiasing)
};

// Training buffers of the MFMA path (T = the operand type, [dim][n] unless noted).
struct TrainMfmaWS {
    void* x0;        // [enc_padded][n] layer-0 inputs (T)
    void* xh;        // [H][W][n] inputs of layers 1..H, post-ReLU (T)
    void* dyh;       // [H][W][n] gradients w.r.t. the pre-activations of hidden layers 0..H-1 (T)
    void* dyo;       // [n_out_pad][n] gradients w.r.t. the raw outputs (T)
    uint32_t* mask;  // [ceil(n / 64)][H][W / 32][64] ReLU sign bits in accumulator order
    float* raw;      // [n_out_pad][n] raw outputs (FP32)
    float* dgrid;    // [L F][n] dL/d(grid features) (FP32)
};

// FP32 fused query: encode -> MLP (FP32 VALU) -> mixture -> sample/pdf.
// wt: MLP weights transposed per layer ([in][out]); table: grid parameters.
hipError_t launch_query_fp32(const DevCfg& c, const float* wt, const float* table, const QueryIO& io, uint32_t n,
                             int n_cu, hipStream_t s);
// Consumer kernel: raw outputs -> mixture -> sample/pdf.
hipError_t launch_mixture(const DevCfg& c, const float* raw, const QueryIO& io, uint32_t n, hipStream_t s);
hipError_t launch_encode(const DevCfg& c, const float* table, const float* in, float* enc, uint32_t n, hipStream_t s);
hipError_t launch_grid_corners(const DevCfg& c, const float* in, uint32_t* idx, float* w, uint32_t n, hipStream_t s);

// Training (FP32): forward storing activations (+ adjoint when tgt != null).
// tgt_bsdf [n]: the BSDF sampling pdf of each record, for the learned
// selection probability (null otherwise).
hipError_t launch_train_forward_fp32(const DevCfg& c, const float* wt, const float* table, const float* in,
                                     const float* tgt_dir, const float* tgt_rgb, const float* tgt_pdf,
                                     const float* tgt_bsdf, LossScale ls, const TrainWS& ws, uint32_t n, int n_cu,
                                     hipStream_t s);
// Backward through the MLP: g of each hidden layer into ws.g, dL/d(grid
// features) into dgrid [L F][n], dL/denc into denc (optional).
hipError_t launch_train_backward_fp32(const DevCfg& c, const float* w, const TrainWS& ws, float* denc, float* dgrid,
                                      uint32_t n, int n_cu, hipStream_t s);
// Grid-gradient scatter (tiny-cuda-nn kernel_grid_backward, grid.h:386-509),
// binned by table tile and reduced in LDS; every word of grid_grad is written.
// Deterministic mode sums exactly in int64 fixed point. Scratch: at least
// grid_scatter_bytes; its first grid_scatter_clear_bytes must be zero before
// the first call (the scatter leaves them zero again). A non-finite dL/dy
// contributes zero and raises *overflow when it is given. The binning needs
// only the positions: launch_grid_scatter_prepare (in, n) may run before
// dgrid exists, then launch_grid_scatter_finish with the same in, n and
// scratch; launch_grid_scatter is both.
size_t grid_scatter_bytes(const DevCfg& c, uint32_t n, bool deterministic);
// Wavefront packers (train.cu:14-67, 174-191). Scene box as min xyz, max xyz.
struct SceneBox {
    float lo[3], hi[3];
};
// The guided pixel state of a frame, SoA per depth (workqueue.h:26-69,
// guideditem.h:23-44): field[d][c][p] at ((d * C) + c) * n_pixels + p.
struct PathRecords {
    const float* L;          // [D][3][P] radiance gathered after the vertex
    const float* thp;        // [D][3][P] path throughput at the vertex
    const float* pos;        // [D][3][P] world position
    const float* dir;        // [D][2][P] scattered direction (theta, phi)
    const float* wi_pdf;     // [D][P] effective pdf of the scattered direction
    const float* mis_weight; // [D][P]
    const uint8_t* delta;    // [D][P] sampled from a delta lobe
    const float* bsdf_val;   // [D][3][P]
    const float* wo;         // [D][2][P] outgoing direction (theta, phi)
    const float* normal;     // [D][2][P] shading normal (theta, phi)
    const float* roughness;  // [D][P]
    const uint32_t* depth;   // [P] recorded vertices (curDepth)
    const float* bsdf_pdf;   // [D][P] BSDF pdf of the scattered direction, or null
    uint32_t max_depth, n_pixels;
};
// Training pixels: thread t -> pixel (t % res[0], t / res[0]) * stride + offset.
struct TrainTile {
    uint32_t offset[2], stride[2], res[2], full[2];
};
hipError_t launch_pack_queries(const float* pos, const float* wo, const float* normal, const float* roughness,
                               const SceneBox& box, uint32_t n, float* in, hipStream_t s);
// Scratch: pack_training_scratch(tile) uint32 values.
size_t pack_training_scratch(const TrainTile& t);
hipError_t launch_pack_training(const PathRecords& r, const TrainTile& t, const SceneBox& box, uint32_t capacity,
                                float* in, float* tgt_dir, float* tgt_rgb, float* tgt_pdf, float* tgt_bsdf,
                                uint32_t* count, uint32_t* scratch, hipStream_t s);

// FP8 inference: fragments with per-layer weight scales and the fused kernel's
// multipliers (fp8 = wscale [L], ascale [L], qscale [L + 1]); calibration maxima.
hipError_t launch_pack_frags_fp8(const DevCfg& c, const float* w, void* frag, const float* amax, float* fp8,
                                 hipStream_t s);
hipError_t launch_fp8_amax(const DevCfg& c, const float* enc, const float* acts, uint32_t m, float* amax,
                           hipStream_t s);
// Zero fill by a kernel (safe inside captured graphs, unlike a memset node).
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);
size_t grid_scatter_clear_bytes(const DevCfg& c, bool deterministic);
hipError_t launch_grid_scatter(const DevCfg& c, const float* in, const float* dgrid, float* grid_grad, uint32_t n,
                               bool deterministic, void* scratch, size_t scratch_bytes, uint32_t* overflow,
                               hipStream_t s);
hipError_t launch_grid_scatter_prepare(const DevCfg& c, const float* in, uint32_t n, bool deterministic,
                                       void* scratch, size_t scratch_bytes, hipStream_t s);
hipError_t launch_grid_scatter_finish(const DevCfg& c, const float* in, const float* dgrid, float* grid_grad,
                                      uint32_t n, bool deterministic, void* scratch, size_t scratch_bytes,
                                      uint32_t* overflow, hipStream_t s);
// Weight gradients dW_l = sum_q g_l x_l^T, deterministic split-K reduction.
// Non-finite sums raise *overflow when it is given.
hipError_t launch_wgrad_fp32(const DevCfg& c, const TrainWS& ws, float* partials, size_t partials_cap, float* dw,
                             uint32_t n, uint32_t* overflow, hipStream_t s);
// Adam (+EMA) over all parameters.
// The gradient is divided by the loss scale (st->scale, or loss_scale without
// state); with check_overflow a raised st->overflow skips the whole step. The
// EMA debiasing reads decay^t from ema_pow[min(t, ema_pow_len - 1)] with
// t = st->step + 1 (tcnn computes std::pow in double on the host).
struct AdamArgs {
    size_t n, n_matrix;
    float loss_scale, lr, beta1, beta2, eps, l2_reg;
    bool ema, scale_from_state, check_overflow;
    // EMA recurrence in FP16 (tcnn's ema_step_half_precision with __half
    // weights, optimizers/ema.h:62-76): the input weight and the stored
    // average are rounded to FP16 each step
    bool ema_half;
    float ema_decay;
    const ScaleState* st;
    const float* ema_pow;
    uint32_t ema_pow_len;
    // Adam's bias correction sqrt(1 - beta2^t) / (1 - beta1^t) for t < adam_bias_len
    // (launch_adam_bias_table; later steps compute it)
    const float* adam_bias;
    uint32_t adam_bias_len;
    // operand-type copies of the grid table (parameters n_matrix..n) written in
    // the same pass: precision PREC_BF16 / PREC_FP16, or none for PREC_FP32
    int half_precision;
    void* half_train;  // from the updated parameters
    void* half_infer;  // from the EMA (the updated parameters without EMA)
};
hipError_t launch_adam(const AdamArgs& a, float* master, const float* grad, float* m, float* v, uint32_t* steps,
                       float* ema, hipStream_t s);
// out[t] = sqrt(1 - beta2^t) / (1 - beta1^t) for t in [0, len), with the
// operations k_adam uses, so a table entry equals the value computed in place.
hipError_t launch_adam_bias_table(float beta1, float beta2, float* out, uint32_t len, hipStream_t s);
// After the optimiser: counts the step and, with dynamic scaling, halves the
// scale on overflow or doubles it after `window` finite steps; clears the flag.
hipError_t launch_scale_update(ScaleState* st, bool dynamic, uint32_t window, float max_scale, hipStream_t s);
// Transposes the MLP weights of every layer ([out][in] -> [in][out]).
hipError_t launch_transpose_mlp(const DevCfg& c, const float* w, float* wt, hipStream_t s);
// Mean of n values into out[0] (fixed summation order). partials: 256 floats.
hipError_t launch_mean(const float* x, uint32_t n, float* partials, float* out, hipStream_t s);
hipError_t launch_adjoint(const DevCfg& c, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                          const float* tgt_pdf, const float* tgt_bsdf, LossScale ls, float* dout, float* loss,
                          uint32_t n, hipStream_t s);

hipError_t launch_synth_queries(uint64_t seed, uint64_t first, uint32_t n, float* in, float* normal, float* u,
                                float* eval_dirs, hipStream_t s);
struct FieldParams {
    float mu0[3][3], axis[3][3], d[3][3], kappa[3], a[3];
};
FieldParams make_field(uint64_t seed);
hipError_t launch_synth_train(uint64_t seed, uint64_t first, uint32_t n, const FieldParams& f, float* in,
                              float* tgt_dir, float* tgt_rgb, float* tgt_pdf, hipStream_t s);
hipError_t launch_debug_math(int fn, const float* x, const float* y, uint32_t n, float* out, hipStream_t s);

}  // namespace npmg
