/* npmg.h -- C ABI of the MI355X neural parametric mixture guiding library.
 *
 * This boundary replaces the calls that the reference's guided integrator
 * makes into tiny-cuda-nn and into the device-side mixture class:
 *   - Network::inference             src/render/guided/integrator.cpp:645
 *   - Network::forward / backward    src/render/guided/integrator.cpp:681, :694
 *   - compute_dL_doutput_divergence  src/render/guided/integrator.cpp:689-691,
 *                                    src/render/guided/train.cu:69-124
 *   - Trainer::optimizer_step        src/render/guided/integrator.cpp:695
 *   - thrust::reduce of the loss     src/render/guided/integrator.cpp:697
 *   - MixedSphericalGaussianDistribution<8>{ctor, applyCosineLobe, sample, pdf}
 *     in the consumer kernels        src/render/guided/integrator.cpp:198-201,
 *                                    :257-260, :297-317
 *   - resetNetwork / initialize_params src/render/guided/integrator.cpp:555-587
 *
 * Conventions
 *   - Every pointer is a device pointer unless documented as host.
 *   - Per-query data is structure-of-arrays: a field of d values is [d][n],
 *     query index fastest (the reference uses column-major [dims][n] network
 *     inputs, tiny-cuda-nn/object.h:147-169).
 *   - n is arbitrary (the reference's multiple-of-128 rule is not required).
 *   - Calls that take a stream enqueue asynchronously on it; a NULL stream is
 *     the default stream. Create, destroy, parameter get/set and size queries
 *     are synchronous.
 *   - Errors are returned as npmg_status; no C++ exception crosses the ABI.
 *     npmg_last_error() returns a message for the last failing call on a
 *     handle (or, with NULL, for the last failing npmg_create).
 *   - A handle is not thread-safe; use one handle per host thread.
 */
#ifndef NPMG_H
#define NPMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct npmg_ctx npmg_ctx;
typedef void* npmg_stream; /* hipStream_t */

typedef enum {
    NPMG_OK = 0,
    NPMG_EINVAL = 1, /* bad argument or configuration */
    NPMG_ENOMEM = 2, /* device allocation failed */
    NPMG_EHIP = 3,   /* HIP runtime error */
    NPMG_ESTATE = 4  /* call not valid in the current state */
} npmg_status;

/* sample/pdf flags */
#define NPMG_COSINE_AWARE 1u /* multiply each lobe by the cosine vMF about `normal`
                                (applyCosineLobe, distribution.h:184-195) */

/* Creates a guiding network from a JSON configuration (host string). The
 * "encoding", "network" and "optimizer" objects use tiny-cuda-nn's keys, so the
 * reference's common/configs/nn/base.json "nn" object is accepted as is;
 * "mixture", "loss", "precision" and "grid_gradient" select the build's modes.
 * Replaces GuidedPathTracer::resetNetwork (integrator.cpp:555-587) including
 * Trainer::initialize_params (trainer.h:53-92) with `seed` in place of 7272. */
npmg_status npmg_create(const char* json_config, uint64_t seed, npmg_ctx** out);
void npmg_destroy(npmg_ctx* ctx);
const char* npmg_last_error(const npmg_ctx* ctx);

/* Parameter counts: MLP weights first, then the grid table (GuidingNetwork
 * parameter order, src/render/guided/network.h:85-100). */
npmg_status npmg_n_params(const npmg_ctx* ctx, size_t* n_mlp, size_t* n_grid);
/* Encoded width (padded to 16) and padded output width of the network. */
npmg_status npmg_widths(const npmg_ctx* ctx, uint32_t* encoded_width, uint32_t* padded_output);
/* Grid level table: offsets[n_levels+1] (entries), scales[n_levels], resolutions[n_levels]
 * (tiny-cuda-nn encodings/grid.h:940-1004). Host arrays. */
npmg_status npmg_grid_layout(const npmg_ctx* ctx, uint32_t* n_levels, uint32_t* offsets, float* scales,
                             uint32_t* resolutions);

/* Full-precision parameters in the tcnn order. set_params writes the training
 * master copy and the inference (EMA) copy (and resets the FP8 calibration);
 * which = 0 reads the training
 * master, which = 1 the inference copy. Trainer::set_params_full_precision /
 * params_inference (trainer.h:314-328). */
npmg_status npmg_set_params(npmg_ctx* ctx, const float* params, size_t n, int on_device);
npmg_status npmg_get_params(const npmg_ctx* ctx, float* params, size_t n, int on_device, int which);
/* Writes one copy only: which = 0 the training master (the parameters that
 * tiny-cuda-nn's PyTorch binding passes to each forward and backward call,
 * bindings/torch/tinycudann/modules.py:82-151, which leave the EMA alone),
 * 1 the inference copy (the EMA; the master when there is no EMA), 2 both
 * (npmg_set_params). Synchronous. A new inference copy resets the FP8
 * calibration (the next inference call recalibrates). */
npmg_status npmg_set_params_which(npmg_ctx* ctx, const float* params, size_t n, int on_device, int which);

/* Network inference with the inference (EMA) weights: in [9][n] -> raw
 * network outputs [n_out][n] (Network::inference, object.h:147-169). */
npmg_status npmg_infer(npmg_ctx* ctx, const float* in, float* raw_out, uint32_t n, npmg_stream stream);

/* Fused query path: encode -> MLP -> mixture -> optional cosine product ->
 * one sample (u [3][n]: lobe choice then the 2D direction sample, the draw
 * order of distribution.h:132-140) -> pdf of the sample and of n_eval given
 * directions eval_dirs [n_eval][3][n]. normal [3][n] is the face-forwarded
 * shading normal used with NPMG_COSINE_AWARE (may be NULL otherwise).
 * Outputs: out_dir [3][n], out_pdf [n], out_eval_pdf [n_eval][n], out_lobe
 * [n] (may be NULL). u may be NULL to skip sampling (then out_dir/out_pdf are
 * unused). raw_out [n_out][n] optionally receives the network outputs. */
npmg_status npmg_sample_pdf(npmg_ctx* ctx, const float* in, const float* normal, const float* u,
                            const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                            float* out_eval_pdf, uint32_t* out_lobe, float* raw_out, uint32_t n,
                            uint32_t flags, npmg_stream stream);

/* Consumer side only: the same mixture work from raw outputs [n_out][n]
 * produced by npmg_infer (the three consumer lambdas of integrator.cpp:158-340). */
npmg_status npmg_mixture_sample_pdf(npmg_ctx* ctx, const float* raw, const float* normal, const float* u,
                                    const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                                    float* out_eval_pdf, uint32_t* out_lobe, float* lobes_out, uint32_t n,
                                    uint32_t flags, npmg_stream stream);

/* One training step on a batch (integrator.cpp:679-699): forward with the
 * training weights, KL loss and mixture adjoint (train.cu:69-124), MLP and
 * grid backward, Adam (+EMA) with the context's loss scale, and the mean loss
 * written to the device scalar loss_out (may be NULL). Targets: tgt_dir [2][n]
 * (theta, phi), tgt_rgb [3][n], tgt_pdf [n] (GuidedOutput, guideditem.h:57-64).
 * No host synchronisation. */
npmg_status npmg_train_step(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                            const float* tgt_pdf, uint32_t n, float* loss_out, npmg_stream stream);
/* The training step of a network with a learned selection probability
 * ("mixture": {"n_lobes": 7, "learn_selection": true}; parameters.h:6-9):
 * the loss and adjoint of compute_dL_doutput_with_selection (train.cu:126-172)
 * need each record's BSDF pdf, tgt_bsdf_pdf [n]. use_graph: replay as the
 * captured graph, as npmg_train_step_graph. */
npmg_status npmg_train_step_selection(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                                      const float* tgt_pdf, const float* tgt_bsdf_pdf, uint32_t n, int use_graph,
                                      float* loss_out, npmg_stream stream);
/* npmg_train_step as one HIP graph: captured on the first call for these
 * buffers and this n (the workspace is sized before the capture), replayed by
 * every later call with the same arguments; other arguments capture anew.
 * Same results as npmg_train_step (the loss scale and step count live on the
 * device). Replaces the per-kernel launches of integrator.cpp:650-702 with one
 * graph launch; the buffers must stay allocated while the graph may replay. */
npmg_status npmg_train_step_graph(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                                  const float* tgt_pdf, uint32_t n, float* loss_out, npmg_stream stream);

/* One queue of a wavefront frame (e.g. one bounce's scatter queue), SoA with
 * row stride `capacity`: in [9][capacity]; normal [3][capacity] (NULL: no
 * cosine product); u [3][capacity] (NULL: no sampling); eval_dirs
 * [n_eval][3][capacity]; outputs out_dir [3][capacity], out_pdf, out_lobe
 * [capacity] (out_lobe may be NULL), out_eval_pdf [n_eval][capacity]. */
typedef struct npmg_queue {
    const float* in;
    const float* normal;
    const float* u;
    const float* eval_dirs;
    float* out_dir;
    float* out_pdf;
    float* out_eval_pdf;
    uint32_t* out_lobe;
    uint32_t capacity;
    float* out_selection; /* [capacity] learned selection probability (mixture.learn_selection), or NULL */
} npmg_queue;

/* npmg_sample_pdf that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). No counterpart in the reference. The level-major grid
 * features of the MFMA path go to a 16 MiB device array that the library
 * holds on each GPU for the life of the process (n x n_levels x
 * features_per_level x 2 bytes at most: 262144 queries of a 16-level,
 * 2-feature grid; larger n is an error). No timing events are recorded, and
 * an FP8 network must be calibrated beforehand (npmg_fp8_calibrate). Every
 * call on one GPU shares that GPU's array and overwrites it: calls that
 * overlap on one GPU (other streams, other contexts, other threads) race. */
npmg_status npmg_sample_pdf_noalloc(npmg_ctx* ctx, const float* in, const float* normal, const float* u,
                                    const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                                    float* out_eval_pdf, uint32_t* out_lobe, float* raw_out, uint32_t n,
                                    uint32_t flags, npmg_stream stream);

/* npmg_sample_pdf with the outputs described by one npmg_queue (row stride n,
 * its capacity is ignored), including out_selection. */
npmg_status npmg_sample_pdf_ex(npmg_ctx* ctx, const npmg_queue* q, uint32_t n, uint32_t n_eval, uint32_t flags,
                               npmg_stream stream);

/* npmg_sample_pdf over n_queues (<= 16) queues of uneven sizes in one
 * persistent launch: workgroups pull 64-query tiles from a device work
 * counter across all queues, so no queue's tail idles the GPU. counts: device
 * array [n_queues] of each queue's length (read on the device, as the
 * wavefront queues' sizes are; values above capacity are clamped). Replaces
 * the per-queue launches over MAX_INFERENCE_NUM threads of the guided passes
 * (integrator.cpp:160-320, 627-648). queues: host array. */
npmg_status npmg_sample_pdf_queues(npmg_ctx* ctx, const npmg_queue* queues, const uint32_t* counts, uint32_t n_queues,
                                   uint32_t n_eval, uint32_t flags, npmg_stream stream);

/* Picks the launch shape of the query path for batches of about n queries by
 * timing the candidates on synthetic queries (pipeline chunks 1/2/4/8 of the
 * MFMA paths and workgroups per CU of the fused and persistent kernels), then
 * the work-counter claim sizes of the persistent queue launches on a
 * synthetic frame of 8 queues decaying by 0.68 per bounce (64-query tiles per
 * claim 1..16, grid-encode items per claim 8..64), and keeps the fastest for
 * later calls. Synchronous; allocates scratch. */
npmg_status npmg_autotune(npmg_ctx* ctx, uint32_t n, npmg_stream stream);
/* The launch shape in use: pipeline chunks and workgroups per CU (0 = as
 * many as fit). */
npmg_status npmg_tuning(const npmg_ctx* ctx, uint32_t* chunks, uint32_t* blocks_per_cu);
/* Sets the launch shape without timing, e.g. one that npmg_tuning reported
 * in an earlier process: chunks 1..8 (used by the BF16/FP16 path on batches
 * of 2^18 queries or more), blocks_per_cu 0..64 (0 = as many as fit; larger
 * values are capped at what fits). Changes no result bit. No reference
 * counterpart (the reference has no launch-shape tuning). */
npmg_status npmg_set_tuning(npmg_ctx* ctx, uint32_t chunks, uint32_t blocks_per_cu);
/* The claim sizes of npmg_sample_pdf_queues: 64-query tiles a wave of the
 * fused kernel takes per work-counter atomic, and (level, 256-query) items a
 * workgroup of the grid encode takes per atomic (every wave shares one
 * counter, whose atomics serialise). Set them without timing: tiles 1..64,
 * items 1..256. Neither changes a result bit. No reference counterpart. */
npmg_status npmg_queue_tuning(const npmg_ctx* ctx, uint32_t* tiles_per_claim, uint32_t* items_per_claim);
npmg_status npmg_set_queue_tuning(npmg_ctx* ctx, uint32_t tiles_per_claim, uint32_t items_per_claim);

/* FP8 inference ("inference_precision": "fp8" with "precision": "bf16"): the
 * query path's MLP runs on E4M3 weights and activations with FP32
 * accumulation (v_mfma_f32_32x32x16_fp8_fp8); training stays BF16. Each layer
 * has a power-of-two weight scale (from max |W|, recomputed whenever the
 * inference parameters change) and input scale (from the maximum activation
 * of an FP32 forward over calibration queries, with one binade of headroom;
 * larger values saturate). The first inference call calibrates on its first
 * 65536 queries unless npmg_fp8_calibrate ran before; npmg_fp8_calibrate
 * recalibrates (in [9][n], at most 65536 queries used; synchronous). No
 * counterpart in the reference, whose inference runs tcnn's FP16 network
 * (integrator.cpp:645). */
npmg_status npmg_fp8_calibrate(npmg_ctx* ctx, const float* in, uint32_t n, npmg_stream stream);
/* The scales in use (host arrays of n_layers = hidden layers + 1 entries):
 * layer l computes (W_l sw_l)(x_l sx_l) in E4M3. Synchronous. */
npmg_status npmg_fp8_scales(npmg_ctx* ctx, float* weight_scales, float* input_scales, uint32_t n_layers);

/* ---- Snapshots ------------------------------------------------------------ */

/* Checkpoint in tcnn's snapshot format: the MessagePack document that
 * nlohmann::json::to_msgpack writes for Trainer::serialize (trainer.h:267-278):
 * n_params, params_type "float", params_binary (the inference parameters) and,
 * with with_optimizer, optimizer = {nested: Adam state (adam.h:282-290:
 * current_step, base_learning_rate, first/second moments, param steps),
 * weights_ema_binary} (ema.h:194-199; Adam alone without EMA). An extra
 * "npmg" map holds the FP32 master weights and the loss-scale state, so that
 * a resume is exact; tcnn ignores it. Host buffer: call with buf = NULL to get
 * the size, then with a buffer of at least that size. Synchronous. */
npmg_status npmg_snapshot_save(npmg_ctx* ctx, int with_optimizer, void* buf, size_t capacity, size_t* size);
/* Restores a snapshot written by npmg_snapshot_save or by tcnn (params_type
 * "float" or "__half"; MessagePack). Without "npmg", the master and EMA
 * weights both take params_binary (Trainer::deserialize, trainer.h:281-307);
 * without "optimizer" the optimiser state is left as it is. Synchronous. */
npmg_status npmg_snapshot_load(npmg_ctx* ctx, const void* buf, size_t size);

/* ---- Wavefront packers (either side of the path) ------------------------ */

/* generate_inference_data (train.cu:174-191): shading points of a queue ->
 * query records. pos [3][n] world positions; wo [3][n] outgoing directions
 * and normal [3][n] shading normals (any length, normalised here);
 * roughness [n]; scene_box: host float[6] = min xyz, max xyz of the scene.
 * in [9][n]: position normalised to the box inflated by 0.5 % of its
 * diagonal (normalizeSpatialCoord, train.h:173-177), (w_o + 1) / 2,
 * normal (theta / pi, phi / 2 pi), 1 - e^-roughness. No context needed. */
npmg_status npmg_pack_queries(const float* pos, const float* wo, const float* normal, const float* roughness,
                              const float* scene_box, uint32_t n, float* in, npmg_stream stream);

/* The guided pixel state of a frame (GuidedPixelStateBuffer,
 * src/render/guided/workqueue.h:15-69, RadianceRecordItem guideditem.h:23-39),
 * SoA per depth: field[d][c][p] at ((d * C) + c) * n_pixels + p. Angles are
 * (theta, phi) as cartesianToSpherical writes them. Device pointers. */
typedef struct npmg_path_records {
    const float* L;          /* [D][3][P] radiance gathered after the vertex (recordRadiance) */
    const float* thp;        /* [D][3][P] throughput at the vertex */
    const float* pos;        /* [D][3][P] world position (ray origin) */
    const float* dir;        /* [D][2][P] scattered direction */
    const float* wi_pdf;     /* [D][P] effective pdf of the scattered direction */
    const float* mis_weight; /* [D][P] */
    const uint8_t* delta;    /* [D][P] nonzero: sampled from a delta lobe */
    const float* bsdf_val;   /* [D][3][P] */
    const float* wo;         /* [D][2][P] outgoing direction */
    const float* normal;     /* [D][2][P] shading normal */
    const float* roughness;  /* [D][P] */
    const uint32_t* depth;   /* [P] recorded vertices per pixel (curDepth) */
    const float* bsdf_pdf;   /* [D][P] BSDF pdf of the scattered direction (learned selection), or NULL */
    uint32_t max_depth;      /* D (MAX_TRAIN_DEPTH, parameters.h:8) */
    uint32_t n_pixels;       /* P */
} npmg_path_records;

/* Training pixels of a frame (integrator.cpp:666-670): thread t covers pixel
 * (t % res[0], t / res[0]) * stride + offset, skipped outside resolution. */
typedef struct npmg_train_tile {
    uint32_t offset[2], stride[2], res[2], resolution[2];
} npmg_train_tile;

/* generate_training_data (train.cu:14-67): recorded path vertices -> training
 * records in (training pixel, depth) order (the reference's atomic push,
 * train.h:117-124, has no fixed order). Skips delta vertices and records with
 * a NaN position/direction/target or a zero or NaN pdf; target radiance
 * (L / thp where thp > 1e-5, else 0) * misWeight * bsdfVal. Outputs with row
 * stride capacity: in [9], tgt_dir [2] (theta, phi), tgt_rgb [3], tgt_pdf;
 * tgt_bsdf_pdf (may be NULL; needs records->bsdf_pdf) the BSDF pdf of the
 * scattered direction (GuidedOutput::bsdfPdf, learned selection);
 * *count (device) = records kept; those past capacity are dropped, so train
 * on min(count, capacity). records, tile, scene_box: host pointers. */
npmg_status npmg_pack_training(npmg_ctx* ctx, const npmg_path_records* records, const npmg_train_tile* tile,
                               const float* scene_box, uint32_t capacity, float* in, float* tgt_dir, float* tgt_rgb,
                               float* tgt_pdf, float* tgt_bsdf_pdf, uint32_t* count, npmg_stream stream);

/* ---- Training policy (host side of trainStep, integrator.cpp:595-702) ---- */

/* updateTrainingStride (integrator.cpp:595-625): the pixel stride of the next
 * frame's training tile from this frame's record count, so that about
 * batch_size * batches_per_frame records are produced (the target tile area
 * is capped at 2^22 pixels). cur, min_stride,
 * out: host uint32[2]. Returns 1 if the reference would accept the change
 * (out differs from cur as its unsigned test decides), else 0 (out = cur). */
int npmg_training_stride(const uint32_t* cur, const uint32_t* min_stride, uint32_t batch_size,
                         uint32_t batches_per_frame, uint32_t n_training_samples, uint32_t* out);
/* The batch loop of trainStep (integrator.cpp:666-672): min(n / batch_size + 1,
 * batches_per_frame) batches at offsets i * batch_size of min(n - offset,
 * batch_size) records each, rounded down to a multiple of granularity (the
 * reference's 128, tcnn's batch rule; 1 here needs none), stopping at the
 * first batch below min_batch (MIN_TRAIN_BATCH_SIZE, parameters.h:19).
 * offsets, sizes: host arrays of max_batches; *n_batches: batches written. */
npmg_status npmg_train_batches(uint32_t n_samples, uint32_t batch_size, uint32_t batches_per_frame,
                               uint32_t granularity, uint32_t min_batch, uint32_t* offsets, uint32_t* sizes,
                               uint32_t max_batches, uint32_t* n_batches);
/* trainStep's loop over packed records (npmg_pack_training output, row
 * stride `stride`, n_records kept): each batch of npmg_train_batches is
 * copied to a contiguous staging buffer and trained with npmg_train_step
 * (npmg_train_step_selection when tgt_bsdf_pdf is given); batch b's mean loss
 * goes to losses[b] (device, may be NULL). *n_batches (host): batches run. */
npmg_status npmg_train_records(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                               const float* tgt_pdf, const float* tgt_bsdf_pdf, uint32_t stride, uint32_t n_records,
                               uint32_t batch_size, uint32_t batches_per_frame, uint32_t granularity,
                               uint32_t min_batch, float* losses, uint32_t* n_batches, npmg_stream stream);

/* ---- Split passes (test harness and autograd wrapping) ------------------ */

/* Encoding only: in [9][n] -> enc [encoded_width][n]. */
npmg_status npmg_encode(npmg_ctx* ctx, const float* in, float* enc_out, uint32_t n, int use_inference_params,
                        npmg_stream stream);
/* Per (level, corner) hash-grid table index relative to the level and
 * trilinear weight: idx [n_levels][8][n], w [n_levels][8][n]. */
npmg_status npmg_grid_corners(npmg_ctx* ctx, const float* in, uint32_t* idx, float* w, uint32_t n,
                              npmg_stream stream);
/* Forward with the training (0) or inference (1) weights: raw [n_out][n]. */
npmg_status npmg_forward(npmg_ctx* ctx, const float* in, float* raw_out, uint32_t n, int use_inference_params,
                         npmg_stream stream);
/* Loss and adjoint w.r.t. the raw outputs, scaled by loss_scale / n like the
 * reference: dout [n_out][n], loss [n] (per-sample -f/q log p). */
npmg_status npmg_mixture_adjoint(npmg_ctx* ctx, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                                 const float* tgt_pdf, uint32_t n, float loss_scale, float* dout, float* loss,
                                 npmg_stream stream);
/* The same with the learned selection probability (compute_dL_doutput_with_
 * selection, train.cu:126-172): tgt_bsdf_pdf [n]; loss -f/P log P with
 * P = s (p + 1e-5) + (1 - s) bsdfPdf. */
npmg_status npmg_mixture_adjoint_selection(npmg_ctx* ctx, const float* raw, const float* tgt_dir,
                                           const float* tgt_rgb, const float* tgt_pdf, const float* tgt_bsdf_pdf,
                                           uint32_t n, float loss_scale, float* dout, float* loss,
                                           npmg_stream stream);
/* Selection probability s = logistic(raw[4 K]) of each query
 * (MixedSGWithSelection::extractSelectionProbability, distribution.h:233-237;
 * the integrator's bsdfSamplingFraction is 1 - s, integrator.cpp:172-176). */
npmg_status npmg_selection_probability(npmg_ctx* ctx, const float* raw, float* out, uint32_t n, npmg_stream stream);
/* Backward of the network with the training weights given dL/draw [n_out][n]:
 * writes the gradient of every parameter (tcnn order) to dparams [n_params]
 * and, if denc is not NULL, dL/denc [encoded_width][n]. */
npmg_status npmg_backward(npmg_ctx* ctx, const float* in, const float* dout, uint32_t n, float* dparams,
                          float* denc, npmg_stream stream);
/* Applies Adam (+EMA) to the training parameters with gradients dparams
 * [n_params] scaled by loss_scale (Trainer::optimizer_step, trainer.h:153-155). */
npmg_status npmg_optimizer_step(npmg_ctx* ctx, const float* dparams, float loss_scale, npmg_stream stream);
/* Current loss scale and optimizer step count (host outputs). */
npmg_status npmg_train_state(const npmg_ctx* ctx, float* loss_scale, uint32_t* step);

/* ---- Synthetic wavefront records (hit records are synthetic) ------------ */

/* Queries keyed by (seed, first + i) with pcg32: in [9][n] packed like
 * generate_inference_data (train.cu:174-191), face-forwarded normal [3][n],
 * sample uniforms u [3][n], eval_dirs [4][3][n] uniform on the sphere. */
npmg_status npmg_synth_queries(uint64_t seed, uint64_t first, uint32_t n, float* in, float* normal, float* u,
                               float* eval_dirs, npmg_stream stream);
/* Training records: cosine-weighted direction about the face-forwarded
 * normal, its pdf, and the radiance of an analytic field of three vMF lobes
 * whose axes rotate with position (the contract of generate_training_data,
 * train.cu:14-67). */
npmg_status npmg_synth_train(uint64_t seed, uint64_t first, uint32_t n, float* in, float* tgt_dir, float* tgt_rgb,
                             float* tgt_pdf, npmg_stream stream);

/* Evaluates the library's device elementary functions over x (and y):
 * fn 0 exp, 1 log, 2 log1p, 3 expm1, 4 sin, 5 cos, 6 acos, 7 atan2(x, y),
 * 8 sqrt, 9 logistic, 10 x / y, 11 x rounded to FP8 E4M3 and back (the
 * conversion of the FP8 inference path), 12 x rounded to FP16 (round to nearest
 * even, the rounding of the FP16 adjoint intermediates and EMA). For bitwise
 * host/device comparisons. */
npmg_status npmg_debug_math(int fn, const float* x, const float* y, uint32_t n, float* out, npmg_stream stream);

/* ---- Kernel timing ---------------------------------------------------------
 * The library-side counterpart of the reference's GPU profiler scopes
 * (PROFILE("Network inference"), integrator.cpp:640-646; profiler.h:229-233):
 * while enabled, HIP events bracket every launch of the tagged kernels on the
 * stream it runs on. Tags: NPMG_TIME_GRID_ENCODE (level-major grid encoding
 * of the MFMA path), NPMG_TIME_QUERY (the fused query kernel, FP32 or MFMA),
 * NPMG_TIME_TRAIN_STEP (one whole npmg_train_step / npmg_train_step_graph). */
#define NPMG_TIME_GRID_ENCODE 0u
#define NPMG_TIME_QUERY 1u
#define NPMG_TIME_TRAIN_STEP 2u
npmg_status npmg_timing_enable(npmg_ctx* ctx, int enable);
/* Waits for the recorded launches, returns their total time (ms) and count
 * for one tag since the last read, and resets that tag. Host-synchronous. */
npmg_status npmg_timing_read(npmg_ctx* ctx, uint32_t tag, double* total_ms, uint32_t* launches);

/* ---- FP64 validation kernels ---------------------------------------------
 * No counterpart in the reference, which has no tests: these are the checks
 * its pdf (distribution.h:100-119) and loss adjoint (train.cu:69-124) are
 * held to, run on the GPU in FP64. Synchronous with respect to the host
 * (the quadrature rule is computed and uploaded on the first call). */

/* Sphere integral of each query's mixture pdf, as the library evaluates it in
 * FP32 from raw outputs raw [n_out][n] (with the cosine product about normal
 * [3][n] when normal is not NULL), accumulated in FP64 over an n_z-point
 * Gauss-Legendre rule in cos(theta) times n_phi uniform points in phi.
 * out: [n] device doubles, 1 for a normalised pdf. */
npmg_status npmg_validate_pdf_integral(npmg_ctx* ctx, const float* raw, const float* normal, uint32_t n, uint32_t n_z,
                                       uint32_t n_phi, double* out, npmg_stream stream);
/* Central finite differences with step h, in FP64, of the per-sample loss
 * -(f/q) log(p + 1e-5) with respect to each raw output (f, q, direction as in
 * npmg_mixture_adjoint). out: [n_out][n] device doubles; rows past 4 K are 0.
 * npmg_mixture_adjoint with loss_scale = n returns the same derivatives. */
npmg_status npmg_validate_loss_fd(npmg_ctx* ctx, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                                  const float* tgt_pdf, uint32_t n, double h, double* out, npmg_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* NPMG_H */
