/* oracle.h -- CPU oracle for the neural parametric mixture guiding path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, on the CPU, the reference's
 * algorithm for the guiding hot path so that the HIP implementation can be
 * checked against it. Only the test suite, the smoke check and the bench's
 * CPU-baseline leg load it; the product library never links or calls it.
 *
 * Every routine is instantiated twice: "_f" runs the reference's FP32
 * semantics with the deterministic elementary functions of npmg_math.h (so
 * that index, lobe-selection and FP32 results can be compared bit for bit),
 * "_d" runs the same algorithm in FP64 with libm (the gold for error bounds).
 *
 * Layouts are structure-of-arrays: a field of d values per query is stored as
 * [d][n], query index fastest.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ENC_GRID = 0, ORC_ENC_ONEBLOB = 1, ORC_ENC_SH = 2, ORC_ENC_IDENTITY = 3 };
enum { ORC_GRID_HASH = 0, ORC_GRID_DENSE = 1 };
enum { ORC_HASH_COHERENT_PRIME = 0, ORC_HASH_PRIME = 1, ORC_HASH_REVERSED_PRIME = 2 };
enum { ORC_INTERP_LINEAR = 0, ORC_INTERP_SMOOTHSTEP = 1, ORC_INTERP_NEAREST = 2 };
enum { ORC_MAP_REFERENCE = 0, ORC_MAP_SOFTMAX = 1 };
enum { ORC_NUM_REFERENCE = 0, ORC_NUM_STABLE = 1 };
enum { ORC_MIX_VMF = 0, ORC_MIX_ASG = 1 };

#define ORC_MAX_ENC 4
#define ORC_MAX_LEVELS 32

typedef struct {
    int n_input_dims;
    int n_enc;
    int enc_type[ORC_MAX_ENC];
    int enc_begin[ORC_MAX_ENC];
    int enc_ndims[ORC_MAX_ENC];
    int enc_param[ORC_MAX_ENC];     /* OneBlob: n_bins; SH: degree */
    int grid_type, hash_type;
    int n_levels, n_features, log2_hashmap_size, base_resolution;
    float per_level_scale;
    int interpolation;
    int width, n_hidden_layers, n_output_dims;
    int mixture, n_lobes, mapping, numerics;
    float cosine_kappa;
    float grad_clamp;               /* 0 disables the upper clamp */
    float clamp_scale;              /* loss scale at which grad_clamp applies (the reference's 128):
                                       at loss scale s the threshold is grad_clamp * s / clamp_scale */
    int learn_selection;            /* 1: output 4 K is the selection logit (MixedSGWithSelection) */
    int fp16_intermediates;         /* 1: the adjoint rounds the raw outputs, the per-lobe dp/d(param)
                                       and dL/draw to FP16 like the reference's __half buffers */
} orc_cfg;

/* Layout of the network: sizes of the padded encoding, MLP and grid. */
int orc_encoded_width(const orc_cfg* c);           /* padded to a multiple of 16 */
int orc_padded_output(const orc_cfg* c);           /* padded to a multiple of 16 */
void orc_n_params(const orc_cfg* c, size_t* n_mlp, size_t* n_grid);
size_t orc_grid_layout(const orc_cfg* c, uint32_t* offsets, float* scales, uint32_t* resolutions);

/* Per (level, corner, query) table index (relative to the level) and weight. */
void orc_grid_corners(const orc_cfg* c, const float* in, int n, uint32_t* idx, float* w);

void orc_encode_f(const orc_cfg* c, const float* grid, const float* in, int n, float* enc);
void orc_encode_d(const orc_cfg* c, const float* grid, const float* in, int n, double* enc);

void orc_mlp_forward_f(const orc_cfg* c, const float* w, const float* enc, int n, float* raw, float* acts);
void orc_mlp_forward_d(const orc_cfg* c, const float* w, const double* enc, int n, double* raw, double* acts);

/* Consumer side: raw network outputs -> mixture -> (cosine product) -> sample + pdfs.
 * normal may be NULL (no cosine product). lobes_out (optional) receives
 * [K][5][n] = lambda', kappa', theta', phi', weight after the cosine product. */
void orc_mixture_f(const orc_cfg* c, const float* raw, const float* normal, const float* u,
                   const float* eval_dirs, int n_eval, int n, float* out_dir, float* out_pdf,
                   float* out_eval_pdf, uint32_t* out_lobe, float* lobes_out);
void orc_mixture_d(const orc_cfg* c, const double* raw, const float* normal, const float* u,
                   const float* eval_dirs, int n_eval, int n, double* out_dir, double* out_pdf,
                   double* out_eval_pdf, uint32_t* out_lobe, double* lobes_out);

/* Full query: encode -> MLP -> mixture, multithreaded over queries. */
void orc_query_f(const orc_cfg* c, const float* params, const float* in, const float* normal,
                 const float* u, const float* eval_dirs, int n_eval, int n, float* out_dir,
                 float* out_pdf, float* out_eval_pdf, uint32_t* out_lobe, float* raw_out, int n_threads);
void orc_query_d(const orc_cfg* c, const float* params, const float* in, const float* normal,
                 const float* u, const float* eval_dirs, int n_eval, int n, double* out_dir,
                 double* out_pdf, double* out_eval_pdf, uint32_t* out_lobe, double* raw_out, int n_threads);

/* Loss and adjoint with the learned selection probability (train.cu:126-172):
 * tgt_bsdf [n] BSDF pdfs. */
void orc_adjoint_sel_f(const orc_cfg* c, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                       const float* tgt_pdf, const float* tgt_bsdf, int n, float loss_scale, float* dout, float* loss);
void orc_adjoint_sel_d(const orc_cfg* c, const double* raw, const float* tgt_dir, const float* tgt_rgb,
                       const float* tgt_pdf, const float* tgt_bsdf, int n, double loss_scale, double* dout,
                       double* loss);
/* Selection probability logistic(raw[4 K]) (distribution.h:233-237). */
void orc_selection_f(const orc_cfg* c, const float* raw, int n, float* out);
/* KL loss and its adjoint with respect to the raw outputs (loss_scale already
 * divided by nothing: the routine divides by n like the reference). */
void orc_adjoint_f(const orc_cfg* c, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                   const float* tgt_pdf, int n, float loss_scale, float* dout, float* loss);
void orc_adjoint_d(const orc_cfg* c, const double* raw, const float* tgt_dir, const float* tgt_rgb,
                   const float* tgt_pdf, int n, double loss_scale, double* dout, double* loss);

void orc_mlp_backward_f(const orc_cfg* c, const float* w, const float* enc, const float* acts,
                        const float* dout, int n, float* dw, float* denc);
void orc_mlp_backward_d(const orc_cfg* c, const float* w, const double* enc, const double* acts,
                        const double* dout, int n, double* dw, double* denc);

void orc_grid_backward_f(const orc_cfg* c, const float* in, const float* denc, int n, float* grid_grad);
void orc_grid_backward_d(const orc_cfg* c, const float* in, const double* denc, int n, double* grid_grad);

/* Adam as tcnn's adam_step (per-parameter step counter, zero-gradient skip for
 * non-matrix parameters, L2 on matrix parameters only). */
void orc_adam_step(size_t n, size_t n_matrix, float loss_scale, float lr, float beta1, float beta2,
                   float eps, float l2_reg, float* master, const float* grad, float* m, float* v,
                   uint32_t* steps);
/* EMA as tcnn's Ema optimizer; step is the nested optimizer's step count after the update. */
void orc_ema_step(size_t n, float decay, uint32_t step, const float* weights, float* ema);
/* The same with tcnn's FP16 recurrence (ema_step_half_precision over __half
 * weights): input weight and stored average rounded to FP16. */
void orc_ema_step_half(size_t n, float decay, uint32_t step, const float* weights, float* ema);
/* One FP32 training step (forward, adjoint, MLP and grid backward, Adam) on n
 * records with n_threads host threads (the CPU baseline); updates params, m,
 * v, steps in place and returns the mean loss. */
float orc_train_step_f(const orc_cfg* c, float* params, float* m, float* v, uint32_t* steps, const float* in,
                       const float* tgt_dir, const float* tgt_rgb, const float* tgt_pdf, int n, float loss_scale,
                       float lr, float beta1, float beta2, float eps, float l2_reg, int n_threads);

/* Parameter initialisation of tcnn's Trainer: std::seed_seq{seed} -> pcg32,
 * Xavier-uniform matrices then the grid table U(-1e-4, 1e-4); repeated `passes` times. */
void orc_init_params(const orc_cfg* c, uint32_t seed, int passes, float* params);

/* Synthetic queries and training records keyed by (seed, query index). */
void orc_synth_queries(uint64_t seed, uint64_t first, int n, float* in, float* normal, float* u,
                       float* eval_dirs);
void orc_synth_train(uint64_t seed, uint64_t first, int n, float* in, float* tgt_dir,
                     float* tgt_rgb, float* tgt_pdf);
/* Wavefront packers: generate_inference_data (train.cu:174-191) and
 * generate_training_data (train.cu:14-67) in (training pixel, depth) order.
 * box: min xyz, max xyz; tile: offset xy, stride xy, training resolution xy,
 * image resolution xy. Returns the number of records kept. */
void orc_pack_queries_f(const float* box, const float* pos, const float* wo, const float* normal,
                        const float* roughness, int n, float* in);
uint32_t orc_pack_training_f(const float* L, const float* thp, const float* pos, const float* dir,
                             const float* wi_pdf, const float* mis, const uint8_t* delta, const float* bsdf,
                             const float* wo, const float* normal, const float* rough, const uint32_t* depth,
                             const float* bsdf_pdf, uint32_t max_depth, uint32_t n_pixels, const uint32_t* tile,
                             const float* box, uint32_t capacity, float* in, float* tgt_dir, float* tgt_rgb,
                             float* tgt_pdf, float* tgt_bsdf_pdf);
/* The analytic radiance field f(x, w) used for the training targets. */
void orc_field_eval(uint64_t seed, const float* pos, const float* dir, int n, double* f);

/* Raw pcg32 outputs (O'Neill's pcg32_srandom_r(initstate, initseq) + pcg32_random_r). */
void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out);

/* Elementary functions of npmg_math.h, for bitwise host/device comparison. */
void orc_math(int fn, const float* x, const float* y, int n, float* out);

#ifdef __cplusplus
}
#endif
