// config.h -- network configuration (host) and its kernel-argument image.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>

namespace npmg {

enum EncType : int { ENC_GRID = 0, ENC_ONEBLOB = 1, ENC_SH = 2, ENC_IDENTITY = 3 };
enum GridType : int { GRID_HASH = 0, GRID_DENSE = 1 };
enum HashType : int { HASH_COHERENT_PRIME = 0, HASH_PRIME = 1, HASH_REVERSED_PRIME = 2 };
enum Interp : int { INTERP_LINEAR = 0, INTERP_SMOOTHSTEP = 1, INTERP_NEAREST = 2 };
enum Mapping : int { MAP_REFERENCE = 0, MAP_SOFTMAX = 1 };
// fast: the stable formulas with hardware transcendentals on the query path
// and in the vMF training adjoint (the ASG adjoint keeps the exact functions).
enum Numerics : int { NUM_REFERENCE = 0, NUM_STABLE = 1, NUM_FAST = 2 };
enum MixtureType : int { MIX_VMF = 0, MIX_ASG = 1 };
// PREC_BF16X3: MFMA on split BF16 operands (hi + lo, three products) with the
// grid table and level-major features in FP16.
enum Precision : int { PREC_FP32 = 0, PREC_BF16 = 1, PREC_FP16 = 2, PREC_FP8 = 3, PREC_BF16X3 = 4 };
// Which grid levels a level-major encode launch covers: all, only the coarse
// levels whose tables are staged in LDS, or only the gathered ones.
enum EncodeLevels : int { ENC_ALL = 0, ENC_LDS_LEVELS = 1, ENC_GATHER_LEVELS = 2 };

constexpr int kMaxEnc = 4;
constexpr int kMaxLevels = 32;
constexpr int kMaxLayers = 9;
constexpr int kMaxLobes = 16;

// Everything a kernel needs to know about the network, passed by value.
struct DevCfg {
    int n_in;  // input dims of a query record (9)
    int n_enc;
    int enc_type[kMaxEnc], enc_begin[kMaxEnc], enc_ndims[kMaxEnc], enc_param[kMaxEnc], enc_offset[kMaxEnc];
    int enc_width, enc_padded;
    int grid_enc;  // index of the grid encoding, -1 if none
    // Layer-0 layout "grid, then OneBlob(16)": when ob_nkb > 0, the input
    // features are the grid's (K-blocks [0, ob_kb0) of 16 features) followed by
    // 16 OneBlob bins of each input dim ob_in .. ob_in + ob_nkb - 1 (one
    // K-block per dim) and nothing else, so the MFMA query kernels can encode
    // them without the general per-feature dispatch.
    int ob_kb0, ob_nkb, ob_in;
    int grid_type, hash_type, n_levels, F, interp;
    uint32_t offsets[kMaxLevels + 1];
    float scales[kMaxLevels];
    uint32_t res[kMaxLevels];
    int width, n_hidden, n_layers, n_out, n_out_pad;
    int layer_in[kMaxLayers], layer_out[kMaxLayers];
    uint32_t layer_off[kMaxLayers];  // offset of each weight matrix in the parameter vector
    int K, mapping, numerics, mixture;
    int lobe_params;  // raw outputs per lobe: 4 (vMF), 6 (anisotropic)
    int sel;          // 1: learned selection probability in output lobe_params * K (MixedSGWithSelection)
    float cos_kappa, grad_clamp;
    float clamp_scale;  // loss scale at which grad_clamp applies (threshold grad_clamp * scale / clamp_scale)
    // 1: the adjoint keeps the reference's FP16 intermediates ("loss": {"fp16_intermediates": true}):
    // raw outputs read as FP16, per-lobe dp/d(param) stored in FP16, dL/draw stored in FP16
    int half_adjoint;
};

struct Config {
    DevCfg d{};
    Precision precision = PREC_FP32;
    // inference (query path with the inference parameters) in FP8 E4M3 with
    // per-layer scales; training stays in `precision` (BF16)
    bool infer_fp8 = false;
    // Adam (tiny-cuda-nn optimizers/adam.h) and EMA (optimizers/ema.h)
    float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-15f, l2_reg = 1e-6f;
    bool ema = true;
    float ema_decay = 0.95f;
    // EMA recurrence in FP32 or FP16: "full_precision" true / false; absent, it
    // runs in the parameter type like tcnn's default (FP16 for precision fp16)
    int ema_full = -1;
    // loss scale: static (reference, parameters.h:24) or dynamic
    float loss_scale = 128.0f;
    bool dynamic_scale = false;
    uint32_t scale_window = 1000;
    bool deterministic_grid = false;
    int init_passes = 2;
    size_t n_mlp = 0, n_grid = 0, n_params = 0;
};

// Parses the JSON configuration; throws std::runtime_error on invalid input.
Config parse_config(const std::string& json);

}  // namespace npmg
