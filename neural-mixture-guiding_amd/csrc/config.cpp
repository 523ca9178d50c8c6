// config.cpp -- JSON configuration -> Config, with tiny-cuda-nn's key names
// for "encoding", "network" and "optimizer" (so the reference's
// common/configs/nn/base.json "nn" object is accepted unchanged).
#include "config.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "json.h"

namespace npmg {
namespace {

bool ieq(const std::string& a, const std::string& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (std::tolower((unsigned char)a[i]) != std::tolower((unsigned char)b[i])) return false;
    return true;
}

[[noreturn]] void bad(const std::string& m) { throw std::runtime_error(m); }

// tiny-cuda-nn encodings/grid.h:195-204 (grid_scale / grid_resolution) and the
// per-level table sizes of the GridEncodingTemplated constructor (:958-991).
void grid_layout(DevCfg& d, uint32_t log2_hashmap_size, uint32_t base_resolution, float per_level_scale) {
    const float log2_pls = std::log2(per_level_scale);
    uint32_t offset = 0;
    for (int i = 0; i < d.n_levels; ++i) {
        const float scale = std::exp2((float)(uint32_t)i * log2_pls) * (float)base_resolution - 1.0f;
        const uint32_t res = (uint32_t)std::ceil(scale) + 1;
        const uint32_t max_params = 0xffffffffu / 2;
        uint32_t params = std::pow((float)res, 3) > (float)max_params ? max_params : res * res * res;
        params = (params + 7u) / 8u * 8u;
        if (d.grid_type == GRID_HASH) params = std::min(params, 1u << log2_hashmap_size);
        d.offsets[i] = offset;
        d.scales[i] = scale;
        d.res[i] = res;
        offset += params;
    }
    d.offsets[d.n_levels] = offset;
}

}  // namespace

Config parse_config(const std::string& text) {
    const Json root = parse_json(text);
    if (root.kind != Json::Object) bad("configuration must be a JSON object");
    Config c;
    DevCfg& d = c.d;
    d.n_in = (int)root.number("n_input_dims", 9);
    if (d.n_in < 3 || d.n_in > 16) bad("n_input_dims must be in [3, 16]");

    // ---- encoding (tiny-cuda-nn encodings/composite.h:135-205) ----
    if (!root.contains("encoding")) bad("missing 'encoding'");
    const Json& enc = root.at("encoding");
    std::vector<const Json*> nested;
    if (ieq(enc.string("otype", "Composite"), "Composite")) {
        if (!enc.contains("nested") || enc.at("nested").kind != Json::Array) bad("Composite encoding needs 'nested'");
        for (const Json& e : enc.at("nested").arr) nested.push_back(&e);
    } else {
        nested.push_back(&enc);
    }
    if (nested.empty() || nested.size() > (size_t)kMaxEnc) bad("1 to 4 nested encodings are supported");
    d.n_enc = (int)nested.size();
    d.grid_enc = -1;
    int begin = 0, specified = 0, unspecified = 0;
    for (const Json* e : nested) {
        if (e->contains("n_dims_to_encode")) specified += (int)e->number("n_dims_to_encode", 0);
        else ++unspecified;
    }
    if (unspecified > 1) bad("only one nested encoding may leave n_dims_to_encode unspecified");
    if (specified > d.n_in) bad("nested encodings encode more dims than the input has");
    int off = 0;
    uint32_t log2_T = 19, base_res = 16;
    float pls = 2.0f;
    for (int i = 0; i < d.n_enc; ++i) {
        const Json& e = *nested[i];
        const std::string ot = e.string("otype", "");
        if (e.contains("dims_to_encode_begin")) begin = (int)e.number("dims_to_encode_begin", 0);
        const int nd = e.contains("n_dims_to_encode") ? (int)e.number("n_dims_to_encode", 0) : d.n_in - specified;
        if (nd <= 0 || begin + nd > d.n_in) bad("encoding " + std::to_string(i) + ": bad n_dims_to_encode");
        d.enc_begin[i] = begin;
        d.enc_ndims[i] = nd;
        d.enc_offset[i] = off;
        d.enc_param[i] = 0;
        int width = 0;
        if (ieq(ot, "HashGrid") || ieq(ot, "DenseGrid") || ieq(ot, "Grid")) {
            if (d.grid_enc >= 0) bad("at most one grid encoding is supported");
            if (nd != 3) bad("grid encoding must encode 3 dims");
            d.enc_type[i] = ENC_GRID;
            d.grid_enc = i;
            const std::string deft = ieq(ot, "DenseGrid") ? "Dense" : "Hash";
            const std::string gt = e.string("type", deft);
            if (ieq(gt, "Hash")) d.grid_type = GRID_HASH;
            else if (ieq(gt, "Dense")) d.grid_type = GRID_DENSE;
            else bad("grid type must be Hash or Dense");
            d.F = (int)e.number("n_features_per_level", 2);
            if (d.F != 1 && d.F != 2 && d.F != 4 && d.F != 8) bad("n_features_per_level must be 1, 2, 4 or 8");
            if (e.contains("n_features") || e.contains("n_grid_features")) {
                const int nf = (int)(e.contains("n_features") ? e.number("n_features", 0) : e.number("n_grid_features", 0));
                if (nf % d.F) bad("n_features must be a multiple of n_features_per_level");
                d.n_levels = nf / d.F;
            } else {
                d.n_levels = (int)e.number("n_levels", 16);
            }
            if (d.n_levels < 1 || d.n_levels > kMaxLevels) bad("n_levels must be in [1, 32]");
            log2_T = (uint32_t)e.number("log2_hashmap_size", 19);
            if (log2_T < 3 || log2_T > 30) bad("log2_hashmap_size must be in [3, 30]");
            base_res = (uint32_t)e.number("base_resolution", 16);
            pls = (float)e.number("per_level_scale", 2.0);
            const std::string it = e.string("interpolation", "Linear");
            if (ieq(it, "Linear")) d.interp = INTERP_LINEAR;
            else if (ieq(it, "Smoothstep")) d.interp = INTERP_SMOOTHSTEP;
            else if (ieq(it, "Nearest")) d.interp = INTERP_NEAREST;
            else bad("interpolation must be Linear, Smoothstep or Nearest");
            const std::string ht = e.string("hash", "CoherentPrime");
            if (ieq(ht, "CoherentPrime")) d.hash_type = HASH_COHERENT_PRIME;
            else if (ieq(ht, "Prime")) d.hash_type = HASH_PRIME;
            else if (ieq(ht, "ReversedPrime")) d.hash_type = HASH_REVERSED_PRIME;
            else bad("hash must be CoherentPrime, Prime or ReversedPrime");
            if (e.boolean("stochastic_interpolation", false)) bad("stochastic_interpolation is not supported");
            width = d.n_levels * d.F;
        } else if (ieq(ot, "OneBlob")) {
            d.enc_type[i] = ENC_ONEBLOB;
            const int nb = (int)e.number("n_bins", 16);
            if (nb < 1 || nb > 64 || (nb & (nb - 1))) bad("OneBlob n_bins must be a power of two <= 64");
            d.enc_param[i] = nb;
            width = nd * nb;
        } else if (ieq(ot, "SphericalHarmonics")) {
            d.enc_type[i] = ENC_SH;
            const int deg = (int)e.number("degree", 2);
            if (deg < 1 || deg > 2) bad("SphericalHarmonics degree must be 1 or 2");
            if (nd != 3) bad("SphericalHarmonics must encode 3 dims");
            d.enc_param[i] = deg;
            width = deg * deg;
        } else if (ieq(ot, "Identity")) {
            d.enc_type[i] = ENC_IDENTITY;
            width = nd;
        } else {
            bad("unsupported encoding otype '" + ot + "'");
        }
        off += width;
        begin += nd;
    }
    d.enc_width = off;
    d.enc_padded = (off + 15) / 16 * 16;
    if (d.grid_enc >= 0) grid_layout(d, log2_T, base_res, pls);
    d.ob_kb0 = d.ob_nkb = d.ob_in = 0;
    if (d.n_enc == 2 && d.grid_enc == 0 && (d.n_levels * d.F) % 16 == 0 && d.enc_type[1] == ENC_ONEBLOB &&
        d.enc_param[1] == 16 && d.enc_width == d.enc_padded) {
        d.ob_kb0 = d.n_levels * d.F / 16;
        d.ob_nkb = d.enc_ndims[1];
        d.ob_in = d.enc_begin[1];
    }

    // ---- mixture ----
    const Json empty;
    const Json& mix = root.contains("mixture") ? root.at("mixture") : empty;
    const std::string mt = mix.kind == Json::Object ? mix.string("otype", "vMF") : "vMF";
    if (ieq(mt, "vMF")) d.mixture = MIX_VMF;
    else if (ieq(mt, "ASG")) d.mixture = MIX_ASG;
    else bad("mixture otype must be vMF or ASG");
    d.K = mix.kind == Json::Object ? (int)mix.number("n_lobes", d.mixture == MIX_ASG ? 16 : 8) : 8;
    if (d.K < 1 || d.K > kMaxLobes) bad("n_lobes must be in [1, 16]");
    if (d.mixture == MIX_ASG && d.K != 16) bad("the anisotropic mixture has 16 lobes");
    d.lobe_params = d.mixture == MIX_ASG ? 6 : 4;
    const std::string mapping = mix.kind == Json::Object ? mix.string("mapping", "reference") : "reference";
    if (ieq(mapping, "reference")) d.mapping = MAP_REFERENCE;
    else if (ieq(mapping, "softmax")) d.mapping = MAP_SOFTMAX;
    else bad("mixture mapping must be reference or softmax");
    const std::string num = mix.kind == Json::Object ? mix.string("numerics", "stable") : "stable";
    if (ieq(num, "reference")) d.numerics = NUM_REFERENCE;
    else if (ieq(num, "stable")) d.numerics = NUM_STABLE;
    else if (ieq(num, "fast")) d.numerics = NUM_FAST;
    else bad("mixture numerics must be reference, stable or fast");
    d.cos_kappa = mix.kind == Json::Object ? (float)mix.number("cosine_kappa", 2.18853) : 2.18853f;
    // learned selection (distribution.h:204-241, parameters.h:6-9): one more
    // output, the logit of the guiding distribution's selection probability;
    // with 7 lobes it takes output 29 of 32, as in the reference
    d.sel = mix.kind == Json::Object && mix.boolean("learn_selection", false) ? 1 : 0;
    if (d.sel && (d.mixture != MIX_VMF || d.K != 7)) bad("learn_selection needs the vMF mixture with 7 lobes");
    d.n_out = d.lobe_params * d.K + d.sel;
    d.n_out_pad = (d.n_out + 15) / 16 * 16;

    // ---- network (tiny-cuda-nn networks/fully_fused_mlp.h) ----
    if (!root.contains("network")) bad("missing 'network'");
    const Json& net = root.at("network");
    d.width = (int)net.number("n_neurons", 64);
    if (d.width != 64 && d.width != 128) bad("n_neurons must be 64 or 128");
    d.n_hidden = (int)net.number("n_hidden_layers", 2);
    if (d.n_hidden < 1 || d.n_hidden > kMaxLayers - 1) bad("n_hidden_layers must be in [1, 8]");
    if (!ieq(net.string("activation", "ReLU"), "ReLU")) bad("only ReLU activation is supported");
    if (!ieq(net.string("output_activation", "None"), "None")) bad("only output_activation None is supported");
    d.n_layers = d.n_hidden + 1;
    uint32_t loff = 0;
    for (int l = 0; l < d.n_layers; ++l) {
        d.layer_in[l] = l == 0 ? d.enc_padded : d.width;
        d.layer_out[l] = l == d.n_layers - 1 ? d.n_out_pad : d.width;
        d.layer_off[l] = loff;
        loff += (uint32_t)(d.layer_in[l] * d.layer_out[l]);
    }
    c.n_mlp = loff;
    c.n_grid = d.grid_enc >= 0 ? (size_t)d.offsets[d.n_levels] * d.F : 0;
    c.n_params = c.n_mlp + c.n_grid;

    // ---- optimizer (tiny-cuda-nn optimizers/ema.h, optimizers/adam.h) ----
    if (root.contains("optimizer")) {
        const Json* o = &root.at("optimizer");
        if (ieq(o->string("otype", "Adam"), "Ema")) {
            c.ema = true;
            c.ema_decay = (float)o->number("decay", 0.99);
            if (o->contains("full_precision")) c.ema_full = o->boolean("full_precision", true) ? 1 : 0;
            if (!o->contains("nested")) bad("Ema optimizer needs 'nested'");
            o = &o->at("nested");
        } else {
            c.ema = false;
        }
        if (!ieq(o->string("otype", "Adam"), "Adam")) bad("only Adam (optionally wrapped in Ema) is supported");
        c.lr = (float)o->number("learning_rate", 1e-3);
        c.beta1 = (float)o->number("beta1", 0.9);
        c.beta2 = (float)o->number("beta2", 0.999);
        c.eps = (float)o->number("epsilon", 1e-8);
        c.l2_reg = (float)o->number("l2_reg", 1e-8);
        if (o->boolean("adabound", false)) bad("adabound is not supported");
        if (o->number("relative_decay", 0) != 0 || o->number("absolute_decay", 0) != 0)
            bad("weight decay is not supported");
    }

    // ---- loss scaling and gradient clamp (train.cu:106-123, parameters.h:24) ----
    const Json& loss = root.contains("loss") ? root.at("loss") : empty;
    if (loss.kind == Json::Object) {
        c.loss_scale = (float)loss.number("scale", 128.0);
        c.dynamic_scale = loss.boolean("dynamic", false);
        c.scale_window = (uint32_t)loss.number("scale_window", 1000);
        d.grad_clamp = (float)loss.number("grad_clamp", 150.0);
        d.clamp_scale = (float)loss.number("clamp_scale", 128.0);
        d.half_adjoint = loss.boolean("fp16_intermediates", false) ? 1 : 0;
    } else {
        d.grad_clamp = 150.0f;
        d.clamp_scale = 128.0f;
    }
    if (!(c.loss_scale > 0)) bad("loss scale must be positive");
    if (!(d.clamp_scale > 0)) bad("loss clamp_scale must be positive");

    const std::string prec = root.string("precision", "fp32");
    if (ieq(prec, "fp32")) c.precision = PREC_FP32;
    else if (ieq(prec, "bf16")) c.precision = PREC_BF16;
    else if (ieq(prec, "fp16")) c.precision = PREC_FP16;
    else if (ieq(prec, "bf16x3")) c.precision = PREC_BF16X3;
    else if (ieq(prec, "fp8")) bad("fp8 is inference-only: precision bf16 with inference_precision fp8");
    else bad("precision must be fp32, bf16, fp16 or bf16x3");
    const std::string iprec = root.string("inference_precision", "same");
    if (ieq(iprec, "fp8")) {
        if (c.precision != PREC_BF16) bad("inference_precision fp8 needs precision bf16");
        c.infer_fp8 = true;
    } else if (!ieq(iprec, "same") && !ieq(iprec, prec)) {
        bad("inference_precision must be same or fp8");
    }

    const std::string gg = root.string("grid_gradient", "atomic");
    if (ieq(gg, "atomic")) c.deterministic_grid = false;
    else if (ieq(gg, "deterministic")) c.deterministic_grid = true;
    else bad("grid_gradient must be atomic or deterministic");
    c.init_passes = (int)root.number("init_passes", 2);
    return c;
}

}  // namespace npmg
