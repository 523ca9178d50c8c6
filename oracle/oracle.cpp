// oracle.cpp -- CPU restatement of the reference's guiding hot path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Reference files are cited as
// path:line relative to the reference repository root. "tcnn" means
// src/ext/tcnn/include/tiny-cuda-nn.
//
// The FP32 instantiation evaluates every expression in the order the
// reference writes it, with two conventions that the HIP kernels share:
//   * a*b+c is never contracted implicitly (-ffp-contract=off); where the
//     reference's nvcc build would contract an accumulation (the pos_fract
//     offset, grid and mixture accumulations) an explicit fma is used;
//   * transcendental functions come from npmg_math.h (exp, log, sin, cos,
//     acos, atan2, log1p, expm1), which run the same operation sequence on the
//     host and on the GPU.
// The FP64 instantiation runs the same algorithm with libm.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../neural-mixture-guiding_amd/csrc/npmg_math.h"

namespace {

namespace fm = npmg::fm;

// ---------------------------------------------------------------------------
// Precision-dispatched elementary functions.
template <class T> struct M;
template <> struct M<float> {
    static float exp(float x) { return fm::exp(x); }
    static float log(float x) { return fm::log(x); }
    static float log1p(float x) { return fm::log1p(x); }
    static float expm1(float x) { return fm::expm1(x); }
    static float sin(float x) { return fm::sin(x); }
    static float cos(float x) { return fm::cos(x); }
    static float acos(float x) { return fm::acos(x); }
    static float atan2(float y, float x) { return fm::atan2(y, x); }
    static float sqrt(float x) { return fm::sqrt_(x); }
    static float fma(float a, float b, float c) { return fm::fma_(a, b, c); }
    static bool finite(float x) { return fm::isfinite_(x); }
};
template <> struct M<double> {
    static double exp(double x) { return std::exp(x); }
    static double log(double x) { return std::log(x); }
    static double log1p(double x) { return std::log1p(x); }
    static double expm1(double x) { return std::expm1(x); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double acos(double x) { return std::acos(x); }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double fma(double a, double b, double c) { return std::fma(a, b, c); }
    static bool finite(double x) { return std::isfinite(x); }
};

// Constants of krrmath (src/backend/src/core/math/include/krrmath/constants.h:15-25).
template <class T> struct C;
template <> struct C<float> {
    static constexpr float pi = 3.14159265358979323846f, two_pi = 6.28318530717958647693f;
    static constexpr float inv_pi = 0.318309886183790671538f, inv_2pi = 0.15915494309189533577f;
    static constexpr float inv_4pi = 0.07957747154594766788f, eps = 1e-5f;
};
template <> struct C<double> {
    static constexpr double pi = 3.14159265358979323846, two_pi = 6.28318530717958647693;
    static constexpr double inv_pi = 0.318309886183790671538, inv_2pi = 0.15915494309189533577;
    static constexpr double inv_4pi = 0.07957747154594766788, eps = 1e-5;
};

template <class T> struct V3 { T x, y, z; };
template <class T> V3<T> v3(T a, T b, T c) { return {a, b, c}; }
template <class T> V3<T> add(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> mul(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V3<T> divs(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> T length(V3<T> a) { return M<T>::sqrt(dot(a, a)); }
template <class T> V3<T> normalize(V3<T> a) { return divs(a, length(a)); }

// src/backend/src/util/math_utils.h:191-201
template <class T> V3<T> sph_to_cart(T theta, T phi) {
    const T st = M<T>::sin(theta), ct = M<T>::cos(theta), sp = M<T>::sin(phi), cp = M<T>::cos(phi);
    return {st * cp, st * sp, ct};
}
// src/backend/src/util/math_utils.h:176-183
template <class T> void cart_to_sph(V3<T> v, T* theta, T* phi) {
    const V3<T> vn = normalize(v);
    *theta = M<T>::acos(vn.z);
    *phi = M<T>::atan2(vn.y, vn.x);
    if (*phi < 0) *phi += C<T>::two_pi;
}
// src/backend/src/util/math_utils.h:141-151
template <class T> V3<T> get_perpendicular(V3<T> u) {
    const V3<T> a = {std::fabs(u.x), std::fabs(u.y), std::fabs(u.z)};
    const uint32_t uyx = (a.x - a.y) < 0 ? 1 : 0;
    const uint32_t uzx = (a.x - a.z) < 0 ? 1 : 0;
    const uint32_t uzy = (a.y - a.z) < 0 ? 1 : 0;
    const uint32_t xm = uyx & uzx;
    const uint32_t ym = (1 ^ xm) & uzy;
    const uint32_t zm = 1 ^ (xm | ym);
    return normalize(cross(u, v3<T>((T)xm, (T)ym, (T)zm)));
}
// src/backend/src/core/raytracing.h:35-56 (Frame(n).toWorld)
template <class T> V3<T> frame_to_world(V3<T> n, V3<T> v) {
    const V3<T> t = get_perpendicular(n);
    const V3<T> b = normalize(cross(n, t));
    return add(add(mul(t, v.x), mul(b, v.y)), mul(n, v.z));
}
// src/backend/src/render/sampling.h:32-37
template <class T> V3<T> uniform_sample_sphere(T u0, T u1) {
    const T z = (T)1 - (T)2 * u0;
    const T r = M<T>::sqrt(std::max((T)0, (T)1 - z * z));
    const T phi = C<T>::two_pi * u1;
    return {r * M<T>::cos(phi), r * M<T>::sin(phi), z};
}

template <class T> T clampv(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }

// ---------------------------------------------------------------------------
// Encoding layout (tcnn composite.h:135-205, encoding.h:90-92, grid.h:940-1004).
struct EncLayout {
    int n_out[ORC_MAX_ENC];   // output width of each nested encoding
    int offset[ORC_MAX_ENC];  // first output row
    int width;                // concatenated width
    int padded;               // padded to a multiple of 16 (pads are 1.0)
};

int enc_out_width(const orc_cfg* c, int e) {
    switch (c->enc_type[e]) {
    case ORC_ENC_GRID: return c->n_levels * c->n_features;
    case ORC_ENC_ONEBLOB: return c->enc_ndims[e] * c->enc_param[e];
    case ORC_ENC_SH: return c->enc_param[e] * c->enc_param[e];
    default: return c->enc_ndims[e];
    }
}

EncLayout enc_layout(const orc_cfg* c) {
    EncLayout L{};
    int off = 0;
    for (int e = 0; e < c->n_enc; ++e) {
        L.offset[e] = off;
        L.n_out[e] = enc_out_width(c, e);
        off += L.n_out[e];
    }
    L.width = off;
    L.padded = (off + 15) / 16 * 16;
    return L;
}

struct GridLayout {
    uint32_t offsets[ORC_MAX_LEVELS + 1];
    float scales[ORC_MAX_LEVELS];
    uint32_t res[ORC_MAX_LEVELS];
};

// tcnn grid.h:195-204 (grid_scale, grid_resolution) and grid.h:958-991 (offsets).
GridLayout grid_layout(const orc_cfg* c) {
    GridLayout g{};
    const float log2_pls = std::log2(c->per_level_scale);
    uint32_t offset = 0;
    for (int i = 0; i < c->n_levels; ++i) {
        const float scale = std::exp2((float)(uint32_t)i * log2_pls) * (float)(uint32_t)c->base_resolution - 1.0f;
        const uint32_t res = (uint32_t)std::ceil(scale) + 1;
        const uint32_t max_params = 0xffffffffu / 2;
        uint32_t params = std::pow((float)res, 3) > (float)max_params ? max_params : res * res * res;
        params = (params + 7u) / 8u * 8u;
        if (c->grid_type == ORC_GRID_HASH) params = std::min(params, 1u << c->log2_hashmap_size);
        g.offsets[i] = offset;
        g.scales[i] = scale;
        g.res[i] = res;
        offset += params;
    }
    g.offsets[c->n_levels] = offset;
    return g;
}

// tcnn grid.h:111-135 (lcg hashes) and grid.h:170-187 (grid_index).
uint32_t grid_index(const orc_cfg* c, uint32_t hashmap_size, uint32_t res, const uint32_t p[3]) {
    uint32_t stride = 1, index = 0;
    for (uint32_t d = 0; d < 3 && stride <= hashmap_size; ++d) {
        index += p[d] * stride;
        stride *= res;
    }
    if (c->grid_type == ORC_GRID_HASH && hashmap_size < stride) {
        static const uint32_t coherent[3] = {1u, 2654435761u, 805459861u};
        static const uint32_t prime[3] = {1958374283u, 2654435761u, 805459861u};
        static const uint32_t reversed[3] = {2165219737u, 1434869437u, 2097192037u};
        const uint32_t* f = c->hash_type == ORC_HASH_PRIME ? prime
                            : c->hash_type == ORC_HASH_REVERSED_PRIME ? reversed : coherent;
        index = 0;
        for (int d = 0; d < 3; ++d) index ^= p[d] * f[d];
    }
    return index % hashmap_size;
}

// tcnn common_device.h:470-496 (pos_fract), :429-431 (smoothstep). The
// reference's nvcc build contracts input*scale+0.5 into one fma.
void pos_fract(float x, float scale, int interp, float* frac, uint32_t* cell) {
    const float p = fm::fma_(x, scale, 0.5f);
    const float fl = fm::floor_(p);
    *cell = (uint32_t)(int)fl;
    float f = p - fl;
    if (interp == ORC_INTERP_SMOOTHSTEP) f = f * f * (3.0f - 2.0f * f);
    *frac = f;
}

// Corner indices (relative to the level) and interpolation weights of one
// level: tcnn kernel_grid (grid.h:220-384) / kernel_grid_backward (:386-509).
// Returns the number of corners (1 for nearest, 8 otherwise).
template <class T>
int level_corners(const orc_cfg* c, const GridLayout& g, int l, const float pos[3], uint32_t idx[8], T w[8]) {
    const uint32_t hsize = g.offsets[l + 1] - g.offsets[l];
    float frac[3];
    uint32_t cell[3];
    for (int d = 0; d < 3; ++d) pos_fract(pos[d], g.scales[l], c->interpolation, &frac[d], &cell[d]);
    if (c->interpolation == ORC_INTERP_NEAREST) {
        idx[0] = grid_index(c, hsize, g.res[l], cell);
        w[0] = 1;
        return 1;
    }
    for (int k = 0; k < 8; ++k) {
        T wk = 1;
        uint32_t loc[3];
        for (int d = 0; d < 3; ++d) {
            if ((k & (1 << d)) == 0) {
                wk *= (T)1 - (T)frac[d];
                loc[d] = cell[d];
            } else {
                wk *= (T)frac[d];
                loc[d] = cell[d] + 1;
            }
        }
        idx[k] = grid_index(c, hsize, g.res[l], loc);
        w[k] = wk;
    }
    return 8;
}

// Forward value of one level: weighted sum of the corner features,
// accumulated in corner order with fma like the reference's fma(w, v, acc).
template <class T>
void grid_level(const orc_cfg* c, const GridLayout& g, int l, const float* grid, const float pos[3], T* out) {
    const float* table = grid + (size_t)g.offsets[l] * c->n_features;
    uint32_t idx[8];
    T w[8];
    const int nc = level_corners<T>(c, g, l, pos, idx, w);
    T acc[8] = {};
    for (int k = 0; k < nc; ++k)
        for (int f = 0; f < c->n_features; ++f)
            acc[f] = nc == 1 ? (T)table[(size_t)idx[k] * c->n_features + f]
                             : M<T>::fma(w[k], (T)table[(size_t)idx[k] * c->n_features + f], acc[f]);
    for (int f = 0; f < c->n_features; ++f) out[f] = acc[f];
}

// tcnn common_device.h:533-548 (quartic_cdf).
template <class T> T quartic_cdf(T x, T inv_radius) {
    const T u = x * inv_radius;
    const T u2 = u * u;
    const T u4 = u2 * u2;
    const T v = (T)15 / (T)16 * u * ((T)1 - (T)2 / (T)3 * u2 + (T)1 / (T)5 * u4) + (T)0.5;
    return std::max((T)0, std::min((T)1, v));
}

// tcnn encodings/oneblob.h:98-127 (kernel_one_blob_soa), one input dimension.
template <class T> void one_blob(T x, int n_bins, T* out) {
    const T nb = (T)n_bins;
    T left = quartic_cdf(-x, nb) + quartic_cdf(-x - (T)1, nb) + quartic_cdf(-x + (T)1, nb);
    for (int k = 0; k < n_bins; ++k) {
        const T rb = (T)(k + 1) / nb;  // scalbn(k+1, -log2(n_bins)), exact
        const T right = quartic_cdf(rb - x, nb) + quartic_cdf(rb - x - (T)1, nb) + quartic_cdf(rb - x + (T)1, nb);
        out[k] = right - left;
        left = right;
    }
}

// tcnn encodings/spherical_harmonics.h:45-63, 126-153 (degree <= 2).
template <class T> void sh_encode(const float in[3], int degree, T* out) {
    const T x = (T)in[0] * (T)2 - (T)1, y = (T)in[1] * (T)2 - (T)1, z = (T)in[2] * (T)2 - (T)1;
    out[0] = (T)0.28209479177387814;
    if (degree <= 1) return;
    out[1] = (T)-0.48860251190291987 * y;
    out[2] = (T)0.48860251190291987 * z;
    out[3] = (T)-0.48860251190291987 * x;
}

template <class T>
void encode_one(const orc_cfg* c, const EncLayout& L, const GridLayout& g, const float* grid,
                const float* in_q /* [n_input_dims] */, T* out /* [padded] */) {
    for (int e = 0; e < c->n_enc; ++e) {
        const float* x = in_q + c->enc_begin[e];
        T* o = out + L.offset[e];
        switch (c->enc_type[e]) {
        case ORC_ENC_GRID:
            for (int l = 0; l < c->n_levels; ++l) grid_level<T>(c, g, l, grid, x, o + l * c->n_features);
            break;
        case ORC_ENC_ONEBLOB:
            for (int d = 0; d < c->enc_ndims[e]; ++d) one_blob<T>((T)x[d], c->enc_param[e], o + d * c->enc_param[e]);
            break;
        case ORC_ENC_SH: sh_encode<T>(x, c->enc_param[e], o); break;
        default:
            for (int d = 0; d < c->enc_ndims[e]; ++d) o[d] = (T)x[d];
        }
    }
    for (int j = L.width; j < L.padded; ++j) out[j] = (T)1;
}

// ---------------------------------------------------------------------------
// Bias-free MLP with tcnn layer semantics (src/fully_fused_mlp.cu:635-678,
// 499-557; ReLU common_device.h:84-89). Weights row-major [out][in], layers in
// parameter order; accumulation over inputs in increasing order.
struct NetLayout {
    int n_layers;          // hidden layers + 1
    int in_w[8], out_w[8];
    size_t off[8];
    size_t n_params;
};

NetLayout net_layout(const orc_cfg* c) {
    NetLayout N{};
    const EncLayout L = enc_layout(c);
    N.n_layers = c->n_hidden_layers + 1;
    size_t off = 0;
    for (int l = 0; l < N.n_layers; ++l) {
        N.in_w[l] = l == 0 ? L.padded : c->width;
        N.out_w[l] = l == N.n_layers - 1 ? (c->n_output_dims + 15) / 16 * 16 : c->width;
        N.off[l] = off;
        off += (size_t)N.in_w[l] * N.out_w[l];
    }
    N.n_params = off;
    return N;
}

template <class T>
void mlp_one(const NetLayout& N, const float* w, const T* x0, T* raw, T* acts /* [H][W] or null */) {
    std::vector<T> a(x0, x0 + N.in_w[0]), b;
    for (int l = 0; l < N.n_layers; ++l) {
        const float* W = w + N.off[l];
        b.assign(N.out_w[l], (T)0);
        for (int o = 0; o < N.out_w[l]; ++o) {
            T acc = 0;
            for (int i = 0; i < N.in_w[l]; ++i) acc = M<T>::fma((T)W[(size_t)o * N.in_w[l] + i], a[i], acc);
            b[o] = acc;
        }
        if (l + 1 < N.n_layers) {
            for (auto& v : b) v = v > (T)0 ? v : (T)0;
            if (acts) std::copy(b.begin(), b.end(), acts + (size_t)l * N.out_w[l]);
        }
        a.swap(b);
    }
    std::copy(a.begin(), a.end(), raw);
}

// ---------------------------------------------------------------------------
// Mixture (src/render/guided/distribution.h, src/backend/src/util/vmf.h).
template <class T> struct Lobe {
    T lambda, kappa, theta, phi;
    V3<T> mu;      // cartesian mean direction
    T norm;        // stable normaliser C = kappa / (2 pi (-expm1(-2 kappa))), 1/4pi at 0
    T lognorm;     // log C
};

// Reference-form normaliser kappa / (2 pi (1 - e^{-2 kappa})) (vmf.h:22-27)
// with the 1/4pi limit below `thresh`.
template <class T> T ref_norm(T k, T thresh) {
    return k < thresh ? C<T>::inv_4pi : k * C<T>::inv_2pi / ((T)1 - M<T>::exp((T)-2 * k));
}
// Stable form C(k) = k e^{-k} / (4 pi sinh k) = k / (2 pi (-expm1(-2 k))), exact
// limit 1/4pi at k = 0; the lobe pdf is C exp(k (mu.w - 1)).
template <class T> T stable_norm(T k) {
    if (!(k > (T)1e-30)) return C<T>::inv_4pi;
    return k * C<T>::inv_2pi / -M<T>::expm1((T)-2 * k);
}

template <class T> struct Mixture {
    int K;
    Lobe<T> lobe[64];
    T weight[64];
    T total;
};

// train.h:199-210 (network_to_params) and distribution.h:85-98 (constructor);
// mapping "softmax" uses slots (theta, phi, log kappa, logit).
template <class T> void build_mixture(const orc_cfg* c, const T* x /* stride 1 */, Mixture<T>& m) {
    m.K = c->n_lobes;
    m.total = 0;
    if (c->mapping == ORC_MAP_REFERENCE) {
        for (int i = 0; i < m.K; ++i) {
            Lobe<T>& L = m.lobe[i];
            L.lambda = M<T>::exp(clampv(x[4 * i + 0], (T)-10, (T)15));
            L.kappa = M<T>::exp(clampv(x[4 * i + 1], (T)-10, (T)15));
            L.theta = C<T>::pi * ((T)1 / ((T)1 + M<T>::exp(-x[4 * i + 2])));
            L.phi = C<T>::two_pi * ((T)1 / ((T)1 + M<T>::exp(-x[4 * i + 3])));
            m.total += L.lambda;
        }
    } else {
        T mx = x[3];
        for (int i = 1; i < m.K; ++i) mx = std::max(mx, x[4 * i + 3]);
        for (int i = 0; i < m.K; ++i) {
            Lobe<T>& L = m.lobe[i];
            L.theta = C<T>::pi * ((T)1 / ((T)1 + M<T>::exp(-x[4 * i + 0])));
            L.phi = C<T>::two_pi * ((T)1 / ((T)1 + M<T>::exp(-x[4 * i + 1])));
            L.kappa = clampv(M<T>::exp(x[4 * i + 2]), (T)1e-2, (T)1e3);
            L.lambda = M<T>::exp(x[4 * i + 3] - mx);
            m.total += L.lambda;
        }
    }
    for (int i = 0; i < m.K; ++i) {
        Lobe<T>& L = m.lobe[i];
        m.weight[i] = L.lambda / m.total;
        L.mu = sph_to_cart(L.theta, L.phi);
        L.norm = stable_norm(L.kappa);
        L.lognorm = M<T>::log(L.norm);
    }
}

// distribution.h:184-195 (applyCosineLobe) and :39-64 (VMFKernel::product).
template <class T> void apply_cosine(const orc_cfg* c, const float* nrm, Mixture<T>& m) {
    const bool ref = c->numerics == ORC_NUM_REFERENCE;
    T nt, np;
    cart_to_sph(v3<T>((T)nrm[0], (T)nrm[1], (T)nrm[2]), &nt, &np);
    const T kb = (T)c->cosine_kappa;
    const V3<T> mu_b = sph_to_cart(nt, np);
    auto norm = [&](T k) { return ref ? ref_norm(k, (T)1e-3) : stable_norm(k); };
    m.total = 0;
    for (int i = 0; i < m.K; ++i) {
        Lobe<T>& L = m.lobe[i];
        const V3<T> mu_a = L.mu;
        V3<T> nm = add(mul(mu_a, L.kappa), mul(mu_b, kb));
        T nk = length(nm);
        if (nk >= (T)1e-3) {
            nm = divs(nm, nk);
        } else {
            nk = 0;
            nm = mu_a;
        }
        T th, ph;
        cart_to_sph(nm, &th, &ph);
        const T e = M<T>::exp(L.kappa * (dot(nm, mu_a) - (T)1) + kb * (dot(nm, mu_b) - (T)1));
        const T s = (ref ? ref_norm(L.kappa, (T)1e-3) : L.norm) * norm(kb) / norm(nk) * e;
        L.lambda = s * L.lambda * (T)1;
        L.kappa = nk;
        L.theta = th;
        L.phi = ph;
        // Reference numerics take the mean back through spherical coordinates;
        // the stable path keeps the normalised cartesian mean.
        L.mu = ref ? sph_to_cart(th, ph) : nm;
        L.norm = stable_norm(nk);
        L.lognorm = M<T>::log(L.norm);
        m.total += L.lambda;
    }
    m.total = std::max(m.total, C<T>::eps);
    for (int i = 0; i < m.K; ++i) m.weight[i] = m.lobe[i].lambda / m.total;
}

// vmf.h:22-27 (reference) or the log-space form.
template <class T> T lobe_pdf(const orc_cfg* c, const Lobe<T>& L, V3<T> w) {
    const T cth = dot(w, L.mu);
    if (c->numerics == ORC_NUM_REFERENCE) {
        if (L.kappa < C<T>::eps) return C<T>::inv_4pi;
        return M<T>::exp(L.kappa * std::min((T)0, cth - (T)1)) * L.kappa /
               (C<T>::two_pi * ((T)1 - M<T>::exp((T)-2 * L.kappa)));
    }
    return M<T>::exp(M<T>::fma(L.kappa, std::min((T)0, cth - (T)1), L.lognorm));
}

// distribution.h:113-119
template <class T> T mixture_pdf(const orc_cfg* c, const Mixture<T>& m, V3<T> w) {
    T p = 0;
    for (int i = 0; i < m.K; ++i) p = M<T>::fma(m.weight[i], lobe_pdf(c, m.lobe[i], w), p);
    return p;
}

// distribution.h:132-140 and vmf.h:48-71.
template <class T> V3<T> mixture_sample(const orc_cfg* c, const Mixture<T>& m, T u0, T u1, T u2, uint32_t* lobe) {
    int sel = 0;
    T u = u0;
    bool found = false;
    for (int i = 0; i < m.K; ++i) {
        if (u < m.weight[i]) { sel = i; found = true; break; }
        u -= m.weight[i];
    }
    if (!found) sel = 0;
    *lobe = (uint32_t)sel;
    const Lobe<T>& L = m.lobe[sel];
    V3<T> loc;
    const bool ref = c->numerics == ORC_NUM_REFERENCE;
    if (ref ? L.kappa < C<T>::eps : !(L.kappa > (T)1e-30)) {
        loc = uniform_sample_sphere(u1, u2);
    } else {
        T ct;
        if (ref) ct = (T)1 + M<T>::log1p(-u1 + M<T>::exp((T)-2 * L.kappa) * u1) / L.kappa;
        else ct = (T)1 + M<T>::log1p(u1 * M<T>::expm1((T)-2 * L.kappa)) / L.kappa;
        const T st = M<T>::sqrt(std::max((T)0, (T)1 - ct * ct));
        const T ph = C<T>::two_pi * u2;
        loc = {M<T>::cos(ph) * st, M<T>::sin(ph) * st, ct};
    }
    return frame_to_world(L.mu, loc);
}

template <class T>
void mixture_one(const orc_cfg* c, const T* x, const float* nrm, const float* u, const float* eval,
                 int n_eval, T* dir, T* pdf, T* eval_pdf, uint32_t* lobe, T* lobes_out) {
    Mixture<T> m;
    build_mixture(c, x, m);
    if (nrm) apply_cosine(c, nrm, m);
    if (lobes_out) {
        for (int i = 0; i < m.K; ++i) {
            lobes_out[5 * i + 0] = m.lobe[i].lambda;
            lobes_out[5 * i + 1] = m.lobe[i].kappa;
            lobes_out[5 * i + 2] = m.lobe[i].theta;
            lobes_out[5 * i + 3] = m.lobe[i].phi;
            lobes_out[5 * i + 4] = m.weight[i];
        }
    }
    if (u) {
        uint32_t sel;
        const V3<T> d = mixture_sample(c, m, (T)u[0], (T)u[1], (T)u[2], &sel);
        dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
        *lobe = sel;
        *pdf = mixture_pdf(c, m, d);
    }
    for (int e = 0; e < n_eval; ++e) {
        const V3<T> w = v3<T>((T)eval[3 * e + 0], (T)eval[3 * e + 1], (T)eval[3 * e + 2]);
        eval_pdf[e] = mixture_pdf(c, m, w);
    }
}

template <class F> void parallel_for(int n, int n_threads, F&& f) {
    if (n_threads <= 1 || n < 1024) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + n_threads - 1) / n_threads;
    for (int t = 0; t < n_threads; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([b, e, &f] { for (int i = b; i < e; ++i) f(i); });
    }
    for (auto& t : th) t.join();
}

template <class T>
void mixture_batch(const orc_cfg* c, const T* raw, const float* normal, const float* u, const float* eval_dirs,
                   int n_eval, int n, T* out_dir, T* out_pdf, T* out_eval_pdf, uint32_t* out_lobe, T* lobes_out) {
    const int n_out = orc_padded_output(c);
    const int K = c->n_lobes;
    for (int q = 0; q < n; ++q) {
        std::vector<T> x(n_out);
        for (int j = 0; j < n_out; ++j) x[j] = raw[(size_t)j * n + q];
        float nrm[3], uu[3];
        std::vector<float> ev(3 * n_eval);
        if (normal) for (int d = 0; d < 3; ++d) nrm[d] = normal[(size_t)d * n + q];
        if (u) for (int d = 0; d < 3; ++d) uu[d] = u[(size_t)d * n + q];
        for (int e = 0; e < n_eval; ++e)
            for (int d = 0; d < 3; ++d) ev[3 * e + d] = eval_dirs[((size_t)e * 3 + d) * n + q];
        T dir[3] = {}, pdf = 0;
        std::vector<T> epdf(n_eval), lobes(5 * K);
        uint32_t lobe = 0;
        mixture_one<T>(c, x.data(), normal ? nrm : nullptr, u ? uu : nullptr, ev.data(), n_eval, dir, &pdf,
                       epdf.data(), &lobe, lobes.data());
        if (u) {
            for (int d = 0; d < 3; ++d) out_dir[(size_t)d * n + q] = dir[d];
            out_pdf[q] = pdf;
            if (out_lobe) out_lobe[q] = lobe;
        }
        for (int e = 0; e < n_eval; ++e) out_eval_pdf[(size_t)e * n + q] = epdf[e];
        if (lobes_out)
            for (int j = 0; j < 5 * K; ++j) lobes_out[(size_t)j * n + q] = lobes[j];
    }
}

template <class T>
void query_batch(const orc_cfg* c, const float* params, const float* in, const float* normal, const float* u,
                 const float* eval_dirs, int n_eval, int n, T* out_dir, T* out_pdf, T* out_eval_pdf,
                 uint32_t* out_lobe, T* raw_out, int n_threads) {
    const EncLayout L = enc_layout(c);
    const GridLayout g = grid_layout(c);
    const NetLayout N = net_layout(c);
    const float* grid = params + N.n_params;
    const int n_out = N.out_w[N.n_layers - 1];
    parallel_for(n, n_threads, [&](int q) {
        float in_q[16];
        for (int d = 0; d < c->n_input_dims; ++d) in_q[d] = in[(size_t)d * n + q];
        std::vector<T> enc(L.padded), raw(n_out);
        encode_one<T>(c, L, g, grid, in_q, enc.data());
        mlp_one<T>(N, params, enc.data(), raw.data(), nullptr);
        if (raw_out) for (int j = 0; j < n_out; ++j) raw_out[(size_t)j * n + q] = raw[j];
        float nrm[3], uu[3];
        std::vector<float> ev(3 * n_eval);
        if (normal) for (int d = 0; d < 3; ++d) nrm[d] = normal[(size_t)d * n + q];
        if (u) for (int d = 0; d < 3; ++d) uu[d] = u[(size_t)d * n + q];
        for (int e = 0; e < n_eval; ++e)
            for (int d = 0; d < 3; ++d) ev[3 * e + d] = eval_dirs[((size_t)e * 3 + d) * n + q];
        T dir[3] = {}, pdf = 0;
        std::vector<T> epdf(n_eval);
        uint32_t lobe = 0;
        mixture_one<T>(c, raw.data(), normal ? nrm : nullptr, u ? uu : nullptr, ev.data(), n_eval, dir, &pdf,
                       epdf.data(), &lobe, nullptr);
        if (u) {
            for (int d = 0; d < 3; ++d) out_dir[(size_t)d * n + q] = dir[d];
            out_pdf[q] = pdf;
            if (out_lobe) out_lobe[q] = lobe;
        }
        for (int e = 0; e < n_eval; ++e) out_eval_pdf[(size_t)e * n + q] = epdf[e];
    });
}

// ---------------------------------------------------------------------------
// Loss and adjoint: train.cu:69-124 (compute_dL_doutput_divergence),
// distribution.h:143-181 (gradients_probability), train.h:212-223.
// With c->learn_selection: compute_dL_doutput_with_selection (train.cu:126-172)
// and MixedSGWithSelection::gradients_probability (distribution.h:216-231):
// P = s (p + eps) + (1 - s) bsdf, loss -f/P log P, prefix -f/P^2, lobe
// gradients times s, the selection slot LR_PROB (p - bsdf) s (1 - s)
// (train.h:55, 61), and neither NaN zeroing nor the clamp.
template <class T>
void adjoint_one(const orc_cfg* c, const T* x, T th_t, T ph_t, const float rgb[3], float wipdf, float bsdf,
                 T ls_over_n, T clamp_hi, T* g, T* loss) {
    const bool ref = c->numerics == ORC_NUM_REFERENCE;
    // fp16_intermediates: the reference's __half buffers (train.cu:80-81 raw
    // outputs, distribution.h:177-178 per-lobe dp, train.cu:95-99 dL/dx).
    const bool h16 = c->fp16_intermediates != 0;
    auto hq = [&](T v) { return h16 ? (T)fm::round_half((float)v) : v; };
    const int n_raw = orc_padded_output(c);
    std::vector<T> xh;
    if (h16) {
        xh.resize(n_raw);
        for (int j = 0; j < n_raw; ++j) xh[j] = hq(x[j]);
        x = xh.data();
    }
    Mixture<T> m;
    build_mixture(c, x, m);
    const int K = m.K;
    const V3<T> wi = sph_to_cart(th_t, ph_t);
    const T Li = ((T)rgb[0] + (T)rgb[1] + (T)rgb[2]) / (T)3;
    const T wiPdf = (T)wipdf + C<T>::eps;
    // dp/d(lambda, kappa, theta, phi) per lobe.
    T vmf[64];
    for (int k = 0; k < K; ++k) vmf[k] = lobe_pdf(c, m.lobe[k], wi);
    T dp[64][4];
    T prob = 0;
    for (int s = 0; s < K; ++s) {
        const Lobe<T>& L = m.lobe[s];
        prob = M<T>::fma(m.weight[s], vmf[s], prob);
        T dl = vmf[s] * (m.total - L.lambda) / (m.total * m.total);
        for (int k = 0; k < K; ++k)
            if (k != s) dl -= m.weight[k] / m.total * vmf[k];
        T ik;
        if (ref) {
            if (L.kappa < (T)1) ik = (T)0.000926 + (T)-0.344883 * L.kappa + (T)0.030147 * L.kappa * L.kappa;
            else ik = (T)1 / L.kappa - ((T)1 + M<T>::exp((T)-2 * L.kappa)) / ((T)1 - M<T>::exp((T)-2 * L.kappa));
        } else {
            // 1/k - coth k, series below 0.5 (-k/3 + k^3/45 - 2k^5/945 + k^7/4725).
            const T k = L.kappa;
            if (k < (T)0.5) {
                const T k2 = k * k;
                ik = -k * ((T)1 / (T)3 - k2 * ((T)1 / (T)45 - k2 * ((T)2 / (T)945 - k2 * ((T)1 / (T)4725))));
            } else {
                const T e = M<T>::exp((T)-2 * k);
                ik = (T)1 / k - ((T)1 + e) / ((T)1 - e);
            }
        }
        const T st = M<T>::sin(L.theta), ct = M<T>::cos(L.theta), sp = M<T>::sin(L.phi), cp = M<T>::cos(L.phi);
        const T wv = m.weight[s] * vmf[s];
        dp[s][0] = hq(dl);
        dp[s][1] = hq(wv * (wi.x * st * cp + wi.y * st * sp + wi.z * ct + ik));
        dp[s][2] = hq(wv * L.kappa * (ct * (wi.x * cp + wi.y * sp) - wi.z * st));
        dp[s][3] = hq(wv * L.kappa * (wi.y * cp - wi.x * sp) * st);
    }
    const T guide = prob + C<T>::eps;
    const bool sel = c->learn_selection != 0;
    T prefix, s_sel = 1;
    if (sel) {
        s_sel = (T)1 / ((T)1 + M<T>::exp(-x[4 * K]));
        const T P = s_sel * guide + ((T)1 - s_sel) * (T)bsdf;
        prefix = -Li / P / P * ls_over_n;
        *loss = -Li / P * M<T>::log(P);
        for (int s = 0; s < K; ++s)
            for (int j = 0; j < 4; ++j) dp[s][j] = hq(dp[s][j] * s_sel);
    } else {
        prefix = -Li / wiPdf / guide * ls_over_n;
        *loss = -Li / wiPdf * M<T>::log(guide);
    }
    const int n_out = orc_padded_output(c);
    for (int j = 0; j < n_out; ++j) g[j] = 0;
    for (int s = 0; s < K; ++s) {
        const T* xs = x + 4 * s;
        if (c->mapping == ORC_MAP_REFERENCE) {
            // d_network_to_d_params: exp(clamp(x)) for both exponential slots
            // (the reference keeps it outside the clamp; stable numerics use the
            // exact derivative, zero outside), logistic' for theta and phi.
            auto dexp = [&](T v) {
                const T e = M<T>::exp(clampv(v, (T)-10, (T)15));
                return (!ref && (v < (T)-10 || v > (T)15)) ? (T)0 : e;
            };
            const T s2 = (T)1 / ((T)1 + M<T>::exp(-xs[2]));
            const T s3 = (T)1 / ((T)1 + M<T>::exp(-xs[3]));
            g[4 * s + 0] = (T)1 * prefix * dp[s][0] * dexp(xs[0]);
            g[4 * s + 1] = (T)1 * prefix * dp[s][1] * dexp(xs[1]);
            g[4 * s + 2] = (T)1 * prefix * dp[s][2] * C<T>::pi * (s2 * ((T)1 - s2));
            g[4 * s + 3] = (T)1 * prefix * dp[s][3] * C<T>::two_pi * (s3 * ((T)1 - s3));
        } else {
            const T s0 = (T)1 / ((T)1 + M<T>::exp(-xs[0]));
            const T s1 = (T)1 / ((T)1 + M<T>::exp(-xs[1]));
            const T ek = M<T>::exp(xs[2]);
            const T dk = (ek > (T)1e-2 && ek < (T)1e3) ? ek : (T)0;
            g[4 * s + 0] = prefix * dp[s][2] * C<T>::pi * (s0 * ((T)1 - s0));
            g[4 * s + 1] = prefix * dp[s][3] * C<T>::two_pi * (s1 * ((T)1 - s1));
            g[4 * s + 2] = prefix * dp[s][1] * dk;
            // lambda_k = exp(logit_k - max): d lambda_k / d logit_s = delta_ks lambda_s.
            g[4 * s + 3] = prefix * dp[s][0] * m.lobe[s].lambda;
        }
    }
    for (int j = 0; j < 4 * K; ++j) g[j] = hq(g[j]);
    if (sel) {
        g[4 * K] = hq((T)0.01 * prefix * hq(prob - (T)bsdf) * (s_sel * ((T)1 - s_sel)));
        return;
    }
    // train.cu:106-123: zero non-finite gradients, clamp from above.
    for (int j = 0; j < 4 * K; ++j) {
        if (!M<T>::finite(g[j])) g[j] = 0;
        if (clamp_hi > 0) g[j] = std::min(clamp_hi, g[j]);
    }
}

template <class T>
void adjoint_batch(const orc_cfg* c, const T* raw, const float* tgt_dir, const float* tgt_rgb, const float* tgt_pdf,
                   const float* tgt_bsdf, int n, T loss_scale, T* dout, T* loss) {
    const int n_out = orc_padded_output(c);
    const T ls = loss_scale / (T)n;
    // The reference clamps the scaled gradient at loss scale 128 (train.cu:116-122,
    // parameters.h:24); the threshold follows the scale so that the clamp acts on
    // the same unscaled values at any loss scale.
    const T clamp_hi = c->grad_clamp > 0 ? (T)c->grad_clamp * (loss_scale / (T)c->clamp_scale) : (T)0;
    std::vector<T> x(n_out), g(n_out);
    for (int q = 0; q < n; ++q) {
        for (int j = 0; j < n_out; ++j) x[j] = raw[(size_t)j * n + q];
        const float rgb[3] = {tgt_rgb[q], tgt_rgb[(size_t)n + q], tgt_rgb[2 * (size_t)n + q]};
        adjoint_one<T>(c, x.data(), (T)tgt_dir[q], (T)tgt_dir[(size_t)n + q], rgb, tgt_pdf[q],
                       tgt_bsdf ? tgt_bsdf[q] : 0.0f, ls, clamp_hi, g.data(), &loss[q]);
        for (int j = 0; j < n_out; ++j) dout[(size_t)j * n + q] = g[j];
    }
}

// ---------------------------------------------------------------------------
// MLP backward (fully_fused_mlp.cu:735-836): dW_l = dY_l X_l^T, dX = W^T dY
// masked by the ReLU of the stored post-activations.
template <class T>
void mlp_backward(const orc_cfg* c, const float* w, const T* enc, const T* acts, const T* dout, int n, T* dw, T* denc) {
    const NetLayout N = net_layout(c);
    const int H = N.n_layers - 1;
    std::fill(dw, dw + N.n_params, (T)0);
    std::vector<T> g, gprev;
    for (int q = 0; q < n; ++q) {
        g.assign(N.out_w[H], 0);
        for (int o = 0; o < N.out_w[H]; ++o) g[o] = dout[(size_t)o * n + q];
        for (int l = H; l >= 0; --l) {
            const int iw = N.in_w[l], ow = N.out_w[l];
            auto xin = [&](int i) -> T {
                return l == 0 ? enc[(size_t)i * n + q] : acts[((size_t)(l - 1) * c->width + i) * n + q];
            };
            for (int o = 0; o < ow; ++o)
                for (int i = 0; i < iw; ++i) dw[N.off[l] + (size_t)o * iw + i] += g[o] * xin(i);
            gprev.assign(iw, 0);
            for (int i = 0; i < iw; ++i) {
                T acc = 0;
                for (int o = 0; o < ow; ++o) acc = M<T>::fma((T)w[N.off[l] + (size_t)o * iw + i], g[o], acc);
                gprev[i] = (l > 0 && !(xin(i) > (T)0)) ? (T)0 : acc;
            }
            g.swap(gprev);
        }
        if (denc)
            for (int i = 0; i < N.in_w[0]; ++i) denc[(size_t)i * n + q] = g[i];
    }
}

// tcnn grid.h:386-509 (kernel_grid_backward): scatter w_corner * dL/dy,
// queries in order, levels in order, corners in order.
template <class T>
void grid_backward(const orc_cfg* c, const float* in, const T* denc, int n, T* gg) {
    const GridLayout g = grid_layout(c);
    const EncLayout L = enc_layout(c);
    int ge = -1;
    for (int e = 0; e < c->n_enc; ++e) if (c->enc_type[e] == ORC_ENC_GRID) ge = e;
    if (ge < 0) return;
    const int F = c->n_features;
    std::fill(gg, gg + (size_t)g.offsets[c->n_levels] * F, (T)0);
    uint32_t idx[8];
    T w[8];
    for (int q = 0; q < n; ++q) {
        float pos[3];
        for (int d = 0; d < 3; ++d) pos[d] = in[(size_t)(c->enc_begin[ge] + d) * n + q];
        for (int l = 0; l < c->n_levels; ++l) {
            const int nc = level_corners<T>(c, g, l, pos, idx, w);
            for (int k = 0; k < nc; ++k)
                for (int f = 0; f < F; ++f)
                    gg[((size_t)g.offsets[l] + idx[k]) * F + f] += w[k] * denc[(size_t)(L.offset[ge] + l * F + f) * n + q];
        }
    }
}

}  // namespace

namespace {

// PCG32 (O'Neill 2014), as vendored by tcnn (dependencies/pcg32/pcg32.h:44-170).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    explicit Pcg32(uint64_t initstate, uint64_t initseq = 1u) { seed(initstate, initseq); }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    uint32_t next_uint() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float next_float() { return fm::bits_to_float((next_uint() >> 9) | 0x3f800000u) - 1.0f; }
    // Jump ahead by delta steps (Brown, "Random number generation with arbitrary strides", 1994).
    void advance(uint64_t delta) {
        uint64_t cur_mult = 0x5851f42d4c957f2dULL, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta /= 2;
        }
        state = acc_mult * state + acc_plus;
    }
};

// Field parameters of the analytic radiance field (three rotating vMF lobes).
struct Field {
    float mu0[3][3], axis[3][3], d[3][3], kappa[3], a[3];
};

Field make_field(uint64_t seed) {
    Field F{};
    Pcg32 r(seed ^ 0x9E3779B97F4A7C15ULL, 0x5eedULL);
    const float amp[3] = {0.6f, 0.3f, 0.1f};
    for (int j = 0; j < 3; ++j) {
        float u0 = r.next_float(), u1 = r.next_float();
        V3<float> m = uniform_sample_sphere(u0, u1);
        u0 = r.next_float(); u1 = r.next_float();
        V3<float> ax = uniform_sample_sphere(u0, u1);
        F.mu0[j][0] = m.x; F.mu0[j][1] = m.y; F.mu0[j][2] = m.z;
        F.axis[j][0] = ax.x; F.axis[j][1] = ax.y; F.axis[j][2] = ax.z;
        for (int k = 0; k < 3; ++k) F.d[j][k] = (2.0f * r.next_float() - 1.0f) * C<float>::pi;
        F.kappa[j] = fm::exp(5.29831736654804f * r.next_float());  // log-uniform in [1, 200]
        F.a[j] = amp[j];
    }
    return F;
}

template <class T> T field_value(const Field& F, const float pos[3], V3<T> w) {
    T f = 0;
    for (int j = 0; j < 3; ++j) {
        const T ang = (T)F.d[j][0] * (T)pos[0] + (T)F.d[j][1] * (T)pos[1] + (T)F.d[j][2] * (T)pos[2];
        const T s = M<T>::sin(ang), c = M<T>::cos(ang);
        const V3<T> k = v3<T>(F.axis[j][0], F.axis[j][1], F.axis[j][2]);
        const V3<T> v = v3<T>(F.mu0[j][0], F.mu0[j][1], F.mu0[j][2]);
        const V3<T> kxv = cross(k, v);
        const T kdv = dot(k, v);
        const V3<T> mu = add(add(mul(v, c), mul(kxv, s)), mul(k, kdv * ((T)1 - c)));
        const T kap = (T)F.kappa[j];
        const T val = kap * M<T>::exp(kap * (dot(mu, w) - (T)1)) / (C<T>::two_pi * -M<T>::expm1((T)-2 * kap));
        f = M<T>::fma((T)F.a[j], val, f);
    }
    return f;
}

// Shared head of a synthetic record: position, normal, outgoing direction,
// roughness, packed like generate_inference_data (train.cu:174-191).
struct Head {
    float in[9];
    V3<float> n, wo;
};

Head synth_head(Pcg32& r) {
    Head h;
    for (int d = 0; d < 3; ++d) h.in[d] = r.next_float();
    float a0 = r.next_float(), a1 = r.next_float();
    h.n = uniform_sample_sphere(a0, a1);
    a0 = r.next_float(); a1 = r.next_float();
    h.wo = uniform_sample_sphere(a0, a1);
    const float rough = r.next_float();
    h.in[3] = (h.wo.x + 1.0f) * 0.5f;
    h.in[4] = (h.wo.y + 1.0f) * 0.5f;
    h.in[5] = (h.wo.z + 1.0f) * 0.5f;
    float th, ph;
    cart_to_sph(h.n, &th, &ph);
    h.in[6] = th / C<float>::pi;
    h.in[7] = ph / C<float>::two_pi;
    h.in[8] = 1.0f - fm::exp(-rough);
    return h;
}

// ---- Wavefront packers ----------------------------------------------------

// normalizeSpatialCoord (src/render/guided/train.h:173-177): the box inflated
// by 0.5 % of its diagonal's length (AABB::inflate, krrmath/aabb.h:36-39),
// then 0.5 + (x - center) / diagonal per axis (Eigen AlignedBox center and
// diagonal).
template <class T> void box_normalize(const T* lo, const T* hi, const T* x, T* out) {
    const T dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const T inflation = M<T>::sqrt(dx * dx + dy * dy + dz * dz) * (T)0.005;
    for (int a = 0; a < 3; ++a) {
        const T l = lo[a] - inflation, h = hi[a] + inflation;
        const T center = (l + h) * (T)0.5, diag = h - l;
        out[a] = (T)0.5 + (x[a] - center) / diag;
    }
}

// One training record from recorded vertex (pid, d) (train.cu:29-64); false
// when the reference drops it (delta vertex, NaN input/target, zero pdf).
template <class T> struct PackRecords {
    const T *L, *thp, *pos, *dir, *wi_pdf, *mis, *bsdf, *wo, *normal, *rough;
    const uint8_t* delta;
    size_t P;
};
template <class T>
bool pack_train_record(const PackRecords<T>& r, const T* lo, const T* hi, size_t pid, size_t d, T* in, T* dir,
                       T* rgb, T* pdf) {
    const size_t P = r.P, s1 = d * P + pid, s2 = d * 2 * P + pid, s3 = d * 3 * P + pid;
    if (r.delta[s1]) return false;
    const T x[3] = {r.pos[s3], r.pos[s3 + P], r.pos[s3 + 2 * P]};
    box_normalize(lo, hi, x, in);
    const V3<T> w = sph_to_cart(r.wo[s2], r.wo[s2 + P]);  // warp_direction_for_sh (train.cu:6-8)
    in[3] = (w.x + (T)1) * (T)0.5;
    in[4] = (w.y + (T)1) * (T)0.5;
    in[5] = (w.z + (T)1) * (T)0.5;
    in[6] = r.normal[s2] * C<T>::inv_pi;
    in[7] = r.normal[s2 + P] * C<T>::inv_2pi;
    in[8] = (T)1 - M<T>::exp(-r.rough[s1]);  // warp_roughness_for_ob (train.cu:10-12)
    bool nan_l = false;
    for (int ch = 0; ch < 3; ++ch) {
        const T thp = r.thp[s3 + ch * P];
        T L = thp > C<T>::eps ? r.L[s3 + ch * P] / thp : (T)0;
        L = L * r.mis[s1];
        nan_l = nan_l || std::isnan(L);
        rgb[ch] = L * r.bsdf[s3 + ch * P];  // output.L *= bsdfVal after the NaN test's L
    }
    dir[0] = r.dir[s2];
    dir[1] = r.dir[s2 + P];
    *pdf = r.wi_pdf[s1];
    const bool bad = std::isnan(in[0]) || std::isnan(in[1]) || std::isnan(in[2]) || std::isnan(dir[0]) ||
                     std::isnan(dir[1]) || std::isnan(in[3]) || std::isnan(in[4]) || std::isnan(in[5]) ||
                     std::isnan(*pdf) || *pdf == (T)0 || nan_l;
    return !bad;
}

}  // namespace

extern "C" {

void orc_pack_queries_f(const float* box, const float* pos, const float* wo, const float* normal,
                        const float* roughness, int n, float* in) {
    const size_t N = (size_t)n;
    for (size_t i = 0; i < N; ++i) {
        const float x[3] = {pos[i], pos[N + i], pos[2 * N + i]};
        float p[3];
        box_normalize(box, box + 3, x, p);
        const V3<float> w = normalize(v3(wo[i], wo[N + i], wo[2 * N + i]));  // sd.wo.normalized()
        float th, ph;
        cart_to_sph(v3(normal[i], normal[N + i], normal[2 * N + i]), &th, &ph);
        const float v[9] = {p[0], p[1], p[2], (w.x + 1.0f) * 0.5f, (w.y + 1.0f) * 0.5f, (w.z + 1.0f) * 0.5f,
                            th / C<float>::pi, ph / C<float>::two_pi, 1.0f - fm::exp(-roughness[i])};
        for (int d = 0; d < 9; ++d) in[(size_t)d * N + i] = v[d];
    }
}

uint32_t orc_pack_training_f(const float* L, const float* thp, const float* pos, const float* dir,
                             const float* wi_pdf, const float* mis, const uint8_t* delta, const float* bsdf,
                             const float* wo, const float* normal, const float* rough, const uint32_t* depth,
                             const float* bsdf_pdf, uint32_t max_depth, uint32_t n_pixels, const uint32_t* tile,
                             const float* box, uint32_t capacity, float* in, float* tgt_dir, float* tgt_rgb,
                             float* tgt_pdf, float* tgt_bsdf_pdf) {
    const PackRecords<float> r{L, thp, pos, dir, wi_pdf, mis, bsdf, wo, normal, rough, delta, n_pixels};
    // tile: offset xy, stride xy, training resolution xy, image resolution xy
    const uint64_t n_threads = (uint64_t)tile[4] * tile[5];
    uint32_t count = 0;
    const size_t C = capacity;
    for (uint64_t t = 0; t < n_threads; ++t) {
        const uint32_t px = (uint32_t)(t % tile[4]) * tile[2] + tile[0];
        const uint32_t py = (uint32_t)(t / tile[4]) * tile[3] + tile[1];
        if (px >= tile[6] || py >= tile[7]) continue;
        const size_t pid = (size_t)px + (size_t)py * tile[6];
        const uint32_t nd = std::min(depth[pid], max_depth);
        for (uint32_t d = 0; d < nd; ++d) {
            float i9[9], dr[2], rgb[3], pdf;
            if (!pack_train_record(r, box, box + 3, pid, d, i9, dr, rgb, &pdf)) continue;
            if (count < capacity) {
                for (int k = 0; k < 9; ++k) in[(size_t)k * C + count] = i9[k];
                tgt_dir[count] = dr[0];
                tgt_dir[C + count] = dr[1];
                for (int k = 0; k < 3; ++k) tgt_rgb[(size_t)k * C + count] = rgb[k];
                tgt_pdf[count] = pdf;
                if (tgt_bsdf_pdf) tgt_bsdf_pdf[count] = bsdf_pdf ? bsdf_pdf[(size_t)d * n_pixels + pid] : 0.0f;
            }
            ++count;
        }
    }
    return count;
}


int orc_encoded_width(const orc_cfg* c) { return enc_layout(c).padded; }
int orc_padded_output(const orc_cfg* c) { return (c->n_output_dims + 15) / 16 * 16; }

void orc_n_params(const orc_cfg* c, size_t* n_mlp, size_t* n_grid) {
    *n_mlp = net_layout(c).n_params;
    bool has_grid = false;
    for (int e = 0; e < c->n_enc; ++e) has_grid |= c->enc_type[e] == ORC_ENC_GRID;
    *n_grid = has_grid ? (size_t)grid_layout(c).offsets[c->n_levels] * c->n_features : 0;
}

size_t orc_grid_layout(const orc_cfg* c, uint32_t* offsets, float* scales, uint32_t* resolutions) {
    const GridLayout g = grid_layout(c);
    for (int l = 0; l <= c->n_levels; ++l) offsets[l] = g.offsets[l];
    for (int l = 0; l < c->n_levels; ++l) { scales[l] = g.scales[l]; resolutions[l] = g.res[l]; }
    return (size_t)g.offsets[c->n_levels] * c->n_features;
}

void orc_grid_corners(const orc_cfg* c, const float* in, int n, uint32_t* idx, float* w) {
    const GridLayout g = grid_layout(c);
    int ge = 0;
    for (int e = 0; e < c->n_enc; ++e) if (c->enc_type[e] == ORC_ENC_GRID) ge = e;
    for (int q = 0; q < n; ++q) {
        float pos[3];
        for (int d = 0; d < 3; ++d) pos[d] = in[(size_t)(c->enc_begin[ge] + d) * n + q];
        for (int l = 0; l < c->n_levels; ++l) {
            uint32_t ix[8] = {};
            float wk[8] = {};
            level_corners<float>(c, g, l, pos, ix, wk);
            for (int k = 0; k < 8; ++k) {
                idx[((size_t)l * 8 + k) * n + q] = ix[k];
                w[((size_t)l * 8 + k) * n + q] = wk[k];
            }
        }
    }
}

#define ORC_ENCODE(SUF, T)                                                                   \
    void orc_encode_##SUF(const orc_cfg* c, const float* grid, const float* in, int n, T* enc) { \
        const EncLayout L = enc_layout(c);                                                   \
        const GridLayout g = grid_layout(c);                                                 \
        std::vector<T> e(L.padded);                                                          \
        for (int q = 0; q < n; ++q) {                                                        \
            float in_q[16];                                                                  \
            for (int d = 0; d < c->n_input_dims; ++d) in_q[d] = in[(size_t)d * n + q];       \
            encode_one<T>(c, L, g, grid, in_q, e.data());                                    \
            for (int j = 0; j < L.padded; ++j) enc[(size_t)j * n + q] = e[j];                \
        }                                                                                    \
    }
ORC_ENCODE(f, float)
ORC_ENCODE(d, double)

#define ORC_MLP(SUF, T)                                                                                  \
    void orc_mlp_forward_##SUF(const orc_cfg* c, const float* w, const T* enc, int n, T* raw, T* acts) { \
        const NetLayout N = net_layout(c);                                                               \
        const int H = N.n_layers - 1, W = c->width, n_out = N.out_w[H];                                  \
        std::vector<T> x(N.in_w[0]), y(n_out), a((size_t)H * W);                                         \
        for (int q = 0; q < n; ++q) {                                                                    \
            for (int j = 0; j < N.in_w[0]; ++j) x[j] = enc[(size_t)j * n + q];                           \
            mlp_one<T>(N, w, x.data(), y.data(), a.data());                                              \
            for (int j = 0; j < n_out; ++j) raw[(size_t)j * n + q] = y[j];                               \
            if (acts) for (size_t j = 0; j < (size_t)H * W; ++j) acts[j * n + q] = a[j];                 \
        }                                                                                                \
    }
ORC_MLP(f, float)
ORC_MLP(d, double)

void orc_mixture_f(const orc_cfg* c, const float* raw, const float* normal, const float* u, const float* eval_dirs,
                   int n_eval, int n, float* out_dir, float* out_pdf, float* out_eval_pdf, uint32_t* out_lobe,
                   float* lobes_out) {
    mixture_batch<float>(c, raw, normal, u, eval_dirs, n_eval, n, out_dir, out_pdf, out_eval_pdf, out_lobe, lobes_out);
}
void orc_mixture_d(const orc_cfg* c, const double* raw, const float* normal, const float* u, const float* eval_dirs,
                   int n_eval, int n, double* out_dir, double* out_pdf, double* out_eval_pdf, uint32_t* out_lobe,
                   double* lobes_out) {
    mixture_batch<double>(c, raw, normal, u, eval_dirs, n_eval, n, out_dir, out_pdf, out_eval_pdf, out_lobe, lobes_out);
}

void orc_query_f(const orc_cfg* c, const float* params, const float* in, const float* normal, const float* u,
                 const float* eval_dirs, int n_eval, int n, float* out_dir, float* out_pdf, float* out_eval_pdf,
                 uint32_t* out_lobe, float* raw_out, int n_threads) {
    query_batch<float>(c, params, in, normal, u, eval_dirs, n_eval, n, out_dir, out_pdf, out_eval_pdf, out_lobe,
                       raw_out, n_threads);
}
void orc_query_d(const orc_cfg* c, const float* params, const float* in, const float* normal, const float* u,
                 const float* eval_dirs, int n_eval, int n, double* out_dir, double* out_pdf, double* out_eval_pdf,
                 uint32_t* out_lobe, double* raw_out, int n_threads) {
    query_batch<double>(c, params, in, normal, u, eval_dirs, n_eval, n, out_dir, out_pdf, out_eval_pdf, out_lobe,
                        raw_out, n_threads);
}

void orc_adjoint_f(const orc_cfg* c, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                   const float* tgt_pdf, int n, float loss_scale, float* dout, float* loss) {
    adjoint_batch<float>(c, raw, tgt_dir, tgt_rgb, tgt_pdf, nullptr, n, loss_scale, dout, loss);
}
void orc_adjoint_d(const orc_cfg* c, const double* raw, const float* tgt_dir, const float* tgt_rgb,
                   const float* tgt_pdf, int n, double loss_scale, double* dout, double* loss) {
    adjoint_batch<double>(c, raw, tgt_dir, tgt_rgb, tgt_pdf, nullptr, n, loss_scale, dout, loss);
}

void orc_adjoint_sel_f(const orc_cfg* c, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                       const float* tgt_pdf, const float* tgt_bsdf, int n, float loss_scale, float* dout, float* loss) {
    adjoint_batch<float>(c, raw, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, n, loss_scale, dout, loss);
}
void orc_adjoint_sel_d(const orc_cfg* c, const double* raw, const float* tgt_dir, const float* tgt_rgb,
                       const float* tgt_pdf, const float* tgt_bsdf, int n, double loss_scale, double* dout,
                       double* loss) {
    adjoint_batch<double>(c, raw, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, n, loss_scale, dout, loss);
}
void orc_selection_f(const orc_cfg* c, const float* raw, int n, float* out) {
    const size_t j = 4 * (size_t)c->n_lobes;
    for (int q = 0; q < n; ++q) out[q] = 1.0f / (1.0f + fm::exp(-raw[j * n + q]));
}

void orc_mlp_backward_f(const orc_cfg* c, const float* w, const float* enc, const float* acts, const float* dout,
                        int n, float* dw, float* denc) {
    mlp_backward<float>(c, w, enc, acts, dout, n, dw, denc);
}
void orc_mlp_backward_d(const orc_cfg* c, const float* w, const double* enc, const double* acts, const double* dout,
                        int n, double* dw, double* denc) {
    mlp_backward<double>(c, w, enc, acts, dout, n, dw, denc);
}

void orc_grid_backward_f(const orc_cfg* c, const float* in, const float* denc, int n, float* gg) {
    grid_backward<float>(c, in, denc, n, gg);
}
void orc_grid_backward_d(const orc_cfg* c, const float* in, const double* denc, int n, double* gg) {
    grid_backward<double>(c, in, denc, n, gg);
}

// tcnn optimizers/adam.h:47-119 with the reference's hyper-parameters.
void orc_adam_step(size_t n, size_t n_matrix, float loss_scale, float lr, float beta1, float beta2, float eps,
                   float l2_reg, float* master, const float* grad, float* m, float* v, uint32_t* steps) {
    for (size_t i = 0; i < n; ++i) {
        float g = grad[i] / loss_scale;
        if (i >= n_matrix && g == 0) continue;
        const float w = master[i];
        if (i < n_matrix) g += l2_reg * w;
        const float g2 = g * g;
        const float m1 = m[i] = beta1 * m[i] + (1 - beta1) * g;
        const float v1 = v[i] = beta2 * v[i] + (1 - beta2) * g2;
        const uint32_t t = ++steps[i];
        // beta^t evaluated as exp(t log beta) with the deterministic elementary functions.
        const float b2t = fm::exp((float)t * fm::log(beta2)), b1t = fm::exp((float)t * fm::log(beta1));
        const float lr_t = lr * (fm::sqrt_(1 - b2t) / (1 - b1t));
        const float eff = std::min(std::max(lr_t / (fm::sqrt_(v1) + eps), 0.0f), 3.402823466e38f);
        master[i] = w - eff * m1;
    }
}

// One FP32 training step on n records with n_threads host threads, the CPU
// baseline of the train step (integrator.cpp:679-699): forward (encode, MLP
// keeping the activations), KL loss and adjoint (train.cu:69-124), MLP
// backward with per-thread weight-gradient partials summed in thread order,
// grid backward with levels spread over the threads (disjoint table ranges,
// so no two threads write one entry), then Adam (adam.h:47-119). Parameters,
// moments and step counts are updated in place; returns the mean loss.
float orc_train_step_f(const orc_cfg* c, float* params, float* m, float* v, uint32_t* steps, const float* in,
                       const float* tgt_dir, const float* tgt_rgb, const float* tgt_pdf, int n, float loss_scale,
                       float lr, float beta1, float beta2, float eps, float l2_reg, int n_threads) {
    const EncLayout L = enc_layout(c);
    const GridLayout g = grid_layout(c);
    const NetLayout N = net_layout(c);
    const int H = N.n_layers - 1, W = c->width, n_out = N.out_w[H];
    const size_t n_mlp = N.n_params, n_grid = (size_t)g.offsets[c->n_levels] * c->n_features;
    const float* grid = params + n_mlp;
    const int T = std::max(1, n_threads);
    std::vector<float> enc((size_t)L.padded * n), acts((size_t)H * W * n), dout((size_t)n_out * n), loss(n);
    const float ls = loss_scale / (float)n;
    const float clamp_hi = c->grad_clamp > 0 ? (float)c->grad_clamp * (loss_scale / (float)c->clamp_scale) : 0.0f;
    // forward + adjoint, samples spread over the threads
    parallel_for(n, T, [&](int q) {
        float in_q[16];
        for (int d = 0; d < c->n_input_dims; ++d) in_q[d] = in[(size_t)d * n + q];
        std::vector<float> e(L.padded), raw(n_out), a((size_t)H * W), gq(n_out);
        encode_one<float>(c, L, g, grid, in_q, e.data());
        mlp_one<float>(N, params, e.data(), raw.data(), a.data());
        for (int i = 0; i < L.padded; ++i) enc[(size_t)i * n + q] = e[i];
        for (size_t i = 0; i < a.size(); ++i) acts[i * n + q] = a[i];
        const float rgb[3] = {tgt_rgb[q], tgt_rgb[(size_t)n + q], tgt_rgb[2 * (size_t)n + q]};
        adjoint_one<float>(c, raw.data(), tgt_dir[q], tgt_dir[(size_t)n + q], rgb, tgt_pdf[q], 0.0f, ls, clamp_hi,
                           gq.data(), &loss[q]);
        for (int j = 0; j < n_out; ++j) dout[(size_t)j * n + q] = gq[j];
    });
    // MLP backward: thread t owns a contiguous range of samples and a partial dW
    std::vector<float> grad(n_mlp + n_grid, 0.0f), denc((size_t)L.padded * n);
    std::vector<std::vector<float>> part(T, std::vector<float>(n_mlp, 0.0f));
    {
        std::vector<std::thread> th;
        const int chunk = (n + T - 1) / T;
        for (int t = 0; t < T; ++t) {
            const int b = t * chunk, e = std::min(n, b + chunk);
            if (b >= e) break;
            th.emplace_back([&, t, b, e] {
                float* dw = part[t].data();
                std::vector<float> gg, gprev;
                for (int q = b; q < e; ++q) {
                    gg.assign(n_out, 0.0f);
                    for (int o = 0; o < n_out; ++o) gg[o] = dout[(size_t)o * n + q];
                    for (int l = H; l >= 0; --l) {
                        const int iw = N.in_w[l], ow = N.out_w[l];
                        auto xin = [&](int i) -> float {
                            return l == 0 ? enc[(size_t)i * n + q] : acts[((size_t)(l - 1) * W + i) * n + q];
                        };
                        for (int o = 0; o < ow; ++o)
                            for (int i = 0; i < iw; ++i) dw[N.off[l] + (size_t)o * iw + i] += gg[o] * xin(i);
                        gprev.assign(iw, 0.0f);
                        for (int i = 0; i < iw; ++i) {
                            float acc = 0;
                            for (int o = 0; o < ow; ++o) acc = fm::fma_(params[N.off[l] + (size_t)o * iw + i], gg[o], acc);
                            gprev[i] = (l > 0 && !(xin(i) > 0.0f)) ? 0.0f : acc;
                        }
                        gg.swap(gprev);
                    }
                    for (int i = 0; i < N.in_w[0]; ++i) denc[(size_t)i * n + q] = gg[i];
                }
            });
        }
        for (auto& x : th) x.join();
    }
    for (int t = 0; t < T; ++t)
        for (size_t i = 0; i < n_mlp; ++i) grad[i] += part[t][i];
    // grid backward: level l on thread l mod T
    int ge = -1;
    for (int e = 0; e < c->n_enc; ++e) if (c->enc_type[e] == ORC_ENC_GRID) ge = e;
    if (ge >= 0) {
        std::vector<std::thread> th;
        const int F = c->n_features;
        for (int t = 0; t < std::min(T, c->n_levels); ++t) {
            th.emplace_back([&, t] {
                uint32_t idx[8];
                float w[8];
                for (int l = t; l < c->n_levels; l += T) {
                    float* gl = grad.data() + n_mlp + (size_t)g.offsets[l] * F;
                    for (int q = 0; q < n; ++q) {
                        float pos[3];
                        for (int d = 0; d < 3; ++d) pos[d] = in[(size_t)(c->enc_begin[ge] + d) * n + q];
                        const int nc = level_corners<float>(c, g, l, pos, idx, w);
                        for (int k = 0; k < nc; ++k)
                            for (int f = 0; f < F; ++f)
                                gl[(size_t)idx[k] * F + f] += w[k] * denc[(size_t)(L.offset[ge] + l * F + f) * n + q];
                    }
                }
            });
        }
        for (auto& x : th) x.join();
    }
    // Adam over the parameters, in T contiguous ranges
    const size_t np = n_mlp + n_grid;
    {
        std::vector<std::thread> th;
        const size_t chunk = (np + T - 1) / T;
        for (int t = 0; t < T; ++t) {
            const size_t b = t * chunk, e = std::min(np, b + chunk);
            if (b >= e) break;
            th.emplace_back([&, b, e] {
                // the matrix/grid boundary relative to this range
                const size_t nm = n_mlp > b ? std::min(n_mlp, e) - b : 0;
                orc_adam_step(e - b, nm, loss_scale, lr, beta1, beta2, eps, l2_reg, params + b, grad.data() + b,
                              m + b, v + b, steps + b);
            });
        }
        for (auto& x : th) x.join();
    }
    double sum = 0;
    for (int q = 0; q < n; ++q) sum += loss[q];
    return (float)(sum / n);
}

// tcnn optimizers/ema.h:62-76 and :102-136.
void orc_ema_step(size_t n, float decay, uint32_t step, const float* weights, float* ema) {
    const float deb_old = 1 - (float)std::pow(decay, step - 1);
    const float deb_new = 1.0f / (1 - (float)std::pow(decay, step));
    for (size_t i = 0; i < n; ++i) ema[i] = (ema[i] * decay * deb_old + weights[i] * (1 - decay)) * deb_new;
}

// ema.h:62-76: weights and weights_ema are __half; each step reads both as
// float and stores the result as __half.
void orc_ema_step_half(size_t n, float decay, uint32_t step, const float* weights, float* ema) {
    const float deb_old = 1 - (float)std::pow(decay, step - 1);
    const float deb_new = 1.0f / (1 - (float)std::pow(decay, step));
    for (size_t i = 0; i < n; ++i)
        ema[i] = fm::round_half((ema[i] * decay * deb_old + fm::round_half(weights[i]) * (1 - decay)) * deb_new);
}

// tcnn trainer.h:53-92, gpu_matrix.h:291-306, grid.h:1333-1336, random.h:66-97.
void orc_init_params(const orc_cfg* c, uint32_t seed, int passes, float* params) {
    std::seed_seq seq{seed};
    std::vector<uint32_t> seeds(2);
    seq.generate(seeds.begin(), seeds.end());
    Pcg32 rng(seeds.front());
    const NetLayout N = net_layout(c);
    size_t n_mlp, n_grid;
    orc_n_params(c, &n_mlp, &n_grid);
    for (int p = 0; p < passes; ++p) {
        for (int l = 0; l < N.n_layers; ++l) {
            const float scale = std::sqrt(6.0f / (float)(N.in_w[l] + N.out_w[l]));
            const size_t ne = (size_t)N.in_w[l] * N.out_w[l];
            for (size_t i = 0; i < ne; ++i) params[N.off[l] + i] = rng.next_float() * 2.0f * scale - scale;
        }
        if (n_grid) {
            const size_t n_threads = (n_grid + 3) / 4;
            Pcg32 r = rng;
            float* out = params + n_mlp;
            for (size_t d = 0; d < 4 * n_threads; ++d) {
                const float val = r.next_float();
                const size_t idx = d / 4 + n_threads * (d % 4);
                if (idx < n_grid) out[idx] = fm::fma_(val, 1e-4f - -1e-4f, -1e-4f);
            }
            rng.advance(n_grid);
        }
    }
}

void orc_synth_queries(uint64_t seed, uint64_t first, int n, float* in, float* normal, float* u, float* eval_dirs) {
    for (int i = 0; i < n; ++i) {
        Pcg32 r(seed, first + (uint64_t)i);
        const Head h = synth_head(r);
        for (int d = 0; d < 9; ++d) in[(size_t)d * n + i] = h.in[d];
        const bool front = dot(h.n, h.wo) > 0.0f;
        normal[i] = front ? h.n.x : -h.n.x;
        normal[(size_t)n + i] = front ? h.n.y : -h.n.y;
        normal[2 * (size_t)n + i] = front ? h.n.z : -h.n.z;
        for (int d = 0; d < 3; ++d) u[(size_t)d * n + i] = r.next_float();
        for (int e = 0; e < 4; ++e) {
            const float e0 = r.next_float(), e1 = r.next_float();
            const V3<float> w = uniform_sample_sphere(e0, e1);
            eval_dirs[((size_t)e * 3 + 0) * n + i] = w.x;
            eval_dirs[((size_t)e * 3 + 1) * n + i] = w.y;
            eval_dirs[((size_t)e * 3 + 2) * n + i] = w.z;
        }
    }
}

void orc_synth_train(uint64_t seed, uint64_t first, int n, float* in, float* tgt_dir, float* tgt_rgb, float* tgt_pdf) {
    const Field F = make_field(seed);
    for (int i = 0; i < n; ++i) {
        Pcg32 r(seed, first + (uint64_t)i);
        const Head h = synth_head(r);
        for (int d = 0; d < 9; ++d) in[(size_t)d * n + i] = h.in[d];
        const V3<float> nf = dot(h.n, h.wo) > 0.0f ? h.n : v3(-h.n.x, -h.n.y, -h.n.z);
        const float c0 = r.next_float(), c1 = r.next_float();
        const float rr = fm::sqrt_(c0), ph = C<float>::two_pi * c1;
        const V3<float> loc = {rr * fm::cos(ph), rr * fm::sin(ph), fm::sqrt_(std::max(0.0f, 1.0f - c0))};
        const V3<float> w = frame_to_world(nf, loc);
        tgt_pdf[i] = std::max(dot(nf, w), 0.0f) * C<float>::inv_pi;
        float th, phi;
        cart_to_sph(w, &th, &phi);
        tgt_dir[i] = th;
        tgt_dir[(size_t)n + i] = phi;
        const float f = field_value<float>(F, h.in, w);
        tgt_rgb[i] = f;
        tgt_rgb[(size_t)n + i] = f;
        tgt_rgb[2 * (size_t)n + i] = f;
    }
}

void orc_field_eval(uint64_t seed, const float* pos, const float* dir, int n, double* f) {
    const Field F = make_field(seed);
    for (int i = 0; i < n; ++i) {
        const float p[3] = {pos[i], pos[(size_t)n + i], pos[2 * (size_t)n + i]};
        const V3<double> w = v3<double>(dir[i], dir[(size_t)n + i], dir[2 * (size_t)n + i]);
        f[i] = field_value<double>(F, p, w);
    }
}

void orc_pcg32(uint64_t initstate, uint64_t initseq, int n, uint32_t* out) {
    Pcg32 r(initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = r.next_uint();
}

void orc_math(int fn, const float* x, const float* y, int n, float* out) {
    for (int i = 0; i < n; ++i) {
        const float a = x[i], b = y ? y[i] : 0.0f;
        float r = 0;
        switch (fn) {
        case 0: r = fm::exp(a); break;
        case 1: r = fm::log(a); break;
        case 2: r = fm::log1p(a); break;
        case 3: r = fm::expm1(a); break;
        case 4: r = fm::sin(a); break;
        case 5: r = fm::cos(a); break;
        case 6: r = fm::acos(a); break;
        case 7: r = fm::atan2(a, b); break;
        case 8: r = fm::sqrt_(a); break;
        case 9: r = fm::logistic(a); break;
        case 10: r = a / b; break;
        case 12: r = fm::round_half(a); break;
        default: r = 0;
        }
        out[i] = r;
    }
}

}  // extern "C"
