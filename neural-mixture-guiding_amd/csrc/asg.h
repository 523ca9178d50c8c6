// asg.h -- the anisotropic lobe mixture. The reference has only isotropic
// vMF lobes (distribution.h:75-202); this variant is this library's own.
//
// Each lobe is a vMF pushed through a linear map (a linearly transformed
// distribution): w = normalize(M w0), w0 ~ vMF(+z, kappa) in the lobe frame
// (t1, t2, z), M = diag(a, 1/a, 1). Its pdf is exact and normalised,
//   D(w) = vMF(w0'; kappa) / |M^-1 w_l|^3,  w0' = M^-1 w_l / |M^-1 w_l|,
//   w_l = (w.t1, w.t2, w.z), det M = 1,
// sampling is exact (draw w0, transform), and a = 1 is the vMF lobe.
// Six raw outputs per lobe, slots (theta, phi, psi, log kappa, log a, logit):
//   z = (sin th cos ph, sin th sin ph, cos th), th = pi s(x0), ph = 2 pi s(x1);
//   e1 = dz/dth, e2 = (-sin ph, cos ph, 0); t1 = cos psi e1 + sin psi e2,
//   t2 = -sin psi e1 + cos psi e2, psi = pi s(x2);
//   kappa = clamp(e^x3, 1e-2, 1e3); a = e^clamp(x4, -ln 8, ln 8);
//   weights = softmax(x5)
// (s the logistic function). The lobe evaluation is written once over a
// scalar type, so forward-mode dual numbers give the exact derivatives of
// the training adjoint. There is no cosine product for these lobes.
#pragma once

#include "device.h"
#include "kernels.h"

namespace npmg {
namespace dev {

constexpr float kAsgLogAMax = 2.07944154f;  // ln 8

// Dual number with N partial derivatives.
template <int N>
struct Dual {
    float v;
    float d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(float v) {
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.0f;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dchain(float v, float dv, const Dual<N>& a) {  // f(a) with f'(a) = dv
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = dv * a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(float s, const Dual<N>& a) {
    Dual<N> r;
    r.v = s * a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const float ib = 1.0f / b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}

// Scalar functions over float and Dual, same values for both.
__device__ __forceinline__ float s_exp(float x) { return fm::exp(x); }
__device__ __forceinline__ float s_log(float x) { return fm::log(x); }
__device__ __forceinline__ float s_expm1(float x) { return fm::expm1(x); }
__device__ __forceinline__ float s_sqrt(float x) { return fm::sqrt_(x); }
__device__ __forceinline__ float s_logistic(float x) { return 1.0f / (1.0f + fm::exp(-x)); }
__device__ __forceinline__ void s_sincos(float x, float* s, float* c) { fm::sincos(x, s, c); }
__device__ __forceinline__ float s_val(float x) { return x; }
template <int N> __device__ __forceinline__ float s_val(const Dual<N>& x) { return x.v; }
// A constant of the same type as x (zero derivatives).
__device__ __forceinline__ float s_lift(float, float v) { return v; }
template <int N> __device__ __forceinline__ Dual<N> s_lift(const Dual<N>&, float v) { return dconst<N>(v); }

template <int N> __device__ __forceinline__ Dual<N> s_exp(const Dual<N>& x) {
    const float e = fm::exp(x.v);
    return dchain(e, e, x);
}
template <int N> __device__ __forceinline__ Dual<N> s_log(const Dual<N>& x) { return dchain(fm::log(x.v), 1.0f / x.v, x); }
template <int N> __device__ __forceinline__ Dual<N> s_expm1(const Dual<N>& x) {
    return dchain(fm::expm1(x.v), fm::exp(x.v), x);
}
template <int N> __device__ __forceinline__ Dual<N> s_sqrt(const Dual<N>& x) {
    const float r = fm::sqrt_(x.v);
    return dchain(r, 0.5f / r, x);
}
template <int N> __device__ __forceinline__ Dual<N> s_logistic(const Dual<N>& x) {
    const float s = 1.0f / (1.0f + fm::exp(-x.v));
    return dchain(s, s * (1.0f - s), x);
}
template <int N> __device__ __forceinline__ void s_sincos(const Dual<N>& x, Dual<N>* s, Dual<N>* c) {
    float sv, cv;
    fm::sincos(x.v, &sv, &cv);
    *s = dchain(sv, cv, x);
    *c = dchain(cv, -sv, x);
}

// The scalar functions of the lobe code under a math policy: float values
// through MT (ExactMath or FastMath, device.h), dual numbers always exact.
template <class MT>
struct Sx {
    static __device__ __forceinline__ float exp(float x) { return MT::exp(x); }
    static __device__ __forceinline__ float log(float x) { return MT::log(x); }
    static __device__ __forceinline__ float expm1(float x) { return MT::expm1(x); }
    static __device__ __forceinline__ float sqrt(float x) { return MT::sqrt(x); }
    static __device__ __forceinline__ float logistic(float x) { return MT::div(1.0f, 1.0f + MT::exp(-x)); }
    static __device__ __forceinline__ void sincos(float x, float* s, float* c) { MT::sincos(x, s, c); }
    static __device__ __forceinline__ float div(float a, float b) { return MT::div(a, b); }
    template <int N> static __device__ __forceinline__ Dual<N> exp(const Dual<N>& x) { return s_exp(x); }
    template <int N> static __device__ __forceinline__ Dual<N> log(const Dual<N>& x) { return s_log(x); }
    template <int N> static __device__ __forceinline__ Dual<N> expm1(const Dual<N>& x) { return s_expm1(x); }
    template <int N> static __device__ __forceinline__ Dual<N> sqrt(const Dual<N>& x) { return s_sqrt(x); }
    template <int N> static __device__ __forceinline__ Dual<N> logistic(const Dual<N>& x) { return s_logistic(x); }
    template <int N> static __device__ __forceinline__ void sincos(const Dual<N>& x, Dual<N>* s, Dual<N>* c) {
        s_sincos(x, s, c);
    }
    template <int N> static __device__ __forceinline__ Dual<N> div(const Dual<N>& a, const Dual<N>& b) { return a / b; }
};

template <class S>
struct AsgLobe {
    S z[3], t1[3], t2[3], kappa, a, lognorm;
};

// Lobe from its five shape outputs r = (x0 .. x4); S = float or Dual<5>.
template <class S, class MT = ExactMath>
__device__ __forceinline__ AsgLobe<S> asg_build(const S (&r)[5]) {
    using F = Sx<MT>;
    AsgLobe<S> L;
    S sth, cth, sph, cph, sps, cps;
    F::sincos(kPi * F::logistic(r[0]), &sth, &cth);
    F::sincos(kTwoPi * F::logistic(r[1]), &sph, &cph);
    F::sincos(kPi * F::logistic(r[2]), &sps, &cps);
    L.z[0] = sth * cph;
    L.z[1] = sth * sph;
    L.z[2] = cth;
    const S e1[3] = {cth * cph, cth * sph, -sth};
    const S e2[3] = {-sph, cph, s_lift(sph, 0.0f)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        L.t1[i] = cps * e1[i] + sps * e2[i];
        L.t2[i] = cps * e2[i] - sps * e1[i];
    }
    // clamps: the clamped value with a zero derivative outside the range
    const S ek = F::exp(r[3]);
    const float kv = s_val(ek);
    L.kappa = ek;
    if (!(kv >= 1e-2f)) L.kappa = s_lift(ek, 1e-2f);
    if (kv > 1e3f) L.kappa = s_lift(ek, 1e3f);
    S la = r[4];
    if (s_val(la) > kAsgLogAMax) la = s_lift(la, kAsgLogAMax);
    if (s_val(la) < -kAsgLogAMax) la = s_lift(la, -kAsgLogAMax);
    L.a = F::exp(la);
    // log C(kappa) = log kappa + log(1/2pi) - log(-expm1(-2 kappa)), kappa >= 1e-2
    L.lognorm = F::log(L.kappa) + s_lift(la, fm::log(kInv2Pi)) - F::log(-F::expm1((-2.0f) * L.kappa));
    return L;
}

// D(w) of one lobe at a unit direction w.
template <class S, class MT = ExactMath>
__device__ __forceinline__ S asg_eval(const AsgLobe<S>& L, V3 w) {
    using F = Sx<MT>;
    const S xl = w.x * L.t1[0] + w.y * L.t1[1] + w.z * L.t1[2];
    const S yl = w.x * L.t2[0] + w.y * L.t2[1] + w.z * L.t2[2];
    const S zl = w.x * L.z[0] + w.y * L.z[1] + w.z * L.z[2];
    const S ux = F::div(xl, L.a), uy = yl * L.a;
    const S s2 = ux * ux + uy * uy + zl * zl;
    if constexpr (__is_same(S, float) && __is_same(MT, FastMath)) {
        // one reciprocal square root r = |u|^-1 gives the cosine and the
        // Jacobian |u|^-3 = r^3 (s2 lies in [1/64, 64] for a in [1/8, 8])
        const float r = __builtin_amdgcn_rsqf(s2);
        float cm1 = zl * r - 1.0f;
        if (cm1 > 0.0f) cm1 = 0.0f;
        return MT::exp(L.kappa * cm1 + L.lognorm) * (r * r * r);
    } else {
        S cm1 = F::div(zl, F::sqrt(s2)) - s_lift(zl, 1.0f);  // cos - 1 in the unstretched frame, at most 0
        if (s_val(cm1) > 0.0f) cm1 = s_lift(cm1, 0.0f);
        return F::exp(L.kappa * cm1 + L.lognorm - 1.5f * F::log(s2));
    }
}

// The shape outputs of lobe i as S, from a raw-output accessor.
template <class Raw>
__device__ __forceinline__ void asg_raw(Raw&& x, int i, float (&r)[5]) {
#pragma unroll
    for (int j = 0; j < 5; ++j) r[j] = x(6 * i + j);
}

// Mixture weights: softmax over the logits (slot 5 of each lobe).
template <int K, class MT = ExactMath, class Raw>
__device__ __forceinline__ void asg_weights(Raw&& x, float (&w)[K]) {
    float mx = x(5);
#pragma unroll
    for (int i = 1; i < K; ++i) mx = smax(mx, x(6 * i + 5));
    float total = 0.0f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        w[i] = MT::exp(x(6 * i + 5) - mx);
        total += w[i];
    }
#pragma unroll
    for (int i = 0; i < K; ++i) w[i] = MT::div(w[i], total);
}

// Lobe choice by running subtraction over the weights (the vMF mixture's rule,
// distribution.h:132-140), then the lobe's own direction.
template <int K, class MT = ExactMath, class Raw>
__device__ __forceinline__ V3 asg_sample(Raw&& x, const float (&w)[K], float u0, float u1, float u2,
                                         uint32_t* lobe) {
    int sel = -1;
    float u = u0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (sel < 0) {
            if (u < w[i]) sel = i;
            else u -= w[i];
        }
    }
    if (sel < 0) sel = 0;
    *lobe = (uint32_t)sel;
    float r[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        uint32_t b = fm::float_to_bits(x(j));
#pragma unroll
        for (int i = 1; i < K; ++i) {
            const uint32_t mask = 0u - (uint32_t)(i == sel);
            b = (b & ~mask) | (fm::float_to_bits(x(6 * i + j)) & mask);
        }
        r[j] = fm::bits_to_float(b);
    }
    const AsgLobe<float> L = asg_build<float, MT>(r);
    const float ct = 1.0f + MT::div(fm::log1p(u1 * fm::expm1(-2.0f * L.kappa)), L.kappa);
    const float st = MT::sqrt(smax(0.0f, 1.0f - ct * ct));
    float sp, cp;
    MT::sincos(kTwoPi * u2, &sp, &cp);
    const float vx = L.a * (cp * st), vy = MT::div(sp * st, L.a), vz = ct;
    const float inv = MT::div(1.0f, MT::sqrt(vx * vx + vy * vy + vz * vz));
    const float lx = vx * inv, ly = vy * inv, lz = vz * inv;
    return mk(lx * L.t1[0] + ly * L.t2[0] + lz * L.z[0], lx * L.t1[1] + ly * L.t2[1] + lz * L.z[1],
              lx * L.t1[2] + ly * L.t2[2] + lz * L.z[2]);
}

// Mixture pdf at up to E directions at once (bit e of `valid` set), one lobe
// built at a time, so only one lobe's state is live.
template <int K, int E, class MT = ExactMath, class Raw>
__device__ __forceinline__ void asg_pdf(Raw&& x, const float (&w)[K], const V3 (&dirs)[E], uint32_t valid,
                                        float (&p)[E]) {
#pragma unroll
    for (int e = 0; e < E; ++e) p[e] = 0.0f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        float r[5];
        asg_raw(x, i, r);
        const AsgLobe<float> L = asg_build<float, MT>(r);
#pragma unroll
        for (int e = 0; e < E; ++e)
            if ((valid >> e) & 1u) p[e] = fm::fma_(w[i], asg_eval<float, MT>(L, dirs[e]), p[e]);
    }
}

// One query's sample and pdfs from its raw outputs: direction, its pdf and
// the lobe when io.u is set, and the pdf at each evaluation direction (four
// per pass over the lobes; the sample direction joins the first pass).
template <int K, class MT, class Raw>
__device__ __forceinline__ void asg_query_m(Raw&& x, const QueryIO& io, uint32_t q, uint32_t n) {
    float w[K];
    asg_weights<K, MT>(x, w);
    V3 dirs[5];
    uint32_t valid = 0u;
    if (io.u) {
        uint32_t lobe;
        dirs[0] = asg_sample<K, MT>(x, w, io.u[q], io.u[(size_t)n + q], io.u[2 * (size_t)n + q], &lobe);
        io.out_dir[q] = dirs[0].x;
        io.out_dir[(size_t)n + q] = dirs[0].y;
        io.out_dir[2 * (size_t)n + q] = dirs[0].z;
        if (io.out_lobe) io.out_lobe[q] = lobe;
        valid = 1u;
    } else {
        dirs[0] = mk(0.0f, 0.0f, 1.0f);
    }
    for (int base = 0; base == 0 || base < io.n_eval; base += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = base + k;
            if (e < io.n_eval) {
                const float* ed = io.eval_dirs + (size_t)e * 3 * n;
                dirs[1 + k] = mk(ed[q], ed[(size_t)n + q], ed[2 * (size_t)n + q]);
                valid |= 2u << k;
            } else {
                dirs[1 + k] = mk(0.0f, 0.0f, 1.0f);
            }
        }
        float p[5];
        asg_pdf<K, 5, MT>(x, w, dirs, valid, p);
        if (valid & 1u) io.out_pdf[q] = p[0];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if ((valid >> (1 + k)) & 1u) io.out_eval_pdf[(size_t)(base + k) * n + q] = p[1 + k];
        valid = 0u;
    }
}

// asg_query_m with the math policy of the configuration's numerics.
template <int K, class Raw>
__device__ __forceinline__ void asg_query(const DevCfg& c, Raw&& x, const QueryIO& io, uint32_t q, uint32_t n) {
    if (c.numerics == NUM_FAST) asg_query_m<K, FastMath>(x, io, q, n);
    else asg_query_m<K, ExactMath>(x, io, q, n);
}

// Loss -(f/q) log(p + eps) and its adjoint for the anisotropic mixture (the
// training step of train.cu:69-124 with this lobe): dp/dlogit_s =
// w_s (D_s - p); the shape derivatives come from dual numbers.
template <int K, class Raw, class Out>
__device__ __forceinline__ float asg_adjoint(const DevCfg& c, Raw&& x, float th_t, float ph_t, float r, float g,
                                             float b, float wipdf, float ls_over_n, float clamp_hi, Out&& out) {
    const V3 wi = sph_to_cart(th_t, ph_t);
    float w[K], D[K];
    asg_weights<K>(x, w);
    float p = 0.0f;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        float rr[5];
        asg_raw(x, i, rr);
        D[i] = asg_eval<float>(asg_build<float>(rr), wi);
        p = fm::fma_(w[i], D[i], p);
    }
    const float Li = (r + g + b) / 3.0f;
    const float wiPdf = wipdf + kEps;
    const float guide = p + kEps;
    const float prefix = -Li / wiPdf / guide * ls_over_n;
    const float loss = -Li / wiPdf * fm::log(guide);
    auto finish = [&](int j, float v) {
        if (!fm::isfinite_(v)) v = 0.0f;
        if (clamp_hi > 0.0f) v = smin(clamp_hi, v);
        out(j, v);
    };
#pragma unroll
    for (int i = 0; i < K; ++i) {
        Dual<5> rd[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            rd[j] = dconst<5>(x(6 * i + j));
            rd[j].d[j] = 1.0f;
        }
        const Dual<5> Dd = asg_eval<Dual<5>>(asg_build<Dual<5>>(rd), wi);
#pragma unroll
        for (int j = 0; j < 5; ++j) finish(6 * i + j, prefix * w[i] * Dd.d[j]);
        finish(6 * i + 5, prefix * w[i] * (D[i] - p));
    }
    for (int j = 6 * K; j < c.n_out_pad; ++j) out(j, 0.0f);
    return loss;
}

// The loss adjoint of either mixture (P raw outputs per lobe). The vMF
// adjoint takes the math policy of the configuration's numerics (hardware
// transcendentals with "fast"); the anisotropic one stays exact.
template <int K, int P, class Raw, class Out>
__device__ __forceinline__ float lobe_adjoint(const DevCfg& c, Raw&& x, float th_t, float ph_t, float r, float g,
                                              float b, float wipdf, float bsdf_pdf, float ls_over_n, float clamp_hi,
                                              Out&& out) {
    if constexpr (P == 6) {
        return asg_adjoint<K>(c, x, th_t, ph_t, r, g, b, wipdf, ls_over_n, clamp_hi, out);
    } else {
        if (c.numerics == NUM_FAST)
            return mixture_adjoint<K, FastMath>(c, x, th_t, ph_t, r, g, b, wipdf, bsdf_pdf, ls_over_n, clamp_hi, out);
        return mixture_adjoint<K, ExactMath>(c, x, th_t, ph_t, r, g, b, wipdf, bsdf_pdf, ls_over_n, clamp_hi, out);
    }
}

}  // namespace dev
}  // namespace npmg
