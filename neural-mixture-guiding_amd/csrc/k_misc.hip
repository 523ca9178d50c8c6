// k_misc.hip -- synthetic wavefront records and the analytic radiance field,
// the device elementary-function probe, and small utility kernels.
//
// Synthetic records are keyed by (seed, query index): each query owns a pcg32
// stream (initstate = seed, initseq = index), so any subset can be
// regenerated on the host without copying.
#include "device.h"
#include "kernels.h"

namespace npmg {

namespace {

using namespace dev;

// Position, normal, outgoing direction and roughness of one record, packed
// like generate_inference_data (train.cu:174-191): pos in [0,1]^3,
// (w_o + 1) / 2, normal as (theta / pi, phi / 2 pi), 1 - exp(-roughness).
struct Head {
    float in[9];
    V3 n, wo;
};

__device__ __forceinline__ Head synth_head(Pcg32& r) {
    Head h;
    h.in[0] = r.next_float();
    h.in[1] = r.next_float();
    h.in[2] = r.next_float();
    float a0 = r.next_float(), a1 = r.next_float();
    h.n = uniform_sample_sphere(a0, a1);
    a0 = r.next_float();
    a1 = r.next_float();
    h.wo = uniform_sample_sphere(a0, a1);
    const float rough = r.next_float();
    h.in[3] = (h.wo.x + 1.0f) * 0.5f;
    h.in[4] = (h.wo.y + 1.0f) * 0.5f;
    h.in[5] = (h.wo.z + 1.0f) * 0.5f;
    float th, ph;
    cart_to_sph(h.n, &th, &ph);
    h.in[6] = th / kPi;
    h.in[7] = ph / kTwoPi;
    h.in[8] = 1.0f - fm::exp(-rough);
    return h;
}

__global__ void __launch_bounds__(256) k_synth_queries(uint64_t seed, uint64_t first, uint32_t n, float* in,
                                                       float* normal, float* u, float* eval_dirs) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    Pcg32 r(seed, first + i);
    const Head h = synth_head(r);
    for (int d = 0; d < 9; ++d) in[(size_t)d * n + i] = h.in[d];
    const bool front = dot(h.n, h.wo) > 0.0f;
    normal[i] = front ? h.n.x : -h.n.x;
    normal[(size_t)n + i] = front ? h.n.y : -h.n.y;
    normal[2 * (size_t)n + i] = front ? h.n.z : -h.n.z;
    for (int d = 0; d < 3; ++d) u[(size_t)d * n + i] = r.next_float();
    for (int e = 0; e < 4; ++e) {
        const float e0 = r.next_float(), e1 = r.next_float();
        const V3 w = uniform_sample_sphere(e0, e1);
        eval_dirs[((size_t)e * 3 + 0) * n + i] = w.x;
        eval_dirs[((size_t)e * 3 + 1) * n + i] = w.y;
        eval_dirs[((size_t)e * 3 + 2) * n + i] = w.z;
    }
}

// f(x, w) = sum_j a_j vMF(w; mu_j(x), kappa_j), mu_j rotated about a fixed
// axis by an angle linear in x (Rodrigues' formula).
__device__ __forceinline__ float field_value(const FieldParams& F, const float pos[3], V3 w) {
    float f = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float ang = F.d[j][0] * pos[0] + F.d[j][1] * pos[1] + F.d[j][2] * pos[2];
        float s, c;
        fm::sincos(ang, &s, &c);
        const V3 k = mk(F.axis[j][0], F.axis[j][1], F.axis[j][2]);
        const V3 v = mk(F.mu0[j][0], F.mu0[j][1], F.mu0[j][2]);
        const V3 kxv = cross(k, v);
        const float kdv = dot(k, v);
        const V3 mu = add(add(mul(v, c), mul(kxv, s)), mul(k, kdv * (1.0f - c)));
        const float kap = F.kappa[j];
        const float val = kap * fm::exp(kap * (dot(mu, w) - 1.0f)) / (kTwoPi * -fm::expm1(-2.0f * kap));
        f = fm::fma_(F.a[j], val, f);
    }
    return f;
}

// Training records: direction cosine-distributed about the face-forwarded
// normal, its pdf max(n.w, 0)/pi, and the field's radiance in all channels.
__global__ void __launch_bounds__(256) k_synth_train(uint64_t seed, uint64_t first, uint32_t n, const FieldParams F,
                                                     float* in, float* tgt_dir, float* tgt_rgb, float* tgt_pdf) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    Pcg32 r(seed, first + i);
    const Head h = synth_head(r);
    for (int d = 0; d < 9; ++d) in[(size_t)d * n + i] = h.in[d];
    const V3 nf = dot(h.n, h.wo) > 0.0f ? h.n : mk(-h.n.x, -h.n.y, -h.n.z);
    const float c0 = r.next_float(), c1 = r.next_float();
    const float rr = fm::sqrt_(c0);
    float sp, cp;
    fm::sincos(kTwoPi * c1, &sp, &cp);
    const V3 loc = mk(rr * cp, rr * sp, fm::sqrt_(smax(0.0f, 1.0f - c0)));
    const V3 w = frame_to_world(nf, loc);
    tgt_pdf[i] = smax(dot(nf, w), 0.0f) * kInvPi;
    float th, ph;
    cart_to_sph(w, &th, &ph);
    tgt_dir[i] = th;
    tgt_dir[(size_t)n + i] = ph;
    const float f = field_value(F, h.in, w);
    tgt_rgb[i] = f;
    tgt_rgb[(size_t)n + i] = f;
    tgt_rgb[2 * (size_t)n + i] = f;
}

__global__ void __launch_bounds__(256) k_debug_math(int fn, const float* x, const float* y, uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = y ? y[i] : 0.0f;
    float r = 0.0f;
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
    case 11: r = __builtin_amdgcn_cvt_f32_fp8(__builtin_amdgcn_cvt_pk_fp8_f32(a, 0.0f, 0, false), 0); break;
    case 12: r = fm::round_half(a); break;
    default: r = 0.0f;
    }
    out[i] = r;
}

// Host-side pcg32 for the field parameters (same recurrence as the device one).
struct HostPcg {
    uint64_t state, inc;
    HostPcg(uint64_t initstate, uint64_t initseq) {
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
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    float next_float() { return fm::bits_to_float((next_uint() >> 9) | 0x3f800000u) - 1.0f; }
};

void host_sphere(float u0, float u1, float out[3]) {
    const float z = 1.0f - 2.0f * u0;
    const float v = 1.0f - z * z;
    const float r = fm::sqrt_(0.0f < v ? v : 0.0f);
    float s, c;
    fm::sincos(kTwoPi * u1, &s, &c);
    out[0] = r * c;
    out[1] = r * s;
    out[2] = z;
}

}  // namespace

FieldParams make_field(uint64_t seed) {
    FieldParams F{};
    HostPcg r(seed ^ 0x9E3779B97F4A7C15ULL, 0x5eedULL);
    const float amp[3] = {0.6f, 0.3f, 0.1f};
    for (int j = 0; j < 3; ++j) {
        float u0 = r.next_float(), u1 = r.next_float();
        host_sphere(u0, u1, F.mu0[j]);
        u0 = r.next_float();
        u1 = r.next_float();
        host_sphere(u0, u1, F.axis[j]);
        for (int k = 0; k < 3; ++k) F.d[j][k] = (2.0f * r.next_float() - 1.0f) * kPi;
        F.kappa[j] = fm::exp(5.29831736654804f * r.next_float());  // log-uniform in [1, 200]
        F.a[j] = amp[j];
    }
    return F;
}

hipError_t launch_synth_queries(uint64_t seed, uint64_t first, uint32_t n, float* in, float* normal, float* u,
                                float* eval_dirs, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_synth_queries, dim3((n + 255) / 256), dim3(256), 0, s, seed, first, n, in, normal, u,
                       eval_dirs);
    return hipGetLastError();
}

hipError_t launch_synth_train(uint64_t seed, uint64_t first, uint32_t n, const FieldParams& f, float* in,
                              float* tgt_dir, float* tgt_rgb, float* tgt_pdf, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_synth_train, dim3((n + 255) / 256), dim3(256), 0, s, seed, first, n, f, in, tgt_dir, tgt_rgb,
                       tgt_pdf);
    return hipGetLastError();
}

hipError_t launch_debug_math(int fn, const float* x, const float* y, uint32_t n, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, s, fn, x, y, n, out);
    return hipGetLastError();
}

namespace {
// Grid-stride fill with 16-byte stores: the `head` floats before the first
// 16-byte boundary of p and the `tail` floats after the last whole float4 are
// stored singly by the first threads.
__global__ void __launch_bounds__(256) k_fill(float* p, float v, size_t n) {
    const size_t tid = (size_t)blockIdx.x * 256u + threadIdx.x, stride = (size_t)gridDim.x * 256u;
    const size_t to_boundary = (4u - (((uintptr_t)p >> 2) & 3u)) & 3u;
    const size_t head = to_boundary < n ? to_boundary : n;
    const size_t n4 = (n - head) / 4u, tail = n - head - 4u * n4;
    if (tid < head) p[tid] = v;
    if (tid < tail) p[head + 4u * n4 + tid] = v;
    float4* q = reinterpret_cast<float4*>(p + head);
    const float4 v4 = make_float4(v, v, v, v);
    for (size_t i = tid; i < n4; i += stride) q[i] = v4;
}
}  // namespace

// Fill kernel used instead of hipMemsetAsync so that the train step stays
// safe to capture in a HIP graph. At most 4096 workgroups (16 per CU), each
// thread storing every stride-th float4.
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = (n / 4u + 255u) / 256u;
    const unsigned grid = (unsigned)(blocks < 1u ? 1u : (blocks > 4096u ? 4096u : blocks));
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace npmg
