// k_validate.hip -- the FP64 validation kernels (the only FP64 work on the GPU):
//   * sphere quadrature of the production FP32 mixture pdf (with or without the
//     cosine product), accumulated in FP64: a Gauss-Legendre rule in z = cos
//     theta times the uniform rule in phi, one workgroup per query;
//   * central finite differences, in FP64, of the per-sample mixture-KL loss
//     -(f/q) log(p + 1e-5) (train.cu:86-93) with respect to the raw outputs,
//     from an FP64 restatement of the mapping and the pdf (distribution.h,
//     vmf.h) -- the check of the analytic adjoint.
#include <hip/hip_runtime.h>

#include "asg.h"
#include "device.h"
#include "kernels.h"

namespace npmg {
namespace {

using namespace dev;

template <int K>
__global__ void __launch_bounds__(256) k_pdf_integral(const DevCfg c, const float* __restrict__ raw,
                                                      const float* __restrict__ normal, const double* __restrict__ gz,
                                                      const double* __restrict__ gw, uint32_t n_z, uint32_t n_phi,
                                                      double* __restrict__ out, uint32_t n) {
    __shared__ double red[256];
    const uint32_t q = blockIdx.x;
    Mixture<K> m;
    build_mixture<K>(c, [&](int j) { return raw[(size_t)j * n + q]; }, m);
    if (normal) {
        float th[K], ph[K];
        apply_cosine<K>(c, mk(normal[q], normal[(size_t)n + q], normal[2 * (size_t)n + q]), m, false, th, ph);
    }
    const double dphi = 2.0 * 3.14159265358979323846 / (double)n_phi;
    double acc = 0.0;
    const uint64_t total = (uint64_t)n_z * n_phi;
    for (uint64_t i = threadIdx.x; i < total; i += 256u) {
        const uint32_t iz = (uint32_t)(i / n_phi), ip = (uint32_t)(i % n_phi);
        const double z = gz[iz], r = sqrt(fmax(0.0, 1.0 - z * z));
        const double ph = ((double)ip + 0.5) * dphi;
        const V3 w = mk((float)(r * cos(ph)), (float)(r * sin(ph)), (float)z);
        acc += gw[iz] * dphi * (double)mixture_pdf<K>(c, m, w);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// Anisotropic mixture: the lobes are built once per query into LDS, every
// quadrature point reads them back.
__global__ void __launch_bounds__(256) k_pdf_integral_asg(const DevCfg c, const float* __restrict__ raw,
                                                          const double* __restrict__ gz,
                                                          const double* __restrict__ gw, uint32_t n_z,
                                                          uint32_t n_phi, double* __restrict__ out, uint32_t n) {
    constexpr int K = 16;
    __shared__ AsgLobe<float> lobes[K];
    __shared__ float weight[K];
    __shared__ double red[256];
    const uint32_t q = blockIdx.x;
    auto x = [&](int j) { return raw[(size_t)j * n + q]; };
    if (threadIdx.x < (unsigned)K) {
        float r[5];
        asg_raw(x, (int)threadIdx.x, r);
        lobes[threadIdx.x] = asg_build<float>(r);
    }
    if (threadIdx.x == 0) {
        float w[K];
        asg_weights<K>(x, w);
        for (int i = 0; i < K; ++i) weight[i] = w[i];
    }
    __syncthreads();
    const double dphi = 2.0 * 3.14159265358979323846 / (double)n_phi;
    double acc = 0.0;
    const uint64_t total = (uint64_t)n_z * n_phi;
    for (uint64_t i = threadIdx.x; i < total; i += 256u) {
        const uint32_t iz = (uint32_t)(i / n_phi), ip = (uint32_t)(i % n_phi);
        const double z = gz[iz], r = sqrt(fmax(0.0, 1.0 - z * z));
        const double ph = ((double)ip + 0.5) * dphi;
        const V3 w = mk((float)(r * cos(ph)), (float)(r * sin(ph)), (float)z);
        float p = 0.0f;
        for (int k = 0; k < K; ++k) p = fm::fma_(weight[k], asg_eval<float>(lobes[k], w), p);
        acc += gw[iz] * dphi * (double)p;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
}

// FP64 restatement of the anisotropic mixture's per-sample loss (asg.h).
template <int K>
__device__ double asg_loss_fp64(const double* x, double th_t, double ph_t, double li, double wipdf) {
    const double pi = 3.14159265358979323846, ln8 = 2.0794415416798357;
    auto sig = [](double v) { return 1.0 / (1.0 + exp(-v)); };
    const double wx = sin(th_t) * cos(ph_t), wy = sin(th_t) * sin(ph_t), wz = cos(th_t);
    double mx = x[5];
    for (int i = 1; i < K; ++i) mx = fmax(mx, x[6 * i + 5]);
    double total = 0.0, lam[K];
    for (int i = 0; i < K; ++i) {
        lam[i] = exp(x[6 * i + 5] - mx);
        total += lam[i];
    }
    double p = 0.0;
    for (int i = 0; i < K; ++i) {
        const double th = pi * sig(x[6 * i]), ph = 2.0 * pi * sig(x[6 * i + 1]), ps = pi * sig(x[6 * i + 2]);
        const double kap = fmin(fmax(exp(x[6 * i + 3]), 1e-2), 1e3);
        const double a = exp(fmin(fmax(x[6 * i + 4], -ln8), ln8));
        const double z[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
        const double e1[3] = {cos(th) * cos(ph), cos(th) * sin(ph), -sin(th)};
        const double e2[3] = {-sin(ph), cos(ph), 0.0};
        double t1[3], t2[3];
        for (int k = 0; k < 3; ++k) {
            t1[k] = cos(ps) * e1[k] + sin(ps) * e2[k];
            t2[k] = cos(ps) * e2[k] - sin(ps) * e1[k];
        }
        const double xl = wx * t1[0] + wy * t1[1] + wz * t1[2];
        const double yl = wx * t2[0] + wy * t2[1] + wz * t2[2];
        const double zl = wx * z[0] + wy * z[1] + wz * z[2];
        const double s2 = (xl / a) * (xl / a) + (yl * a) * (yl * a) + zl * zl;
        const double cm1 = fmin(0.0, zl / sqrt(s2) - 1.0);
        const double lognorm = log(kap / (2.0 * pi * -expm1(-2.0 * kap)));
        p += lam[i] / total * exp(kap * cm1 + lognorm - 1.5 * log(s2));
    }
    return -li / (wipdf + 1e-5) * log(p + 1e-5);
}

// FP64 restatement of the per-sample loss (both mappings and both numerics,
// matching build_mixture / lobe_pdf / mixture_adjoint in their exact forms).
template <int K>
__device__ double loss_fp64(const DevCfg& c, const double* x, double th_t, double ph_t, double li, double wipdf) {
    const double pi = 3.14159265358979323846;
    const bool ref = c.numerics == NUM_REFERENCE;
    double lam[K], kap[K], th[K], ph[K], total = 0.0;
    auto sig = [](double v) { return 1.0 / (1.0 + exp(-v)); };
    auto clampd = [](double v, double lo, double hi) { return fmin(fmax(v, lo), hi); };
    if (c.mapping == MAP_REFERENCE) {
        for (int i = 0; i < K; ++i) {
            lam[i] = exp(clampd(x[4 * i + 0], -10.0, 15.0));
            kap[i] = exp(clampd(x[4 * i + 1], -10.0, 15.0));
            th[i] = pi * sig(x[4 * i + 2]);
            ph[i] = 2.0 * pi * sig(x[4 * i + 3]);
            total += lam[i];
        }
    } else {
        double mx = x[3];
        for (int i = 1; i < K; ++i) mx = fmax(mx, x[4 * i + 3]);
        for (int i = 0; i < K; ++i) {
            th[i] = pi * sig(x[4 * i + 0]);
            ph[i] = 2.0 * pi * sig(x[4 * i + 1]);
            kap[i] = clampd(exp(x[4 * i + 2]), 1e-2, 1e3);
            lam[i] = exp(x[4 * i + 3] - mx);
            total += lam[i];
        }
    }
    const double wx = sin(th_t) * cos(ph_t), wy = sin(th_t) * sin(ph_t), wz = cos(th_t);
    double p = 0.0;
    for (int i = 0; i < K; ++i) {
        const double cth = sin(th[i]) * cos(ph[i]) * wx + sin(th[i]) * sin(ph[i]) * wy + cos(th[i]) * wz;
        const double d = fmin(0.0, cth - 1.0), k = kap[i];
        double v;
        if (ref) v = k < 1e-5 ? 1.0 / (4.0 * pi) : exp(k * d) * k / (2.0 * pi * (1.0 - exp(-2.0 * k)));
        else v = k > 1e-30 ? k / (2.0 * pi * -expm1(-2.0 * k)) * exp(k * d) : 1.0 / (4.0 * pi);
        p += lam[i] / total * v;
    }
    return -li / (wipdf + 1e-5) * log(p + 1e-5);
}

// Thread (q, j): (L(x + h e_j) - L(x - h e_j)) / 2h.
template <int K, int P>
__global__ void __launch_bounds__(256) k_loss_fd(const DevCfg c, const float* __restrict__ raw,
                                                 const float* __restrict__ tgt_dir, const float* __restrict__ tgt_rgb,
                                                 const float* __restrict__ tgt_pdf, double h, double* __restrict__ out,
                                                 uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const int j = blockIdx.y;
    if (q >= n) return;
    if (j >= P * K) {
        out[(size_t)j * n + q] = 0.0;
        return;
    }
    double x[P * K];
    for (int i = 0; i < P * K; ++i) x[i] = (double)raw[(size_t)i * n + q];
    const double th = tgt_dir[q], ph = tgt_dir[(size_t)n + q];
    const double li = ((double)tgt_rgb[q] + (double)tgt_rgb[(size_t)n + q] + (double)tgt_rgb[2 * (size_t)n + q]) / 3.0;
    const double qp = tgt_pdf[q];
    const double x0 = x[j];
    auto loss = [&]() {
        if constexpr (P == 6) return asg_loss_fp64<K>(x, th, ph, li, qp);
        else return loss_fp64<K>(c, x, th, ph, li, qp);
    };
    x[j] = x0 + h;
    const double lp = loss();
    x[j] = x0 - h;
    const double lm = loss();
    out[(size_t)j * n + q] = (lp - lm) / (2.0 * h);
}

}  // namespace

hipError_t launch_pdf_integral(const DevCfg& c, const float* raw, const float* normal, const double* gz,
                               const double* gw, uint32_t n_z, uint32_t n_phi, double* out, uint32_t n,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (c.lobe_params == 6) {
        if (normal) return hipErrorInvalidValue;  // no cosine product for anisotropic lobes
        hipLaunchKernelGGL(k_pdf_integral_asg, dim3(n), dim3(256), 0, s, c, raw, gz, gw, n_z, n_phi, out, n);
    } else if (c.K == 8) hipLaunchKernelGGL((k_pdf_integral<8>), dim3(n), dim3(256), 0, s, c, raw, normal, gz, gw, n_z, n_phi, out, n);
    else if (c.K == 16) hipLaunchKernelGGL((k_pdf_integral<16>), dim3(n), dim3(256), 0, s, c, raw, normal, gz, gw, n_z, n_phi, out, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_loss_fd(const DevCfg& c, const float* raw, const float* td, const float* tr, const float* tp,
                          double h, double* out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256, c.n_out_pad);
    if (c.lobe_params == 6) hipLaunchKernelGGL((k_loss_fd<16, 6>), grid, dim3(256), 0, s, c, raw, td, tr, tp, h, out, n);
    else if (c.K == 8) hipLaunchKernelGGL((k_loss_fd<8, 4>), grid, dim3(256), 0, s, c, raw, td, tr, tp, h, out, n);
    else if (c.K == 16) hipLaunchKernelGGL((k_loss_fd<16, 4>), grid, dim3(256), 0, s, c, raw, td, tr, tp, h, out, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace npmg
