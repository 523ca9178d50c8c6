// k_scatter.hip -- grid-gradient scatter, the adjoint of the grid encoding
// (tiny-cuda-nn kernel_grid_backward, encodings/grid.h:386-509).
//
// Atomic mode: grid (ceil(n / 256), n_levels); blocks are dispatched in
// linear order, so all samples of one level are scattered before the next
// level starts and the level's gradient slice (at most 2^19 x F floats) stays
// resident in L2 while it is hit.
//
// Deterministic mode: every (sample, level, corner) becomes a pair (global
// table entry, pair index) plus its F contributions; a stable LSD radix sort
// (rocPRIM) groups the pairs by entry keeping the pair order (level, sample,
// corner); each run is then summed in that order by one thread. The result is
// independent of scheduling.
#include <rocprim/device/device_radix_sort.hpp>

#include "device.h"
#include "kernels.h"

namespace npmg {
namespace {

using namespace dev;

__global__ void __launch_bounds__(256) k_scatter_atomic(const DevCfg c, const float* __restrict__ in,
                                                        const float* __restrict__ dgrid, float* __restrict__ grad,
                                                        uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const int l = blockIdx.y;
    if (q >= n) return;
    const int b = c.enc_begin[c.grid_enc];
    const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
    uint32_t idx[8];
    float w[8];
    const int nc = level_corners(c, l, x0, x1, x2, idx, w);
    const int F = c.F;
    float* gl = grad + (size_t)c.offsets[l] * F;
    for (int f = 0; f < F; ++f) {
        const float d = dgrid[(size_t)(l * F + f) * n + q];
        if (d == 0.0f) continue;
        for (int k = 0; k < nc; ++k) atomicAdd(gl + (size_t)idx[k] * F + f, w[k] * d);
    }
}

constexpr uint32_t kNoEntry = 0xffffffffu;

// Pair p = (l n + q) 8 + k.
__global__ void __launch_bounds__(256) k_scatter_pairs(const DevCfg c, const float* __restrict__ in,
                                                       const float* __restrict__ dgrid, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals, float* __restrict__ contrib,
                                                       uint32_t n) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const int l = blockIdx.y;
    if (q >= n) return;
    const int b = c.enc_begin[c.grid_enc];
    const float x0 = in[(size_t)b * n + q], x1 = in[(size_t)(b + 1) * n + q], x2 = in[(size_t)(b + 2) * n + q];
    uint32_t idx[8];
    float w[8];
    const int nc = level_corners(c, l, x0, x1, x2, idx, w);
    const int F = c.F;
    const size_t p0 = ((size_t)l * n + q) * 8;
    for (int k = 0; k < 8; ++k) {
        keys[p0 + k] = k < nc ? c.offsets[l] + idx[k] : kNoEntry;
        vals[p0 + k] = (uint32_t)(p0 + k);
        for (int f = 0; f < F; ++f)
            contrib[(p0 + k) * F + f] = k < nc ? w[k] * dgrid[(size_t)(l * F + f) * n + q] : 0.0f;
    }
}

__global__ void __launch_bounds__(256) k_scatter_runs(const DevCfg c, const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ vals,
                                                      const float* __restrict__ contrib, float* __restrict__ grad,
                                                      size_t n_pairs) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n_pairs) return;
    const uint32_t key = keys[i];
    if (key == kNoEntry || (i > 0 && keys[i - 1] == key)) return;
    const int F = c.F;
    float acc[8] = {};
    for (size_t j = i; j < n_pairs && keys[j] == key; ++j) {
        const size_t p = vals[j];
        for (int f = 0; f < F; ++f) acc[f] += contrib[p * F + f];
    }
    for (int f = 0; f < F; ++f) grad[(size_t)key * F + f] = acc[f];
}

int key_bits(const DevCfg& c) {
    const uint32_t total = c.offsets[c.n_levels];
    int bits = 1;
    while (bits < 32 && (1ull << bits) <= (unsigned long long)total) ++bits;
    return bits;
}

}  // namespace

hipError_t launch_grid_scatter_atomic(const DevCfg& c, const float* in, const float* dgrid, float* grad, uint32_t n,
                                      hipStream_t s) {
    if (n == 0 || c.grid_enc < 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_atomic, dim3((n + 255) / 256, c.n_levels), dim3(256), 0, s, c, in, dgrid, grad, n);
    return hipGetLastError();
}

// Scratch layout: keys_in, keys_out, vals_in, vals_out (n_pairs u32 each),
// contributions (n_pairs F floats), then the sort's temporary storage.
size_t grid_scatter_sorted_bytes(const DevCfg& c, uint32_t n) {
    if (c.grid_enc < 0) return 0;
    const size_t np = (size_t)n * c.n_levels * 8;
    size_t tmp = 0;
    (void)rocprim::radix_sort_pairs((void*)nullptr, tmp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                              (uint32_t*)nullptr, np, 0, key_bits(c) + 1, (hipStream_t)0);
    return np * 4 * 4 + np * c.F * 4 + ((tmp + 255) / 256) * 256 + 1024;
}

hipError_t launch_grid_scatter_sorted(const DevCfg& c, const float* in, const float* dgrid, float* grad, uint32_t n,
                                      void* scratch, size_t scratch_bytes, hipStream_t s) {
    if (n == 0 || c.grid_enc < 0) return hipSuccess;
    const size_t np = (size_t)n * c.n_levels * 8;
    if (np >= 0xffffffffull) return hipErrorInvalidValue;
    if (scratch_bytes < grid_scatter_sorted_bytes(c, n)) return hipErrorInvalidValue;
    char* p = static_cast<char*>(scratch);
    auto carve = [&](size_t bytes) { char* r = p; p += (bytes + 255) / 256 * 256; return r; };
    uint32_t* k_in = reinterpret_cast<uint32_t*>(carve(np * 4));
    uint32_t* k_out = reinterpret_cast<uint32_t*>(carve(np * 4));
    uint32_t* v_in = reinterpret_cast<uint32_t*>(carve(np * 4));
    uint32_t* v_out = reinterpret_cast<uint32_t*>(carve(np * 4));
    float* contrib = reinterpret_cast<float*>(carve(np * c.F * 4));
    size_t tmp = 0;
    // bit key_bits(c) distinguishes the sentinel 0xffffffff from every entry
    const unsigned end_bit = (unsigned)key_bits(c) + 1;
    hipError_t e = rocprim::radix_sort_pairs((void*)nullptr, tmp, k_in, k_out, v_in, v_out, np, 0u, end_bit, s);
    if (e != hipSuccess) return e;
    void* sort_tmp = carve(tmp);
    if ((size_t)(p - static_cast<char*>(scratch)) > scratch_bytes) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scatter_pairs, dim3((n + 255) / 256, c.n_levels), dim3(256), 0, s, c, in, dgrid, k_in, v_in,
                       contrib, n);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(sort_tmp, tmp, k_in, k_out, v_in, v_out, np, 0u, end_bit, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter_runs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, c, k_out, v_out, contrib,
                       grad, np);
    return hipGetLastError();
}

}  // namespace npmg
