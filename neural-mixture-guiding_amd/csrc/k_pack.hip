// k_pack.hip -- the wavefront packers on either side of the guiding path:
// shading points -> query records (generate_inference_data, train.cu:174-191)
// and recorded path vertices -> training records (generate_training_data,
// train.cu:14-67). The reference pushes training records with an atomic
// counter (train.h:117-124), so their order changes from run to run; here a
// count / scan / write sequence places them in (training pixel, depth) order,
// the same for every run.
#include "device.h"
#include "kernels.h"

namespace npmg {
namespace {

using namespace dev;

// normalizeSpatialCoord (train.h:173-177): the box inflated by 0.5 % of its
// diagonal's length, then 0.5 + (x - center) / diagonal per axis.
__device__ __forceinline__ void box_normalize(const SceneBox& b, const float* x, float* out) {
    const float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    const float inflation = fm::sqrt_(dx * dx + dy * dy + dz * dz) * 0.005f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float lo = b.lo[a] - inflation, hi = b.hi[a] + inflation;
        const float center = (lo + hi) * 0.5f, diag = hi - lo;
        out[a] = 0.5f + (x[a] - center) / diag;
    }
}

__global__ void __launch_bounds__(256) k_pack_queries(const float* __restrict__ pos, const float* __restrict__ wo,
                                                      const float* __restrict__ normal,
                                                      const float* __restrict__ roughness, const SceneBox box,
                                                      uint32_t n, float* __restrict__ in) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    const float x[3] = {pos[i], pos[N + i], pos[2 * N + i]};
    float p[3];
    box_normalize(box, x, p);
    const V3 w = normalize(mk(wo[i], wo[N + i], wo[2 * N + i]));  // sd.wo.normalized()
    float th, ph;
    cart_to_sph(mk(normal[i], normal[N + i], normal[2 * N + i]), &th, &ph);  // cartesianToSphericalNormalized
    const float v[9] = {p[0], p[1], p[2], (w.x + 1.0f) * 0.5f, (w.y + 1.0f) * 0.5f, (w.z + 1.0f) * 0.5f,
                        th / kPi, ph / kTwoPi, 1.0f - fm::exp(-roughness[i])};
#pragma unroll
    for (int d = 0; d < 9; ++d) in[(size_t)d * N + i] = v[d];
}

struct TrainRecord {
    float in[9], dir[2], rgb[3], pdf, bsdf_pdf;
};

// One recorded vertex -> one training record, and whether the reference keeps it.
__device__ __forceinline__ bool train_record(const PathRecords& r, const SceneBox& box, uint32_t pid, uint32_t d,
                                             TrainRecord& t) {
    const size_t P = r.n_pixels;
    const size_t s1 = (size_t)d * P + pid, s2 = (size_t)d * 2 * P + pid, s3 = (size_t)d * 3 * P + pid;
    if (r.delta[s1]) return false;
    const float x[3] = {r.pos[s3], r.pos[s3 + P], r.pos[s3 + 2 * P]};
    box_normalize(box, x, t.in);
    const V3 w = sph_to_cart(r.wo[s2], r.wo[s2 + P]);  // warp_direction_for_sh(sphericalToCartesian(wo))
    t.in[3] = (w.x + 1.0f) * 0.5f;
    t.in[4] = (w.y + 1.0f) * 0.5f;
    t.in[5] = (w.z + 1.0f) * 0.5f;
    t.in[6] = r.normal[s2] * kInvPi;
    t.in[7] = r.normal[s2 + P] * kInv2Pi;
    t.in[8] = 1.0f - fm::exp(-r.roughness[s1]);
    float L[3];
    bool nan_l = false;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float thp = r.thp[s3 + ch * P];
        L[ch] = thp > kEps ? r.L[s3 + ch * P] / thp : 0.0f;
        L[ch] = L[ch] * r.mis_weight[s1];
        nan_l |= __builtin_isnan(L[ch]);
        t.rgb[ch] = L[ch] * r.bsdf_val[s3 + ch * P];
    }
    t.dir[0] = r.dir[s2];
    t.dir[1] = r.dir[s2 + P];
    t.pdf = r.wi_pdf[s1];
    t.bsdf_pdf = r.bsdf_pdf ? r.bsdf_pdf[s1] : 0.0f;
    const bool bad = __builtin_isnan(t.in[0]) || __builtin_isnan(t.in[1]) || __builtin_isnan(t.in[2]) ||
                     __builtin_isnan(t.dir[0]) || __builtin_isnan(t.dir[1]) || __builtin_isnan(t.in[3]) ||
                     __builtin_isnan(t.in[4]) || __builtin_isnan(t.in[5]) || __builtin_isnan(t.pdf) ||
                     t.pdf == 0.0f || nan_l;
    return !bad;
}

// Pixel of training thread t, or false when it falls outside the image.
__device__ __forceinline__ bool train_pixel(const TrainTile& tt, uint32_t t, uint32_t* pid) {
    const uint32_t px = (t % tt.res[0]) * tt.stride[0] + tt.offset[0];
    const uint32_t py = (t / tt.res[0]) * tt.stride[1] + tt.offset[1];
    if (px >= tt.full[0] || py >= tt.full[1]) return false;
    *pid = px + py * tt.full[0];
    return true;
}

__device__ __forceinline__ uint32_t pixel_count(const PathRecords& r, const TrainTile& tt, const SceneBox& box,
                                                uint32_t t, uint32_t n_threads) {
    uint32_t pid;
    if (t >= n_threads || !train_pixel(tt, t, &pid)) return 0u;
    const uint32_t depth = min(r.depth[pid], r.max_depth);
    uint32_t c = 0;
    TrainRecord rec;
    for (uint32_t d = 0; d < depth; ++d) c += train_record(r, box, pid, d, rec) ? 1u : 0u;
    return c;
}

// Exclusive scan of v over the 256 threads of a block (4 waves of 64).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (int w = 0; w < wave; ++w) base += wsum[w];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return base + incl - v;
}

__global__ void __launch_bounds__(256) k_pack_train_count(const PathRecords r, const TrainTile tt,
                                                          const SceneBox box, uint32_t n_threads,
                                                          uint32_t* __restrict__ block_total) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t total;
    (void)block_exclusive_scan(pixel_count(r, tt, box, t, n_threads), &total);
    if (threadIdx.x == 0) block_total[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the block totals in place, grand total to *count.
__global__ void __launch_bounds__(256) k_pack_train_scan(uint32_t* __restrict__ block_total, uint32_t n_blocks,
                                                         uint32_t* __restrict__ count) {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < n_blocks; b0 += 256u) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < n_blocks ? block_total[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, &total);
        if (i < n_blocks) block_total[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *count = carry;
}

__global__ void __launch_bounds__(256) k_pack_train_write(const PathRecords r, const TrainTile tt, const SceneBox box,
                                                          uint32_t n_threads, const uint32_t* __restrict__ block_base,
                                                          uint32_t capacity, float* __restrict__ in,
                                                          float* __restrict__ tgt_dir, float* __restrict__ tgt_rgb,
                                                          float* __restrict__ tgt_pdf, float* __restrict__ tgt_bsdf) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint32_t total;
    uint32_t idx = block_base[blockIdx.x] + block_exclusive_scan(pixel_count(r, tt, box, t, n_threads), &total);
    uint32_t pid;
    if (t >= n_threads || !train_pixel(tt, t, &pid)) return;
    const uint32_t depth = min(r.depth[pid], r.max_depth);
    const size_t C = capacity;
    for (uint32_t d = 0; d < depth; ++d) {
        TrainRecord rec;
        if (!train_record(r, box, pid, d, rec)) continue;
        if (idx < capacity) {
#pragma unroll
            for (int k = 0; k < 9; ++k) in[(size_t)k * C + idx] = rec.in[k];
            tgt_dir[idx] = rec.dir[0];
            tgt_dir[C + idx] = rec.dir[1];
#pragma unroll
            for (int k = 0; k < 3; ++k) tgt_rgb[(size_t)k * C + idx] = rec.rgb[k];
            tgt_pdf[idx] = rec.pdf;
            if (tgt_bsdf) tgt_bsdf[idx] = rec.bsdf_pdf;
        }
        ++idx;
    }
}

}  // namespace

hipError_t launch_pack_queries(const float* pos, const float* wo, const float* normal, const float* roughness,
                               const SceneBox& box, uint32_t n, float* in, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_queries, dim3((n + 255u) / 256u), dim3(256), 0, s, pos, wo, normal, roughness, box, n,
                       in);
    return hipGetLastError();
}

size_t pack_training_scratch(const TrainTile& t) {
    const uint64_t n_threads = (uint64_t)t.res[0] * t.res[1];
    return (size_t)((n_threads + 255u) / 256u) + 1u;
}

hipError_t launch_pack_training(const PathRecords& r, const TrainTile& t, const SceneBox& box, uint32_t capacity,
                                float* in, float* tgt_dir, float* tgt_rgb, float* tgt_pdf, float* tgt_bsdf,
                                uint32_t* count, uint32_t* scratch, hipStream_t s) {
    const uint64_t nt = (uint64_t)t.res[0] * t.res[1];
    if (nt >= (1ull << 32)) return hipErrorInvalidValue;
    const uint32_t n_threads = (uint32_t)nt, n_blocks = (n_threads + 255u) / 256u;
    if (n_blocks)
        hipLaunchKernelGGL(k_pack_train_count, dim3(n_blocks), dim3(256), 0, s, r, t, box, n_threads, scratch);
    hipLaunchKernelGGL(k_pack_train_scan, dim3(1), dim3(256), 0, s, scratch, n_blocks, count);
    if (n_blocks)
        hipLaunchKernelGGL(k_pack_train_write, dim3(n_blocks), dim3(256), 0, s, r, t, box, n_threads, scratch,
                           capacity, in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf);
    return hipGetLastError();
}

}  // namespace npmg
