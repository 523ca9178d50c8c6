// k_scatter.hip -- grid-gradient scatter, the adjoint of the grid encoding
// (tiny-cuda-nn kernel_grid_backward, encodings/grid.h:386-509), without
// global atomics on the hot path.
//
// Every (sample, level, corner) contributes w_k * dL/dy_{l,f} to one table
// entry. The contributions are binned by (level, tile of 8192 / F entries),
// so that one tile of gradients fits in LDS:
//   1. count    per (bin, block of 512 samples), histogram in LDS;
//   2. scan     per bin over its blocks, then over bins: each (bin, block)
//               owns a contiguous run of records, in block order; the bins
//               are cut into chunks of at most kChunk records;
//   3. place    the records (local word index, F contributions; 1 + F
//               words) at their slots (slot order inside a (bin, block) run
//               is arbitrary), and max |dL/dy| of each level. For F = 2 the
//               block's records are first sorted into their runs in LDS and
//               then written run by run with consecutive lanes on
//               consecutive records, instead of three scattered 4-byte stores
//               per record;
// Steps 1-2 read only the sample positions, so the training step runs them
// (launch_grid_scatter_prepare) beside the forward pass, and only 3-4 wait
// for dL/dy (launch_grid_scatter_finish).
//   4. reduce   one workgroup per chunk accumulates its records into the
//               tile in LDS, then writes the tile: plain stores when the bin
//               has one chunk (the usual case at fine levels), atomics for
//               the few coarse bins with several chunks.
// The LDS accumulators are int64 fixed point with a per-level scale 2^S chosen
// from max |dL/dy| so that no sum can overflow (|sum| <= 8 n max < 2^62): the
// sums are exact, so the order of the LDS atomics does not matter, and the
// resolution is 2^-(62 - log2(8 n)) max |dL/dy| (2^-41 of the level's largest
// contribution at n = 2^18). Each contribution is rounded to the nearest step
// (to_fixed), so one below half a step -- under 2^-42 of the level's largest
// |dL/dy| (rounded up to a power of two) at n = 2^18 -- flushes to zero, in
// every order. An FP32 atomic sum like the reference's drops a term once the
// running sum exceeds 2^25 times it, so the fixed-point sum keeps every term
// that FP32 accumulation would keep in practice. Integer LDS atomics are also several times faster
// than FP32 ones on gfx950. The modes differ only for the coarse bins that
// span several chunks: atomic mode adds each chunk's converted tile to the
// gradient with FP32 global atomics (order-dependent rounding, like the
// reference's atomics); deterministic mode adds the int64 tiles into an int64
// buffer that a final pass converts (and clears for the next call), so every
// schedule gives the same bits.
#include <algorithm>

#include "device.h"
#include "kernels.h"

namespace npmg {
namespace {

using namespace dev;

constexpr uint32_t kSamplesPerBlock = 512;
constexpr uint32_t kTileWords = 8192;  // words of one tile: tile entries x F
constexpr uint32_t kChunk = 65536;     // records reduced by one workgroup
constexpr uint32_t kMaxBinsPerLevel = 8192;

struct BinLayout {
    uint32_t bin_off[33];  // first bin of each level, bin_off[L] = number of bins
    uint32_t shift;        // log2(tile entries)
    uint32_t nblk;         // sample blocks
    uint32_t max_chunks;   // capacity of the chunk table
    uint32_t bin_off_max;  // bins of the largest level
};

struct Scratch {
    uint32_t* counts;      // [bins][nblk] counts, then exclusive offsets within the bin
    uint32_t* bin_start;   // [bins] first record of the bin
    uint32_t* bin_chunks;  // [bins]
    uint32_t* bin_count;   // [bins] records of the bin
    uint4* chunks;         // [max_chunks] (bin, begin, end, chunks of the bin); chunks[-1].x = count
    uint32_t* maxbits;     // [L] max |dL/dy| of each level (float bits)
    uint32_t* rec;         // [R][1 + F] records: word index within the tile, then F contributions
    unsigned long long* wide;  // [n_grid F] int64 accumulators of multi-chunk bins (deterministic)
};

__device__ __forceinline__ uint32_t level_of_bin(const BinLayout& b, int L, uint32_t bin) {
    int l = 0;
    while (l + 1 < L && bin >= b.bin_off[l + 1]) ++l;
    return (uint32_t)l;
}

// Fixed-point scale exponent of level l: |sum| <= n_rec max < 2^62.
__device__ __forceinline__ int fixed_shift(uint32_t maxbits, uint32_t n_rec) {
    if (maxbits == 0u) return 0;
    const int e = (int)((maxbits >> 23) & 0xffu) - 126;  // max < 2^e
    int r = 0;
    while (r < 32 && (1ull << r) < (unsigned long long)n_rec) ++r;  // n_rec <= 2^r
    return 62 - r - e;
}

__device__ __forceinline__ long long to_fixed(float v, int s) { return __float2ll_rn(ldexpf(v, s)); }

constexpr uint32_t kPerThread = kSamplesPerBlock / 256u;

// Positions of the thread's samples (sample i: block base + 256 i + thread), all
// loads issued before any use; samples past n read sample n - 1.
__device__ __forceinline__ void load_positions(const float* __restrict__ in, int g, uint32_t n, float* px, float* py,
                                               float* pz) {
#pragma unroll
    for (int i = 0; i < (int)kPerThread; ++i) {
        uint32_t q = blockIdx.x * kSamplesPerBlock + (uint32_t)i * 256u + threadIdx.x;
        q = q < n ? q : n - 1u;
        px[i] = in[(size_t)g * n + q];
        py[i] = in[(size_t)(g + 1) * n + q];
        pz[i] = in[(size_t)(g + 2) * n + q];
    }
}

__global__ void __launch_bounds__(256) k_bin_count(const DevCfg c, const BinLayout b, const float* __restrict__ in,
                                                   Scratch sc, uint32_t n) {
    extern __shared__ uint32_t hist[];
    const int l = blockIdx.y;
    const uint32_t nb = b.bin_off[l + 1] - b.bin_off[l];
    for (uint32_t i = threadIdx.x; i < nb; i += 256u) hist[i] = 0u;
    __syncthreads();
    const int g = c.enc_begin[c.grid_enc];
    float px[kPerThread], py[kPerThread], pz[kPerThread];
    load_positions(in, g, n, px, py, pz);
#pragma unroll
    for (int i = 0; i < (int)kPerThread; ++i) {
        const uint32_t q = blockIdx.x * kSamplesPerBlock + (uint32_t)i * 256u + threadIdx.x;
        if (q >= n) break;
        uint32_t idx[8];
        float w[8];
        const int nc = level_corners(c, l, px[i], py[i], pz[i], idx, w);
        for (int k = 0; k < nc; ++k) atomicAdd(&hist[idx[k] >> b.shift], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += 256u) sc.counts[(size_t)(b.bin_off[l] + i) * b.nblk + blockIdx.x] = hist[i];
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the total.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* red, uint32_t& excl) {
    const uint32_t t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < 256u; d <<= 1) {
        const uint32_t o = t >= d ? red[t - d] : 0u;
        __syncthreads();
        red[t] += o;
        __syncthreads();
    }
    excl = red[t] - v;
    const uint32_t total = red[255];
    __syncthreads();
    return total;
}

// Exclusive scan of each bin's per-block counts; the bin total goes to bin_start.
__global__ void __launch_bounds__(256) k_bin_scan(const BinLayout b, Scratch sc) {
    __shared__ uint32_t red[256];
    uint32_t* row = sc.counts + (size_t)blockIdx.x * b.nblk;
    const uint32_t per = (b.nblk + 255u) / 256u;
    const uint32_t lo = threadIdx.x * per, hi = lo + per < b.nblk ? lo + per : b.nblk;
    uint32_t sum = 0u;
    for (uint32_t i = lo; i < hi; ++i) sum += row[i];
    uint32_t run;
    const uint32_t total = block_scan(sum, red, run);
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) sc.bin_start[blockIdx.x] = total;
}

// One workgroup: bin totals -> bin starts, chunks per bin, and the chunk table.
__global__ void __launch_bounds__(256) k_bin_chunks(const BinLayout b, uint32_t n_bins, Scratch sc) {
    __shared__ uint32_t red[256];
    const uint32_t per = (n_bins + 255u) / 256u;
    const uint32_t lo = threadIdx.x * per, hi = lo + per < n_bins ? lo + per : n_bins;
    uint32_t s_rec = 0u, s_ch = 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t t = sc.bin_start[i];
        s_rec += t;
        s_ch += (t + kChunk - 1u) / kChunk;
    }
    uint32_t rec, ch;
    block_scan(s_rec, red, rec);
    const uint32_t total_ch = block_scan(s_ch, red, ch);
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t t = sc.bin_start[i];
        const uint32_t k = (t + kChunk - 1u) / kChunk;
        sc.bin_start[i] = rec;
        sc.bin_chunks[i] = k;
        sc.bin_count[i] = t;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t be = rec + j * kChunk;
            const uint32_t en = be + kChunk < rec + t ? be + kChunk : rec + t;
            if (ch + j < b.max_chunks) sc.chunks[ch + j] = make_uint4(i, be, en, k);
        }
        rec += t;
        ch += k;
    }
    if (threadIdx.x == 0) sc.chunks[-1] = make_uint4(total_ch < b.max_chunks ? total_ch : b.max_chunks, 0u, 0u, 0u);
}

// Contribution of one record: w dL/dy, or zero for a non-finite dL/dy (the
// overflow flag then skips the step under dynamic loss scaling).
__device__ __forceinline__ float contrib(float w, float d) { return __builtin_isfinite(d) ? w * d : 0.0f; }

// Place for F != 2: each record stored where its bin's cursor points.
__global__ void __launch_bounds__(256) k_bin_place(const DevCfg c, const BinLayout b, const float* __restrict__ in,
                                                   const float* __restrict__ dgrid, Scratch sc, uint32_t n,
                                                   uint32_t* __restrict__ overflow) {
    extern __shared__ uint32_t cursor[];
    __shared__ float wave_max[4];
    const int l = blockIdx.y;
    const uint32_t b0 = b.bin_off[l], nb = b.bin_off[l + 1] - b0;
    for (uint32_t i = threadIdx.x; i < nb; i += 256u)
        cursor[i] = sc.bin_start[b0 + i] + sc.counts[(size_t)(b0 + i) * b.nblk + blockIdx.x];
    __syncthreads();
    const int g = c.enc_begin[c.grid_enc];
    const int F = c.F;
    const uint32_t mask = (1u << b.shift) - 1u;
    float mx = 0.0f;
    bool bad = false;
    auto seen = [&](float d) {
        if (__builtin_isfinite(d)) mx = smax(mx, fabsf(d));
        else bad = true;
    };
    float px[kPerThread], py[kPerThread], pz[kPerThread];
    load_positions(in, g, n, px, py, pz);
#pragma unroll
    for (int i = 0; i < (int)kPerThread; ++i) {
        const uint32_t q = blockIdx.x * kSamplesPerBlock + (uint32_t)i * 256u + threadIdx.x;
        if (q >= n) break;
        uint32_t idx[8];
        float w[8];
        const int nc = level_corners(c, l, px[i], py[i], pz[i], idx, w);
        for (int f = 0; f < F; ++f) seen(dgrid[(size_t)(l * F + f) * n + q]);
        for (int k = 0; k < nc; ++k) {
            uint32_t* r = sc.rec + (size_t)atomicAdd(&cursor[idx[k] >> b.shift], 1u) * (1u + F);
            r[0] = (idx[k] & mask) * (uint32_t)F;
            for (int f = 0; f < F; ++f)
                r[1 + f] = __float_as_uint(contrib(w[k], dgrid[(size_t)(l * F + f) * n + q]));
        }
    }
    // max |dL/dy| of the block, then one global atomic per block (bits of
    // non-negative floats order like the values); the reduce derives the
    // level's fixed-point scale from it
    for (int o = 32; o > 0; o >>= 1) mx = smax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63u) == 0) wave_max[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(&sc.maxbits[l], __float_as_uint(smax(smax(wave_max[0], wave_max[1]), smax(wave_max[2], wave_max[3]))));
    if (bad && overflow) atomicOr(overflow, 1u);
}

// Records of one block for F = 2 (at most 8 per sample).
constexpr uint32_t kStage = kSamplesPerBlock * 8u;

// Place for F = 2, staged: the block's count in each bin follows from the
// scanned offsets (the next block's offset, or the bin's total), so the
// block's records can be sorted into their runs in LDS as they are made:
// position loc[bin] + arrival within the bin. Position r of the staged block
// then goes to global slot run[bin] + r - loc[bin], so consecutive lanes
// write consecutive records of a run.
__global__ void __launch_bounds__(256) k_bin_place2(const DevCfg c, const BinLayout b, const float* __restrict__ in,
                                                    const float* __restrict__ dgrid, Scratch sc, uint32_t n,
                                                    uint32_t* __restrict__ overflow) {
    extern __shared__ uint32_t sm[];
    __shared__ float wave_max[4];
    __shared__ uint32_t red[256];
    const int l = blockIdx.y;
    const uint32_t b0 = b.bin_off[l], nb = b.bin_off[l + 1] - b0;
    const uint32_t nbp = (b.bin_off_max + 3u) & ~3u;  // the arrays' stride: the largest level's bins
    uint32_t* cur = sm;             // [nb] local cursor of each bin, from loc[bin]
    uint32_t* off = sm + nbp;       // [nb] run[bin] - loc[bin]
    uint32_t* tw = sm + 2u * nbp;   // [kStage] bin << 16 | word index
    float* s0 = reinterpret_cast<float*>(tw + kStage);
    float* s1 = s0 + kStage;
    const uint32_t blk = blockIdx.x;
    // positions and dL/dy first: their latency overlaps the prologue below
    const int g = c.enc_begin[c.grid_enc];
    float px[kPerThread], py[kPerThread], pz[kPerThread], d0[kPerThread], d1[kPerThread];
    load_positions(in, g, n, px, py, pz);
#pragma unroll
    for (int i = 0; i < (int)kPerThread; ++i) {
        uint32_t q = blockIdx.x * kSamplesPerBlock + (uint32_t)i * 256u + threadIdx.x;
        q = q < n ? q : n - 1u;
        d0[i] = dgrid[(size_t)(2 * l) * n + q];
        d1[i] = dgrid[(size_t)(2 * l + 1) * n + q];
    }
    // the block's count per bin and their exclusive scan over the level's bins
    const uint32_t per = (nb + 255u) / 256u;
    const uint32_t lo = threadIdx.x * per, hi = lo + per < nb ? lo + per : nb;
    uint32_t sum = 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t* offs = sc.counts + (size_t)(b0 + i) * b.nblk;
        const uint32_t cnt = (blk + 1u < b.nblk ? offs[blk + 1u] : sc.bin_count[b0 + i]) - offs[blk];
        sum += cnt;
    }
    uint32_t run;
    block_scan(sum, red, run);
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t* offs = sc.counts + (size_t)(b0 + i) * b.nblk;
        const uint32_t cnt = (blk + 1u < b.nblk ? offs[blk + 1u] : sc.bin_count[b0 + i]) - offs[blk];
        cur[i] = run;
        off[i] = sc.bin_start[b0 + i] + offs[blk] - run;
        run += cnt;
    }
    __syncthreads();
    const uint32_t mask = (1u << b.shift) - 1u;
    float mx = 0.0f;
    bool bad = false;
    auto seen = [&](float d) {
        if (__builtin_isfinite(d)) mx = smax(mx, fabsf(d));
        else bad = true;
    };
#pragma unroll
    for (int i = 0; i < (int)kPerThread; ++i) {
        const uint32_t q = blockIdx.x * kSamplesPerBlock + (uint32_t)i * 256u + threadIdx.x;
        if (q >= n) break;
        uint32_t idx[8];
        float w[8];
        const int nc = level_corners(c, l, px[i], py[i], pz[i], idx, w);
        seen(d0[i]);
        seen(d1[i]);
        for (int k = 0; k < nc; ++k) {
            const uint32_t bin = idx[k] >> b.shift;
            const uint32_t p = atomicAdd(&cur[bin], 1u);
            tw[p] = (bin << 16) | ((idx[k] & mask) * 2u);
            s0[p] = contrib(w[k], d0[i]);
            s1[p] = contrib(w[k], d1[i]);
        }
    }
    __syncthreads();
    // every bin's cursor now ends its range: the last one ends the block's records
    const uint32_t total = nb ? cur[nb - 1u] : 0u;
    for (uint32_t r = threadIdx.x; r < total; r += 256u) {
        const uint32_t t = tw[r];
        reinterpret_cast<uint3*>(sc.rec)[off[t >> 16] + r] =
            make_uint3(t & 0xffffu, __float_as_uint(s0[r]), __float_as_uint(s1[r]));
    }
    for (int o = 32; o > 0; o >>= 1) mx = smax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63u) == 0) wave_max[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicMax(&sc.maxbits[l], __float_as_uint(smax(smax(wave_max[0], wave_max[1]), smax(wave_max[2], wave_max[3]))));
    if (bad && overflow) atomicOr(overflow, 1u);
}

// 512 threads per workgroup: the 64 KB tile allows two workgroups per CU,
// and the record loads are latency bound, so more of them are kept in flight.
constexpr uint32_t kReduceThreads = 512;

template <bool DET>
__global__ void __launch_bounds__(kReduceThreads) k_bin_reduce(const DevCfg c, const BinLayout b, Scratch sc,
                                                               float* __restrict__ grad, uint32_t n) {
    using Acc = unsigned long long;
    __shared__ Acc tile[kTileWords];
    if (blockIdx.x >= sc.chunks[-1].x) return;
    const uint4 ch = sc.chunks[blockIdx.x];
    const int F = c.F;
    const uint32_t l = level_of_bin(b, c.n_levels, ch.x);
    const uint32_t t = ch.x - b.bin_off[l];
    const uint32_t size = c.offsets[l + 1] - c.offsets[l];
    const uint32_t e0 = t << b.shift;
    const uint32_t entries = size - e0 < (1u << b.shift) ? size - e0 : (1u << b.shift);
    const uint32_t words = entries * (uint32_t)F;
    const int s = fixed_shift(sc.maxbits[l], n * 8u);
    for (uint32_t i = threadIdx.x; i < kTileWords; i += kReduceThreads) tile[i] = (Acc)0;
    __syncthreads();
    auto add = [&](uint32_t wi, float v) { atomicAdd(&tile[wi], (Acc)to_fixed(v, s)); };
    if (F == 2) {
        // 8 records per thread in flight: all loads, then the LDS atomics.
        constexpr uint32_t U = 8;
        const uint3* rec = reinterpret_cast<const uint3*>(sc.rec);
        uint32_t r0 = ch.y + threadIdx.x;
        for (; r0 + kReduceThreads * (U - 1u) < ch.z; r0 += kReduceThreads * U) {
            uint3 rv[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) rv[u] = rec[r0 + kReduceThreads * u];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                add(rv[u].x, __uint_as_float(rv[u].y));
                add(rv[u].x + 1u, __uint_as_float(rv[u].z));
            }
        }
        for (; r0 < ch.z; r0 += kReduceThreads) {
            const uint3 r = rec[r0];
            add(r.x, __uint_as_float(r.y));
            add(r.x + 1u, __uint_as_float(r.z));
        }
    } else {
        for (uint32_t r = ch.y + threadIdx.x; r < ch.z; r += kReduceThreads) {
            const uint32_t* rp = sc.rec + (size_t)r * (1u + F);
            for (int f = 0; f < F; ++f) add(rp[0] + f, __uint_as_float(rp[1 + f]));
        }
    }
    __syncthreads();
    const size_t gbase = ((size_t)c.offsets[l] + e0) * F;
    if (ch.w == 1u) {
        for (uint32_t i = threadIdx.x; i < words; i += kReduceThreads) grad[gbase + i] = ldexpf((float)(long long)tile[i], -s);
    } else {
        for (uint32_t i = threadIdx.x; i < words; i += kReduceThreads) {
            if (tile[i] == (Acc)0) continue;
            if (DET) atomicAdd(&sc.wide[gbase + i], tile[i]);
            else atomicAdd(&grad[gbase + i], ldexpf((float)(long long)tile[i], -s));
        }
    }
}

// Deterministic mode: multi-chunk bins, int64 -> float, and clear for the next call.
__global__ void __launch_bounds__(256) k_bin_finish(const DevCfg c, const BinLayout b, Scratch sc,
                                                    float* __restrict__ grad, uint32_t n) {
    const uint32_t bin = blockIdx.x;
    if (sc.bin_chunks[bin] < 2u) return;
    const int F = c.F;
    const uint32_t l = level_of_bin(b, c.n_levels, bin);
    const uint32_t t = bin - b.bin_off[l];
    const uint32_t size = c.offsets[l + 1] - c.offsets[l];
    const uint32_t e0 = t << b.shift;
    const uint32_t entries = size - e0 < (1u << b.shift) ? size - e0 : (1u << b.shift);
    const int s = fixed_shift(sc.maxbits[l], n * 8u);
    const size_t gbase = ((size_t)c.offsets[l] + e0) * F;
    for (uint32_t i = threadIdx.x; i < entries * (uint32_t)F; i += 256u) {
        grad[gbase + i] = ldexpf((float)(long long)sc.wide[gbase + i], -s);
        sc.wide[gbase + i] = 0ull;
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

BinLayout make_layout(const DevCfg& c, uint32_t n, uint32_t* n_bins) {
    BinLayout b{};
    int fl = 0;
    while ((1 << (fl + 1)) <= c.F) ++fl;
    b.shift = 13u - (uint32_t)fl;  // (8192 / F) entries per tile
    uint32_t bins = 0;
    for (int l = 0; l < c.n_levels; ++l) {
        b.bin_off[l] = bins;
        const uint32_t size = c.offsets[l + 1] - c.offsets[l];
        bins += (size + (1u << b.shift) - 1u) >> b.shift;
    }
    b.bin_off[c.n_levels] = bins;
    for (int l = 0; l < c.n_levels; ++l) b.bin_off_max = std::max(b.bin_off_max, b.bin_off[l + 1] - b.bin_off[l]);
    b.nblk = (n + kSamplesPerBlock - 1u) / kSamplesPerBlock;
    const size_t recs = (size_t)n * c.n_levels * 8u;
    b.max_chunks = (uint32_t)(recs / kChunk + bins + 1u);
    *n_bins = bins;
    return b;
}

struct Carve {
    char* p;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        T* r = reinterpret_cast<T*>(p ? p + used : nullptr);
        used += (count * sizeof(T) + 255u) / 256u * 256u;
        return r;
    }
};

Scratch carve(const DevCfg& c, const BinLayout& b, uint32_t n_bins, uint32_t n, bool det, void* base, size_t* bytes) {
    Carve cv{static_cast<char*>(base)};
    Scratch sc{};
    const size_t recs = (size_t)n * c.n_levels * 8u;
    // int64 accumulators first: they are cleared once at allocation and kept clear.
    sc.wide = det ? cv.take<unsigned long long>((size_t)c.offsets[c.n_levels] * c.F) : nullptr;
    sc.counts = cv.take<uint32_t>((size_t)n_bins * b.nblk);
    sc.bin_start = cv.take<uint32_t>(n_bins);
    sc.bin_chunks = cv.take<uint32_t>(n_bins);
    sc.bin_count = cv.take<uint32_t>(n_bins);
    uint4* ct = cv.take<uint4>((size_t)b.max_chunks + 1u);
    sc.chunks = ct ? ct + 1 : nullptr;
    sc.maxbits = cv.take<uint32_t>(32);
    sc.rec = cv.take<uint32_t>(recs * (1u + c.F));
    *bytes = cv.used;
    return sc;
}

}  // namespace

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255u) / 256u), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

size_t grid_scatter_bytes(const DevCfg& c, uint32_t n, bool deterministic) {
    if (c.grid_enc < 0 || n == 0) return 0;
    uint32_t n_bins;
    const BinLayout b = make_layout(c, n, &n_bins);
    size_t bytes = 0;
    carve(c, b, n_bins, n, deterministic, nullptr, &bytes);
    return bytes;
}

size_t grid_scatter_clear_bytes(const DevCfg& c, bool deterministic) {
    if (c.grid_enc < 0 || !deterministic) return 0;
    return ((size_t)c.offsets[c.n_levels] * c.F * 8u + 255u) / 256u * 256u;
}

namespace {
// Layout and scratch of one call, or an error for a batch beyond the limits.
hipError_t scatter_setup(const DevCfg& c, uint32_t n, bool deterministic, void* scratch, size_t scratch_bytes,
                         BinLayout* b, uint32_t* n_bins, Scratch* sc) {
    if ((size_t)n * c.n_levels * 8u >= 0xffffffffull || c.F > 8) return hipErrorInvalidValue;
    *b = make_layout(c, n, n_bins);
    for (int l = 0; l < c.n_levels; ++l)
        if (b->bin_off[l + 1] - b->bin_off[l] > kMaxBinsPerLevel) return hipErrorInvalidValue;
    size_t need = 0;
    *sc = carve(c, *b, *n_bins, n, deterministic, scratch, &need);
    if (!scratch || need > scratch_bytes) return hipErrorInvalidValue;
    return hipSuccess;
}

size_t hist_bytes(const DevCfg& c, const BinLayout& b) {
    uint32_t bmax = 0;
    for (int l = 0; l < c.n_levels; ++l) bmax = std::max(bmax, b.bin_off[l + 1] - b.bin_off[l]);
    return (size_t)bmax * 4u;
}
}  // namespace

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s);

hipError_t launch_grid_scatter_prepare(const DevCfg& c, const float* in, uint32_t n, bool deterministic,
                                       void* scratch, size_t scratch_bytes, hipStream_t s) {
    if (n == 0 || c.grid_enc < 0) return hipSuccess;
    BinLayout b;
    uint32_t n_bins;
    Scratch sc;
    hipError_t e = scatter_setup(c, n, deterministic, scratch, scratch_bytes, &b, &n_bins, &sc);
    if (e != hipSuccess) return e;
    const dim3 grid(b.nblk, c.n_levels);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, sc.maxbits, 32u);
    hipLaunchKernelGGL(k_bin_count, grid, dim3(256), hist_bytes(c, b), s, c, b, in, sc, n);
    hipLaunchKernelGGL(k_bin_scan, dim3(n_bins), dim3(256), 0, s, b, sc);
    hipLaunchKernelGGL(k_bin_chunks, dim3(1), dim3(256), 0, s, b, n_bins, sc);
    return hipGetLastError();
}

hipError_t launch_grid_scatter_finish(const DevCfg& c, const float* in, const float* dgrid, float* grad, uint32_t n,
                                      bool deterministic, void* scratch, size_t scratch_bytes, uint32_t* overflow,
                                      hipStream_t s) {
    if (c.grid_enc < 0) return hipSuccess;
    hipError_t e = launch_fill(grad, 0.0f, (size_t)c.offsets[c.n_levels] * c.F, s);
    if (e != hipSuccess || n == 0) return e;
    BinLayout b;
    uint32_t n_bins;
    Scratch sc;
    e = scatter_setup(c, n, deterministic, scratch, scratch_bytes, &b, &n_bins, &sc);
    if (e != hipSuccess) return e;
    const dim3 grid(b.nblk, c.n_levels);
    if (c.F == 2) {
        const size_t lds = (2u * (size_t)((b.bin_off_max + 3u) & ~3u) + 3u * kStage) * 4u;
        if (lds > 65536u) {
            e = hipFuncSetAttribute((const void*)k_bin_place2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_bin_place2, grid, dim3(256), lds, s, c, b, in, dgrid, sc, n, overflow);
    } else {
        hipLaunchKernelGGL(k_bin_place, grid, dim3(256), hist_bytes(c, b), s, c, b, in, dgrid, sc, n, overflow);
    }
    if (deterministic) {
        hipLaunchKernelGGL(k_bin_reduce<true>, dim3(b.max_chunks), dim3(kReduceThreads), 0, s, c, b, sc, grad, n);
        hipLaunchKernelGGL(k_bin_finish, dim3(n_bins), dim3(256), 0, s, c, b, sc, grad, n);
    } else {
        hipLaunchKernelGGL(k_bin_reduce<false>, dim3(b.max_chunks), dim3(kReduceThreads), 0, s, c, b, sc, grad, n);
    }
    return hipGetLastError();
}

hipError_t launch_grid_scatter(const DevCfg& c, const float* in, const float* dgrid, float* grad, uint32_t n,
                               bool deterministic, void* scratch, size_t scratch_bytes, uint32_t* overflow,
                               hipStream_t s) {
    hipError_t e = launch_grid_scatter_prepare(c, in, n, deterministic, scratch, scratch_bytes, s);
    if (e != hipSuccess) return e;
    return launch_grid_scatter_finish(c, in, dgrid, grad, n, deterministic, scratch, scratch_bytes, overflow, s);
}

}  // namespace npmg
