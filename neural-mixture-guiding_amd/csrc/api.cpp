// api.cpp -- the C ABI of include/npmg.h: context, parameter state and the
// host-side sequencing of the kernels.
//
// The context owns every device buffer: full-precision parameters (tcnn
// order: MLP matrices, then the grid table), the inference copy (EMA),
// Adam moments and per-parameter step counters, gradients, the kernel-layout
// (transposed) MLP weights for training and inference, and a training
// workspace grown on demand. Calls enqueue on the caller's stream and never
// synchronise the host, except create/destroy/get/set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/npmg.h"
#include "msgpack.h"
#include "config.h"
#include "kernels.h"
#include "npmg_math.h"

namespace npmg {
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s);
size_t wgrad_partials_needed(const DevCfg& c, uint32_t n, uint32_t* n_splits, uint32_t* q_per_split);
hipError_t launch_query_mfma(const DevCfg& c, int precision, const void* wfrag, const void* table, const void* genc,
                             const QueryIO& io, uint32_t n, uint32_t q_begin, uint32_t q_end, int n_cu,
                             int blocks_per_cu, const float* fp8_scale, hipStream_t s);
hipError_t launch_query_mfma_queues(const DevCfg& c, int precision, const void* wfrag, const void* table,
                                    const QueueSet& qs, int n_cu, int blocks_per_cu, const float* fp8_scale,
                                    hipStream_t s);
hipError_t launch_grid_encode_queues(const DevCfg& c, int precision, const void* table, const QueueSet& qs, int n_cu,
                                     uint64_t capacity, hipStream_t s);
hipError_t launch_query_fp32_queues(const DevCfg& c, const float* wt, const float* table, const QueueSet& qs,
                                    int n_cu, hipStream_t s);
hipError_t launch_grid_encode_range(const DevCfg& c, int precision, const void* table, const float* in, void* genc,
                                    uint32_t n, uint32_t q_begin, uint32_t q_end, hipStream_t s, uint32_t lds_reserve,
                                    int n_lds, int levels);
int grid_encode_lds_levels(const DevCfg& c, const void* table, uint32_t batch);
bool mfma_split_grid(const DevCfg& c);
hipError_t launch_pack_frags(const DevCfg& c, int precision, const float* w, void* frag, hipStream_t s);
hipError_t launch_to_half(int precision, const float* src, void* dst, size_t n, hipStream_t s);
bool mfma_train_supported(const DevCfg& c);
hipError_t launch_train_forward_mfma(const DevCfg& c, int precision, const void* wfrag, const void* table, void* genc,
                                     const float* in, const TrainMfmaWS& ws, uint32_t n, int n_cu, hipStream_t s);
hipError_t launch_train_backward_mfma(const DevCfg& c, int precision, const void* bfrag, const float* dout,
                                      const TrainMfmaWS& ws, float* denc, uint32_t n, uint32_t* overflow, int n_cu,
                                      hipStream_t s);
hipError_t launch_wgrad_mfma(const DevCfg& c, int precision, const TrainMfmaWS& ws, const void* genc, const float* in,
                             float* partials, size_t partials_cap, float* dw, uint32_t n, uint32_t* overflow,
                             hipStream_t s);
size_t wgrad_mfma_partials_needed(const DevCfg& c, uint32_t n);
size_t train_mask_words(const DevCfg& c, uint32_t n);
size_t bfrag_elements(const DevCfg& c);
hipError_t launch_pdf_integral(const DevCfg& c, const float* raw, const float* normal, const double* gz,
                               const double* gw, uint32_t n_z, uint32_t n_phi, double* out, uint32_t n,
                               hipStream_t s);
hipError_t launch_loss_fd(const DevCfg& c, const float* raw, const float* td, const float* tr, const float* tp,
                          double h, double* out, uint32_t n, hipStream_t s);
hipError_t launch_pack_bfrags(const DevCfg& c, int precision, const float* w, void* frag, hipStream_t s);
}  // namespace npmg

using namespace npmg;

namespace {

thread_local std::string g_create_error;

constexpr float kMaxLossScale = 16777216.0f;  // upper bound of the dynamic loss scale, 2^24

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) return nullptr;
    const hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) throw std::bad_alloc();
    return static_cast<T*>(p);
}

// PCG32 with jump-ahead (tiny-cuda-nn dependencies/pcg32/pcg32.h:44-170).
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    explicit Pcg32(uint64_t initstate, uint64_t initseq = 1u) {
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

// Trainer's initialisation (tiny-cuda-nn trainer.h:53-92): std::seed_seq{seed}
// -> pcg32; Xavier-uniform MLP matrices drawn in order on the host
// (gpu_matrix.h:291-306); the grid table U(-1e-4, 1e-4) as generated by
// generate_random_uniform's 4-values-per-thread kernel (random.h:66-97,
// grid.h:1333-1336), after which the generator advances by the table size.
void init_params(const Config& cfg, uint32_t seed, std::vector<float>& p) {
    const DevCfg& d = cfg.d;
    std::seed_seq seq{seed};
    std::vector<uint32_t> seeds(2);
    seq.generate(seeds.begin(), seeds.end());
    Pcg32 rng(seeds.front());
    p.assign(cfg.n_params, 0.0f);
    for (int pass = 0; pass < std::max(1, cfg.init_passes); ++pass) {
        for (int l = 0; l < d.n_layers; ++l) {
            const float scale = std::sqrt(6.0f / (float)(d.layer_in[l] + d.layer_out[l]));
            const size_t ne = (size_t)d.layer_in[l] * d.layer_out[l];
            float* w = p.data() + d.layer_off[l];
            for (size_t i = 0; i < ne; ++i) w[i] = rng.next_float() * 2.0f * scale - scale;
        }
        if (cfg.n_grid) {
            const size_t n = cfg.n_grid, n_threads = (n + 3) / 4;
            Pcg32 r = rng;
            float* out = p.data() + cfg.n_mlp;
            for (size_t k = 0; k < 4 * n_threads; ++k) {
                const float val = r.next_float();
                const size_t idx = k / 4 + n_threads * (k % 4);
                if (idx < n) out[idx] = fm::fma_(val, 1e-4f - -1e-4f, -1e-4f);
            }
            rng.advance(n);
        }
    }
}

void need(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}

}  // namespace

// Scratch of npmg_sample_pdf_noalloc: the level-major grid features (16-bit,
// [L F][n]) of the MFMA query path in a file-scope device array, so that the
// call allocates nothing. Each GPU holds its own copy for the life of the
// process; its address is looked up once per device, with that device
// current, when the first context on it is created (scratch_on_device).
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device dev (current), looked up on first use.
void* scratch_on_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

// Unused dynamic LDS per grid-encode workgroup in the pipelined query path
// (encode of chunk c + 1 beside the fused kernel of chunk c): caps the encode
// at 5 workgroups per CU so that the fused kernel's workgroups find room.
constexpr uint32_t kEncodeLdsReserve = 27u * 1024u;

struct npmg_ctx {
    Config cfg;
    std::string err;
    int n_cu = 256;
    float* master = nullptr;  // training parameters, full precision
    float* ema = nullptr;     // inference parameters (EMA), or null without EMA
    float* m = nullptr;
    float* v = nullptr;
    uint32_t* steps = nullptr;
    float* grad = nullptr;
    float* wt_train = nullptr;  // transposed MLP weights
    float* wt_infer = nullptr;
    // half-precision copies for the MFMA path: weight fragments and grid tables
    void* frag_train = nullptr;
    void* frag_infer = nullptr;
    void* htab_train = nullptr;
    void* htab_infer = nullptr;
    // training workspace
    uint32_t ws_cap = 0;
    TrainWS ws{};
    float* partials = nullptr;
    size_t partials_cap = 0;
    void* scatter_scratch = nullptr;  // grid-gradient binning
    size_t scatter_cap = 0;
    void* genc = nullptr;  // level-major grid features [L F][n] of the MFMA paths
    uint32_t genc_cap = 0;
    // second stream and events: the grid encode of chunk c + 1 overlaps the fused kernel of chunk c
    hipStream_t aux = nullptr;
    // the captured train step (npmg_train_step_graph) and what it was captured for
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void* graph_key[7] = {};
    static constexpr int kPipe = 8;
    // launch shape of the query path (npmg_autotune): pipeline chunks, workgroups per CU (0: all that fit)
    uint32_t tune_chunks = 4;
    int tune_bpc = 0;
    // persistent queue launches: 64-query tiles per work-counter claim (fused kernel)
    // and (level, 256-query) items per claim (grid encode)
    uint32_t tune_tiles = 4, tune_items = 16;
    uint32_t* queue_counter = nullptr;  // [2] work counters of the persistent queue launches
    void* noalloc_scratch = nullptr;    // this device's `scratch` (npmg_sample_pdf_noalloc)
    uint32_t* pack_scratch = nullptr;   // block totals of the training packer
    size_t pack_cap = 0;
    float* stage = nullptr;  // contiguous batch of npmg_train_records: in [9], dir [2], rgb [3], pdf, bsdf pdf
    uint32_t stage_cap = 0;
    // FP8 inference: calibrated input maxima per layer and the scales derived from them
    // (wscale [L], ascale [L], qscale [L + 1]); calibration uses at most kFp8CalibMax queries
    static constexpr uint32_t kFp8CalibMax = 1u << 16;
    float* fp8_amax = nullptr;
    float* fp8_scales = nullptr;
    bool fp8_calibrated = false;
    // kernel timing: event pairs per tag, recorded while enabled, read by npmg_timing_read
    static constexpr int kTags = 3;
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> t_pending[kTags], t_free;
    double t_ms[kTags] = {};
    uint32_t t_count[kTags] = {};
    hipEvent_t ev_fork = nullptr, ev_chunk[kPipe] = {};
    hipEvent_t ev_bwd = nullptr, ev_wgrad = nullptr;  // fork/join of the weight gradients (train_backward)
    hipEvent_t ev_prep0 = nullptr, ev_prep = nullptr;  // fork/join of the grid-gradient binning (train_forward)
    // MFMA training: transposed weight fragments and the operand-type buffers
    void* bfrag_train = nullptr;
    TrainMfmaWS mws{};
    // loss scale, overflow flag and step count on the device; decay^t table of the EMA debiasing
    ScaleState* sstate = nullptr;
    float* mean_partials = nullptr;  // [256]
    double* gl_nodes = nullptr;      // Gauss-Legendre rule of the quadrature check: nodes, then weights
    uint32_t gl_n = 0;
    float* ema_pow = nullptr;
    uint32_t ema_pow_len = 0;
    float* adam_bias = nullptr;  // Adam's bias correction per step count (k_adam)
    static constexpr uint32_t kAdamBiasLen = 1u << 16;

    const float* infer_params() const { return ema ? ema : master; }

    ~npmg_ctx() {
        drop_graph();
        for (auto& v : t_pending)
            for (auto& e : v) t_free.push_back(e);
        for (auto& e : t_free) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (capture) (void)hipStreamDestroy(capture);
        if (aux) (void)hipStreamDestroy(aux);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_bwd) (void)hipEventDestroy(ev_bwd);
        if (ev_wgrad) (void)hipEventDestroy(ev_wgrad);
        if (ev_prep0) (void)hipEventDestroy(ev_prep0);
        if (ev_prep) (void)hipEventDestroy(ev_prep);
        for (hipEvent_t e : ev_chunk)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)master, (void*)ema, (void*)m, (void*)v, (void*)steps, (void*)grad, (void*)wt_train,
                        (void*)wt_infer, (void*)ws.enc, (void*)ws.acts, (void*)ws.g, (void*)ws.dout, (void*)ws.loss,
                        (void*)partials, frag_train, frag_infer, htab_train, htab_infer, (void*)ws.dgrid,
                        scatter_scratch, genc, bfrag_train, mws.x0, mws.xh, mws.dyh, mws.dyo, (void*)mws.mask,
                        (void*)mws.raw, (void*)sstate, (void*)ema_pow, (void*)adam_bias, (void*)mean_partials, (void*)gl_nodes,
                        (void*)queue_counter, (void*)fp8_amax, (void*)fp8_scales, (void*)pack_scratch,
                        (void*)stage})
            if (p) (void)hipFree(p);
    }

    bool half() const {
        return cfg.precision == PREC_BF16 || cfg.precision == PREC_FP16 || cfg.precision == PREC_BF16X3;
    }
    // planes of an MFMA operand buffer: 2 (hi, lo) for the split BF16 operands
    size_t planes() const { return cfg.precision == PREC_BF16X3 ? 2u : 1u; }

    // Brackets the launches enqueued by f on stream s with a timing event pair (tag).
    template <class F>
    void timed(int tag, hipStream_t s, F&& f) {
        if (!timing) {
            f();
            return;
        }
        std::pair<hipEvent_t, hipEvent_t> e{};
        if (!t_free.empty()) {
            e = t_free.back();
            t_free.pop_back();
        } else {
            hip_check(hipEventCreate(&e.first), "event");
            hip_check(hipEventCreate(&e.second), "event");
        }
        hip_check(hipEventRecord(e.first, s), "event");
        f();
        hip_check(hipEventRecord(e.second, s), "event");
        t_pending[tag].push_back(e);
    }

    void timing_collect(int tag) {
        for (auto& e : t_pending[tag]) {
            hip_check(hipEventSynchronize(e.second), "event sync");
            float ms = 0.0f;
            hip_check(hipEventElapsedTime(&ms, e.first, e.second), "event time");
            t_ms[tag] += ms;
            ++t_count[tag];
            t_free.push_back(e);
        }
        t_pending[tag].clear();
    }

    void check_fused() const {
        if (cfg.d.mixture == MIX_ASG && cfg.d.n_hidden < 1) throw std::invalid_argument("invalid network");
        if (cfg.d.K != 7 && cfg.d.K != 8 && cfg.d.K != 16)
            throw std::invalid_argument("the fused kernels support 7, 8 or 16 lobes");
    }

    void check_train() const {
        check_fused();
        if (half() && !mfma_train_supported(cfg.d))
            throw std::invalid_argument("MFMA training supports width 64 with 2-4 hidden layers or width 128 with 2-3");
    }

    void ensure_genc(uint32_t n) {
        if (!mfma_split_grid(cfg.d) || n <= genc_cap) return;
        drop_graph();
        if (genc) hip_check(hipFree(genc), "hipFree");
        genc = nullptr;
        genc_cap = 0;
        genc = dalloc<uint16_t>((size_t)cfg.d.n_levels * cfg.d.F * n);
        genc_cap = n;
    }

    // Kernel layouts of the parameters: transposed FP32 weights, MFMA fragments and,
    // unless the optimiser already wrote them, the operand-type grid tables.
    void refresh_layouts(hipStream_t s, bool tables = true) {
        // The transposed FP32 weights feed the FP32 kernels only: in the 16-bit
        // precisions, just the FP8 calibration's FP32 forward (inference weights).
        if (!half()) hip_check(launch_transpose_mlp(cfg.d, master, wt_train, s), "transpose");
        if (!half() || cfg.infer_fp8) hip_check(launch_transpose_mlp(cfg.d, infer_params(), wt_infer, s), "transpose");
        if (half()) {
            hip_check(launch_pack_frags(cfg.d, cfg.precision, master, frag_train, s), "pack");
            if (bfrag_train) hip_check(launch_pack_bfrags(cfg.d, cfg.precision, master, bfrag_train, s), "pack");
            if (cfg.infer_fp8)
                hip_check(launch_pack_frags_fp8(cfg.d, infer_params(), frag_infer, fp8_amax, fp8_scales, s), "pack");
            else
                hip_check(launch_pack_frags(cfg.d, cfg.precision, infer_params(), frag_infer, s), "pack");
            if (tables) {
                hip_check(launch_to_half(cfg.precision, master + cfg.n_mlp, htab_train, cfg.n_grid, s), "table");
                hip_check(launch_to_half(cfg.precision, infer_params() + cfg.n_mlp, htab_infer, cfg.n_grid, s),
                          "table");
            }
        }
    }

    // Fused query with the training (0) or inference (1) parameters in the configured precision.
    // genc_buf: a caller-provided buffer for the level-major grid features (at least
    // L F n 16-bit values); then the call allocates nothing: no workspace growth, no
    // FP8 calibration (an uncalibrated FP8 network is an error) and no timing events.
    void query(const QueryIO& io, uint32_t n, bool inference, hipStream_t s, void* genc_buf = nullptr) {
        const bool no_alloc = genc_buf != nullptr;
        auto run = [&](int tag, hipStream_t st, auto&& f) {
            if (no_alloc) f();
            else timed(tag, st, f);
        };
        if (half()) {
            const void* frag = inference ? frag_infer : frag_train;
            const void* tab = inference ? htab_infer : htab_train;
            const float* q8 = nullptr;
            if (inference && cfg.infer_fp8) {
                if (!fp8_calibrated && n) {
                    if (no_alloc) throw StateError("FP8 inference is not calibrated: call npmg_fp8_calibrate first");
                    calibrate_fp8(io.in, n, std::min(n, kFp8CalibMax), s);
                }
                q8 = fp8_scales + 2 * cfg.d.n_layers;
            }
            if (!mfma_split_grid(cfg.d)) {
                run(NPMG_TIME_QUERY, s, [&] {
                    hip_check(launch_query_mfma(cfg.d, cfg.precision, frag, tab, nullptr, io, n, 0u, n, n_cu, tune_bpc, q8, s),
                              "query (mfma)");
                });
                return;
            }
            if (!no_alloc) ensure_genc(n);
            void* g = no_alloc ? genc_buf : genc;
            // Large batches: tune_chunks (<= kPipe) chunks; the encodes run on aux (after everything
            // already queued on s, which may still read g), chunk c's fused
            // kernel on s after chunk c's encode.
            const uint32_t chunks = n >= (1u << 18) ? tune_chunks : 1u;
            const uint32_t m = ((n + chunks - 1u) / chunks + 63u) / 64u * 64u;
            const int n_lds = grid_encode_lds_levels(cfg.d, tab, n);
            if (chunks == 1u) {
                run(NPMG_TIME_GRID_ENCODE, s, [&] {
                    hip_check(launch_grid_encode_range(cfg.d, cfg.precision, tab, io.in, g, n, 0u, n, s, 0u, n_lds, ENC_ALL),
                              "grid encode");
                });
                run(NPMG_TIME_QUERY, s, [&] {
                    hip_check(launch_query_mfma(cfg.d, cfg.precision, frag, tab, g, io, n, 0u, n, n_cu, tune_bpc, q8, s),
                              "query (mfma)");
                });
                return;
            }
            hip_check(hipEventRecord(ev_fork, s), "event");
            hip_check(hipStreamWaitEvent(aux, ev_fork, 0), "stream wait");
            // the LDS-staged coarse levels of the whole batch first, then the gathered levels chunk by chunk
            if (n_lds > 0)
                run(NPMG_TIME_GRID_ENCODE, aux, [&] {
                    hip_check(launch_grid_encode_range(cfg.d, cfg.precision, tab, io.in, g, n, 0u, n, aux, 0u, n_lds,
                                                       ENC_LDS_LEVELS),
                              "grid encode");
                });
            for (uint32_t k = 0; k < chunks; ++k) {
                const uint32_t b = std::min(n, k * m), e = std::min(n, b + m);
                run(NPMG_TIME_GRID_ENCODE, aux, [&] {
                    hip_check(launch_grid_encode_range(cfg.d, cfg.precision, tab, io.in, g, n, b, e, aux,
                                                       kEncodeLdsReserve, n_lds, ENC_GATHER_LEVELS),
                              "grid encode");
                });
                hip_check(hipEventRecord(ev_chunk[k], aux), "event");
            }
            for (uint32_t k = 0; k < chunks; ++k) {
                const uint32_t b = std::min(n, k * m), e = std::min(n, b + m);
                hip_check(hipStreamWaitEvent(s, ev_chunk[k], 0), "stream wait");
                run(NPMG_TIME_QUERY, s, [&] {
                    hip_check(launch_query_mfma(cfg.d, cfg.precision, frag, tab, g, io, n, b, e, n_cu, tune_bpc, q8, s),
                              "query (mfma)");
                });
            }
        } else {
            const float* p = inference ? infer_params() : master;
            run(NPMG_TIME_QUERY, s, [&] {
                hip_check(launch_query_fp32(cfg.d, inference ? wt_infer : wt_train, p + cfg.n_mlp, io, n, n_cu, s),
                          "query (fp32)");
            });
        }
    }

    // Several queues in one persistent launch (inference parameters): the grid
    // encode pulls (level, tile) items over all queues, then the fused kernel
    // pulls 64-query tiles; each launch has its own work counter.
    void query_queues(const npmg_queue* qv, const uint32_t* counts, uint32_t nq, uint32_t n_eval, uint32_t flags,
                      hipStream_t s) {
        QueueSet qs{};
        qs.n_queues = nq;
        qs.counts = counts;
        uint64_t total_cap = 0;
        for (uint32_t i = 0; i < nq; ++i) {
            const npmg_queue& q = qv[i];
            need(q.capacity == 0 || q.in, "null queue input");
            need(!q.u || (q.out_dir && q.out_pdf), "sampling needs out_dir and out_pdf");
            need(n_eval == 0 || (q.eval_dirs && q.out_eval_pdf), "evaluation needs eval_dirs and out_eval_pdf");
            need(!(flags & NPMG_COSINE_AWARE) || q.normal, "NPMG_COSINE_AWARE needs normal");
            QueryIO& io = qs.io[i];
            io.in = q.in;
            io.normal = (flags & NPMG_COSINE_AWARE) ? q.normal : nullptr;
            io.u = q.u;
            io.eval_dirs = q.eval_dirs;
            io.n_eval = (int)n_eval;
            io.out_dir = q.out_dir;
            io.out_pdf = q.out_pdf;
            io.out_eval_pdf = q.out_eval_pdf;
            io.out_lobe = q.out_lobe;
            io.out_sel = q.out_selection;
            qs.stride[i] = q.capacity;
            total_cap += q.capacity;
        }
        need(total_cap < (1ull << 31), "total queue capacity must stay below 2^31");
        if (total_cap == 0) return;
        hip_check(launch_zero_u32(queue_counter, 2u, s), "zero");
        timed(NPMG_TIME_QUERY, s, [&] {
            if (half()) {
                const bool split = mfma_split_grid(cfg.d);
                if (split) {
                    ensure_genc((uint32_t)total_cap);
                    const size_t per_q = (size_t)cfg.d.n_levels * cfg.d.F;
                    size_t off = 0;
                    for (uint32_t i = 0; i < nq; ++i) {
                        qs.genc[i] = static_cast<uint16_t*>(genc) + per_q * off;
                        off += qv[i].capacity;
                    }
                    qs.counter = queue_counter;
                    qs.batch = tune_items;
                    hip_check(launch_grid_encode_queues(cfg.d, cfg.precision, htab_infer, qs, n_cu, total_cap, s),
                              "grid encode (queues)");
                }
                qs.counter = queue_counter + 1;
                qs.batch = tune_tiles;
                need(!cfg.infer_fp8 || fp8_calibrated,
                     "FP8 inference is not calibrated: call npmg_fp8_calibrate (or one plain query) first");
                hip_check(launch_query_mfma_queues(cfg.d, cfg.precision, frag_infer, htab_infer, qs, n_cu, tune_bpc,
                                                   cfg.infer_fp8 ? fp8_scales + 2 * cfg.d.n_layers : nullptr, s),
                          "query (mfma, queues)");
            } else {
                qs.counter = queue_counter + 1;
                qs.batch = 1u;
                hip_check(launch_query_fp32_queues(cfg.d, wt_infer, infer_params() + cfg.n_mlp, qs, n_cu, s),
                          "query (fp32, queues)");
            }
        });
    }

    // Times the launch shapes of the query path on n synthetic queries and keeps
    // the fastest: pipeline chunks (split-grid MFMA path) x workgroups per CU.
    void autotune(uint32_t n, hipStream_t s) {
        if (!half() || n == 0) return;
        const bool split = mfma_split_grid(cfg.d);
        // synthetic queries: in [9][n], normal [3][n], u [3][n], eval_dirs [4][3][n]; outputs dir [3][n], pdf [n],
        // eval pdfs [4][n] (the workload of a guided bounce: sample + 4 pdfs)
        size_t bytes = sizeof(float) * (size_t)n * (9 + 3 + 3 + 12 + 3 + 1 + 4) + 8 * sizeof(uint32_t);
        float* buf = nullptr;
        hip_check(hipMalloc((void**)&buf, bytes), "hipMalloc");
        uint32_t* qcounts = reinterpret_cast<uint32_t*>(buf + (size_t)n * (9 + 3 + 3 + 12 + 3 + 1 + 4));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        // restored on every exit: timing is off while tuning, and the FP8 path is
        // timed without calibrating it on the synthetic queries
        const bool was_timing = timing, was_calibrated = fp8_calibrated;
        const uint32_t old_c = tune_chunks, old_t = tune_tiles, old_i = tune_items;
        const int old_b = tune_bpc;
        try {
            float *in = buf, *nrm = in + 9 * (size_t)n, *u = nrm + 3 * (size_t)n, *ed = u + 3 * (size_t)n;
            float *od = ed + 12 * (size_t)n, *op = od + 3 * (size_t)n, *oe = op + (size_t)n;
            hip_check(launch_synth_queries(0x7a11u, 0u, n, in, nrm, u, ed, s), "synth");
            QueryIO io{};
            io.in = in;
            io.normal = cfg.d.mixture == MIX_ASG ? nullptr : nrm;
            io.u = u;
            io.out_dir = od;
            io.out_pdf = op;
            io.eval_dirs = ed;
            io.n_eval = 4;
            io.out_eval_pdf = oe;
            hip_check(hipEventCreate(&e0), "event");
            hip_check(hipEventCreate(&e1), "event");
            timing = false;
            fp8_calibrated = true;
            float best = 1e30f;
            uint32_t best_c = tune_chunks;
            int best_b = tune_bpc;
            for (uint32_t ch : {1u, 2u, 4u, 8u}) {
                if (!split && ch > 1u) continue;
                if (split && n < (1u << 18) && ch > 1u) continue;  // query() does not pipeline these
                for (int b : {0, 1, 2, 3}) {
                    tune_chunks = ch;
                    tune_bpc = b;
                    query(io, n, true, s);  // warm-up
                    hip_check(hipEventRecord(e0, s), "event");
                    for (int r = 0; r < 5; ++r) query(io, n, true, s);
                    hip_check(hipEventRecord(e1, s), "event");
                    hip_check(hipEventSynchronize(e1), "event");
                    float ms = 0.0f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
                    if (ms < best * 0.98f) {  // prefer the earlier (simpler) shape unless clearly faster
                        best = ms;
                        best_c = ch;
                        best_b = b;
                    }
                }
            }
            tune_chunks = best_c;
            tune_bpc = best_b;
            // Queue launches: the claim sizes of the work counters, timed on a
            // frame of 8 queues whose lengths decay by 0.68 per bounce (the
            // queues share the synthetic arrays; identical queries write
            // identical results, so the overlap changes nothing).
            npmg_queue qv[8];
            uint32_t cnt_h[8];
            double wsum = 0.0;
            for (int b = 0; b < 8; ++b) wsum += std::pow(0.68, b);
            for (int b = 0; b < 8; ++b) {
                qv[b] = npmg_queue{in, cfg.d.mixture == MIX_ASG ? nullptr : nrm, u, ed, od, op, oe, nullptr, n, nullptr};
                cnt_h[b] = (uint32_t)std::max(1.0, std::floor(n * std::pow(0.68, b) / wsum));
            }
            hip_check(hipMemcpyAsync(qcounts, cnt_h, sizeof(cnt_h), hipMemcpyHostToDevice, s), "upload");
            const uint32_t flags = cfg.d.mixture == MIX_ASG ? 0u : NPMG_COSINE_AWARE;
            best = 1e30f;
            uint32_t best_t = tune_tiles, best_i = tune_items;
            for (uint32_t tl : {4u, 1u, 2u, 8u, 16u}) {
                for (uint32_t it : {16u, 8u, 32u, 64u}) {
                    if (!split && it != 16u) continue;  // no grid encode launch to tune
                    tune_tiles = tl;
                    tune_items = it;
                    query_queues(qv, qcounts, 8, 4, flags, s);  // warm-up
                    hip_check(hipEventRecord(e0, s), "event");
                    for (int r = 0; r < 5; ++r) query_queues(qv, qcounts, 8, 4, flags, s);
                    hip_check(hipEventRecord(e1, s), "event");
                    hip_check(hipEventSynchronize(e1), "event");
                    float ms = 0.0f;
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
                    if (ms < best * 0.98f) {
                        best = ms;
                        best_t = tl;
                        best_i = it;
                    }
                }
            }
            tune_tiles = best_t;
            tune_items = best_i;
            timing = was_timing;
            fp8_calibrated = was_calibrated;
        } catch (...) {
            timing = was_timing;
            fp8_calibrated = was_calibrated;
            tune_chunks = old_c;
            tune_bpc = old_b;
            tune_tiles = old_t;
            tune_items = old_i;
            (void)hipStreamSynchronize(s);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            (void)hipFree(buf);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        hip_check(hipFree(buf), "hipFree");
    }

    // Activation maxima of the inference parameters' FP32 forward over m queries of
    // `in` (row stride n), then the FP8 fragments and scales. Synchronous.
    void calibrate_fp8(const float* in, uint32_t n, uint32_t m, hipStream_t s) {
        if (!cfg.infer_fp8) throw std::invalid_argument("inference_precision is not fp8");
        need(m > 0 && m <= n && in, "calibration needs queries");
        const DevCfg& d = cfg.d;
        float* buf = nullptr;
        const size_t n_tin = (size_t)d.n_in * m, n_enc = (size_t)d.enc_padded * m;
        const size_t n_act = (size_t)(d.n_layers - 1) * d.width * m;
        hip_check(hipMalloc((void**)&buf, (n_tin + n_enc + n_act) * sizeof(float)), "hipMalloc");
        try {
            float* tin = buf;
            TrainWS t{};
            t.enc = buf + n_tin;
            t.acts = t.enc + n_enc;
            hip_check(hipMemcpy2DAsync(tin, (size_t)m * sizeof(float), in, (size_t)n * sizeof(float),
                                       (size_t)m * sizeof(float), (size_t)d.n_in, hipMemcpyDeviceToDevice, s),
                      "copy calibration queries");
            hip_check(launch_train_forward_fp32(d, wt_infer, infer_params() + cfg.n_mlp, tin, nullptr, nullptr,
                                                nullptr, nullptr, LossScale{}, t, m, n_cu, s),
                      "calibration forward");
            hip_check(launch_fp8_amax(d, t.enc, t.acts, m, fp8_amax, s), "calibration maxima");
            hip_check(launch_pack_frags_fp8(d, infer_params(), frag_infer, fp8_amax, fp8_scales, s), "pack");
            hip_check(hipStreamSynchronize(s), "synchronize");
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(buf);
            throw;
        }
        hip_check(hipFree(buf), "hipFree");
        fp8_calibrated = true;
    }

    void free_ws() {
        drop_graph();  // the captured step points into the workspace
        for (void* p : {(void*)ws.enc, (void*)ws.acts, (void*)ws.g, (void*)ws.dout, (void*)ws.loss, (void*)ws.dgrid,
                        mws.x0, mws.xh, mws.dyh, mws.dyo, (void*)mws.mask, (void*)mws.raw})
            if (p) hip_check(hipFree(p), "hipFree");
        ws = TrainWS{};
        mws = TrainMfmaWS{};
        ws_cap = 0;
    }

    // Training workspace for n samples: FP32 activations for the FP32 path, the
    // operand-type buffers and sign masks for the MFMA path.
    void ensure_ws(uint32_t n) {
        if (n <= ws_cap) return;
        free_ws();
        const DevCfg& d = cfg.d;
        const size_t hw = (size_t)d.n_hidden * d.width * n;
        ws.dout = dalloc<float>((size_t)d.n_out_pad * n);
        ws.loss = dalloc<float>(n);
        if (d.grid_enc >= 0) ws.dgrid = dalloc<float>((size_t)d.n_levels * d.F * n);
        size_t need;
        if (half()) {
            mws.x0 = dalloc<uint16_t>((size_t)d.enc_padded * n * planes());
            mws.xh = dalloc<uint16_t>(hw * planes());
            mws.dyh = dalloc<uint16_t>(hw * planes());
            mws.dyo = dalloc<uint16_t>((size_t)d.n_out_pad * n * planes());
            mws.mask = dalloc<uint32_t>(train_mask_words(d, n));
            mws.raw = dalloc<float>((size_t)d.n_out_pad * n);
            mws.dgrid = ws.dgrid;
            ensure_genc(n);
            need = wgrad_mfma_partials_needed(d, n);
        } else {
            ws.enc = dalloc<float>((size_t)d.enc_padded * n);
            ws.acts = dalloc<float>(hw);
            ws.g = dalloc<float>(hw);
            uint32_t splits, qps;
            need = wgrad_partials_needed(d, n, &splits, &qps);
        }
        const size_t need_scatter = grid_scatter_bytes(d, n, cfg.deterministic_grid);
        if (need_scatter > scatter_cap) {
            if (scatter_scratch) hip_check(hipFree(scatter_scratch), "hipFree");
            scatter_scratch = nullptr;
            scatter_cap = 0;
            scatter_scratch = dalloc<char>(need_scatter);
            scatter_cap = need_scatter;
            const size_t clear = grid_scatter_clear_bytes(d, cfg.deterministic_grid);
            if (clear) hip_check(hipMemset(scatter_scratch, 0, clear), "memset");
        }
        if (need > partials_cap) {
            if (partials) hip_check(hipFree(partials), "hipFree");
            partials = nullptr;
            partials_cap = 0;
            partials = dalloc<float>(need);
            partials_cap = need;
        }
        ws_cap = n;
    }

    LossScale state_scale(uint32_t n) const { return LossScale{&sstate->scale, 0.0f, 1.0f / (float)n}; }

    // Forward with the training weights keeping what the backward pass needs;
    // with targets, the loss and dL/draw (scaled by ls) go to ws.loss / ws.dout.
    // tb: BSDF pdfs of the records (learned selection only).
    void train_forward(const float* in, const float* td, const float* tr, const float* tp, const float* tb, uint32_t n,
                       LossScale ls, hipStream_t s) {
        ensure_ws(n);
        // The grid-gradient binning needs only the positions: it runs on aux
        // beside the forward pass, and train_backward joins it before the scatter.
        if (cfg.n_grid) {
            hip_check(hipEventRecord(ev_prep0, s), "event");
            hip_check(hipStreamWaitEvent(aux, ev_prep0, 0), "stream wait");
            hip_check(launch_grid_scatter_prepare(cfg.d, in, n, cfg.deterministic_grid, scatter_scratch, scatter_cap,
                                                  aux),
                      "grid scatter (binning)");
            hip_check(hipEventRecord(ev_prep, aux), "event");
        }
        if (half()) {
            hip_check(launch_train_forward_mfma(cfg.d, cfg.precision, frag_train, htab_train, genc, in, mws, n, n_cu, s),
                      "train forward (mfma)");
            if (td) hip_check(launch_adjoint(cfg.d, mws.raw, td, tr, tp, tb, ls, ws.dout, ws.loss, n, s), "adjoint");
        } else {
            hip_check(launch_train_forward_fp32(cfg.d, wt_train, master + cfg.n_mlp, in, td, tr, tp, tb, ls, ws, n, n_cu,
                                                s),
                      "train forward");
        }
    }

    // Backward from dL/draw (dout: ws.dout, or a caller buffer on the MFMA path):
    // parameter gradients into dparams (tcnn order), dL/denc into denc when given.
    // Non-finite gradients raise the overflow flag when `overflow` is set.
    void train_backward(const float* in, const float* dout, float* dparams, float* denc, uint32_t n, bool overflow,
                        hipStream_t s) {
        uint32_t* of = overflow ? &sstate->overflow : nullptr;
        if (half()) {
            hip_check(launch_train_backward_mfma(cfg.d, cfg.precision, bfrag_train, dout, mws, denc, n, of, n_cu, s),
                      "train backward (mfma)");
        } else {
            if (dout != ws.dout)
                hip_check(hipMemcpyAsync(ws.dout, dout, (size_t)cfg.d.n_out_pad * n * sizeof(float),
                                         hipMemcpyDeviceToDevice, s),
                          "copy dout");
            hip_check(launch_train_backward_fp32(cfg.d, master, ws, denc, ws.dgrid, n, n_cu, s), "train backward");
        }
        // The weight gradients (from the stored layer inputs and output
        // gradients) and the grid-gradient scatter (from dL/d(grid features))
        // are independent and bound by different units (MFMA operand loads vs
        // record bytes and LDS atomics): the weight gradients run on aux beside
        // the scatter, and s waits for them before anything after the backward.
        const bool fork = cfg.n_grid != 0;
        hipStream_t sw = s;
        if (fork) {
            hip_check(hipEventRecord(ev_bwd, s), "event");
            hip_check(hipStreamWaitEvent(aux, ev_bwd, 0), "stream wait");
            sw = aux;
        }
        if (half())
            hip_check(launch_wgrad_mfma(cfg.d, cfg.precision, mws, genc, in, partials, partials_cap, dparams, n, of, sw),
                      "weight gradient (mfma)");
        else
            hip_check(launch_wgrad_fp32(cfg.d, ws, partials, partials_cap, dparams, n, of, sw), "weight gradient");
        if (fork) {
            hip_check(hipEventRecord(ev_wgrad, aux), "event");
            float* gg = dparams + cfg.n_mlp;
            hip_check(hipStreamWaitEvent(s, ev_prep, 0), "stream wait");
            hip_check(launch_grid_scatter_finish(cfg.d, in, ws.dgrid, gg, n, cfg.deterministic_grid, scatter_scratch,
                                                 scatter_cap, of, s),
                      "grid scatter");
            hip_check(hipStreamWaitEvent(s, ev_wgrad, 0), "stream wait");
        }
    }

    // One optimisation step on a batch: forward + adjoint, backward, Adam/EMA,
    // loss mean. Enqueues kernels only (no allocation once ensure_ws(n) ran, no
    // host synchronisation, no memset nodes), so it can be captured.
    void train_step(const float* in, const float* td, const float* tr, const float* tp, const float* tb, uint32_t n,
                    float* loss_out, hipStream_t s) {
        train_forward(in, td, tr, tp, tb, n, state_scale(n), s);
        train_backward(in, ws.dout, grad, nullptr, n, cfg.dynamic_scale, s);
        optimizer_step(grad, 0.0f, s);
        if (loss_out) hip_check(launch_mean(ws.loss, n, mean_partials, loss_out, s), "loss mean");
    }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        std::fill(graph_key, graph_key + 7, nullptr);
    }

    // Adam + EMA. ls > 0: the gradient is divided by ls (split-pass API);
    // ls == 0: by the device loss scale, and a raised overflow flag skips the step.
    void optimizer_step(const float* dparams, float ls, hipStream_t s) {
        AdamArgs a{};
        a.n = cfg.n_params;
        a.n_matrix = cfg.n_mlp;
        a.loss_scale = ls;
        a.scale_from_state = !(ls > 0.0f);
        a.check_overflow = a.scale_from_state && cfg.dynamic_scale;
        a.st = sstate;
        a.ema_pow = ema_pow;
        a.ema_pow_len = ema_pow_len;
        a.adam_bias = adam_bias;
        a.adam_bias_len = kAdamBiasLen;
        a.lr = cfg.lr;
        a.beta1 = cfg.beta1;
        a.beta2 = cfg.beta2;
        a.eps = cfg.eps;
        a.l2_reg = cfg.l2_reg;
        a.ema = ema != nullptr;
        a.ema_decay = cfg.ema_decay;
        a.ema_half = cfg.ema_full < 0 ? cfg.precision == PREC_FP16 : cfg.ema_full == 0;
        a.half_precision = half() ? cfg.precision : PREC_FP32;
        a.half_train = htab_train;
        a.half_infer = htab_infer;
        hip_check(launch_adam(a, master, dparams, m, v, steps, ema, s), "adam");
        hip_check(launch_scale_update(sstate, a.check_overflow, cfg.scale_window, kMaxLossScale, s), "loss scale");
        refresh_layouts(s, false);
    }
};

namespace {

template <class F>
npmg_status guarded(npmg_ctx* ctx, F&& f) {
    try {
        f();
        return NPMG_OK;
    } catch (const HipError& e) {
        if (ctx) ctx->err = e.what();
        return NPMG_EHIP;
    } catch (const StateError& e) {
        if (ctx) ctx->err = e.what();
        return NPMG_ESTATE;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "device or host allocation failed";
        return NPMG_ENOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return NPMG_EINVAL;
    } catch (...) {
        if (ctx) ctx->err = "unknown error";
        return NPMG_EINVAL;
    }
}

hipStream_t S(npmg_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Gauss-Legendre rule on [-1, 1]: n nodes then n weights (Newton iteration on
// P_n from the Chebyshev initial guesses, the textbook construction).
std::vector<double> gauss_legendre(uint32_t n) {
    std::vector<double> r(2 * (size_t)n);
    const double pi = 3.14159265358979323846;
    for (uint32_t i = 0; i < (n + 1) / 2; ++i) {
        double x = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (uint32_t k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        double p0 = 1.0, p1 = x;
        for (uint32_t k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        const double w = 2.0 / ((1.0 - x * x) * dp * dp);
        r[i] = -x;
        r[n - 1 - i] = x;
        r[n + i] = w;
        r[n + n - 1 - i] = w;
    }
    return r;
}

}  // namespace

extern "C" {

npmg_status npmg_create(const char* json_config, uint64_t seed, npmg_ctx** out) {
    g_create_error.clear();
    if (!out || !json_config) {
        g_create_error = "null argument";
        return NPMG_EINVAL;
    }
    *out = nullptr;
    npmg_ctx* ctx = new (std::nothrow) npmg_ctx();
    if (!ctx) return NPMG_ENOMEM;
    const npmg_status st = guarded(ctx, [&] {
        ctx->cfg = parse_config(json_config);
        const int K = ctx->cfg.d.K;
        need(K == 1 || K == 2 || K == 4 || K == 7 || K == 8 || K == 16, "n_lobes must be 1, 2, 4, 7, 8 or 16");
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hip_check(hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
        ctx->noalloc_scratch = scratch_on_device(dev);
        hip_check(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking), "stream");
        hip_check(hipStreamCreateWithFlags(&ctx->capture, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ctx->ev_bwd, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ctx->ev_wgrad, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ctx->ev_prep0, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ctx->ev_prep, hipEventDisableTiming), "event");
        for (hipEvent_t& e : ctx->ev_chunk) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        const Config& c = ctx->cfg;
        ctx->master = dalloc<float>(c.n_params);
        if (c.ema) ctx->ema = dalloc<float>(c.n_params);
        ctx->m = dalloc<float>(c.n_params);
        ctx->v = dalloc<float>(c.n_params);
        ctx->steps = dalloc<uint32_t>(c.n_params);
        ctx->grad = dalloc<float>(c.n_params);
        ctx->wt_train = dalloc<float>(c.n_mlp);
        ctx->wt_infer = dalloc<float>(c.n_mlp);
        if (ctx->half()) {
            need(c.n_mlp % 512 == 0, "half precision needs layer sizes that are multiples of 32 x 16");
            // the split operands keep two planes of weight fragments in LDS
            need(c.precision != PREC_BF16X3 || c.d.width == 64, "precision bf16x3 needs a network of width 64");
            ctx->frag_train = dalloc<uint16_t>(c.n_mlp * ctx->planes());
            if (c.infer_fp8) ctx->frag_infer = dalloc<uint8_t>(c.n_mlp);
            else ctx->frag_infer = dalloc<uint16_t>(c.n_mlp * ctx->planes());
            ctx->htab_train = dalloc<uint16_t>(std::max<size_t>(c.n_grid, 1));
            ctx->htab_infer = dalloc<uint16_t>(std::max<size_t>(c.n_grid, 1));
            if (mfma_train_supported(c.d))
                ctx->bfrag_train = dalloc<uint16_t>(bfrag_elements(c.d) * ctx->planes());
        }
        ctx->sstate = dalloc<ScaleState>(1);
        ctx->mean_partials = dalloc<float>(256);
        ctx->queue_counter = dalloc<uint32_t>(2);
        if (c.infer_fp8) {
            // until calibrated, unit maxima (valid scales; the first query calibrates)
            ctx->fp8_amax = dalloc<float>(kMaxLayers);
            ctx->fp8_scales = dalloc<float>(3 * kMaxLayers + 1);
            const std::vector<float> ones(kMaxLayers, 1.0f);
            hip_check(hipMemcpy(ctx->fp8_amax, ones.data(), kMaxLayers * sizeof(float), hipMemcpyHostToDevice),
                      "upload");
        }
        const ScaleState st0{c.loss_scale, 0u, 0u, 0u};
        hip_check(hipMemcpy(ctx->sstate, &st0, sizeof(st0), hipMemcpyHostToDevice), "upload");
        // decay^t for the EMA debiasing (tcnn ema.h:110-111 evaluates std::pow in
        // double): tabulated until 1 - decay^t rounds to 1 in FP32.
        std::vector<float> pw;
        for (uint32_t t = 0; t < (1u << 22); ++t) {
            const float v = (float)std::pow((double)c.ema_decay, (double)t);
            pw.push_back(v);
            if (1.0f - v == 1.0f) break;
        }
        ctx->ema_pow_len = (uint32_t)pw.size();
        ctx->ema_pow = dalloc<float>(pw.size());
        hip_check(hipMemcpy(ctx->ema_pow, pw.data(), pw.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
        ctx->adam_bias = dalloc<float>(npmg_ctx::kAdamBiasLen);
        hip_check(launch_adam_bias_table(c.beta1, c.beta2, ctx->adam_bias, npmg_ctx::kAdamBiasLen, nullptr),
                  "adam bias table");
        hip_check(hipMemset(ctx->m, 0, c.n_params * sizeof(float)), "memset");
        hip_check(hipMemset(ctx->v, 0, c.n_params * sizeof(float)), "memset");
        hip_check(hipMemset(ctx->steps, 0, c.n_params * sizeof(uint32_t)), "memset");
        hip_check(hipMemset(ctx->grad, 0, c.n_params * sizeof(float)), "memset");
        std::vector<float> p;
        init_params(c, (uint32_t)seed, p);
        hip_check(hipMemcpy(ctx->master, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
        if (ctx->ema) hip_check(hipMemcpy(ctx->ema, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice), "upload");
        ctx->refresh_layouts(nullptr);
        hip_check(hipDeviceSynchronize(), "synchronize");
    });
    if (st != NPMG_OK) {
        g_create_error = ctx->err;
        delete ctx;
        return st;
    }
    *out = ctx;
    return NPMG_OK;
}

void npmg_destroy(npmg_ctx* ctx) {
    if (!ctx) return;
    (void)hipDeviceSynchronize();
    delete ctx;
}

const char* npmg_last_error(const npmg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

npmg_status npmg_n_params(const npmg_ctx* ctx, size_t* n_mlp, size_t* n_grid) {
    if (!ctx || !n_mlp || !n_grid) return NPMG_EINVAL;
    *n_mlp = ctx->cfg.n_mlp;
    *n_grid = ctx->cfg.n_grid;
    return NPMG_OK;
}

npmg_status npmg_widths(const npmg_ctx* ctx, uint32_t* encoded_width, uint32_t* padded_output) {
    if (!ctx || !encoded_width || !padded_output) return NPMG_EINVAL;
    *encoded_width = (uint32_t)ctx->cfg.d.enc_padded;
    *padded_output = (uint32_t)ctx->cfg.d.n_out_pad;
    return NPMG_OK;
}

npmg_status npmg_grid_layout(const npmg_ctx* ctx, uint32_t* n_levels, uint32_t* offsets, float* scales,
                             uint32_t* resolutions) {
    if (!ctx || !n_levels) return NPMG_EINVAL;
    const DevCfg& d = ctx->cfg.d;
    *n_levels = d.grid_enc >= 0 ? (uint32_t)d.n_levels : 0u;
    if (d.grid_enc < 0) return NPMG_OK;
    for (int l = 0; l <= d.n_levels; ++l) if (offsets) offsets[l] = d.offsets[l];
    for (int l = 0; l < d.n_levels; ++l) {
        if (scales) scales[l] = d.scales[l];
        if (resolutions) resolutions[l] = d.res[l];
    }
    return NPMG_OK;
}

npmg_status npmg_set_params(npmg_ctx* ctx, const float* params, size_t n, int on_device) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(params && n == ctx->cfg.n_params, "params must hold exactly n_params values");
        const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        hip_check(hipMemcpy(ctx->master, params, n * sizeof(float), k), "copy params");
        if (ctx->ema) hip_check(hipMemcpy(ctx->ema, params, n * sizeof(float), k), "copy params");
        ctx->fp8_calibrated = false;  // new inference weights: the next query recalibrates
        ctx->refresh_layouts(nullptr);
        hip_check(hipDeviceSynchronize(), "synchronize");
    });
}

npmg_status npmg_set_params_which(npmg_ctx* ctx, const float* params, size_t n, int on_device, int which) {
    if (!ctx) return NPMG_EINVAL;
    if (which == 2) return npmg_set_params(ctx, params, n, on_device);
    return guarded(ctx, [&] {
        need(params && n == ctx->cfg.n_params, "params must hold exactly n_params values");
        need(which == 0 || which == 1, "which must be 0 (training), 1 (inference) or 2 (both)");
        const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        hip_check(hipDeviceSynchronize(), "synchronize");
        // without an EMA the inference copy is the training master
        float* dst = which == 1 && ctx->ema ? ctx->ema : ctx->master;
        hip_check(hipMemcpy(dst, params, n * sizeof(float), k), "copy params");
        if (ctx->cfg.infer_fp8 && dst == ctx->infer_params()) ctx->fp8_calibrated = false;
        ctx->refresh_layouts(nullptr);
        hip_check(hipDeviceSynchronize(), "synchronize");
    });
}

npmg_status npmg_get_params(const npmg_ctx* ctx, float* params, size_t n, int on_device, int which) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(const_cast<npmg_ctx*>(ctx), [&] {
        need(params && n == ctx->cfg.n_params, "params must hold exactly n_params values");
        need(which == 0 || which == 1, "which must be 0 (training) or 1 (inference)");
        hip_check(hipDeviceSynchronize(), "synchronize");
        const float* src = which == 0 ? ctx->master : ctx->infer_params();
        hip_check(hipMemcpy(params, src, n * sizeof(float), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost),
                  "copy params");
    });
}

npmg_status npmg_infer(npmg_ctx* ctx, const float* in, float* raw_out, uint32_t n, npmg_stream stream) {
    return npmg_forward(ctx, in, raw_out, n, 1, stream);
}

npmg_status npmg_forward(npmg_ctx* ctx, const float* in, float* raw_out, uint32_t n, int use_inference_params,
                         npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        need(n == 0 || (in && raw_out), "null buffer");
        QueryIO io{};
        io.in = in;
        io.raw_out = raw_out;
        ctx->query(io, n, use_inference_params != 0, S(stream));
    });
}

npmg_status npmg_sample_pdf(npmg_ctx* ctx, const float* in, const float* normal, const float* u,
                            const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                            float* out_eval_pdf, uint32_t* out_lobe, float* raw_out, uint32_t n, uint32_t flags,
                            npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        need(n == 0 || in, "null input");
        need(!u || (out_dir && out_pdf), "sampling needs out_dir and out_pdf");
        need(n_eval == 0 || (eval_dirs && out_eval_pdf), "evaluation needs eval_dirs and out_eval_pdf");
        need(!(flags & NPMG_COSINE_AWARE) || normal, "NPMG_COSINE_AWARE needs normal");
        need(!(flags & NPMG_COSINE_AWARE) || ctx->cfg.d.mixture != MIX_ASG,
             "the anisotropic mixture has no cosine product");
        QueryIO io{};
        io.in = in;
        io.normal = (flags & NPMG_COSINE_AWARE) ? normal : nullptr;
        io.u = u;
        io.eval_dirs = eval_dirs;
        io.n_eval = (int)n_eval;
        io.out_dir = out_dir;
        io.out_pdf = out_pdf;
        io.out_eval_pdf = out_eval_pdf;
        io.out_lobe = out_lobe;
        io.raw_out = raw_out;
        ctx->query(io, n, true, S(stream));
    });
}

npmg_status npmg_sample_pdf_noalloc(npmg_ctx* ctx, const float* in, const float* normal, const float* u,
                                    const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                                    float* out_eval_pdf, uint32_t* out_lobe, float* raw_out, uint32_t n,
                                    uint32_t flags, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        need(n == 0 || in, "null input");
        need(!u || (out_dir && out_pdf), "sampling needs out_dir and out_pdf");
        need(n_eval == 0 || (eval_dirs && out_eval_pdf), "evaluation needs eval_dirs and out_eval_pdf");
        need(!(flags & NPMG_COSINE_AWARE) || normal, "NPMG_COSINE_AWARE needs normal");
        need(!(flags & NPMG_COSINE_AWARE) || ctx->cfg.d.mixture != MIX_ASG,
             "the anisotropic mixture has no cosine product");
        const DevCfg& d = ctx->cfg.d;
        const size_t bytes = ctx->half() && mfma_split_grid(d) ? (size_t)d.n_levels * d.F * n * 2u : 0u;
        need(bytes <= kNoAllocScratchBytes,
             "npmg_sample_pdf_noalloc: n x grid features x 2 bytes exceeds its 16 MiB scratch (262144 queries of "
             "a 32-feature grid); split the batch");
        QueryIO io{};
        io.in = in;
        io.normal = (flags & NPMG_COSINE_AWARE) ? normal : nullptr;
        io.u = u;
        io.eval_dirs = eval_dirs;
        io.n_eval = (int)n_eval;
        io.out_dir = out_dir;
        io.out_pdf = out_pdf;
        io.out_eval_pdf = out_eval_pdf;
        io.out_lobe = out_lobe;
        io.raw_out = raw_out;
        ctx->query(io, n, true, S(stream), ctx->noalloc_scratch);
    });
}

npmg_status npmg_sample_pdf_ex(npmg_ctx* ctx, const npmg_queue* q, uint32_t n, uint32_t n_eval, uint32_t flags,
                               npmg_stream stream) {
    if (!ctx || !q) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        need(n == 0 || q->in, "null input");
        need(!q->u || (q->out_dir && q->out_pdf), "sampling needs out_dir and out_pdf");
        need(n_eval == 0 || (q->eval_dirs && q->out_eval_pdf), "evaluation needs eval_dirs and out_eval_pdf");
        need(!(flags & NPMG_COSINE_AWARE) || q->normal, "NPMG_COSINE_AWARE needs normal");
        need(!(flags & NPMG_COSINE_AWARE) || ctx->cfg.d.mixture != MIX_ASG,
             "the anisotropic mixture has no cosine product");
        need(!q->out_selection || ctx->cfg.d.sel, "out_selection needs mixture.learn_selection");
        QueryIO io{};
        io.in = q->in;
        io.normal = (flags & NPMG_COSINE_AWARE) ? q->normal : nullptr;
        io.u = q->u;
        io.eval_dirs = q->eval_dirs;
        io.n_eval = (int)n_eval;
        io.out_dir = q->out_dir;
        io.out_pdf = q->out_pdf;
        io.out_eval_pdf = q->out_eval_pdf;
        io.out_lobe = q->out_lobe;
        io.out_sel = q->out_selection;
        ctx->query(io, n, true, S(stream));
    });
}

npmg_status npmg_sample_pdf_queues(npmg_ctx* ctx, const npmg_queue* queues, const uint32_t* counts, uint32_t n_queues,
                                   uint32_t n_eval, uint32_t flags, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        need(n_queues <= (uint32_t)kMaxQueues, "at most 16 queues per launch");
        need(n_queues == 0 || (queues && counts), "null queues or counts");
        need(!(flags & NPMG_COSINE_AWARE) || ctx->cfg.d.mixture != MIX_ASG,
             "the anisotropic mixture has no cosine product");
        if (n_queues) ctx->query_queues(queues, counts, n_queues, n_eval, flags, S(stream));
    });
}

npmg_status npmg_autotune(npmg_ctx* ctx, uint32_t n, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_fused();
        ctx->autotune(n, S(stream));
    });
}

}  // extern "C"

namespace {

// IEEE binary16 -> binary32 (tcnn's "__half" snapshots).
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t bits;
    if (exp == 0x1fu) bits = sign | 0x7f800000u | (man << 13);
    else if (exp != 0) bits = sign | ((exp + 112u) << 23) | (man << 13);
    else if (man == 0) bits = sign;
    else {  // subnormal: normalise
        int e = -1;
        uint32_t m = man;
        do {
            ++e;
            m <<= 1;
        } while (!(m & 0x400u));
        bits = sign | ((uint32_t)(112 - e) << 23) | ((m & 0x3ffu) << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// n float values from a "float" (4 n bytes) or "__half" (2 n bytes) blob.
std::vector<float> blob_floats(const npmg::mp::Value* v, size_t n, const char* what) {
    if (!v || v->kind != npmg::mp::Value::BIN) throw std::invalid_argument(std::string("snapshot: missing ") + what);
    std::vector<float> out(n);
    if (v->bn == n * 4) {
        std::memcpy(out.data(), v->b, n * 4);
    } else if (v->bn == n * 2) {
        for (size_t i = 0; i < n; ++i) {
            uint16_t h;
            std::memcpy(&h, v->b + 2 * i, 2);
            out[i] = half_to_float(h);
        }
    } else {
        throw std::invalid_argument(std::string("snapshot: ") + what + " has the wrong size");
    }
    return out;
}

// tcnn Trainer::serialize (trainer.h:267-278) with the optimizer state of
// Adam (adam.h:282-290) nested in Ema (ema.h:194-199), keys in nlohmann's
// sorted order, plus an "npmg" map (FP32 master weights and the loss-scale
// state) that tcnn ignores and that makes a resume exact. With dry, blobs are
// counted instead of written.
std::vector<uint8_t> build_snapshot(const npmg_ctx* ctx, bool with_optimizer, bool dry, size_t* total) {
    const Config& c = ctx->cfg;
    const size_t n = c.n_params;
    npmg::mp::Writer w;
    size_t skipped = 0;
    auto blob = [&](const void* dev, size_t bytes) {
        if (dry) {
            w.bin_header(bytes);
            skipped += bytes;
            return;
        }
        std::vector<uint8_t> h(bytes);
        hip_check(hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost), "download");
        w.bin(h.data(), bytes);
    };
    ScaleState st{};
    if (!dry) hip_check(hipMemcpy(&st, ctx->sstate, sizeof(st), hipMemcpyDeviceToHost), "download");
    w.map(with_optimizer ? 5u : 4u);
    w.str("n_params");
    w.uint(n);
    w.str("npmg");
    w.map(3);
    w.str("loss_scale");
    w.f64(st.scale);
    w.str("master_binary");
    blob(ctx->master, n * 4);
    w.str("scale_good");
    w.uint(st.good);
    if (with_optimizer) {
        w.str("optimizer");
        auto adam = [&] {
            w.map(5);
            w.str("base_learning_rate");
            w.f64(c.lr);
            w.str("current_step");
            w.uint(st.step);
            w.str("first_moments_binary");
            blob(ctx->m, n * 4);
            w.str("param_steps_binary");
            blob(ctx->steps, n * 4);
            w.str("second_moments_binary");
            blob(ctx->v, n * 4);
        };
        if (ctx->ema) {
            w.map(2);
            w.str("nested");
            adam();
            w.str("weights_ema_binary");
            blob(ctx->ema, n * 4);
        } else {
            adam();
        }
    }
    w.str("params_binary");
    blob(ctx->infer_params(), n * 4);
    w.str("params_type");
    w.str("float");
    *total = w.out.size() + skipped;
    return std::move(w.out);
}

}  // namespace

extern "C" {

npmg_status npmg_snapshot_save(npmg_ctx* ctx, int with_optimizer, void* buf, size_t capacity, size_t* size) {
    if (!ctx || !size) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        hip_check(hipDeviceSynchronize(), "synchronize");
        size_t total = 0;
        if (!buf) {
            build_snapshot(ctx, with_optimizer != 0, true, &total);
            *size = total;
            return;
        }
        std::vector<uint8_t> doc = build_snapshot(ctx, with_optimizer != 0, false, &total);
        *size = total;
        need(capacity >= total, "snapshot buffer too small (query the size with buf = NULL)");
        std::memcpy(buf, doc.data(), total);
    });
}

npmg_status npmg_snapshot_load(npmg_ctx* ctx, const void* buf, size_t size) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(buf && size, "empty snapshot");
        const npmg::mp::Value doc = npmg::mp::Reader(static_cast<const uint8_t*>(buf), size).parse();
        need(doc.kind == npmg::mp::Value::MAP, "snapshot: not a map");
        const size_t n = ctx->cfg.n_params;
        const npmg::mp::Value* np = doc.get("n_params");
        need(np && np->is_number() && (size_t)np->number() == n, "snapshot: n_params differs from this network");
        const npmg::mp::Value* ty = doc.get("params_type");
        if (ty) need(ty->kind == npmg::mp::Value::STR && (ty->s == "float" || ty->s == "__half"),
                     "snapshot: params_type must be float or __half");
        const std::vector<float> params = blob_floats(doc.get("params_binary"), n, "params_binary");
        const npmg::mp::Value* ext = doc.get("npmg");
        const std::vector<float> master =
            ext && ext->get("master_binary") ? blob_floats(ext->get("master_binary"), n, "master_binary") : params;
        // optimizer state (all of it parsed before anything is written)
        const npmg::mp::Value* opt = doc.get("optimizer");
        std::vector<float> ema, m, v;
        std::vector<uint32_t> steps;
        uint32_t cur_step = 0;
        if (opt) {
            need(opt->kind == npmg::mp::Value::MAP, "snapshot: optimizer must be a map");
            const npmg::mp::Value* adam = opt->get("nested") ? opt->get("nested") : opt;
            if (ctx->ema) ema = opt->get("weights_ema_binary") ? blob_floats(opt->get("weights_ema_binary"), n,
                                                                             "weights_ema_binary")
                                                                 : params;
            m = blob_floats(adam->get("first_moments_binary"), n, "first_moments_binary");
            v = blob_floats(adam->get("second_moments_binary"), n, "second_moments_binary");
            steps.assign(n, 0u);
            if (const npmg::mp::Value* ps = adam->get("param_steps_binary")) {
                need(ps->kind == npmg::mp::Value::BIN && ps->bn == n * 4, "snapshot: param_steps_binary size");
                std::memcpy(steps.data(), ps->b, n * 4);
            }
            const npmg::mp::Value* cs = adam->get("current_step");
            need(cs && cs->is_number() && cs->number() >= 0, "snapshot: current_step");
            cur_step = (uint32_t)cs->number();
        } else if (ctx->ema) {
            ema = params;  // set_params_full_precision sets every copy (trainer.h:288)
        }
        hip_check(hipDeviceSynchronize(), "synchronize");
        hip_check(hipMemcpy(ctx->master, master.data(), n * 4, hipMemcpyHostToDevice), "upload");
        if (ctx->ema) hip_check(hipMemcpy(ctx->ema, ema.data(), n * 4, hipMemcpyHostToDevice), "upload");
        ScaleState st{};
        hip_check(hipMemcpy(&st, ctx->sstate, sizeof(st), hipMemcpyDeviceToHost), "download");
        if (opt) {
            hip_check(hipMemcpy(ctx->m, m.data(), n * 4, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(ctx->v, v.data(), n * 4, hipMemcpyHostToDevice), "upload");
            hip_check(hipMemcpy(ctx->steps, steps.data(), n * 4, hipMemcpyHostToDevice), "upload");
            st.step = cur_step;
        }
        if (ext) {
            const npmg::mp::Value* ls = ext->get("loss_scale");
            const npmg::mp::Value* gd = ext->get("scale_good");
            if (ls && ls->is_number() && ls->number() > 0) st.scale = (float)ls->number();
            if (gd && gd->is_number()) st.good = (uint32_t)gd->number();
        }
        st.overflow = 0u;
        hip_check(hipMemcpy(ctx->sstate, &st, sizeof(st), hipMemcpyHostToDevice), "upload");
        ctx->fp8_calibrated = false;  // new inference weights: the next query recalibrates
        ctx->refresh_layouts(nullptr);
        hip_check(hipDeviceSynchronize(), "synchronize");
    });
}

npmg_status npmg_pack_queries(const float* pos, const float* wo, const float* normal, const float* roughness,
                              const float* scene_box, uint32_t n, float* in, npmg_stream stream) {
    if (!scene_box || (n && !(pos && wo && normal && roughness && in))) return NPMG_EINVAL;
    SceneBox box{};
    for (int a = 0; a < 3; ++a) {
        box.lo[a] = scene_box[a];
        box.hi[a] = scene_box[3 + a];
    }
    return launch_pack_queries(pos, wo, normal, roughness, box, n, in, S(stream)) == hipSuccess ? NPMG_OK : NPMG_EHIP;
}

npmg_status npmg_pack_training(npmg_ctx* ctx, const npmg_path_records* records, const npmg_train_tile* tile,
                               const float* scene_box, uint32_t capacity, float* in, float* tgt_dir, float* tgt_rgb,
                               float* tgt_pdf, float* tgt_bsdf_pdf, uint32_t* count, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(records && tile && scene_box && count, "null argument");
        const npmg_path_records& r = *records;
        need(capacity == 0 || (in && tgt_dir && tgt_rgb && tgt_pdf), "null output");
        need(tile->res[0] > 0 || tile->res[1] == 0, "training resolution");
        need(tile->stride[0] > 0 && tile->stride[1] > 0, "stride must be positive");
        need((uint64_t)tile->resolution[0] * tile->resolution[1] <= r.n_pixels, "resolution exceeds n_pixels");
        need(r.max_depth <= 64, "max_depth above 64");
        const uint64_t n_threads = (uint64_t)tile->res[0] * tile->res[1];
        need(n_threads * r.max_depth < (1ull << 32), "too many records");
        need(n_threads == 0 || (r.L && r.thp && r.pos && r.dir && r.wi_pdf && r.mis_weight && r.delta &&
                                r.bsdf_val && r.wo && r.normal && r.roughness && r.depth),
             "null record field");
        need(!tgt_bsdf_pdf || r.bsdf_pdf, "tgt_bsdf_pdf needs records->bsdf_pdf");
        PathRecords pr{r.L, r.thp, r.pos, r.dir, r.wi_pdf, r.mis_weight, r.delta, r.bsdf_val, r.wo, r.normal,
                       r.roughness, r.depth, r.bsdf_pdf, r.max_depth, r.n_pixels};
        TrainTile tt{};
        for (int a = 0; a < 2; ++a) {
            tt.offset[a] = tile->offset[a];
            tt.stride[a] = tile->stride[a];
            tt.res[a] = tile->res[a];
            tt.full[a] = tile->resolution[a];
        }
        SceneBox box{};
        for (int a = 0; a < 3; ++a) {
            box.lo[a] = scene_box[a];
            box.hi[a] = scene_box[3 + a];
        }
        const size_t need_words = pack_training_scratch(tt);
        if (need_words > ctx->pack_cap) {
            if (ctx->pack_scratch) hip_check(hipFree(ctx->pack_scratch), "hipFree");
            ctx->pack_scratch = nullptr;
            ctx->pack_cap = 0;
            ctx->pack_scratch = dalloc<uint32_t>(need_words);
            ctx->pack_cap = need_words;
        }
        hip_check(launch_pack_training(pr, tt, box, capacity, in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf_pdf, count,
                                       ctx->pack_scratch, S(stream)),
                  "pack training");
    });
}

npmg_status npmg_fp8_calibrate(npmg_ctx* ctx, const float* in, uint32_t n, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n > 0 && in, "calibration needs queries");
        ctx->calibrate_fp8(in, n, std::min(n, npmg_ctx::kFp8CalibMax), S(stream));
    });
}

npmg_status npmg_fp8_scales(npmg_ctx* ctx, float* weight_scales, float* input_scales, uint32_t n_layers) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        if (!ctx->cfg.infer_fp8) throw std::invalid_argument("inference_precision is not fp8");
        const int L = ctx->cfg.d.n_layers;
        need(n_layers == (uint32_t)L && weight_scales && input_scales, "need n_layers entries per array");
        hip_check(hipDeviceSynchronize(), "synchronize");
        hip_check(hipMemcpy(weight_scales, ctx->fp8_scales, L * sizeof(float), hipMemcpyDeviceToHost), "download");
        hip_check(hipMemcpy(input_scales, ctx->fp8_scales + L, L * sizeof(float), hipMemcpyDeviceToHost), "download");
    });
}

npmg_status npmg_set_tuning(npmg_ctx* ctx, uint32_t chunks, uint32_t blocks_per_cu) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(chunks >= 1u && chunks <= (uint32_t)npmg_ctx::kPipe, "pipeline chunks must be 1..8");
        need(blocks_per_cu <= 64u, "workgroups per CU must be 0..64");
        ctx->tune_chunks = chunks;
        ctx->tune_bpc = (int)blocks_per_cu;
    });
}

npmg_status npmg_queue_tuning(const npmg_ctx* ctx, uint32_t* tiles_per_claim, uint32_t* items_per_claim) {
    if (!ctx || !tiles_per_claim || !items_per_claim) return NPMG_EINVAL;
    *tiles_per_claim = ctx->tune_tiles;
    *items_per_claim = ctx->tune_items;
    return NPMG_OK;
}

npmg_status npmg_set_queue_tuning(npmg_ctx* ctx, uint32_t tiles_per_claim, uint32_t items_per_claim) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(tiles_per_claim >= 1u && tiles_per_claim <= 64u, "tiles per claim must be 1..64");
        need(items_per_claim >= 1u && items_per_claim <= 256u, "items per claim must be 1..256");
        ctx->tune_tiles = tiles_per_claim;
        ctx->tune_items = items_per_claim;
    });
}

npmg_status npmg_tuning(const npmg_ctx* ctx, uint32_t* chunks, uint32_t* blocks_per_cu) {
    if (!ctx || !chunks || !blocks_per_cu) return NPMG_EINVAL;
    *chunks = ctx->half() && mfma_split_grid(ctx->cfg.d) ? ctx->tune_chunks : 1u;
    *blocks_per_cu = (uint32_t)ctx->tune_bpc;
    return NPMG_OK;
}

npmg_status npmg_mixture_sample_pdf(npmg_ctx* ctx, const float* raw, const float* normal, const float* u,
                                    const float* eval_dirs, uint32_t n_eval, float* out_dir, float* out_pdf,
                                    float* out_eval_pdf, uint32_t* out_lobe, float* lobes_out, uint32_t n,
                                    uint32_t flags, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n == 0 || raw, "null raw outputs");
        need(!u || (out_dir && out_pdf), "sampling needs out_dir and out_pdf");
        need(n_eval == 0 || (eval_dirs && out_eval_pdf), "evaluation needs eval_dirs and out_eval_pdf");
        need(!(flags & NPMG_COSINE_AWARE) || normal, "NPMG_COSINE_AWARE needs normal");
        need(!(flags & NPMG_COSINE_AWARE) || ctx->cfg.d.mixture != MIX_ASG,
             "the anisotropic mixture has no cosine product");
        need(ctx->cfg.d.K == 7 || ctx->cfg.d.K == 8 || ctx->cfg.d.K == 16, "the mixture kernel supports 7, 8 or 16 lobes");
        need(!lobes_out || ctx->cfg.d.mixture != MIX_ASG, "lobe parameters are written for vMF lobes only");
        QueryIO io{};
        io.normal = (flags & NPMG_COSINE_AWARE) ? normal : nullptr;
        io.u = u;
        io.eval_dirs = eval_dirs;
        io.n_eval = (int)n_eval;
        io.out_dir = out_dir;
        io.out_pdf = out_pdf;
        io.out_eval_pdf = out_eval_pdf;
        io.out_lobe = out_lobe;
        io.lobes_out = lobes_out;
        hip_check(launch_mixture(ctx->cfg.d, raw, io, n, S(stream)), "mixture");
    });
}

npmg_status npmg_encode(npmg_ctx* ctx, const float* in, float* enc_out, uint32_t n, int use_inference_params,
                        npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n == 0 || (in && enc_out), "null buffer");
        const float* p = use_inference_params ? ctx->infer_params() : ctx->master;
        hip_check(launch_encode(ctx->cfg.d, p + ctx->cfg.n_mlp, in, enc_out, n, S(stream)), "encode");
    });
}

npmg_status npmg_grid_corners(npmg_ctx* ctx, const float* in, uint32_t* idx, float* w, uint32_t n,
                              npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(ctx->cfg.d.grid_enc >= 0, "configuration has no grid encoding");
        need(n == 0 || (in && idx && w), "null buffer");
        hip_check(launch_grid_corners(ctx->cfg.d, in, idx, w, n, S(stream)), "grid corners");
    });
}

npmg_status npmg_mixture_adjoint(npmg_ctx* ctx, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                                 const float* tgt_pdf, uint32_t n, float loss_scale, float* dout, float* loss,
                                 npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n == 0 || (raw && tgt_dir && tgt_rgb && tgt_pdf && dout && loss), "null buffer");
        need(!ctx->cfg.d.sel, "learned selection: use npmg_mixture_adjoint_selection");
        hip_check(launch_adjoint(ctx->cfg.d, raw, tgt_dir, tgt_rgb, tgt_pdf, nullptr,
                                 LossScale{nullptr, loss_scale, n ? 1.0f / (float)n : 0.0f}, dout, loss, n, S(stream)),
                  "adjoint");
    });
}

npmg_status npmg_mixture_adjoint_selection(npmg_ctx* ctx, const float* raw, const float* tgt_dir,
                                           const float* tgt_rgb, const float* tgt_pdf, const float* tgt_bsdf_pdf,
                                           uint32_t n, float loss_scale, float* dout, float* loss,
                                           npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(ctx->cfg.d.sel, "the configuration has no learned selection (mixture.learn_selection)");
        need(n == 0 || (raw && tgt_dir && tgt_rgb && tgt_pdf && tgt_bsdf_pdf && dout && loss), "null buffer");
        hip_check(launch_adjoint(ctx->cfg.d, raw, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf_pdf,
                                 LossScale{nullptr, loss_scale, n ? 1.0f / (float)n : 0.0f}, dout, loss, n, S(stream)),
                  "adjoint");
    });
}

npmg_status npmg_selection_probability(npmg_ctx* ctx, const float* raw, float* out, uint32_t n, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(ctx->cfg.d.sel, "the configuration has no learned selection (mixture.learn_selection)");
        need(n == 0 || (raw && out), "null buffer");
        QueryIO io{};
        io.out_sel = out;
        hip_check(launch_mixture(ctx->cfg.d, raw, io, n, S(stream)), "selection");
    });
}

npmg_status npmg_backward(npmg_ctx* ctx, const float* in, const float* dout, uint32_t n, float* dparams, float* denc,
                          npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        ctx->check_train();
        need(n > 0 && in && dout && dparams, "null buffer or empty batch");
        ctx->train_forward(in, nullptr, nullptr, nullptr, nullptr, n, LossScale{nullptr, 1.0f, 1.0f / (float)n},
                           S(stream));
        ctx->train_backward(in, dout, dparams, denc, n, false, S(stream));
    });
}

npmg_status npmg_optimizer_step(npmg_ctx* ctx, const float* dparams, float loss_scale, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(dparams != nullptr, "null gradients");
        need(loss_scale > 0, "loss_scale must be positive");
        ctx->optimizer_step(dparams, loss_scale, S(stream));
    });
}

}  // extern "C"

namespace {

// One train step, eager or as the captured graph (replayed while the buffers
// and n are those it was captured for).
void train_step_entry(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb, const float* tgt_pdf,
                      const float* tgt_bsdf, uint32_t n, float* loss_out, bool use_graph, hipStream_t stream) {
    ctx->check_train();
    need(n > 0 && in && tgt_dir && tgt_rgb && tgt_pdf, "null buffer or empty batch");
    need(!ctx->cfg.d.sel || tgt_bsdf, "learned selection: train with npmg_train_step_selection (BSDF pdfs)");
    if (!use_graph) {
        ctx->timed(NPMG_TIME_TRAIN_STEP, stream,
                   [&] { ctx->train_step(in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, n, loss_out, stream); });
        return;
    }
    const void* key[7] = {in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, loss_out, (const void*)(uintptr_t)n};
    if (!ctx->graph_exec || !std::equal(key, key + 7, ctx->graph_key)) {
        ctx->drop_graph();
        ctx->ensure_ws(n);  // every allocation happens before the capture
        hip_check(hipStreamBeginCapture(ctx->capture, hipStreamCaptureModeRelaxed), "begin capture");
        hipGraph_t g = nullptr;
        try {
            ctx->train_step(in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf, n, loss_out, ctx->capture);
        } catch (...) {
            (void)hipStreamEndCapture(ctx->capture, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(ctx->capture, &g), "end capture");
        ctx->graph = g;
        hip_check(hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0), "instantiate graph");
        std::copy(key, key + 7, ctx->graph_key);
    }
    ctx->timed(NPMG_TIME_TRAIN_STEP, stream,
               [&] { hip_check(hipGraphLaunch(ctx->graph_exec, stream), "launch graph"); });
}

}  // namespace

extern "C" {

npmg_status npmg_train_step(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                            const float* tgt_pdf, uint32_t n, float* loss_out, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        train_step_entry(ctx, in, tgt_dir, tgt_rgb, tgt_pdf, nullptr, n, loss_out, false, S(stream));
    });
}

int npmg_training_stride(const uint32_t* cur, const uint32_t* min_stride, uint32_t batch_size,
                         uint32_t batches_per_frame, uint32_t n_training_samples, uint32_t* out) {
    if (!cur || !min_stride || !out) return 0;
    // the reference's float arithmetic and unsigned comparisons, in its order
    const float tile = float(cur[0] * cur[1]);
    const float utilization = float(batch_size * batches_per_frame) / float(n_training_samples);
    // (bounded: with no batch budget the reference's loop below would not end)
    const float period = std::min(tile / utilization, 4194304.0f);
    const uint32_t side = std::max(1u, uint32_t(std::sqrt(std::max(0.f, period))));
    uint32_t x = std::max(side, min_stride[0]), y = std::max(side, min_stride[1]);
    while (period - float((x + 1) * y) >= 0) x += 1;
    bool accept = x * y < cur[0] * cur[1];
    accept |= std::max(x - cur[0], 0u) + std::max(y - cur[1], 0u) >= 1;
    out[0] = accept ? x : cur[0];
    out[1] = accept ? y : cur[1];
    return accept ? 1 : 0;
}

npmg_status npmg_train_batches(uint32_t n_samples, uint32_t batch_size, uint32_t batches_per_frame,
                               uint32_t granularity, uint32_t min_batch, uint32_t* offsets, uint32_t* sizes,
                               uint32_t max_batches, uint32_t* n_batches) {
    if (!n_batches || batch_size == 0 || granularity == 0 || (max_batches && !(offsets && sizes))) return NPMG_EINVAL;
    const uint32_t count = std::min(n_samples / batch_size + 1u, batches_per_frame);
    uint32_t k = 0;
    for (uint32_t it = 0; it < count; ++it) {
        uint32_t local = std::min(n_samples - it * batch_size, batch_size);
        local -= local % granularity;
        if (local < min_batch || local == 0) break;
        if (k >= max_batches) return NPMG_EINVAL;
        offsets[k] = it * batch_size;
        sizes[k] = local;
        ++k;
    }
    *n_batches = k;
    return NPMG_OK;
}

npmg_status npmg_train_records(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                               const float* tgt_pdf, const float* tgt_bsdf_pdf, uint32_t stride, uint32_t n_records,
                               uint32_t batch_size, uint32_t batches_per_frame, uint32_t granularity,
                               uint32_t min_batch, float* losses, uint32_t* n_batches, npmg_stream stream) {
    if (!ctx || !n_batches) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n_records <= stride, "n_records exceeds the row stride");
        need(in && tgt_dir && tgt_rgb && tgt_pdf, "null records");
        need(!ctx->cfg.d.sel || tgt_bsdf_pdf, "learned selection needs tgt_bsdf_pdf");
        std::vector<uint32_t> off(batches_per_frame ? batches_per_frame : 1), len(off.size());
        uint32_t nb = 0;
        need(npmg_train_batches(n_records, batch_size, batches_per_frame, granularity, min_batch, off.data(),
                                len.data(), (uint32_t)off.size(), &nb) == NPMG_OK,
             "invalid batch policy");
        *n_batches = nb;
        if (nb == 0) return;
        const uint32_t cap = *std::max_element(len.begin(), len.begin() + nb);
        if (cap > ctx->stage_cap) {
            if (ctx->stage) hip_check(hipFree(ctx->stage), "hipFree");
            ctx->stage = nullptr;
            ctx->stage_cap = 0;
            ctx->stage = dalloc<float>((size_t)16 * cap);
            ctx->stage_cap = cap;
        }
        const hipStream_t s = S(stream);
        for (uint32_t b = 0; b < nb; ++b) {
            const uint32_t m = len[b], o = off[b];
            float* st_in = ctx->stage;
            float* st_dir = st_in + (size_t)9 * m;
            float* st_rgb = st_dir + (size_t)2 * m;
            float* st_pdf = st_rgb + (size_t)3 * m;
            float* st_bsdf = st_pdf + m;
            auto rows = [&](float* dst, const float* src, int r) {
                hip_check(hipMemcpy2DAsync(dst, (size_t)m * 4, src + o, (size_t)stride * 4, (size_t)m * 4, (size_t)r,
                                           hipMemcpyDeviceToDevice, s),
                          "stage batch");
            };
            rows(st_in, in, 9);
            rows(st_dir, tgt_dir, 2);
            rows(st_rgb, tgt_rgb, 3);
            rows(st_pdf, tgt_pdf, 1);
            if (tgt_bsdf_pdf) rows(st_bsdf, tgt_bsdf_pdf, 1);
            train_step_entry(ctx, st_in, st_dir, st_rgb, st_pdf, ctx->cfg.d.sel ? st_bsdf : nullptr, m,
                             losses ? losses + b : nullptr, false, s);
        }
    });
}

npmg_status npmg_train_step_selection(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                                      const float* tgt_pdf, const float* tgt_bsdf_pdf, uint32_t n, int use_graph,
                                      float* loss_out, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(ctx->cfg.d.sel, "the configuration has no learned selection (mixture.learn_selection)");
        need(tgt_bsdf_pdf != nullptr, "null BSDF pdfs");
        train_step_entry(ctx, in, tgt_dir, tgt_rgb, tgt_pdf, tgt_bsdf_pdf, n, loss_out, use_graph != 0, S(stream));
    });
}

npmg_status npmg_timing_enable(npmg_ctx* ctx, int enable) {
    if (!ctx) return NPMG_EINVAL;
    ctx->timing = enable != 0;
    return NPMG_OK;
}

npmg_status npmg_timing_read(npmg_ctx* ctx, uint32_t tag, double* total_ms, uint32_t* launches) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(tag < (uint32_t)npmg_ctx::kTags, "unknown timing tag");
        ctx->timing_collect((int)tag);
        if (total_ms) *total_ms = ctx->t_ms[tag];
        if (launches) *launches = ctx->t_count[tag];
        ctx->t_ms[tag] = 0.0;
        ctx->t_count[tag] = 0;
    });
}

npmg_status npmg_train_step_graph(npmg_ctx* ctx, const float* in, const float* tgt_dir, const float* tgt_rgb,
                                  const float* tgt_pdf, uint32_t n, float* loss_out, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        train_step_entry(ctx, in, tgt_dir, tgt_rgb, tgt_pdf, nullptr, n, loss_out, true, S(stream));
    });
}

npmg_status npmg_train_state(const npmg_ctx* ctx, float* loss_scale, uint32_t* step) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(const_cast<npmg_ctx*>(ctx), [&] {
        ScaleState st{};
        hip_check(hipDeviceSynchronize(), "synchronize");
        hip_check(hipMemcpy(&st, ctx->sstate, sizeof(st), hipMemcpyDeviceToHost), "copy state");
        if (loss_scale) *loss_scale = st.scale;
        if (step) *step = st.step;
    });
}

npmg_status npmg_validate_pdf_integral(npmg_ctx* ctx, const float* raw, const float* normal, uint32_t n, uint32_t n_z,
                                       uint32_t n_phi, double* out, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n == 0 || (raw && out), "null buffer");
        need(n_z >= 2 && n_z <= 1u << 16 && n_phi >= 1 && n_phi <= 1u << 20, "quadrature size out of range");
        need(!normal || ctx->cfg.d.mixture != MIX_ASG, "the anisotropic mixture has no cosine product");
        need(ctx->cfg.d.K == 8 || ctx->cfg.d.K == 16, "the validation kernels support 8 or 16 lobes");
        if (ctx->gl_n != n_z) {
            std::vector<double> r = gauss_legendre(n_z);
            if (ctx->gl_nodes) hip_check(hipFree(ctx->gl_nodes), "hipFree");
            ctx->gl_nodes = nullptr;
            ctx->gl_n = 0;
            ctx->gl_nodes = dalloc<double>(r.size());
            hip_check(hipMemcpy(ctx->gl_nodes, r.data(), r.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
            ctx->gl_n = n_z;
        }
        hip_check(launch_pdf_integral(ctx->cfg.d, raw, normal, ctx->gl_nodes, ctx->gl_nodes + n_z, n_z, n_phi, out, n,
                                      S(stream)),
                  "pdf integral");
        hip_check(hipStreamSynchronize(S(stream)), "synchronize");
    });
}

npmg_status npmg_validate_loss_fd(npmg_ctx* ctx, const float* raw, const float* tgt_dir, const float* tgt_rgb,
                                  const float* tgt_pdf, uint32_t n, double h, double* out, npmg_stream stream) {
    if (!ctx) return NPMG_EINVAL;
    return guarded(ctx, [&] {
        need(n == 0 || (raw && tgt_dir && tgt_rgb && tgt_pdf && out), "null buffer");
        need(h > 0, "step must be positive");
        need(ctx->cfg.d.K == 8 || ctx->cfg.d.K == 16, "the validation kernels support 8 or 16 lobes");
        hip_check(launch_loss_fd(ctx->cfg.d, raw, tgt_dir, tgt_rgb, tgt_pdf, h, out, n, S(stream)), "loss fd");
        hip_check(hipStreamSynchronize(S(stream)), "synchronize");
    });
}

npmg_status npmg_synth_queries(uint64_t seed, uint64_t first, uint32_t n, float* in, float* normal, float* u,
                               float* eval_dirs, npmg_stream stream) {
    if (n && !(in && normal && u && eval_dirs)) return NPMG_EINVAL;
    return launch_synth_queries(seed, first, n, in, normal, u, eval_dirs, S(stream)) == hipSuccess ? NPMG_OK
                                                                                                   : NPMG_EHIP;
}

npmg_status npmg_synth_train(uint64_t seed, uint64_t first, uint32_t n, float* in, float* tgt_dir, float* tgt_rgb,
                             float* tgt_pdf, npmg_stream stream) {
    if (n && !(in && tgt_dir && tgt_rgb && tgt_pdf)) return NPMG_EINVAL;
    const FieldParams f = make_field(seed);
    return launch_synth_train(seed, first, n, f, in, tgt_dir, tgt_rgb, tgt_pdf, S(stream)) == hipSuccess ? NPMG_OK
                                                                                                       : NPMG_EHIP;
}

npmg_status npmg_debug_math(int fn, const float* x, const float* y, uint32_t n, float* out, npmg_stream stream) {
    if (n && !(x && out)) return NPMG_EINVAL;
    if (fn < 0 || fn > 12) return NPMG_EINVAL;
    return launch_debug_math(fn, x, y, n, out, S(stream)) == hipSuccess ? NPMG_OK : NPMG_EHIP;
}

}  // extern "C"
