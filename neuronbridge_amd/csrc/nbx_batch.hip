// Batched multi-instance engine: hundreds of independent LIF networks of uneven
// size and parameters in one launch.
//
// Workgroup w owns instance order[w] for the whole launch. Each thread owns NPT (8 or 16)
// consecutive neurons, whose state lives in VGPRs; the spike list lives in LDS. Per step
// the workgroup gathers the BF16 weight rows of the neurons that fired in the previous
// step (dense all-to-all, delay 1): each thread reads the 16-byte pieces of a fired row
// that cover its neurons, with kRows rows in flight, and accumulates them as int32 fixed
// point (the same integers the FP32 per-instance path produces from the same BF16-rounded
// weights, so rasters are identical). It then integrates and appends its new spikes to the
// next list (one LDS atomic per thread that fired). Instances never communicate, so a step
// needs only workgroup barriers, never a grid barrier. Chunks of steps are launched from a
// captured hipGraph; the workgroup size is picked by a timing autotuner.
//
// Weight rows are padded to a multiple of 16 elements (32 bytes), so every thread's
// 16-byte loads are aligned and stay inside its row.
//
// Delivery is a column gather, not a GEMM: each instance has its own W and a
// binary spike vector with a handful of ones per step, so reading only the
// fired rows (2 B per synaptic event) beats any dense product by orders of
// magnitude in bytes; MFMA has no multi-right-hand-side product to work on here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "nbx_internal.h"

using namespace nbx;

namespace {

constexpr int kMaxPoisTab = 32;
constexpr int kMaxBatchN = 16384;

constexpr int kRowPad = 16;  // weight row stride: a multiple of 16 elements

struct InstDev {
    int32_t n;
    int32_t ld;          // weight row stride (elements), n rounded up to kRowPad
    int32_t words;       // raster words per step
    int64_t w_off;       // element offset of W[pre][post] (bf16), row stride ld
    int64_t s_off;       // offset into the state arrays
    int64_t r_off;       // offset of this instance's raster rows (words)
    float tau;
    float scale, inv_scale;
    int32_t q_bg;
    uint32_t seed_lo, seed_hi;
    int32_t kmax;
    uint32_t tab[kMaxPoisTab];
};

struct BatchArgs {
    ModelParams P;  // shared parameters (tau taken per instance)
    const InstDev* inst;
    const int32_t* order;
    const uint16_t* w;
    float* v;
    int32_t* last;
    uint16_t* spk;        // [sum n] previous-step spike ids per instance (persisted between launches)
    int32_t* spk_cnt;     // [instances]
    uint32_t* raster;
    int32_t rec_t0, rec_n;
    unsigned long long* events;
    int32_t t0, nsteps;
    int32_t max_n;        // spike-list stride in LDS
};

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

template <int NPT>
__global__ __launch_bounds__(1024) void nbx_batch_kernel(BatchArgs B) {
    static_assert(NPT % 8 == 0, "a thread owns whole 16-byte pieces of a row");
    extern __shared__ uint16_t s_ids[];  // [2][max_n] spike lists, then [blockDim] raster bits
    __shared__ int32_t s_cnt[2];
    const int32_t inst_id = B.order[blockIdx.x];
    const InstDev& I = B.inst[inst_id];
    const int32_t n = I.n;
    const int32_t T = blockDim.x, tid = threadIdx.x;
    const int32_t c0 = tid * NPT;  // first neuron of this thread
    const bool own = c0 < n;
    uint16_t* s_rb = s_ids + 2 * B.max_n;
    ModelParams P = B.P;
    P.tau = I.tau;
    float V[NPT];
    int32_t last[NPT];
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        const int32_t i = c0 + j;
        V[j] = 0.0f;
        last[j] = 0;
        if (i < n) {
            V[j] = B.v[I.s_off + i];
            last[j] = B.last[I.s_off + i];
        }
    }
    const int32_t cnt0 = B.spk_cnt[inst_id];
    for (int32_t j = tid; j < cnt0; j += T) s_ids[j] = B.spk[I.s_off + j];
    if (tid == 0) s_cnt[0] = cnt0;
    __syncthreads();
    unsigned long long ev = 0;
    int cur = 0;
    const uint16_t* wbase = B.w + I.w_off + c0;
    for (int32_t k = 0; k < B.nsteps; ++k) {
        const int32_t t = B.t0 + k;
        const int32_t cnt = s_cnt[cur];
        const uint16_t* ids = s_ids + cur * B.max_n;
        // gather the fired rows over this thread's neurons, kRows rows in flight
        int32_t acc[NPT];
#pragma unroll
        for (int j = 0; j < NPT; ++j) acc[j] = 0;
        constexpr int kRows = (NPT == 8) ? 16 : 4;
        for (int32_t s0 = 0; s0 < cnt; s0 += kRows) {
            uint4 wv[kRows][NPT / 8];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const bool ok = own && s0 + r < cnt;
                const uint4* row = reinterpret_cast<const uint4*>(wbase + (int64_t)(ok ? ids[s0 + r] : 0) * I.ld);
#pragma unroll
                for (int h = 0; h < NPT / 8; ++h) wv[r][h] = ok ? row[h] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r)
#pragma unroll
                for (int h = 0; h < NPT / 8; ++h) {
                    const uint32_t q4[4] = {wv[r][h].x, wv[r][h].y, wv[r][h].z, wv[r][h].w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        acc[h * 8 + 2 * e] += quantize(bf16_to_f32((uint16_t)(q4[e] & 0xFFFFu)), I.scale);
                        acc[h * 8 + 2 * e + 1] += quantize(bf16_to_f32((uint16_t)(q4[e] >> 16)), I.scale);
                    }
                }
        }
        if (tid == 0) s_cnt[cur ^ 1] = 0;
        __syncthreads();
        uint32_t fired_bits = 0;
#pragma unroll
        for (int j = 0; j < NPT; ++j) {
            const int32_t i = c0 + j;
            if (i < n) {
                const int32_t a0 = acc[j] + poisson_count(I.seed_lo, I.seed_hi, i, t, I.tab, I.kmax) * I.q_bg;
                if (last[j] > P.t_ref) V[j] += (float)a0 * I.inv_scale;
                float g = 0.0f;
                if (lif_step<false>(P, V[j], g, 0.0f, last[j])) fired_bits |= 1u << j;
            }
        }
        uint16_t* nids = s_ids + (cur ^ 1) * B.max_n;
        if (fired_bits) {
            int32_t pos = atomicAdd(&s_cnt[cur ^ 1], __popc(fired_bits));
            for (uint32_t m = fired_bits; m; m &= m - 1u) nids[pos++] = (uint16_t)(c0 + __builtin_ctz(m));
        }
        const bool rec = B.raster && t >= B.rec_t0 && t < B.rec_t0 + B.rec_n;
        if (rec) s_rb[tid] = (uint16_t)fired_bits;
        __syncthreads();
        if (rec) {
            // raster word w holds the bits of threads [w * 32/NPT, (w + 1) * 32/NPT)
            uint32_t* row = B.raster + I.r_off + (int64_t)(t - B.rec_t0) * I.words;
            constexpr int kPer = 32 / NPT;
            for (int32_t w = tid; w < I.words; w += T) {
                uint32_t word = 0;
#pragma unroll
                for (int q = 0; q < kPer; ++q) {
                    const int32_t th = w * kPer + q;
                    if (th * NPT < n) word |= (uint32_t)s_rb[th] << (q * NPT);
                }
                row[w] = word;
            }
        }
        ev += (unsigned long long)s_cnt[cur ^ 1] * (unsigned long long)n;
        cur ^= 1;  // s_rb and the old list are rewritten only after the next step's first barrier
    }
    // persist state and the last step's spikes for the next launch
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        const int32_t i = c0 + j;
        if (i < n) {
            B.v[I.s_off + i] = V[j];
            B.last[I.s_off + i] = last[j];
        }
    }
    const int32_t cl = s_cnt[cur];
    for (int32_t j = tid; j < cl; j += T) B.spk[I.s_off + j] = s_ids[cur * B.max_n + j];
    if (tid == 0) {
        B.spk_cnt[inst_id] = cl;
        if (ev) atomicAdd(B.events, ev);
    }
}

void* batch_kernel(int npt) { return npt == 8 ? (void*)nbx_batch_kernel<8> : (void*)nbx_batch_kernel<16>; }

}  // namespace

struct nbx_batch {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    int32_t count = 0;
    ModelParams P{};
    std::vector<InstDev> h_inst;
    int64_t total_n = 0, total_w = 0, total_words = 0;
    int32_t max_n = 0;
    std::vector<void*> allocs;
    InstDev* d_inst = nullptr;
    int32_t* d_order = nullptr;
    uint16_t* d_w = nullptr;
    float* d_v = nullptr;
    int32_t* d_last = nullptr;
    uint16_t* d_spk = nullptr;
    int32_t* d_spk_cnt = nullptr;
    uint32_t* d_raster = nullptr;
    unsigned long long* d_events = nullptr;
    int32_t rec_t0 = 0, rec_n = 0;
    int32_t t = 0;
    int32_t threads = 0;  // 0 = not tuned yet
    int32_t chunk = 500;
    hipGraphExec_t gexec = nullptr;
    int32_t g_steps = 0;

    template <typename T>
    int32_t alloc(T** p, size_t count_) {
        void* q = nullptr;
        NBX_HIP_TRY(hipMalloc(&q, std::max<size_t>(1, count_) * sizeof(T)));
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return NBX_OK;
    }
    void release() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
    // neurons per thread: 8 or 16 (whole 16-byte row pieces); 32 means the size does not fit
    int npt_for(int32_t thr) const {
        const int need = (max_n + thr - 1) / thr;
        return need <= 8 ? 8 : (need <= 16 ? 16 : 32);
    }
    hipError_t launch(int32_t t0, int32_t nsteps, int32_t thr) {
        BatchArgs A{};
        A.P = P;
        A.inst = d_inst;
        A.order = d_order;
        A.w = d_w;
        A.v = d_v;
        A.last = d_last;
        A.spk = d_spk;
        A.spk_cnt = d_spk_cnt;
        A.raster = d_raster;
        A.rec_t0 = rec_t0;
        A.rec_n = rec_n;
        A.events = d_events;
        A.t0 = t0;
        A.nsteps = nsteps;
        A.max_n = max_n;
        void* fn = batch_kernel(npt_for(thr));
        const size_t lds = ((size_t)2 * max_n + (size_t)thr) * sizeof(uint16_t);
        void* args[] = {&A};
        return hipLaunchKernel(fn, dim3(count), dim3(thr), args, lds, stream);
    }
};

namespace {

int32_t poisson_tab(double lambda, uint32_t* tab, int32_t cap) {
    double p = std::exp(-lambda), cdf = 0.0;
    for (int k = 0; k < cap; ++k) {
        cdf += p;
        const double v = std::floor(cdf * 4294967296.0);
        const uint32_t th = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
        tab[k] = th;
        if (th == 0xFFFFFFFFu) return k;
        p = p * lambda / (double)(k + 1);
    }
    return cap - 1;
}

}  // namespace

extern "C" {

int32_t nbx_batch_create(const nbx_batch_desc* d, int32_t device, nbx_batch** out) {
    if (!d || !out || d->count <= 0 || !d->inst) {
        set_error("nbx_batch_create: bad arguments");
        return NBX_ERR_INVALID;
    }
    *out = nullptr;
    if (d->integrator != NBX_EULER && d->integrator != NBX_RK4) {
        set_error("nbx_batch_create: unknown integrator");
        return NBX_ERR_INVALID;
    }
    std::unique_ptr<nbx_batch> b(new nbx_batch());
    b->device = device;
    b->count = d->count;
    ModelParams& P = b->P;
    P.model = NBX_MODEL_LIF_VJ;
    P.integrator = d->integrator;
    P.t_ref = d->t_ref;
    P.dt = d->dt;
    P.hdt = d->dt * 0.5f;
    P.dt6 = d->dt / 6.0f;
    P.v_rest = d->v_rest;
    P.v_th = d->v_th;
    P.r = d->r;
    P.v_reset = d->v_reset;
    P.g_decay = 1.0f;
    for (int32_t k = 0; k < d->count; ++k) {
        const nbx_batch_instance& s = d->inst[k];
        if (s.n <= 0 || s.n > kMaxBatchN || !s.w_bf16 || s.tau == 0.0f || s.fixed_shift < 0 || s.fixed_shift > 30 ||
            s.pois_lambda < 0.0 || s.pois_lambda > 8.0) {
            set_error("nbx_batch_create: instance " + std::to_string(k) + " invalid (n in [1, 16384], weights, tau != 0, "
                      "shift in [0, 30], lambda in [0, 8])");
            return NBX_ERR_INVALID;
        }
        InstDev I{};
        I.n = s.n;
        I.ld = (s.n + kRowPad - 1) / kRowPad * kRowPad;
        I.words = (s.n + 31) / 32;
        I.w_off = b->total_w;
        I.s_off = b->total_n;
        I.r_off = 0;
        I.tau = s.tau;
        I.scale = std::ldexp(1.0f, s.fixed_shift);
        I.inv_scale = 1.0f / I.scale;
        I.q_bg = (int32_t)std::lrint(d->w_bg * I.scale);
        I.seed_lo = s.seed_lo;
        I.seed_hi = s.seed_hi;
        I.kmax = s.pois_lambda > 0.0 ? poisson_tab(s.pois_lambda, I.tab, kMaxPoisTab) : 0;
        b->h_inst.push_back(I);
        b->total_w += (int64_t)s.n * I.ld;
        b->total_n += s.n;
        b->total_words += I.words;
        b->max_n = std::max(b->max_n, s.n);
    }
    NBX_HIP_TRY(hipSetDevice(device));
    NBX_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int32_t rc = b->alloc(&b->d_w, (size_t)b->total_w);
    if (rc) return rc;
    NBX_HIP_TRY(hipMemset(b->d_w, 0, sizeof(uint16_t) * (size_t)b->total_w));  // row padding reads as 0
    std::vector<float> v((size_t)b->total_n);
    for (int32_t k = 0; k < d->count; ++k) {
        const nbx_batch_instance& s = d->inst[k];
        const InstDev& I = b->h_inst[(size_t)k];
        NBX_HIP_TRY(hipMemcpy2D(b->d_w + I.w_off, sizeof(uint16_t) * (size_t)I.ld, s.w_bf16,
                                sizeof(uint16_t) * (size_t)s.n, sizeof(uint16_t) * (size_t)s.n, (size_t)s.n,
                                hipMemcpyHostToDevice));
        for (int32_t i = 0; i < s.n; ++i) v[(size_t)(I.s_off + i)] = s.v0 ? s.v0[i] : d->v_reset;
    }
    std::vector<int32_t> last((size_t)b->total_n, 10000);
    if ((rc = b->alloc(&b->d_v, (size_t)b->total_n))) return rc;
    if ((rc = b->alloc(&b->d_last, (size_t)b->total_n))) return rc;
    if ((rc = b->alloc(&b->d_spk, (size_t)b->total_n))) return rc;
    if ((rc = b->alloc(&b->d_spk_cnt, (size_t)b->count))) return rc;
    if ((rc = b->alloc(&b->d_events, 1))) return rc;
    if ((rc = b->alloc(&b->d_inst, b->h_inst.size()))) return rc;
    if ((rc = b->alloc(&b->d_order, (size_t)b->count))) return rc;
    NBX_HIP_TRY(hipMemcpy(b->d_v, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_last, last.data(), sizeof(int32_t) * last.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemset(b->d_spk_cnt, 0, sizeof(int32_t) * b->count));
    NBX_HIP_TRY(hipMemset(b->d_events, 0, sizeof(unsigned long long)));
    NBX_HIP_TRY(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(InstDev) * b->h_inst.size(), hipMemcpyHostToDevice));
    // largest instances first: they bound the launch, so they start on the first wave of CUs
    std::vector<int32_t> order((size_t)b->count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return b->h_inst[(size_t)x].n > b->h_inst[(size_t)y].n; });
    NBX_HIP_TRY(hipMemcpy(b->d_order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice));
    *out = b.release();
    return NBX_OK;
}

int32_t nbx_batch_destroy(nbx_batch* b) {
    if (!b) return NBX_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->release();
    delete b;
    return NBX_OK;
}

int32_t nbx_batch_record(nbx_batch* b, int32_t t_begin, int32_t nsteps) {
    if (!b || nsteps < 0 || t_begin < 0) {
        set_error("nbx_batch_record: bad arguments");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->d_raster) {
        auto it = std::find(b->allocs.begin(), b->allocs.end(), (void*)b->d_raster);
        if (it != b->allocs.end()) b->allocs.erase(it);
        NBX_HIP_TRY(hipFree(b->d_raster));
        b->d_raster = nullptr;
    }
    b->rec_t0 = t_begin;
    b->rec_n = nsteps;
    int64_t off = 0;
    for (auto& I : b->h_inst) {
        I.r_off = off;
        off += (int64_t)I.words * nsteps;
    }
    NBX_HIP_TRY(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(InstDev) * b->h_inst.size(), hipMemcpyHostToDevice));
    if (nsteps) {
        int32_t rc = b->alloc(&b->d_raster, (size_t)off);
        if (rc) return rc;
        NBX_HIP_TRY(hipMemset(b->d_raster, 0, sizeof(uint32_t) * (size_t)off));
    }
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return NBX_OK;
}

int32_t nbx_batch_set_schedule(nbx_batch* b, int32_t threads, int32_t chunk) {
    if (!b || (threads != 256 && threads != 512 && threads != 1024) || chunk <= 0) {
        set_error("nbx_batch_set_schedule: threads must be 256, 512 or 1024 and chunk > 0");
        return NBX_ERR_INVALID;
    }
    if (b->npt_for(threads) > 16) {
        set_error("nbx_batch_set_schedule: too many neurons per thread for this workgroup size");
        return NBX_ERR_INVALID;
    }
    b->threads = threads;
    b->chunk = chunk;
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    return NBX_OK;
}

int32_t nbx_batch_get_schedule(const nbx_batch* b, int32_t* threads, int32_t* chunk) {
    if (!b || !threads || !chunk) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    *threads = b->threads;
    *chunk = b->chunk;
    return NBX_OK;
}

int32_t nbx_batch_run_async(nbx_batch* b, int32_t nsteps);

// Autotuner: times `steps` steps for each workgroup size on a scratch copy of the state
// (restored afterwards) and keeps the fastest.
int32_t nbx_batch_tune(nbx_batch* b, int32_t steps, int32_t* chosen) {
    if (!b || steps <= 0) {
        set_error("nbx_batch_tune: bad arguments");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    const size_t n = (size_t)b->total_n;
    std::vector<float> v(n);
    std::vector<int32_t> last(n), cnt((size_t)b->count);
    std::vector<uint16_t> spk(n);
    unsigned long long ev = 0;
    NBX_HIP_TRY(hipMemcpy(v.data(), b->d_v, sizeof(float) * n, hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(last.data(), b->d_last, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(spk.data(), b->d_spk, sizeof(uint16_t) * n, hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(cnt.data(), b->d_spk_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(&ev, b->d_events, sizeof(ev), hipMemcpyDeviceToHost));
    const int32_t saved_rec = b->rec_n;
    b->rec_n = 0;  // no raster writes while tuning
    float best = 1e30f;
    int32_t best_thr = 0;
    hipEvent_t e0, e1;
    NBX_HIP_TRY(hipEventCreate(&e0));
    NBX_HIP_TRY(hipEventCreate(&e1));
    for (int32_t thr : {256, 512, 1024}) {
        if (b->npt_for(thr) > 16) continue;
        NBX_HIP_TRY(b->launch(b->t, std::min(steps, 20), thr));  // warm
        NBX_HIP_TRY(hipEventRecord(e0, b->stream));
        NBX_HIP_TRY(b->launch(b->t, steps, thr));
        NBX_HIP_TRY(hipEventRecord(e1, b->stream));
        NBX_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.0f;
        NBX_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_thr = thr;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->rec_n = saved_rec;
    NBX_HIP_TRY(hipMemcpy(b->d_v, v.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_last, last.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_spk, spk.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_spk_cnt, cnt.data(), sizeof(int32_t) * cnt.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_events, &ev, sizeof(ev), hipMemcpyHostToDevice));
    if (best_thr == 0) {
        set_error("nbx_batch_tune: no workgroup size fits the largest instance");
        return NBX_ERR_UNSUPPORTED;
    }
    b->threads = best_thr;
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (chosen) *chosen = best_thr;
    return NBX_OK;
}

// Runs nsteps as a captured graph of `chunk`-step launches (the graph is rebuilt when the
// number of steps, the schedule or the recording window changes).
int32_t nbx_batch_run_async(nbx_batch* b, int32_t nsteps) {
    if (!b || nsteps < 0) {
        set_error("nbx_batch_run: bad arguments");
        return NBX_ERR_INVALID;
    }
    if (nsteps == 0) return NBX_OK;
    NBX_HIP_TRY(hipSetDevice(b->device));
    if (b->threads == 0) {  // untuned: the smallest workgroup that gives each thread at most 8 neurons
        int32_t thr = 1024;
        while (thr > 256 && b->max_n <= thr / 2 * 8) thr /= 2;
        b->threads = thr;
    }
    // step numbers are baked into the launches, so the graph is specific to (t, nsteps)
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    NBX_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int32_t done = 0; done < nsteps && e == hipSuccess; done += b->chunk)
        e = b->launch(b->t + done, std::min(b->chunk, nsteps - done), b->threads);
    hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error(std::string("batch graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : ec));
        return NBX_ERR_HIP;
    }
    hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NBX_HIP_TRY(ei);
    NBX_HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
    b->t += nsteps;
    return NBX_OK;
}

int32_t nbx_batch_sync(nbx_batch* b) {
    if (!b) {
        set_error("null batch");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    return NBX_OK;
}

int32_t nbx_batch_run(nbx_batch* b, int32_t nsteps) {
    int32_t rc = nbx_batch_run_async(b, nsteps);
    if (rc) return rc;
    return nbx_batch_sync(b);
}

int32_t nbx_batch_get_raster(nbx_batch* b, int32_t instance, uint32_t* raster, int64_t nwords) {
    if (!b || instance < 0 || instance >= b->count || !raster || !b->d_raster) {
        set_error("nbx_batch_get_raster: bad arguments or no recording");
        return NBX_ERR_INVALID;
    }
    const InstDev& I = b->h_inst[(size_t)instance];
    if (nwords != (int64_t)I.words * b->rec_n) {
        set_error("nbx_batch_get_raster: buffer must hold rec_steps * ceil(n/32) words");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    NBX_HIP_TRY(hipMemcpy(raster, b->d_raster + I.r_off, sizeof(uint32_t) * (size_t)nwords, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_batch_get_state(nbx_batch* b, int32_t instance, float* v, int32_t* last) {
    if (!b || instance < 0 || instance >= b->count || !v || !last) {
        set_error("nbx_batch_get_state: bad arguments");
        return NBX_ERR_INVALID;
    }
    const InstDev& I = b->h_inst[(size_t)instance];
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    NBX_HIP_TRY(hipMemcpy(v, b->d_v + I.s_off, sizeof(float) * I.n, hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(last, b->d_last + I.s_off, sizeof(int32_t) * I.n, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_batch_events(nbx_batch* b, int64_t* events) {
    if (!b || !events) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
    unsigned long long e = 0;
    NBX_HIP_TRY(hipMemcpy(&e, b->d_events, sizeof(e), hipMemcpyDeviceToHost));
    *events = (int64_t)e;
    return NBX_OK;
}

void* nbx_batch_stream(nbx_batch* b) { return b ? (void*)b->stream : nullptr; }

}  // extern "C"
