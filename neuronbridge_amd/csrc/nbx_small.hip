// Single-workgroup engine for small networks with dense all-to-all delivery (the reference's
// 1000-neuron Izhikevich network, configs up to 1024 neurons).
//
// A network this size fits one workgroup: thread i owns neuron i, its state stays in registers
// for the whole launch and the delay ring lives in LDS ([slot][channel][neuron], each thread
// reading and clearing only its own column, so the ring needs no barrier). Per step: the update
// (neuron_tick, as nbx_update), wave ballots compacted through LDS into the step's ascending spike
// list (the firing history, counts and raster are written as nbx_update writes them), then the
// dense delivery of those spikes (as nbx_deliver_dense: each thread sums the quantized column
// entries of its neuron). Two workgroup barriers per step replace the two kernel launches per
// step of the graph engine; the ring and state are read from and written back to the network's
// buffers at the launch's ends, so the engines interoperate.
#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

constexpr int kSmallThreads = 1024;
constexpr int kSmallWaves = kSmallThreads / 64;
constexpr int64_t kSmallRingBytes = 96 * 1024;

struct SmallArgs {
    StepArgs A;
    int32_t t0, nsteps;
};

template <int MODEL>
__global__ __launch_bounds__(kSmallThreads) void nbx_small_dense(SmallArgs X) {
    extern __shared__ int32_t s_ring[];  // [D][nch][n]
    __shared__ int32_t s_ids[kSmallThreads];
    __shared__ int32_t s_wcnt[kSmallWaves];
    const StepArgs& A = X.A;
    const int32_t n = A.n, D = A.D, nch = A.nch, words = D * nch * n;
    const int32_t i = (int32_t)threadIdx.x, lane = i & 63, wave = i >> 6;
    const bool own = i < n;
    for (int32_t q = i; q < words; q += kSmallThreads) s_ring[q] = A.ring[q];
    float V = 0.0f, S1 = 0.0f, GA = 0.0f, GG = 0.0f, I = 0.0f;
    int32_t last = 0;
    if (own) {
        V = A.v[i];
        if (model_uses_s1(MODEL)) S1 = A.s1[i];
        if (model_uses_g2(MODEL)) {
            GA = A.ga[i];
            GG = A.gg[i];
        }
        I = A.iext[i];
        last = A.last[i];
    }
    unsigned long long ev = 0;
    __syncthreads();
    for (int32_t k = 0; k < X.nsteps; ++k) {
        const int32_t t = X.t0 + k, slot = t % D;
        bool fired = false;
        if (own) {  // nbx_update: this step's arrivals, cleared, then the neuron update
            int32_t* r = s_ring + slot * nch * n;
            const int32_t a0 = r[i];
            r[i] = 0;
            int32_t a1 = 0, a2 = 0;
            if (nch > 1) {
                a1 = r[n + i];
                r[n + i] = 0;
            }
            if (nch > 2) {
                a2 = r[2 * n + i];
                r[2 * n + i] = 0;
            }
            fired = neuron_tick<MODEL>(A, i, t, a0, a1, a2, V, S1, GA, GG, I, last);
        }
        const uint64_t mask = __ballot(fired);
        if (lane == 0) s_wcnt[wave] = __popcll(mask);
        if (A.raster && lane == 0 && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n) {
            const int32_t i0 = wave * 64;  // 64 consecutive neurons per wavefront: two aligned words
            uint32_t* row = A.raster + (int64_t)(t - A.rec_t0) * A.words;
            if (i0 < n) row[i0 >> 5] = (uint32_t)mask;
            if (i0 + 32 < n) row[(i0 >> 5) + 1] = (uint32_t)(mask >> 32);
        }
        __syncthreads();  // wave counts written; the previous step's list read by every thread
        int32_t base = 0, tot = 0;
        for (int w = 0; w < kSmallWaves; ++w) {
            const int32_t c = s_wcnt[w];
            base += w < wave ? c : 0;
            tot += c;
        }
        if (fired) {
            const int32_t pos = base + __popcll(mask & ((1ull << lane) - 1ull));
            s_ids[pos] = i;
            A.hist[(int64_t)slot * n + pos] = i;
        }
        if (i == 0) {
            A.hist_cnt[slot] = tot;
            A.hist_cnt[(t + 1) % D] = 0;  // held step t + 1 - D, which nothing reads any more
            if (tot && A.counts && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n) A.counts[t - A.rec_t0] += tot;
            ev += (unsigned long long)tot * (unsigned long long)n;  // synaptic events, counted at the spike
        }
        __syncthreads();  // the step's spike list is in LDS; the wave counts were read by every thread
        if (own && tot > 0) {  // nbx_deliver_dense
            int32_t pos = 0, neg = 0;
            for (int32_t s = 0; s < tot; ++s) {
                const float w = A.wt[(int64_t)s_ids[s] * n + i];
                if (nch > 1 && w < 0.0f) neg += quantize(-w, A.scale);
                else pos += quantize(w, A.scale);
            }
            int32_t* r = s_ring + ((t + A.dense_delay) % D) * nch * n;
            r[i] += pos;
            if (nch > 1 && neg) r[n + i] += neg;
        }
    }
    if (own) {
        A.v[i] = V;
        if (model_uses_s1(MODEL)) A.s1[i] = S1;
        if (model_uses_g2(MODEL)) {
            A.ga[i] = GA;
            A.gg[i] = GG;
        }
        A.iext[i] = I;
        A.last[i] = last;
    }
    __syncthreads();
    for (int32_t q = i; q < words; q += kSmallThreads) A.ring[q] = s_ring[q];
    if (i == 0 && ev) atomicAdd(A.events, ev);
}

template <int MODEL>
void* small_kernel() {
    return (void*)nbx_small_dense<MODEL>;
}

void* small_kernel_for(int model) {
    switch (model) {
        case NBX_MODEL_LIF_VJ: return small_kernel<NBX_MODEL_LIF_VJ>();
        case NBX_MODEL_LIF_EXP: return small_kernel<NBX_MODEL_LIF_EXP>();
        case NBX_MODEL_IZH: return small_kernel<NBX_MODEL_IZH>();
        case NBX_MODEL_LIF_EXP2: return small_kernel<NBX_MODEL_LIF_EXP2>();
        case NBX_MODEL_LIF_EXP3: return small_kernel<NBX_MODEL_LIF_EXP3>();
        case NBX_MODEL_POISSON: return small_kernel<NBX_MODEL_POISSON>();
        case NBX_MODEL_TRIGGER_RELAY: return small_kernel<NBX_MODEL_TRIGGER_RELAY>();
        default: return small_kernel<NBX_MODEL_IZH_CUR>();
    }
}

}  // namespace

std::string small_dense_reason(const nbx_net* net) {
    if (!net->has_dense || net->has_sparse) return "needs dense delivery only";
    if (net->stdp || !net->rules.empty()) return "no learning on the single-workgroup engine";
    if (net->n > kSmallThreads) return "more than " + std::to_string(kSmallThreads) + " neurons";
    if ((int64_t)net->D * net->nch * net->n * 4 > kSmallRingBytes) return "delay ring larger than 96 KB";
    return "";
}

int32_t run_small_dense(nbx_net* net, int32_t nsteps) {
    if (nsteps == 0) return NBX_OK;
    void* fn = small_kernel_for(net->md.model);
    const size_t lds = (size_t)net->D * net->nch * net->n * sizeof(int32_t);
    if (lds > 64 * 1024) NBX_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SmallArgs X;
    X.A = net->args();
    X.t0 = net->t;
    X.nsteps = nsteps;
    void* kargs[] = {&X};
    NBX_HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(kSmallThreads), kargs, lds, net->stream));
    net->t += nsteps;
    return NBX_OK;
}

}  // namespace nbx
