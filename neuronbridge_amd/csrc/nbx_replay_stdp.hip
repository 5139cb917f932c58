// Replay STDP: the pair-STDP updates that a recorded raster causes, applied to every plastic
// synapse in one pass over the synapses.
//
// In replay mode the spikes do not depend on the weights, so a synapse's trajectory over K steps
// is fixed by two sorted lists: the arrival steps of its presynaptic spikes (spike step + delay)
// and the spike steps of its target. Merging the two and applying STDP.cpp:19-61 at each entry
// (presynaptic arrival first when both fall on one step, as the closed loop delivers before it
// updates the neurons) reproduces, bit for bit, what K closed-loop steps with those spikes leave
// in the weight and the trace record. The replayed schedule is the pattern of
// tests/model_codegen/pair_stdp_gpu_trajectory.cpp:52-66 (pre-side, then post-side, per tick),
// here over a whole network.
//
// Per launch the kernel reads and writes each plastic synapse's weight and trace record once,
// whatever K is, instead of once per update: the HBM stream is the synapse table itself. The
// spike-time lists (a few bytes per neuron) stay in L2.
//
// Steps: per-neuron spike counts (atomics), their exclusive scan, a scatter of the local step
// of each spike into its neuron's list, a per-neuron sort (which also finds a neuron listed
// twice in one step), then the synapse pass: one wave per presynaptic row, a lane per synapse.
#include <hipcub/hipcub.hpp>

#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

constexpr int kRsThreads = 256;
constexpr int32_t kNever = 0x7FFFFFFF;

struct ReplayStdpArgs {
    const int64_t* row_ptr;
    const uint32_t* syn;
    const uint8_t* plastic;
    float* w;
    SynTrace* trace;
    const float* ltp_tab;
    const float* ltd_tab;
    int32_t tab_len;
    float max_ltp, max_ltd;
    const int32_t* st_ptr;  // [n + 1] spike lists per neuron
    const int32_t* st_k;    // local steps, ascending within a neuron
    int32_t n, lead, kend;  // local step k is step base + k; post-sides from k >= lead, arrivals in [lead, kend)
    int32_t base;
    unsigned long long* updates;  // plastic updates applied (or null)
    int32_t* err;
};

// Clears error bits (the raster error of an earlier call; other bits stay).
__global__ void rs_clear_err(int32_t* err, int32_t bits) { atomicAnd(err, ~bits); }

__global__ void rs_zero(int32_t* p, int32_t count) {
    for (int32_t j = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); j < count; j += (int32_t)(gridDim.x * blockDim.x))
        p[j] = 0;
}

// Spikes per neuron. The raster's length is checked against the workspace here; an id out of
// range raises kErrRaster and is not counted.
__global__ void rs_count(const int32_t* __restrict__ ids, const int32_t* __restrict__ tick_ptr, int32_t K, int64_t cap,
                         int32_t n, int32_t* cnt, int32_t* err) {
    const int32_t total = tick_ptr[K];
    if (tick_ptr[0] != 0 || total < 0 || total > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrRaster);
        return;
    }
    for (int32_t f = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); f < total; f += (int32_t)(gridDim.x * blockDim.x)) {
        const int32_t id = ids[f];
        if (id < 0 || id >= n) {
            atomicOr(err, kErrRaster);
            continue;
        }
        atomicAdd(&cnt[id], 1);
    }
}

// Workgroup k places the spikes of local step k into their neurons' lists (any order within a
// list; rs_sort orders them).
__global__ void rs_scatter(const int32_t* __restrict__ ids, const int32_t* __restrict__ tick_ptr, int32_t K, int32_t n,
                           const int32_t* __restrict__ st_ptr, int32_t* fill, int32_t* st_k, int32_t* err) {
    if (*err & kErrRaster) return;
    for (int32_t k = (int32_t)blockIdx.x; k < K; k += (int32_t)gridDim.x) {
        const int32_t b = tick_ptr[k], e = tick_ptr[k + 1];
        for (int32_t f = b + (int32_t)threadIdx.x; f < e; f += (int32_t)blockDim.x) {
            const int32_t id = ids[f];
            if (id < 0 || id >= n) continue;
            const int32_t pos = st_ptr[id] + atomicAdd(&fill[id], 1);
            if (pos < st_ptr[id + 1]) st_k[pos] = k;
            else atomicOr(err, kErrRaster);  // step lists that overlap (tick_ptr not ascending)
        }
    }
}

// Insertion sort of each neuron's list (a few entries); equal steps mean the neuron was listed
// twice in one step.
__global__ void rs_sort(const int32_t* __restrict__ st_ptr, int32_t* st_k, int32_t n, int32_t* err) {
    for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int32_t)(gridDim.x * blockDim.x)) {
        const int32_t b = st_ptr[i], e = st_ptr[i + 1];
        bool dup = false;
        for (int32_t j = b + 1; j < e; ++j) {
            const int32_t v = st_k[j];
            int32_t m = j - 1;
            while (m >= b && st_k[m] > v) {
                st_k[m + 1] = st_k[m];
                --m;
            }
            st_k[m + 1] = v;
            dup = dup || (m >= b && st_k[m] == v);
        }
        if (dup) atomicOr(err, kErrRaster);
    }
}

// Arrival step of list entry i (a synapse of delay d), or kNever past the window.
__device__ inline int32_t arrival_at(const ReplayStdpArgs& X, int32_t i, int32_t e, int32_t d) {
    if (i >= e) return kNever;
    const int32_t a = X.st_k[i] + d;
    return a < X.kend ? a : kNever;
}

// The synapse pass: wave per presynaptic row (grid-stride), lane per synapse.
__global__ __launch_bounds__(kRsThreads) void nbx_replay_stdp(ReplayStdpArgs X) {
    if (*X.err & kErrRaster) return;  // uniform: a bad raster changes no synapse
    constexpr int32_t kWaves = kRsThreads / 64;
    const int32_t lane = (int32_t)threadIdx.x & 63;
    const int32_t nw = (int32_t)gridDim.x * kWaves;
    unsigned long long upd = 0;
    for (int32_t pre = (int32_t)blockIdx.x * kWaves + ((int32_t)threadIdx.x >> 6); pre < X.n; pre += nw) {
        const int64_t rb = X.row_ptr[pre], re = X.row_ptr[pre + 1];
        const int32_t pb = X.st_ptr[pre], pe = X.st_ptr[pre + 1];
        for (int64_t j = rb + lane; j < re; j += 64) {
            if (!X.plastic[j]) continue;
            const uint32_t en = X.syn[j];
            const int32_t d = syn_delay(en), post = syn_post(en);
            int32_t ip = pb;
            while (ip < pe && X.st_k[ip] + d < X.lead) ++ip;  // arrived before the window: applied earlier
            int32_t iq = X.st_ptr[post];
            const int32_t qe = X.st_ptr[post + 1];
            while (iq < qe && X.st_k[iq] < X.lead) ++iq;  // post-sides of the lead steps: applied earlier
            int32_t a = arrival_at(X, ip, pe, d);
            int32_t b = iq < qe ? X.st_k[iq] : kNever;
            if (a == kNever && b == kNever) continue;
            float w = X.w[j];
            TraceHead tr = load_head(X.trace + j);
            do {
                ++upd;
                if (a <= b) {  // STDP.cpp:19-35: decay, Apre += 1, w -= A- * Apost, clip
                    stdp_decay_tab(X.ltp_tab, X.ltd_tab, X.tab_len, tr, X.base + a);
                    tr.apre += 1.0f;
                    w = stdp_clip(w + -X.max_ltd * tr.apost, tr.wmax);
                    a = arrival_at(X, ++ip, pe, d);
                } else {  // STDP.cpp:37-59: decay, Apost += 1, w += A+ * Apre, clip
                    stdp_decay_tab(X.ltp_tab, X.ltd_tab, X.tab_len, tr, X.base + b);
                    tr.apost += 1.0f;
                    w = stdp_clip(w + X.max_ltp * tr.apre, tr.wmax);
                    ++iq;
                    b = iq < qe ? X.st_k[iq] : kNever;
                }
            } while (a != kNever || b != kNever);
            store_head(X.trace + j, tr);
            X.w[j] = w;
        }
    }
    if (X.updates) {
        for (int off = 32; off > 0; off >>= 1) upd += __shfl_xor(upd, off);
        if (lane == 0 && upd) atomicAdd(X.updates, upd);
    }
}

}  // namespace

// Clears the raster error once it has been reported: an atomic on the network's stream, like the
// kernels that raise it (a host copy into the flag could race with the line the atomics hold).
int32_t clear_raster_error(nbx_net* net) {
    rs_clear_err<<<1, 1, 0, net->stream>>>(net->d_err, kErrRaster);
    NBX_HIP_TRY(hipGetLastError());
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    return NBX_OK;
}

// Workspace sized for n neurons and cap spikes (grown on demand, kept with the network).
static int32_t replay_stdp_workspace(nbx_net* net, int64_t cap) {
    ReplayStdpWork& R = net->rsw;
    const int32_t n = net->n;
    if (!R.ptr) {
        NBX_PASS(net->alloc(&R.cnt, (size_t)n + 1));
        NBX_PASS(net->alloc(&R.ptr, (size_t)n + 1));
        size_t bytes = 0;
        NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, R.cnt, R.ptr, n + 1, net->stream));
        NBX_PASS(net->alloc(&R.scan_tmp, bytes));
        R.scan_bytes = bytes;
    }
    if (cap > R.cap) {
        if (R.k) {
            auto it = std::find(net->allocs.begin(), net->allocs.end(), (void*)R.k);
            if (it != net->allocs.end()) net->allocs.erase(it);
            NBX_HIP_TRY(hipStreamSynchronize(net->stream));
            NBX_HIP_TRY(hipFree(R.k));
            R.k = nullptr;
            R.cap = 0;
        }
        NBX_PASS(net->alloc(&R.k, (size_t)cap));
        R.cap = cap;
    }
    return NBX_OK;
}

int32_t launch_replay_stdp(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0, int32_t lead,
                           int32_t nsteps, int64_t nspikes, unsigned long long* d_updates) {
    NBX_PASS(replay_stdp_workspace(net, std::max<int64_t>(nspikes, 1)));
    ReplayStdpWork& R = net->rsw;
    hipStream_t st = net->stream;
    const int32_t n = net->n, K = lead + nsteps;
    rs_clear_err<<<1, 1, 0, st>>>(net->d_err, kErrRaster);
    NBX_HIP_TRY(hipGetLastError());
    rs_zero<<<std::max(1, std::min((n + 256) / 256, 1024)), 256, 0, st>>>(R.cnt, n + 1);
    NBX_HIP_TRY(hipGetLastError());
    rs_count<<<std::max<int64_t>(1, std::min<int64_t>((nspikes + 255) / 256, 2048)), 256, 0, st>>>(
        d_ids, d_tick_ptr, K, R.cap, n, R.cnt, net->d_err);
    NBX_HIP_TRY(hipGetLastError());
    size_t bytes = R.scan_bytes;
    NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(R.scan_tmp, bytes, R.cnt, R.ptr, n + 1, st));
    rs_zero<<<std::max(1, std::min((n + 256) / 256, 1024)), 256, 0, st>>>(R.cnt, n + 1);
    NBX_HIP_TRY(hipGetLastError());
    rs_scatter<<<std::max(1, std::min(K, 4096)), 256, 0, st>>>(d_ids, d_tick_ptr, K, n, R.ptr, R.cnt, R.k, net->d_err);
    NBX_HIP_TRY(hipGetLastError());
    rs_sort<<<std::max(1, std::min((n + 255) / 256, 2048)), 256, 0, st>>>(R.ptr, R.k, n, net->d_err);
    NBX_HIP_TRY(hipGetLastError());

    ReplayStdpArgs X{};
    X.row_ptr = net->d_row_ptr;
    X.syn = net->d_syn;
    X.plastic = net->d_plastic;
    X.w = net->d_w;
    X.trace = net->d_trace;
    X.ltp_tab = net->d_ltp_tab;
    X.ltd_tab = net->d_ltd_tab;
    X.tab_len = net->tab_len;
    X.max_ltp = net->sd.max_ltp;
    X.max_ltd = net->sd.max_ltd;
    X.st_ptr = R.ptr;
    X.st_k = R.k;
    X.n = n;
    X.lead = lead;
    X.kend = K;
    X.base = t0 - lead;
    X.updates = d_updates;
    X.err = net->d_err;
    const int32_t grid = std::max(1, std::min((n + 3) / 4, 8192));
    nbx_replay_stdp<<<grid, kRsThreads, 0, st>>>(X);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

}  // namespace nbx
