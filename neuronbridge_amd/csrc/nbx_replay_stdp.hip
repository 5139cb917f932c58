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
// spike-step lists (a few bytes per neuron) stay in L2.
//
// Spike lists per neuron, sorted, without a sort: each spike sets its bit in a 64-bit word per
// (64-step block, neuron) (a bit already set is a neuron listed twice in one step); a pass per
// neuron turns the words' popcounts into the list position of each block's first spike and the
// neuron's count; a scan gives the list starts; each spike then lands at start + block position +
// popcount of the lower bits of its word.
#include <hipcub/hipcub.hpp>

#include <mutex>
#include <string>

#include "nbx_internal.h"
#include "nbx_net.h"

namespace nbx {

namespace {

constexpr int32_t kNever = 0x7FFFFFFF;

struct ReplayStdpArgs {
    const int64_t* row_ptr;
    const uint32_t* syn;
    const uint8_t* plastic;
    float* w;
    TraceHead* heads;
    int32_t head_stride;
    const float2* tab;  // {exp(-k dt / tau+), exp(-k dt / tau-)}, k < tab_len
    int32_t tab_len;
    float max_ltp, max_ltd;
    const int32_t* st_ptr;  // [n + 1] spike lists per neuron
    const int32_t* st_k;    // local steps, ascending within a neuron
    const uint2* fix;       // [n][kFix / 4]: each neuron's first kFix steps (16-bit), 0xFFFF after the last
    int32_t n, lead, kend;  // local step k is step base + k; post-sides from k >= lead, arrivals in [lead, kend)
    int32_t base;
    unsigned long long* updates;  // plastic updates applied (or null)
    int32_t* err;
};

constexpr int32_t kFix = 20;  // spike steps per neuron in the fixed-stride copy (40 bytes)
constexpr uint16_t kEnd = 0xFFFF;

// Clears error bits (the raster error of an earlier call; other bits stay).
__global__ void rs_clear_err(int32_t* err, int32_t bits) { atomicAnd(err, ~bits); }

__global__ void rs_zero64(unsigned long long* p, int64_t count) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = 0ull;
}

__global__ void rs_fill16(uint16_t* p, int64_t count) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = kEnd;
}

__global__ void rs_interleave(const float* __restrict__ a, const float* __restrict__ b, int32_t len, float2* out) {
    for (int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); k < len; k += (int32_t)(gridDim.x * blockDim.x))
        out[k] = make_float2(a[k], b[k]);
}

// Workgroup k sets the bits of the spikes of local step k. Entries beyond tick_ptr[K] or beyond
// the workspace, ids out of range and bits already set (a neuron twice in one step, or step lists
// that overlap) raise kErrRaster.
__global__ void rs_mark(const int32_t* __restrict__ ids, const int32_t* __restrict__ tick_ptr, int32_t K, int64_t cap,
                        int32_t n, int32_t blocks, unsigned long long* mask, int32_t* err) {
    const int32_t total = tick_ptr[K];
    if (tick_ptr[0] != 0 || total < 0 || total > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, kErrRaster);
        return;
    }
    for (int32_t k = (int32_t)blockIdx.x; k < K; k += (int32_t)gridDim.x) {
        const int32_t b = tick_ptr[k], e = tick_ptr[k + 1];
        if (b < 0 || b > e || e > total) {  // step lists must tile [0, total): at most total bits are set
            if (threadIdx.x == 0) atomicOr(err, kErrRaster);
            continue;
        }
        const unsigned long long bit = 1ull << (k & 63);
        for (int32_t f = b + (int32_t)threadIdx.x; f < e; f += (int32_t)blockDim.x) {
            const int32_t id = ids[f];
            if (id < 0 || id >= n) {
                atomicOr(err, kErrRaster);
                continue;
            }
            if (atomicOr(&mask[(int64_t)(k >> 6) * n + id], bit) & bit) atomicOr(err, kErrRaster);
        }
    }
}

// Per neuron: list position of each block's first spike, and the neuron's spike count.
__global__ void rs_blocks(const unsigned long long* __restrict__ mask, int32_t n, int32_t blocks, int32_t* blk,
                          int32_t* cnt) {
    for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int32_t)(gridDim.x * blockDim.x)) {
        int32_t run = 0;
        for (int32_t q = 0; q < blocks; ++q) {
            blk[(int64_t)q * n + i] = run;
            run += __popcll(mask[(int64_t)q * n + i]);
        }
        cnt[i] = run;
    }
}

// Workgroup k places the spikes of local step k at their sorted list positions (the step lists
// were checked by rs_mark).
__global__ void rs_place(const int32_t* __restrict__ ids, const int32_t* __restrict__ tick_ptr, int32_t K, int32_t n,
                         int32_t blocks, const unsigned long long* __restrict__ mask, const int32_t* __restrict__ blk,
                         const int32_t* __restrict__ st_ptr, int32_t* st_k, uint16_t* fix, const int32_t* err) {
    if (*err & kErrRaster) return;
    for (int32_t k = (int32_t)blockIdx.x; k < K; k += (int32_t)gridDim.x) {
        const int32_t b = tick_ptr[k], e = tick_ptr[k + 1];
        const unsigned long long below = (1ull << (k & 63)) - 1ull;
        for (int32_t f = b + (int32_t)threadIdx.x; f < e; f += (int32_t)blockDim.x) {
            const int32_t id = ids[f];
            if (id < 0 || id >= n) continue;
            const int64_t q = (int64_t)(k >> 6) * n + id;
            const int32_t pos = blk[q] + __popcll(mask[q] & below);
            st_k[st_ptr[id] + pos] = k;
            if (pos < kFix) fix[(int64_t)id * kFix + pos] = (uint16_t)k;
        }
    }
}

// Orders one wave's LDS stores before its later LDS loads (other lanes' data): wavefront-scope
// release / acquire around a wave barrier.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The synapse pass: wave per presynaptic row (grid-stride), lane per synapse. Each plastic synapse
// merges its arrival list (the presynaptic spikes + delay) with its target's spike list and applies
// STDP.cpp:19-61 per entry; an arrival goes first when both fall on one step.
//
// Everything a merge step reads is in LDS, so the chain of dependent steps waits on no memory and
// runs without branches: the decay tables (FULL_TAB: all of them; else their first entries, longer
// step deltas reading the table in memory), the row's presynaptic spike steps once per wave and each
// lane's target spike steps (as [entry][lane]). Steps are 16-bit local steps (a call covers at most
// 65535 steps); each list ends with the sentinel 65535, past every step of the window, so a step
// only compares the two heads with the window end.
//
// The 64 synapses of the next group are loaded while the current group merges (index, plastic
// flag, weight, trace record), and a target's list comes from its fixed-stride copy (one load level
// after the index instead of two), so each group waits on one L2 round trip. A wave whose row or one
// of whose targets has kFix or more spikes runs the same merge from the lists in memory.
constexpr int32_t kPreCap = 64;
constexpr int32_t kTabPart = 2048;  // entries staged when the whole table does not fit
constexpr int32_t kTabFull = 4352;  // entries that fit (34 KB)
constexpr int kRsPassThreads = 1024;
constexpr int32_t kRsPassWaves = kRsPassThreads / 64;

struct MergeState {
    float w;
    TraceHead tr;
    int32_t since;  // last update, local steps
};

// One merge step on the state: the entry at local step t, an arrival when pre_side.
__device__ inline void merge_step(const ReplayStdpArgs& X, MergeState& s, bool pre_side, int32_t t, float2 f) {
    // STDP_State::SetUpdate on the tables (as stdp_decay_tab)
    s.tr.apre = s.tr.apre * f.x;
    s.tr.apost = s.tr.apost * f.y;
    s.since = t;
    // STDP.cpp:19-35 (arrival): Apre += 1, w -= A- * Apost; STDP.cpp:37-59 (spike): Apost += 1,
    // w += A+ * Apre; both clipped to [0, w_max]
    s.tr.apre = pre_side ? s.tr.apre + 1.0f : s.tr.apre;
    s.tr.apost = pre_side ? s.tr.apost : s.tr.apost + 1.0f;
    s.w = stdp_clip(s.w + (pre_side ? -X.max_ltd * s.tr.apost : X.max_ltp * s.tr.apre), s.tr.wmax);
}

// One synapse group's loads: the first level (index, flag, weight, record) of group j0.
struct Group {
    uint32_t en;
    uint32_t pl;
    float w;
    TraceHead tr;
};

__device__ inline void load_group(const ReplayStdpArgs& X, int64_t j, int64_t re, Group& g) {
    const int64_t jc = j < re ? j : re - 1;  // in bounds for every lane (re > rb)
    g.en = X.syn[jc];
    g.pl = j < re ? (uint32_t)X.plastic[jc] : 0u;
    g.w = 0.0f;
    g.tr = TraceHead{};
    if (g.pl) {  // a non-plastic synapse's weight and record stay in memory
        g.w = X.w[jc];
        g.tr = load_head(X.heads, X.head_stride, jc);
    }
}

template <bool FULL_TAB>
__global__ __launch_bounds__(kRsPassThreads) __attribute__((amdgpu_waves_per_eu(8))) void nbx_replay_stdp(
    ReplayStdpArgs X) {
    __shared__ float2 s_tab[FULL_TAB ? kTabFull : kTabPart];
    __shared__ uint16_t s_pre[kRsPassWaves][kPreCap];
    __shared__ uint16_t s_post[kRsPassWaves][kFix * 64];
    if (*X.err & kErrRaster) return;  // the same flag for every thread: a bad raster changes no synapse
    constexpr int32_t kTab = FULL_TAB ? kTabFull : kTabPart;
    const int32_t ntab = min(X.tab_len, kTab);
    for (int32_t k = (int32_t)threadIdx.x; k < ntab; k += kRsPassThreads) s_tab[k] = X.tab[k];
    __syncthreads();
    const int32_t lane = (int32_t)threadIdx.x & 63, wave = (int32_t)threadIdx.x >> 6;
    const int32_t nw = (int32_t)gridDim.x * kRsPassWaves;
    const int32_t last = X.tab_len - 1, kend = X.kend, lead = X.lead;
    uint16_t* sp = s_pre[wave];
    uint16_t* sq = s_post[wave] + lane;  // this lane's list: entry i at sq[64 i]
    auto table = [&](int32_t dt) -> float2 {
        if (FULL_TAB) return s_tab[dt];
        float2 f = s_tab[min(dt, kTabPart - 1)];
        if (dt >= kTabPart) f = X.tab[dt];
        return f;
    };
    unsigned long long upd = 0;
    for (int32_t pre = (int32_t)blockIdx.x * kRsPassWaves + wave; pre < X.n; pre += nw) {
        const int64_t rb = X.row_ptr[pre], re = X.row_ptr[pre + 1];
        if (rb == re) continue;  // wave-uniform
        const int32_t pb = X.st_ptr[pre], na = X.st_ptr[pre + 1] - pb;
        const bool pre_lds = na < kPreCap;  // wave-uniform
        // every lane of the wave is here (the row loop is wave-uniform)
        if (pre_lds) {
            for (int32_t e = lane; e < na; e += 64) sp[e] = (uint16_t)X.st_k[pb + e];
            if (lane == 0) sp[na] = kEnd;
        }
        wave_sync();
        // group j0 with its loads in g; loads group j0 + 64 into gn while it merges
        auto group = [&](int64_t j0, const Group& g, Group& gn) {
            const int64_t j = j0 + lane;
            const int32_t d = syn_delay(g.en), post = syn_post(g.en);
            const bool pl = g.pl != 0;
            const bool any = __any(pl);  // wave-uniform; rows without plastic synapses skip the lists
            // the lists first, then the next group: waiting for the lists leaves the group in flight
            uint2 lst[kFix / 4];
            if (any) {
#pragma unroll
                for (int e = 0; e < kFix / 4; ++e) lst[e] = X.fix[(int64_t)post * (kFix / 4) + e];
            }
            if (j0 + 64 < re) load_group(X, j0 + 64 + lane, re, gn);
            if (!any) return;
            const bool fast = pre_lds && !__any(pl && (lst[kFix / 4 - 1].y >> 16) != kEnd);  // wave-uniform
            if (fast) {
                if (pl) {
#pragma unroll
                    for (int e = 0; e < kFix / 4; ++e) {
                        sq[64 * (4 * e)] = (uint16_t)(lst[e].x & 0xFFFFu);
                        sq[64 * (4 * e + 1)] = (uint16_t)(lst[e].x >> 16);
                        sq[64 * (4 * e + 2)] = (uint16_t)(lst[e].y & 0xFFFFu);
                        sq[64 * (4 * e + 3)] = (uint16_t)(lst[e].y >> 16);
                    }
                    int32_t ia = 0, ib = 0;
                    if (lead > 0) {
                        while ((int32_t)sp[ia] + d < lead) ++ia;  // arrived before the window: applied earlier
                        while ((int32_t)sq[64 * ib] < lead) ++ib;  // post-sides of the lead steps: applied earlier
                    }
                    const int32_t ia0 = ia, ib0 = ib;
                    int32_t a = (int32_t)sp[ia] + d, b = sq[64 * ib];
                    MergeState s{g.w, g.tr, g.tr.slast - X.base};
                    while (min(a, b) < kend) {
                        const bool pre_side = a <= b;
                        const int32_t t = pre_side ? a : b;
                        merge_step(X, s, pre_side, t, table(min(max(t - s.since, 0), last)));
                        ia += pre_side ? 1 : 0;
                        ib += pre_side ? 0 : 1;
                        const int32_t an = (int32_t)sp[ia] + d, bn = sq[64 * ib];
                        a = pre_side ? an : a;
                        b = pre_side ? b : bn;
                    }
                    if (ia != ia0 || ib != ib0) {
                        upd += (unsigned long long)(ia - ia0 + ib - ib0);
                        s.tr.slast = s.since + X.base;
                        store_head(X.heads, X.head_stride, j, s.tr);
                        X.w[j] = s.w;
                    }
                }
            } else if (pl) {  // the same merge with the lists read from memory
                const int32_t qb = X.st_ptr[post], nb = X.st_ptr[post + 1] - qb;
                auto arr = [&](int32_t i) { return i < na ? X.st_k[pb + i] + d : kNever; };
                auto spk = [&](int32_t i) { return i < nb ? X.st_k[qb + i] : kNever; };
                int32_t ia = 0, ib = 0;
                while (arr(ia) < lead) ++ia;
                while (spk(ib) < lead) ++ib;
                const int32_t ia0 = ia, ib0 = ib;
                int32_t a = arr(ia), b = spk(ib);
                MergeState s{g.w, g.tr, g.tr.slast - X.base};
                while (min(a, b) < kend) {
                    const bool pre_side = a <= b;
                    const int32_t t = pre_side ? a : b;
                    const int32_t dt = min(max(t - s.since, 0), last);
                    merge_step(X, s, pre_side, t, X.tab[dt]);
                    if (pre_side) a = arr(++ia);
                    else b = spk(++ib);
                }
                if (ia != ia0 || ib != ib0) {
                    upd += (unsigned long long)(ia - ia0 + ib - ib0);
                    s.tr.slast = s.since + X.base;
                    store_head(X.heads, X.head_stride, j, s.tr);
                    X.w[j] = s.w;
                }
            }
        };
        // groups in pairs, so the two load buffers alternate without copies
        Group g0, g1;
        load_group(X, rb + lane, re, g0);
        for (int64_t j0 = rb; j0 < re; j0 += 128) {  // wave-uniform
            group(j0, g0, g1);
            if (j0 + 64 >= re) break;
            group(j0 + 64, g1, g0);
        }
        wave_sync();  // the next row's staging follows every read of this one
    }
    if (X.updates) {
        for (int off = 32; off > 0; off >>= 1) upd += __shfl_xor(upd, off);
        if (lane == 0 && upd) atomicAdd(X.updates, upd);
    }
}

}  // namespace

// Clears the raster error once it has been reported: an atomic on the network's stream, like the
// kernels that raise it (a host copy into the flag could race with the line the atomics hold).
int32_t clear_raster_error(nbx_net* net) { return clear_error_bits(net, kErrRaster); }

int32_t clear_error_bits(nbx_net* net, int32_t bits) {
    rs_clear_err<<<1, 1, 0, net->stream>>>(net->d_err, bits);
    NBX_HIP_TRY(hipGetLastError());
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    return NBX_OK;
}

template <typename T>
static int32_t regrow(nbx_net* net, T** p, size_t count) {
    if (*p) {
        auto it = std::find(net->allocs.begin(), net->allocs.end(), (void*)*p);
        if (it != net->allocs.end()) net->allocs.erase(it);
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        NBX_HIP_TRY(hipFree(*p));
        *p = nullptr;
    }
    return net->alloc(p, count);
}

// Workspace for n neurons, cap spikes and `blocks` 64-step blocks (grown on demand, kept with the
// network): n * blocks * 12 bytes of bit words and block positions, 4 bytes per spike.
static int32_t replay_stdp_workspace(nbx_net* net, int64_t cap, int32_t blocks) {
    ReplayStdpWork& R = net->rsw;
    const int32_t n = net->n;
    if (!R.ptr) {
        NBX_PASS(net->alloc(&R.cnt, (size_t)n + 1));
        NBX_PASS(net->alloc(&R.ptr, (size_t)n + 1));
        NBX_HIP_TRY(hipMemsetAsync(R.cnt, 0, sizeof(int32_t) * ((size_t)n + 1), net->stream));  // cnt[n] stays 0
        size_t bytes = 0;
        NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, R.cnt, R.ptr, n + 1, net->stream));
        NBX_PASS(net->alloc(&R.scan_tmp, bytes));
        R.scan_bytes = bytes;
        NBX_PASS(net->alloc(&R.tab2, (size_t)net->tab_len));
        rs_interleave<<<std::max(1, std::min((net->tab_len + 255) / 256, 256)), 256, 0, net->stream>>>(
            net->d_ltp_tab, net->d_ltd_tab, net->tab_len, R.tab2);
        NBX_HIP_TRY(hipGetLastError());
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    }
    if (cap > R.cap) {
        NBX_PASS(regrow(net, &R.k, (size_t)cap));
        R.cap = cap;
    }
    if (!R.fix) NBX_PASS(net->alloc(&R.fix, (size_t)n * kFix));
    const int64_t m = (int64_t)n * blocks;
    if (m > R.mcap) {
        NBX_PASS(regrow(net, &R.mask, (size_t)m));
        NBX_PASS(regrow(net, &R.blk, (size_t)m));
        R.mcap = m;
    }
    return NBX_OK;
}

// The launch sequence over a workspace R sized for n neurons, `blocks` 64-step blocks and R.cap
// spikes (cnt[n] must read 0).
static int32_t enqueue_replay_stdp(nbx_net* net, const ReplayStdpWork& R, const int32_t* d_ids,
                                   const int32_t* d_tick_ptr, int32_t t0, int32_t lead, int32_t nsteps,
                                   unsigned long long* d_updates) {
    const int32_t n = net->n, K = lead + nsteps, blocks = (K + 63) / 64;
    hipStream_t st = net->stream;
    const int64_t m = (int64_t)n * blocks;
    const int32_t kgrid = std::max(1, std::min(K, 4096));
    const int32_t ngrid = std::max(1, std::min((n + 255) / 256, 2048));
    rs_clear_err<<<1, 1, 0, st>>>(net->d_err, kErrRaster);
    NBX_HIP_TRY(hipGetLastError());
    rs_zero64<<<(int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 4096)), 256, 0, st>>>(R.mask, m);
    NBX_HIP_TRY(hipGetLastError());
    rs_mark<<<kgrid, 256, 0, st>>>(d_ids, d_tick_ptr, K, R.cap, n, blocks, R.mask, net->d_err);
    NBX_HIP_TRY(hipGetLastError());
    rs_blocks<<<ngrid, 256, 0, st>>>(R.mask, n, blocks, R.blk, R.cnt);
    NBX_HIP_TRY(hipGetLastError());
    size_t bytes = R.scan_bytes;
    NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(R.scan_tmp, bytes, R.cnt, R.ptr, n + 1, st));
    rs_fill16<<<std::max(1, std::min((n * kFix + 255) / 256, 4096)), 256, 0, st>>>(R.fix, (int64_t)n * kFix);
    NBX_HIP_TRY(hipGetLastError());
    rs_place<<<kgrid, 256, 0, st>>>(d_ids, d_tick_ptr, K, n, blocks, R.mask, R.blk, R.ptr, R.k, R.fix, net->d_err);
    NBX_HIP_TRY(hipGetLastError());

    ReplayStdpArgs X{};
    X.row_ptr = net->d_row_ptr;
    X.syn = net->d_syn;
    X.plastic = net->d_plastic;
    X.w = net->d_w;
    X.heads = net->d_heads;
    X.head_stride = net->head_stride;
    X.tab = R.tab2;
    X.tab_len = net->tab_len;
    X.max_ltp = net->sd.max_ltp;
    X.max_ltd = net->sd.max_ltd;
    X.st_ptr = R.ptr;
    X.st_k = R.k;
    X.fix = reinterpret_cast<const uint2*>(R.fix);
    X.n = n;
    X.lead = lead;
    X.kend = K;
    X.base = t0 - lead;
    X.updates = d_updates;
    X.err = net->d_err;
    const int32_t grid = std::max(1, std::min((n + kRsPassWaves - 1) / kRsPassWaves, 4096));
    if (net->tab_len <= kTabFull) nbx_replay_stdp<true><<<grid, kRsPassThreads, 0, st>>>(X);
    else nbx_replay_stdp<false><<<grid, kRsPassThreads, 0, st>>>(X);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

int32_t launch_replay_stdp(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0, int32_t lead,
                           int32_t nsteps, int64_t nspikes, unsigned long long* d_updates) {
    const int32_t K = lead + nsteps, blocks = (K + 63) / 64;
    NBX_PASS(replay_stdp_workspace(net, std::max<int64_t>(nspikes, 1), blocks));
    return enqueue_replay_stdp(net, net->rsw, d_ids, d_tick_ptr, t0, lead, nsteps, d_updates);
}

// ---------------------------------------------------------------- allocation-free variant
// The workspace of one call in a file-scope device array: each GPU holds its own copy for the
// life of the process. Host code never passes the symbol itself (in host code its name is only a
// stand-in): the device address comes from hipGetSymbolAddress, looked up once per device with
// that device current, on the entry point's first call there.
static __device__ __align__(256) unsigned char scratch[kReplayStdpScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int32_t scratch_address(int32_t device, unsigned char** out) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the scratch table");
        return NBX_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        NBX_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("hipGetSymbolAddress returned a null scratch address");
            return NBX_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return NBX_OK;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

int64_t replay_stdp_scratch_bytes(const nbx_net* net, int32_t K, int64_t nspikes, size_t* scan_bytes) {
    const int32_t n = net->n, blocks = (K + 63) / 64;
    size_t sb = 0;
    // a size query only: hipcub computes the temporary storage of the scan on the host
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, sb, (int32_t*)nullptr, (int32_t*)nullptr, n + 1, net->stream) !=
        hipSuccess)
        return -1;
    if (scan_bytes) *scan_bytes = sb;
    const size_t m = (size_t)n * blocks;
    return (int64_t)(round256(sizeof(int32_t) * ((size_t)n + 1)) * 2 +
                     round256(sizeof(int32_t) * (size_t)std::max<int64_t>(nspikes, 1)) +
                     round256(sizeof(unsigned long long) * m) + round256(sizeof(int32_t) * m) +
                     round256(sizeof(uint16_t) * (size_t)n * kFix) + round256(sizeof(float2) * (size_t)net->tab_len) +
                     round256(std::max<size_t>(sb, 1)));
}

int32_t launch_replay_stdp_noalloc(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0,
                                   int32_t lead, int32_t nsteps, int64_t nspikes, unsigned long long* d_updates) {
    const int32_t n = net->n, K = lead + nsteps, blocks = (K + 63) / 64;
    size_t sb = 0;
    const int64_t need = replay_stdp_scratch_bytes(net, K, nspikes, &sb);
    if (need < 0) {
        set_error("replay STDP: scan storage query failed");
        return NBX_ERR_HIP;
    }
    if (need > (int64_t)kReplayStdpScratchBytes) {
        set_error("replay STDP without allocation: the call needs " + std::to_string(need) +
                  " bytes of workspace, more than the static " + std::to_string(kReplayStdpScratchBytes) +
                  " (fewer steps or spikes per call)");
        return NBX_ERR_INVALID;
    }
    unsigned char* base = nullptr;
    NBX_PASS(scratch_address(net->device, &base));
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += round256(bytes);
        return p;
    };
    const size_t m = (size_t)n * blocks;
    ReplayStdpWork R;
    R.cnt = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * ((size_t)n + 1)));
    R.ptr = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * ((size_t)n + 1)));
    R.cap = std::max<int64_t>(nspikes, 1);
    R.k = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * (size_t)R.cap));
    R.mask = reinterpret_cast<unsigned long long*>(take(sizeof(unsigned long long) * m));
    R.blk = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * m));
    R.mcap = (int64_t)m;
    R.fix = reinterpret_cast<uint16_t*>(take(sizeof(uint16_t) * (size_t)n * kFix));
    R.tab2 = reinterpret_cast<float2*>(take(sizeof(float2) * (size_t)net->tab_len));
    R.scan_tmp = take(std::max<size_t>(sb, 1));
    R.scan_bytes = sb;
    // the array keeps the previous call's contents: the scan's last input and the decay table are
    // rewritten here (fill kernels, no memset nodes, so the call can be captured)
    NBX_HIP_TRY(launch_fill_i32(R.cnt + n, 0, 1, net->stream));
    rs_interleave<<<std::max(1, std::min((net->tab_len + 255) / 256, 256)), 256, 0, net->stream>>>(
        net->d_ltp_tab, net->d_ltd_tab, net->tab_len, R.tab2);
    NBX_HIP_TRY(hipGetLastError());
    return enqueue_replay_stdp(net, R, d_ids, d_tick_ptr, t0, lead, nsteps, d_updates);
}

}  // namespace nbx
