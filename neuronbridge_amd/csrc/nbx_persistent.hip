// Persistent closed-loop engine and replay delivery on the post-blocked layout.
//
// One cooperative launch runs many steps. Workgroup b owns the neurons
// [b*P, b*P + P): their state stays in VGPRs and their delay ring
// ring[D][channel][P] stays in LDS for the whole launch. Per step t a workgroup
//   1. delivers the spikes of step t-1 that target its own neurons into its LDS
//      ring (int32 fixed point, ds_add, so the sum is order independent),
//   2. updates its neurons for step t (same neuron_tick as the per-phase engine)
//      and compacts its spikes (wavefront ballot + LDS prefix) into the shared
//      firing-history ring in HBM,
//   3. joins one grid barrier, after which every workgroup sees all spikes of t.
// One barrier per step replaces the reference's separate clear/mark/drain/update
// launches and host round trips (GpuPropagationRuntime.cu:1687-1827,
// DenseUnifiedNeuronRuntime.cu:284-372). Workgroups are numbered so that
// consecutive neuron blocks share an XCD and its L2: a presynaptic row's segments
// for neighbouring blocks are adjacent in memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nbx_net.h"

namespace nbx {

namespace {

constexpr int kXcds = 8;
constexpr int kMaxThreads = 1024;
constexpr int kMaxNpt = 4;
constexpr uint32_t kSpinLimit = 1u << 24;  // ~0.5 s of s_sleep(1) spins before giving up
constexpr int kLanes = 4;                  // lanes sharing one spike's synapse segment


__device__ inline int32_t block_of(int32_t w, int32_t B) {
    const int32_t per = B / kXcds;  // B is a multiple of kXcds
    return (w % kXcds) * per + (w / kXcds);
}

// Grid barrier, two-level: workgroups of one XCD group (blockIdx % 8) arrive on their
// group counter; the last arrival of a group arrives on the top counter; the last top
// arrival raises every group's generation flag. Counters only grow (targets scale with
// the step), each sits on its own 128-byte line, and the spin is bounded: on time-out
// the error flag is raised and every waiter leaves.
//
// Memory ordering: the only data other workgroups read within a launch are the firing
// ids and counts of the previous step. They are written with device-scope atomic
// stores/adds (coherence-point writes, `sc1`) and read with device-scope atomic loads,
// and each arrival is preceded by s_waitcnt 0, so every store of a workgroup has been
// performed before its arrival is counted. No L2 write-back or invalidate is needed.
constexpr int kLine = 32;  // int32 per 128-byte line
struct Barrier {
    int32_t group[kXcds * kLine];
    int32_t top[kLine];
    int32_t gen[kXcds * kLine];
};

struct PersistArgs {
    StepArgs A;
    const int32_t* off;
    const uint2* seg;
    int32_t B, P;
    int32_t t0, nsteps;
    Barrier* bar;
};

__device__ inline int32_t ld_dev(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline bool grid_barrier(Barrier* bar, int32_t k, int32_t nblocks, int32_t* err, int32_t* s_abort) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_s_waitcnt(0);  // this workgroup's published stores are performed
        const int32_t g = blockIdx.x % kXcds;
        const int32_t per = nblocks / kXcds;
        const int32_t old = __hip_atomic_fetch_add(&bar->group[g * kLine], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == per * (k + 1) - 1) {
            const int32_t o2 = __hip_atomic_fetch_add(&bar->top[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (o2 == kXcds * (k + 1) - 1)
                for (int q = 0; q < kXcds; ++q) st_dev(&bar->gen[q * kLine], k + 1);
        }
        uint32_t spins = 0;
        bool abort = false;
        while (ld_dev(&bar->gen[g * kLine]) < k + 1) {
            if (ld_dev(err) != 0) {
                abort = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) {
                atomicExch(err, 1);
                abort = true;
                break;
            }
        }
        *s_abort = abort ? 1 : 0;
    }
    __syncthreads();
    return *s_abort == 0;
}

template <int MODEL, int NPT>
__global__ __launch_bounds__(kMaxThreads) void nbx_persistent(PersistArgs X) {
    extern __shared__ int32_t s_ring[];  // [D][nch][P]
    __shared__ int32_t s_wcnt[kMaxThreads / 64];
    __shared__ int32_t s_base;
    __shared__ int32_t s_abort;
    __shared__ unsigned long long s_ev;
    const StepArgs& A = X.A;
    const int32_t P = X.P, D = A.D, nch = A.nch, n = A.n;
    const int32_t b = block_of(blockIdx.x, X.B);
    const int32_t p0 = b * P;
    const int32_t np = max(0, min(P, n - p0));
    const int32_t T = blockDim.x;
    const int32_t tid = threadIdx.x;
    const int32_t lane = tid & 63, wave = tid >> 6, nwaves = T >> 6;
    const int32_t gl = tid % kLanes, grp = tid / kLanes, ngrp = T / kLanes;

    // load this block's slice of the delay ring and its neurons' state
    for (int32_t j = tid; j < D * nch * P; j += T) {
        const int32_t pl = j % P, sc = j / P;  // sc = slot * nch + ch
        s_ring[j] = (pl < np) ? A.ring[(int64_t)sc * n + p0 + pl] : 0;
    }
    if (tid == 0) s_ev = 0;
    float V[NPT], S1[NPT], GA[NPT], GG[NPT], I[NPT];
    int32_t last[NPT];
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int32_t pl = tid + u * T;
        const int32_t i = p0 + pl;
        V[u] = S1[u] = GA[u] = GG[u] = I[u] = 0.0f;
        last[u] = 0;
        if (pl < np) {
            V[u] = A.v[i];
            if (MODEL != NBX_MODEL_LIF_VJ) S1[u] = A.s1[i];
            if (MODEL == NBX_MODEL_IZH) {
                GA[u] = A.ga[i];
                GG[u] = A.gg[i];
            }
            I[u] = A.iext[i];
            last[u] = A.last[i];
        }
    }
    unsigned long long ev = 0;
    __syncthreads();

    for (int32_t k = 0; k <= X.nsteps; ++k) {
        const int32_t t = X.t0 + k;
        if (k > 0) {
            // spikes of step t-1 were complete when the previous barrier opened
            const int32_t ps = (t - 1) % D;
            const int32_t cnt = ld_dev(&A.hist_cnt[ps]);
            const int32_t* ids = A.hist + (int64_t)ps * n;
            if (b == 0 && tid == 0 && A.counts && t - 1 >= A.rec_t0 && t - 1 < A.rec_t0 + A.rec_n)
                A.counts[t - 1 - A.rec_t0] = cnt;
            // lanes in groups of kLanes share one spike so its segment entries coalesce
            for (int32_t base = grp; base < cnt; base += ngrp * 2) {
                int32_t pre[2], o0[2], o1[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) pre[r] = (base + r * ngrp < cnt) ? ld_dev(&ids[base + r * ngrp]) : -1;
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    o0[r] = o1[r] = 0;
                    if (pre[r] >= 0) {
                        const int32_t* op = X.off + (int64_t)pre[r] * (X.B + 1) + b;
                        o0[r] = op[0];
                        o1[r] = op[1];
                    }
                }
                uint2 e[2][2];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const int32_t j = o0[r] + gl + u * kLanes;
                        e[r][u] = (j < o1[r]) ? X.seg[j] : make_uint2(0u, 0u);
                    }
                auto deposit = [&](const uint2 x) {
                    const int32_t d = (int32_t)((x.x >> 16) & 0xFFu);
                    atomicAdd(&s_ring[(((t - 1 + d) % D) * nch + (int32_t)(x.x >> 24)) * P + (int32_t)(x.x & 0xFFFFu)],
                              (int32_t)x.y);
                };
#pragma unroll
                for (int r = 0; r < 2; ++r) {
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        if (o0[r] + gl + u * kLanes < o1[r]) deposit(e[r][u]);
                    for (int32_t j = o0[r] + gl + 2 * kLanes; j < o1[r]; j += kLanes) deposit(X.seg[j]);
                    if (gl == 0) ev += (unsigned long long)(o1[r] - o0[r]);
                }
            }
        }
        if (k == X.nsteps) break;
        __syncthreads();
        const int32_t slot = t % D;
        const bool rec = A.raster && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n;
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            const int32_t pl = tid + u * T;
            const int32_t i = p0 + pl;
            bool fired = false;
            if (pl < np) {
                const int32_t a0 = s_ring[(slot * nch) * P + pl];
                s_ring[(slot * nch) * P + pl] = 0;
                int32_t a1 = 0;
                if (nch > 1) {
                    a1 = s_ring[(slot * nch + 1) * P + pl];
                    s_ring[(slot * nch + 1) * P + pl] = 0;
                }
                fired = neuron_tick<MODEL>(A, i, t, a0, a1, V[u], S1[u], GA[u], GG[u], I[u], last[u]);
            }
            const uint64_t mask = __ballot(fired);
            if (lane == 0) s_wcnt[wave] = __popcll(mask);
            if (rec && lane == 0) {
                const int32_t pw = u * T + wave * 64;  // first local neuron of this wavefront
                uint32_t* row = A.raster + (int64_t)(t - A.rec_t0) * A.words;
                if (pw < np) row[(p0 + pw) >> 5] = (uint32_t)mask;
                if (pw + 32 < np) row[((p0 + pw) >> 5) + 1] = (uint32_t)(mask >> 32);
            }
            __syncthreads();
            if (tid == 0) {
                int32_t tot = 0;
                for (int w = 0; w < nwaves; ++w) {
                    const int32_t c = s_wcnt[w];
                    s_wcnt[w] = tot;
                    tot += c;
                }
                s_base = tot ? atomicAdd(&A.hist_cnt[slot], tot) : 0;
            }
            __syncthreads();
            if (fired) st_dev(&A.hist[(int64_t)slot * n + s_base + s_wcnt[wave] + __popcll(mask & ((1ull << lane) - 1ull))], i);
            __syncthreads();
        }
        // slot of step t+1 held step t+1-D: nobody reads it in this phase (D >= 3)
        if (b == 0 && tid == 0) st_dev(&A.hist_cnt[(t + 1) % D], 0);
        if (!grid_barrier(X.bar, k, X.B, A.err, &s_abort)) break;
    }
    // write back state and ring; the ring then holds every delivery up to the last step
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int32_t pl = tid + u * T;
        const int32_t i = p0 + pl;
        if (pl < np) {
            A.v[i] = V[u];
            if (MODEL != NBX_MODEL_LIF_VJ) A.s1[i] = S1[u];
            if (MODEL == NBX_MODEL_IZH) {
                A.ga[i] = GA[u];
                A.gg[i] = GG[u];
            }
            A.iext[i] = I[u];
            A.last[i] = last[u];
        }
    }
    __syncthreads();
    for (int32_t j = tid; j < D * nch * P; j += T) {
        const int32_t pl = j % P, sc = j / P;
        if (pl < np) A.ring[(int64_t)sc * n + p0 + pl] = s_ring[j];
    }
    if (ev) atomicAdd(&s_ev, ev);
    __syncthreads();
    if (tid == 0 && s_ev) atomicAdd(A.events, s_ev);
}

// Replay delivery on the post-blocked layout. Block b owns neurons [b*P, b*P + P) and
// streams the injected spikes of all K input steps through an LDS window of
// kReplayWin output steps plus a (D-1)-step carry: contributions that land beyond the
// window stay in LDS and move to the front when the window is flushed, so every
// synapse segment is read exactly once however many steps one launch covers.
// Lanes work in groups of kLanes on one spike, so the consecutive entries of its
// segment coalesce into one memory request; each group keeps kR spikes in flight
// (ids, then segment offsets, then entries) before its LDS atomics.
constexpr int kReplayWin = 64;
constexpr int kR = 4;
constexpr int kU = 2;  // entries per lane loaded ahead (kLanes * kU per segment)

__global__ __launch_bounds__(1024) void nbx_replay_blocked(StepArgs A, const int32_t* __restrict__ off,
                                                           const uint2* __restrict__ seg, int32_t B, int32_t P,
                                                           const int32_t* __restrict__ ids,
                                                           const int32_t* __restrict__ tick_ptr, int32_t K,
                                                           int32_t* __restrict__ out) {
    extern __shared__ int32_t s_acc[];  // [(kReplayWin + D - 1) * nch * P]
    __shared__ int32_t s_tp[kReplayWin + 1];
    const int32_t nch = A.nch, n = A.n, D = A.D;
    const int32_t b = block_of(blockIdx.x, B);
    const int32_t p0 = b * P;
    const int32_t np = max(0, min(P, n - p0));
    const int32_t T = blockDim.x, tid = threadIdx.x;
    const int32_t gl = tid % kLanes, grp = tid / kLanes, ngrp = T / kLanes;
    const int32_t span = (kReplayWin + D - 1) * nch * P;
    const int32_t tick_words = nch * P;
    for (int32_t j = tid; j < span; j += T) s_acc[j] = 0;
    unsigned long long ev = 0;
    for (int32_t w0 = 0; w0 < K; w0 += kReplayWin) {
        const int32_t wl = min(kReplayWin, K - w0);
        for (int32_t j = tid; j <= wl; j += T) s_tp[j] = tick_ptr[w0 + j];
        __syncthreads();
        const int32_t sb = s_tp[0], cnt = s_tp[wl] - sb;
        auto deposit = [&](const uint2 x, int32_t tkr) {
            const int32_t ko = tkr + (int32_t)((x.x >> 16) & 0xFFu) - 1;  // window-relative output step
            if (w0 + ko < K)
                atomicAdd(&s_acc[ko * tick_words + (int32_t)(x.x >> 24) * P + (int32_t)(x.x & 0xFFFFu)], (int32_t)x.y);
        };
        for (int32_t base = grp; base < cnt; base += ngrp * kR) {
            int32_t pre[kR], o0[kR], o1[kR], tk[kR];
#pragma unroll
            for (int r = 0; r < kR; ++r) pre[r] = (base + r * ngrp < cnt) ? ids[sb + base + r * ngrp] : -1;
#pragma unroll
            for (int r = 0; r < kR; ++r) {
                o0[r] = o1[r] = 0;
                if (pre[r] >= 0) {
                    const int32_t* op = off + (int64_t)pre[r] * (B + 1) + b;
                    o0[r] = op[0];
                    o1[r] = op[1];
                }
            }
#pragma unroll
            for (int r = 0; r < kR; ++r) {  // window step of each spike
                const int32_t f = sb + base + r * ngrp;
                int32_t lo = 0, hi = wl;
                while (hi - lo > 1) {
                    const int32_t mid = (lo + hi) >> 1;
                    if (s_tp[mid] <= f) lo = mid;
                    else hi = mid;
                }
                tk[r] = lo;
            }
            uint2 e[kR][kU];
#pragma unroll
            for (int r = 0; r < kR; ++r)
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int32_t j = o0[r] + gl + u * kLanes;
                    e[r][u] = (j < o1[r]) ? seg[j] : make_uint2(0u, 0u);
                }
#pragma unroll
            for (int r = 0; r < kR; ++r) {
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (o0[r] + gl + u * kLanes < o1[r]) deposit(e[r][u], tk[r]);
                for (int32_t j = o0[r] + gl + kU * kLanes; j < o1[r]; j += kLanes) deposit(seg[j], tk[r]);
                if (gl == 0) ev += (unsigned long long)(o1[r] - o0[r]);
            }
        }
        __syncthreads();
        // flush the window's steps, move the carry to the front, clear the rest
        for (int32_t j = tid; j < wl * tick_words; j += T) {
            const int32_t pl = j % P, kc = j / P;
            if (pl < np) out[(int64_t)(w0 * nch + kc) * n + p0 + pl] = s_acc[j];
        }
        if (w0 + kReplayWin >= K) break;  // last window: no carry needed
        __syncthreads();
        const int32_t carry = (D - 1) * tick_words;  // kReplayWin >= D - 1, so source and target do not overlap
        for (int32_t j = tid; j < carry; j += T) s_acc[j] = s_acc[wl * tick_words + j];
        __syncthreads();
        for (int32_t j = carry + tid; j < span; j += T) s_acc[j] = 0;
        __syncthreads();
    }
    if (ev) atomicAdd(A.events, ev);
}

// Replay with global atomics (networks the post-blocked layout does not fit).
__global__ __launch_bounds__(256) void nbx_replay_global(StepArgs A, const int32_t* __restrict__ ids,
                                                         const int32_t* __restrict__ tick_ptr, int32_t K,
                                                         int32_t* __restrict__ out) {
    const int32_t total = tick_ptr[K];
    for (int32_t s = blockIdx.x; s < total; s += gridDim.x) {
        int32_t lo = 0, hi = K;
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (tick_ptr[mid] <= s) lo = mid;
            else hi = mid;
        }
        const int32_t pre = ids[s];
        const int64_t rb = A.row_ptr[pre], re = A.row_ptr[pre + 1];
        for (int64_t j = rb + threadIdx.x; j < re; j += 256) {
            const uint32_t en = A.syn[j];
            const int32_t ko = lo + syn_delay(en) - 1;
            if (ko >= K) continue;
            int32_t q = quantize(A.w[j], A.scale), ch = 0;
            if (A.P.model == NBX_MODEL_IZH) ch = syn_type(en) == 1 ? 1 : 0;
            else if (A.P.model != NBX_MODEL_IZH_CUR && syn_type(en) == 1) q = -q;
            atomicAdd(&out[((int64_t)ko * A.nch + ch) * A.n + syn_post(en)], q);
        }
        if (threadIdx.x == 0) atomicAdd(A.events, (unsigned long long)(re - rb));
    }
}

__global__ void nbx_zero_i32(int32_t* p, int64_t count) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = 0;
}

template <int MODEL>
void* persistent_kernel(int npt) {
    switch (npt) {
        case 1: return (void*)nbx_persistent<MODEL, 1>;
        case 2: return (void*)nbx_persistent<MODEL, 2>;
        default: return (void*)nbx_persistent<MODEL, 4>;
    }
}

void* persistent_kernel_for(int model, int npt) {
    switch (model) {
        case NBX_MODEL_LIF_VJ: return persistent_kernel<NBX_MODEL_LIF_VJ>(npt);
        case NBX_MODEL_LIF_EXP: return persistent_kernel<NBX_MODEL_LIF_EXP>(npt);
        case NBX_MODEL_IZH: return persistent_kernel<NBX_MODEL_IZH>(npt);
        default: return persistent_kernel<NBX_MODEL_IZH_CUR>(npt);
    }
}

struct Geometry {
    int32_t B = 0, P = 0, threads = 0, npt = 0, lds = 0;
    std::string why;
};

Geometry geometry(const nbx_net* net) {
    Geometry g;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device) != hipSuccess || cus <= 0)
        cus = 256;
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, net->device) != hipSuccess ||
        max_lds <= 0)
        max_lds = 64 * 1024;
    const int32_t Bt = std::max(kXcds, (cus / kXcds) * kXcds);
    int32_t P = (net->n + Bt - 1) / Bt;
    P = (P + 31) / 32 * 32;
    int32_t B = (net->n + P - 1) / P;
    B = (B + kXcds - 1) / kXcds * kXcds;
    g.P = P;
    g.B = B;
    int32_t threads = std::min(kMaxThreads, (P + 63) / 64 * 64);
    g.npt = (P + threads - 1) / threads;
    if (g.npt == 3) g.npt = 4;
    g.threads = threads;
    g.lds = net->D * net->nch * P * (int32_t)sizeof(int32_t);
    if (g.npt > kMaxNpt) g.why = "more than 4 neurons per thread";
    else if (g.lds > max_lds - 4096) g.why = "delay ring slice does not fit in LDS";
    if (P > 65535) g.why = "block too large for 16-bit local indices";
    return g;
}

}  // namespace

std::string persistent_reason(const nbx_net* net) {
    if (!net->has_sparse) return "needs sparse synapses";
    if (net->has_dense) return "dense delivery runs on the graph engine";
    if (net->stdp) return "STDP networks use arrival-time delivery";
    if (net->S >= ((int64_t)1 << 31)) return "more than 2^31 synapses";
    Geometry g = geometry(net);
    return g.why;
}

bool persistent_supported(const nbx_net* net) { return persistent_reason(net).empty(); }

int32_t build_persistent_layout(nbx_net* net) {
    PersistentLayout& L = net->pl;
    if (L.built) return NBX_OK;
    Geometry g = geometry(net);
    if (!g.why.empty()) {
        set_error("persistent layout: " + g.why);
        return NBX_ERR_UNSUPPORTED;
    }
    const int32_t n = net->n, B = g.B, P = g.P;
    const float scale = std::ldexp(1.0f, net->shift);
    std::vector<int32_t> off((size_t)n * (B + 1));
    std::vector<uint2> seg((size_t)net->S);
    std::vector<int32_t> cnt((size_t)B);
    int64_t pos = 0;
    for (int32_t p = 0; p < n; ++p) {
        const int64_t rb = net->h_row_ptr[(size_t)p], re = net->h_row_ptr[(size_t)p + 1];
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int64_t j = rb; j < re; ++j) cnt[(size_t)(syn_post(net->h_syn[(size_t)j]) / P)]++;
        int32_t* o = off.data() + (size_t)p * (B + 1);
        o[0] = (int32_t)pos;
        for (int32_t b = 0; b < B; ++b) o[b + 1] = o[b] + cnt[(size_t)b];
        for (int32_t b = 0; b < B; ++b) cnt[(size_t)b] = o[b];
        for (int64_t j = rb; j < re; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            const int32_t post = syn_post(e), ty = syn_type(e);
            int32_t q = (int32_t)std::lrint(net->h_w[(size_t)j] * scale);
            int32_t ch = 0;
            if (net->md.model == NBX_MODEL_IZH) ch = (ty == 1) ? 1 : 0;
            else if (net->md.model != NBX_MODEL_IZH_CUR && ty == 1) q = -q;
            const uint32_t meta = (uint32_t)(post % P) | ((uint32_t)syn_delay(e) << 16) | ((uint32_t)ch << 24);
            seg[(size_t)cnt[(size_t)(post / P)]++] = make_uint2(meta, (uint32_t)q);
        }
        pos = o[B];
    }
    int32_t rc = net->upload(&L.d_off, off.data(), off.size());
    if (rc) return rc;
    rc = net->upload(&L.d_seg, seg.data(), seg.size());
    if (rc) return rc;
    rc = net->alloc(&L.d_bar, sizeof(Barrier) / sizeof(int32_t));
    if (rc) return rc;
    L.B = B;
    L.P = P;
    L.threads = g.threads;
    L.npt = g.npt;
    L.lds_bytes = g.lds;
    L.built = true;
    return NBX_OK;
}

int32_t refresh_persistent_weights(nbx_net* net) {
    PersistentLayout& L = net->pl;
    if (!L.built) return NBX_OK;
    for (void* p : {(void*)L.d_off, (void*)L.d_seg, (void*)L.d_bar}) {
        auto it = std::find(net->allocs.begin(), net->allocs.end(), p);
        if (it != net->allocs.end()) net->allocs.erase(it);
        NBX_HIP_TRY(hipFree(p));
    }
    L = PersistentLayout();
    return NBX_OK;
}

int32_t run_persistent(nbx_net* net, int32_t nsteps) {
    int32_t rc = build_persistent_layout(net);
    if (rc) return rc;
    if (nsteps == 0) return NBX_OK;
    PersistentLayout& L = net->pl;
    void* fn = persistent_kernel_for(net->md.model, L.npt);
    int per_cu = 0, cus = 0;
    NBX_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, L.threads, (size_t)L.lds_bytes));
    NBX_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device));
    if ((int64_t)per_cu * cus < L.B) {
        set_error("persistent engine: " + std::to_string(L.B) + " workgroups cannot be co-resident");
        return NBX_ERR_UNSUPPORTED;
    }
    if (L.lds_bytes > 64 * 1024)
        NBX_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, L.lds_bytes));
    nbx_zero_i32<<<1, 256, 0, net->stream>>>(L.d_bar, (int64_t)(sizeof(Barrier) / sizeof(int32_t)));
    NBX_HIP_TRY(hipGetLastError());
    PersistArgs X;
    X.A = net->args();
    X.off = L.d_off;
    X.seg = L.d_seg;
    X.B = L.B;
    X.P = L.P;
    X.t0 = net->t;
    X.nsteps = nsteps;
    X.bar = reinterpret_cast<Barrier*>(L.d_bar);
    void* kargs[] = {&X};
    // Co-residency of all L.B workgroups is checked above (occupancy x CUs). The cooperative
    // launch additionally has the runtime verify it; NBX_COOPERATIVE=0 selects a regular launch
    // for tools that cannot trace cooperative dispatches (the barrier spin is bounded either way).
    const char* coop = std::getenv("NBX_COOPERATIVE");
    if (coop && coop[0] == '0')
        NBX_HIP_TRY(hipLaunchKernel(fn, dim3(L.B), dim3(L.threads), kargs, (size_t)L.lds_bytes, net->stream));
    else
        NBX_HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(L.B), dim3(L.threads), kargs, (unsigned)L.lds_bytes,
                                               net->stream));
    net->t += nsteps;
    return NBX_OK;
}

int32_t launch_replay(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t nsteps, int32_t* d_out) {
    const StepArgs A = net->args();
    // the blocked layout carries pre-quantized weights, so it is not used while STDP changes them
    if (!net->stdp && build_persistent_layout(net) == NBX_OK) {
        PersistentLayout& L = net->pl;
        const size_t lds = (size_t)(kReplayWin + net->D - 1) * net->nch * L.P * sizeof(int32_t);
        int max_lds = 0;
        NBX_HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, net->device));
        if (lds + 1024 <= (size_t)max_lds) {
            if (lds > 64 * 1024)
                NBX_HIP_TRY(hipFuncSetAttribute((const void*)nbx_replay_blocked,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            nbx_replay_blocked<<<L.B, 1024, lds, net->stream>>>(A, L.d_off, L.d_seg, L.B, L.P, d_ids, d_tick_ptr,
                                                                 nsteps, d_out);
            NBX_HIP_TRY(hipGetLastError());
            return NBX_OK;
        }
    }
    const int64_t count = (int64_t)nsteps * net->nch * net->n;
    nbx_zero_i32<<<(int)std::min<int64_t>((count + 255) / 256, 4096), 256, 0, net->stream>>>(d_out, count);
    NBX_HIP_TRY(hipGetLastError());
    nbx_replay_global<<<2048, 256, 0, net->stream>>>(A, d_ids, d_tick_ptr, nsteps, d_out);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

}  // namespace nbx
