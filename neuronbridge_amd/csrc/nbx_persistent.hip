// Persistent closed-loop engine and replay delivery on the post-blocked layout.
//
// One cooperative launch runs many steps. Workgroup b owns the neurons
// [b*P, b*P + P): their state stays in VGPRs and their delay ring
// ring[D][channel][P] stays in LDS for the whole launch. Per step t a workgroup
//   1. waits for the step counters of step t-1 and delivers its spikes that target
//      its own neurons into its LDS ring (int32 fixed point, ds_add, so the sum is
//      order independent),
//   2. updates its neurons for step t (same neuron_tick as the per-phase engine),
//      compacts its spikes (wavefront ballot + LDS prefix) and publishes them with
//      one atomic on its group's counter and tagged entries in its group's list,
//   3. draws the next step's Poisson counts while the others publish.
// No grid barrier: a workgroup waits only for the counters of the step it reads.
// This replaces the reference's separate clear/mark/drain/update launches and host
// round trips (GpuPropagationRuntime.cu:1687-1827, DenseUnifiedNeuronRuntime.cu:284-372).
// Workgroups are numbered so that consecutive neuron blocks share an XCD and its L2:
// a presynaptic row's segments for neighbouring blocks are adjacent in memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "nbx_net.h"

namespace nbx {

namespace {

constexpr int kXcds = 8;
constexpr int kMaxThreads = 512;  // persistent workgroup bound: 256 VGPRs per thread, no spills
constexpr int kMaxNpt = 4;
constexpr uint32_t kSpinLimit = 1u << 24;  // ~0.5 s of s_sleep(1) spins before giving up
constexpr int kLanes = 4;                  // lanes sharing one spike's synapse segment
constexpr int kTabCap = 64;                // Poisson threshold table entries held in LDS
// Spikes per lane group in flight per delivery round: 4 covers a step of 4 * 112 spikes on a
// 448-thread workgroup in one round (the first, busier steps of a run); 2 is 3% faster once the
// 1e5-neuron network settles at ~160 spikes per step but 9% slower over its first 25 steps.
constexpr int kRound = 4;
constexpr uint32_t kSlotEmpty = 0xFFFFFFFFu;  // replay slot entry: unused
constexpr uint32_t kSlotLink = 0x80000000u;   // replay slot entry: link to the overflow list
constexpr int kSlotCap = 16;                  // entries per replay slot (128 bytes)


// Raw buffer loads: an offset at or past the resource's byte count reads zeros without a memory
// request, so a masked-off lane needs no branch around its load (a load under a branch makes the
// compiler wait for it before the branch ends, serialising the loads a lane keeps in flight).
constexpr uint32_t kBufOff = 0x80000000u;         // offset of a lane with nothing to load
constexpr int32_t kBufDword3 = 0x00020000;       // resource word 3 for gfx9 (raw, no swizzle)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int32_t)min(bytes, 0x7FFFFFFFu), kBufDword3);
}
__device__ inline uint4 buf_load16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int32_t)off, 0, 0);
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline int64_t buf_load_i64(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int32_t)off, 0, 0);
    return (int64_t)(((uint64_t)v.y << 32) | v.x);
}
__device__ inline int32_t buf_load_i32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int32_t)off, 0, 0);
}

__device__ inline int32_t block_of(int32_t w, int32_t B) {
    const int32_t per = B / kXcds;  // B is a multiple of kXcds
    return (w % kXcds) * per + (w / kXcds);
}

// Step synchronisation by counters. Workgroup w belongs to group w % kGroups; per ring slot and
// group one 64-bit counter {spikes << 32 | arrivals} sits on its own 128-byte line. Publishing step
// t, a workgroup adds {its spikes << 32 | 1} to its group's counter of slot t % D with one returning
// atomic. Counters are cumulative over a launch and never cleared while it runs: the k-th use of a
// slot (k = (t - t0) / D + 1) is complete when the group's arrivals reach k times the group's
// workgroup count, and a step's spike count per group is the difference between the spike totals
// after this use and after the previous one, which every workgroup keeps in LDS (it reads every
// step's counters). The atomic's old spike total minus that previous total is where the
// workgroup's spike ids go in its group's list. A workgroup needs step t when the kGroups counters
// of the slot all show their use complete; kGroups lanes poll one counter each per round, so few
// workgroups share one atomic address and nobody polls a line that every workgroup polls twice.
//
// Memory ordering: every exchange is a single-location atomic (the counters: returning
// fetch-adds, polled with relaxed loads; each location's modification order alone makes the
// reader's view consistent) or a self-validating list entry: entries are 64-bit {step tag << 32 |
// id}, written and read with device-scope atomic stores and loads (single-copy atomic), and a reader
// re-reads an entry until it carries the step's tag (tags are unique over the network's lifetime:
// the host advances the epoch by nsteps + 1 per launch). No ordering between different locations is
// needed. Counters come in two banks, one per launch parity: a launch counts in its bank and clears
// the other, which the next launch uses after the kernel boundary. Spins are bounded: on time-out
// the error flag's kErrBarrier bit is raised and every waiter that sees it leaves.
constexpr int kCtrLine = 16;  // uint64 per 128-byte line
constexpr int kGroups = 16;   // counter groups: workgroup w belongs to group w % kGroups

// Residency agreement. The step exchange waits on every workgroup, so all B must be resident at
// once; the kernel is launched as an ordinary dispatch (a cooperative launch, which guarantees
// residency, costs ~12 µs more per launch) and the workgroups agree on residency before any of
// them writes global state. Each adds 1 to its launch's start counter (zero at launch: it sits
// in the counter bank) and waits until the counter reads B. A workgroup that has waited
// start_spins polls replaces the value v < B it last read with v + kStartPoison by a
// compare-and-swap, after which B can never be read: every workgroup then leaves without taking a
// step, and kErrResidency is raised. Once the counter reads B no compare-and-swap can succeed (they
// expect a value below B), so every workgroup reaches the same decision. A launch finding
// kErrResidency already raised takes no step either.
constexpr unsigned long long kStartPoison = 1ull << 40;

struct PersistArgs {
    StepArgs A;
    const int32_t* off;
    const uint2* seg;
    int32_t B, P;
    int32_t rb, gs, os;  // blocks per region, ints per region group, ints per offset row
    int32_t t0, nsteps;
    unsigned long long* gctr;   // [D][kGroups][kCtrLine] group counters {spikes << 32 | arrivals}, this launch's bank
    unsigned long long* gctr_next;  // the other bank, cleared by this launch for the next one
    unsigned long long* gstart;     // this bank's start counter (residency agreement)
    unsigned long long start_want;  // the workgroups that must arrive: B
    uint32_t start_spins;           // polls before a workgroup gives the agreement up
    unsigned long long* glist;  // [D][kGroups][gcap] group spike lists {tag << 32 | id}
    int32_t gcap;               // list capacity: the neurons of one group
    uint32_t epoch;             // tag of step t0 - 1 (step t0 + k carries epoch + 1 + k)
};

__device__ inline int32_t ld_dev(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_dev64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t tag_of(unsigned long long x) { return (uint32_t)(x >> 32); }

// Re-reads every non-null p[r] whose word does not carry tag, all together per round, until all
// do; x[r] holds the first loads on entry. False if the wait was abandoned (kErrBarrier raised by
// another workgroup, or by this one at the spin limit).
template <int N>
__device__ inline bool await_tags(const unsigned long long* const (&p)[N], uint32_t tag, unsigned long long (&x)[N],
                                  int32_t* err) {
    uint32_t pend = 0;
#pragma unroll
    for (int r = 0; r < N; ++r)
        if (p[r] && tag_of(x[r]) != tag) pend |= 1u << r;
    uint32_t spins = 0;
    while (pend) {
        if ((spins & 63u) == 0 && (ld_dev(err) & kErrBarrier) != 0) return false;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit) {
            atomicOr(err, kErrBarrier);
            return false;
        }
#pragma unroll
        for (int r = 0; r < N; ++r)
            if ((pend >> r) & 1u) x[r] = ld_dev64(p[r]);
#pragma unroll
        for (int r = 0; r < N; ++r)
            if (((pend >> r) & 1u) && tag_of(x[r]) == tag) pend &= ~(1u << r);
    }
    return true;
}

// Wave 0 of a workgroup: waits until all groups' workgroups have published step st (ring slot sl),
// records the groups' spike totals after it in prev[sl] (the totals after the slot's previous use
// move to prev_old[sl]) and leaves the groups' list starts (step numbering group-major) in
// s_gstart[0..kGroups), the step's spike count in s_gstart[kGroups]. False if the wait was
// abandoned.
__device__ inline bool await_step(const PersistArgs& X, int32_t st, int32_t sl, int32_t lane, int32_t* s_gstart,
                                  int32_t* prev, int32_t* prev_old, int32_t* err) {
    const int32_t per = X.B / kGroups + ((lane < X.B % kGroups) ? 1 : 0);  // workgroups w with w % kGroups == lane
    const uint32_t want = (uint32_t)((st - X.t0) / X.A.D + 1) * (uint32_t)per;  // arrivals after this use
    unsigned long long c = 0;
    bool ok = true;
    if (lane < kGroups) {
        const unsigned long long* p = X.gctr + ((int64_t)sl * kGroups + lane) * kCtrLine;
        c = ld_dev64(p);
        uint32_t spins = 0;
        while ((uint32_t)c != want) {
            if ((spins & 63u) == 0 && (ld_dev(err) & kErrBarrier) != 0) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) {
                atomicOr(err, kErrBarrier);
                ok = false;
                break;
            }
            c = ld_dev64(p);
        }
    }
    int32_t gc = 0;
    if (lane < kGroups) {
        const int32_t total = (int32_t)(c >> 32), before = prev[sl * kGroups + lane];
        gc = total - before;
        prev_old[sl * kGroups + lane] = before;
        prev[sl * kGroups + lane] = total;
    }
    int32_t gp = gc;  // inclusive prefix over the kGroups lanes
#pragma unroll
    for (int o = 1; o < kGroups; o <<= 1) {
        const int32_t z = __shfl_up(gp, o, 64);
        if (lane >= o) gp += z;
    }
    if (lane < kGroups) s_gstart[lane] = gp - gc;
    if (lane == kGroups - 1) s_gstart[kGroups] = gp;
    return __all(ok ? 1 : 0) != 0;
}

// Where spike number f of slot sl (f < the step's count) is listed; gs holds the group list starts.
__device__ inline const unsigned long long* spike_ptr(const PersistArgs& X, int32_t sl, const int32_t (&gs)[kGroups],
                                                      int32_t f) {
    int32_t g = 0, start = 0;
#pragma unroll
    for (int q = 1; q < kGroups; ++q)
        if (f >= gs[q]) {
            g = q;
            start = gs[q];
        }
    return X.glist + ((int64_t)sl * kGroups + g) * X.gcap + (f - start);
}

template <int MODEL, int NPT>
__global__ __launch_bounds__(kMaxThreads) void nbx_persistent(PersistArgs X) {
    extern __shared__ int32_t s_ring[];  // [D][nch][P]
    __shared__ int32_t s_wc[kMaxNpt * (kMaxThreads / 64)];  // spikes per (neuron chunk, wave) of the update
    __shared__ int32_t s_gstart[kGroups + 1];                 // the step's group list starts and count
    __shared__ int32_t s_base;
    __shared__ int32_t s_abort;
    __shared__ unsigned long long s_ev;
    __shared__ uint32_t s_tab[kTabCap];
    const StepArgs& A = X.A;
    const int32_t P = X.P, D = A.D, nch = A.nch, n = A.n;
    const int32_t b = block_of(blockIdx.x, X.B);
    const int32_t xg = (int32_t)(blockIdx.x % kGroups);  // this workgroup's group
    const int32_t p0 = b * P;
    const int32_t np = max(0, min(P, n - p0));
    const int32_t T = blockDim.x;
    const int32_t tid = threadIdx.x;
    const int32_t lane = tid & 63, wave = tid >> 6, nwaves = T >> 6;
    const int32_t gl = tid % kLanes, grp = tid / kLanes, ngrp = T / kLanes;
    const int32_t i0 = (b / X.rb) * X.gs + b % X.rb;  // this block's entry in an offset row; i0 + 1 ends it

    // residency agreement (above): arrive now, decide after the ring load
    __shared__ int32_t s_go;
    bool stale = false;  // thread 0: kErrResidency was raised before this launch
    unsigned long long arrived = 0;
    if (tid == 0) {
        stale = (ld_dev(A.err) & kErrResidency) != 0;
        if (!stale)
            arrived = __hip_atomic_fetch_add(X.gstart, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1ull;
    }
    // load this block's slice of the delay ring (rows of P neurons, many loads in flight), the
    // Poisson thresholds and its neurons' state
    {
        const int32_t rows = D * nch;
        for (int32_t r0 = 0; r0 < rows; r0 += 8) {
            int32_t v[8][NPT];
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int u = 0; u < NPT; ++u) {
                    const int32_t pl = tid + u * T;
                    v[r][u] = (r0 + r < rows && pl < np) ? A.ring[(int64_t)(r0 + r) * n + p0 + pl] : 0;
                }
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int u = 0; u < NPT; ++u) {
                    const int32_t pl = tid + u * T;
                    if (r0 + r < rows && pl < P) s_ring[(r0 + r) * P + pl] = v[r][u];
                }
        }
    }
    const int32_t kmax = min(A.kmax, kTabCap - 1);
    if (A.pois && tid <= kmax) s_tab[tid] = A.pois_tab[tid];
    if (tid == 0) {
        s_ev = 0;
        s_abort = 0;
    }
    // spike totals of each (slot, group) counter after its last use read by this workgroup, and
    // after the use before that; this launch's bank starts at zero
    int32_t* s_prev = s_ring + D * nch * P;
    int32_t* s_prev_old = s_prev + D * kGroups;
    for (int32_t j = tid; j < 2 * D * kGroups; j += T) s_prev[j] = 0;
    if (tid == 0) {
        int32_t go = 0;
        if (!stale) {
            unsigned long long v = arrived;
            uint32_t spins = 0;
            for (;;) {
                if (v == X.start_want) {
                    go = 1;
                    break;
                }
                if (v >= kStartPoison) break;
                if (++spins > X.start_spins) {  // give up unless everyone has arrived meanwhile
                    unsigned long long seen = v;
                    if (__hip_atomic_compare_exchange_strong(X.gstart, &seen, v + kStartPoison, __ATOMIC_RELAXED,
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        atomicOr(A.err, kErrResidency);
                        break;
                    }
                    v = seen;  // the counter moved: decide on its new value
                    continue;
                }
                __builtin_amdgcn_s_sleep(1);
                v = ld_dev64(X.gstart);
            }
        }
        s_go = go;
    }
    __syncthreads();
    if (!s_go) return;  // uniform; nothing global was written
    // the next launch's bank: nobody reads it during this launch; the kernel boundary orders these
    // stores before the next launch's atomics
    if (b == 0) {
        for (int32_t j = tid; j < D * kGroups; j += T) X.gctr_next[(int64_t)j * kCtrLine] = 0ull;
        if (tid == 0) X.gctr_next[(int64_t)D * kGroups * kCtrLine] = 0ull;  // its start counter
    }
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
            if (model_uses_s1(MODEL)) S1[u] = A.s1[i];
            if (model_uses_g2(MODEL)) {
                GA[u] = A.ga[i];
                GG[u] = A.gg[i];
            }
            I[u] = A.iext[i];
            last[u] = A.last[i];
        }
    }
    unsigned long long ev = 0;
    // Poisson background counts of the next step: they depend on no other workgroup, so they are
    // drawn between this workgroup's publication of a step and its wait for the others'
    int32_t pc[NPT];
    auto draw = [&](int32_t t) {
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            const int32_t pl = tid + u * T;
            pc[u] = (A.pois && pl < np) ? poisson_count(A.seed_lo, A.seed_hi, p0 + pl, t, s_tab, kmax) : 0;
        }
    };
    __syncthreads();  // s_tab and s_ring are loaded
    draw(X.t0);
    bool aborted = false;

    for (int32_t k = 0; k <= X.nsteps; ++k) {
        const int32_t t = X.t0 + k;
        const int32_t sb = t % D;  // ring slot of step t
        if (k > 0) {
            // 1. the spikes of step t-1 (published by every workgroup once the slot's counters are full)
            const int32_t ps = (t - 1) % D;
            const uint32_t want = X.epoch + (uint32_t)k;
            if (wave == 0 && !await_step(X, t - 1, ps, lane, s_gstart, s_prev, s_prev_old, A.err) && lane == 0)
                s_abort = 1;
            __syncthreads();
            if (s_abort) {  // uniform: read after the barrier, set before it
                aborted = true;
                break;
            }
            const int32_t cnt = s_gstart[kGroups];
            int32_t gs[kGroups];
#pragma unroll
            for (int q = 0; q < kGroups; ++q) gs[q] = s_gstart[q];
            if (b == 0 && tid == 0 && A.counts && t - 1 >= A.rec_t0 && t - 1 < A.rec_t0 + A.rec_n)
                A.counts[t - 1 - A.rec_t0] = cnt;
            bool ok = true;
            // lanes in groups of kLanes share one spike so its segment entries coalesce; each round
            // keeps kRound spikes per lane group in flight
            auto deliver = [&](auto rounds) {
                constexpr int R = decltype(rounds)::value;
                unsigned long long x[R];
                const unsigned long long* sp[R];
                for (int32_t base = grp; base < cnt; base += ngrp * R) {
                    int32_t pre[R], o0[R], o1[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        sp[r] = (base + r * ngrp < cnt) ? spike_ptr(X, ps, gs, base + r * ngrp) : nullptr;
                        x[r] = sp[r] ? ld_dev64(sp[r]) : 0ull;
                    }
                    if (ok) ok = await_tags(sp, want, x, A.err);
#pragma unroll
                    for (int r = 0; r < R; ++r) pre[r] = (sp[r] && ok) ? (int32_t)(uint32_t)x[r] : -1;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        o0[r] = o1[r] = 0;
                        if (pre[r] >= 0) {
                            const int32_t* op = X.off + (int64_t)pre[r] * X.os + i0;
                            o0[r] = op[0];
                            o1[r] = op[1];
                        }
                    }
                    uint2 e[R][2];
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const int32_t j = o0[r] + gl + u * kLanes;
                            e[r][u] = (j < o1[r]) ? X.seg[j] : make_uint2(0u, 0u);
                        }
                    auto deposit = [&](const uint2 y) {
                        int32_t s = sb + (int32_t)((y.x >> 16) & 0x3Fu) - 1;  // slot of step t - 1 + delay
                        s -= (s >= D) ? D : 0;
                        atomicAdd(&s_ring[(s * nch + (int32_t)((y.x >> 22) & 3u)) * P + (int32_t)(y.x & 0xFFFFu)],
                                  (int32_t)y.y);
                    };
#pragma unroll
                    for (int r = 0; r < R; ++r) {
#pragma unroll
                        for (int u = 0; u < 2; ++u)
                            if (o0[r] + gl + u * kLanes < o1[r]) deposit(e[r][u]);
                        for (int32_t j = o0[r] + gl + 2 * kLanes; j < o1[r]; j += kLanes) deposit(X.seg[j]);
                        if (gl == 0) ev += (unsigned long long)(o1[r] - o0[r]);
                    }
                }
            };
            deliver(std::integral_constant<int, kRound>{});
            if (!ok) s_abort = 1;
        }
        __syncthreads();  // every delivery into this step's ring slot is in; s_abort is uniform
        if (s_abort) {
            aborted = true;
            break;
        }
        if (k == X.nsteps) break;
        // 2. update this block's neurons for step t and publish their spikes
        const int32_t slot = sb;
        const bool rec = A.raster && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n;
        const uint32_t tag = X.epoch + 1u + (uint32_t)k;
        uint64_t msk[NPT];
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            const int32_t pl = tid + u * T;
            const int32_t i = p0 + pl;
            bool fired = false;
            if (pl < np) {
                const int32_t a0 = s_ring[(slot * nch) * P + pl];
                s_ring[(slot * nch) * P + pl] = 0;
                int32_t a1 = 0, a2 = 0;
                if (nch > 1) {
                    a1 = s_ring[(slot * nch + 1) * P + pl];
                    s_ring[(slot * nch + 1) * P + pl] = 0;
                }
                if (nch > 2) {
                    a2 = s_ring[(slot * nch + 2) * P + pl];
                    s_ring[(slot * nch + 2) * P + pl] = 0;
                }
                fired = neuron_tick<MODEL, true, true>(A, i, t, a0, a1, a2, V[u], S1[u], GA[u], GG[u], I[u], last[u],
                                                       pc[u]);
            }
            msk[u] = __ballot(fired);
            if (lane == 0) s_wc[u * nwaves + wave] = __popcll(msk[u]);
            if (rec && lane == 0) {
                const int32_t pw = u * T + wave * 64;  // first local neuron of this wavefront
                uint32_t* row = A.raster + (int64_t)(t - A.rec_t0) * A.words;
                if (pw < np) row[(p0 + pw) >> 5] = (uint32_t)msk[u];
                if (pw + 32 < np) row[((p0 + pw) >> 5) + 1] = (uint32_t)(msk[u] >> 32);
            }
        }
        __syncthreads();
        {
            // this workgroup's spikes, numbered by (neuron chunk, wave, lane), go into its group's
            // list at the place its arrival atomic returns; each wave scans the per-(chunk, wave) counts
            const int32_t ne = NPT * nwaves;
            const int32_t mine = (lane < ne) ? s_wc[lane] : 0;
            int32_t y = mine;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t z = __shfl_up(y, o, 64);
                if (lane >= o) y += z;
            }
            const int32_t tot = __shfl(y, ne - 1, 64);
            if (tid == 0) {
                const unsigned long long old = __hip_atomic_fetch_add(
                    X.gctr + ((int64_t)slot * kGroups + xg) * kCtrLine, ((unsigned long long)(uint32_t)tot << 32) | 1ull,
                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                s_base = (int32_t)(old >> 32) - s_prev[slot * kGroups + xg];  // within this use of the slot
            }
            __syncthreads();
            unsigned long long* lst = X.glist + ((int64_t)slot * kGroups + xg) * X.gcap + s_base;
            const unsigned long long hi = (unsigned long long)tag << 32;
#pragma unroll
            for (int u = 0; u < NPT; ++u) {
                const int32_t start = __shfl(y - mine, u * nwaves + wave, 64);
                if ((msk[u] >> lane) & 1ull)
                    st_dev64(&lst[start + __popcll(msk[u] & ((1ull << lane) - 1ull))], hi | (uint32_t)(p0 + tid + u * T));
            }
        }
        draw(t + 1);
    }
    // write back state and ring; the ring then holds every delivery up to the last step
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int32_t pl = tid + u * T;
        const int32_t i = p0 + pl;
        if (pl < np) {
            A.v[i] = V[u];
            if (model_uses_s1(MODEL)) A.s1[i] = S1[u];
            if (model_uses_g2(MODEL)) {
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
    // The firing-history ring for the steps it still holds (the launch's last D - 1), from the group
    // lists: the workgroups share them out, B / (D - 1) per step, each taking every (B / (D - 1))-th
    // chunk of the step's list (every workgroup saw every step complete, and an entry not yet seen
    // with its tag is waited for). Each history entry is written by one workgroup.
    if (!aborted) {
        const int32_t nh = min(X.nsteps, D - 1);
        const int32_t parts = max(1, X.B / max(nh, 1));  // workgroups sharing one step's list
        for (int32_t item = b; item < nh * parts; item += X.B) {
            const int32_t j = item % nh, part = item / nh;
            const int32_t s = X.t0 + X.nsteps - 1 - j, sl = s % D;
            const uint32_t tag = X.epoch + (uint32_t)(X.nsteps - j);
            __syncthreads();  // s_gstart of the previous step is read
            if (wave == 0) {  // the step's list starts from the totals its read left (s was the slot's last use)
                const int32_t gc = (lane < kGroups) ? s_prev[sl * kGroups + lane] - s_prev_old[sl * kGroups + lane] : 0;
                int32_t gp = gc;
#pragma unroll
                for (int o = 1; o < kGroups; o <<= 1) {
                    const int32_t z = __shfl_up(gp, o, 64);
                    if (lane >= o) gp += z;
                }
                if (lane < kGroups) s_gstart[lane] = gp - gc;
                if (lane == kGroups - 1) s_gstart[kGroups] = gp;
            }
            __syncthreads();
            const int32_t cnt = s_gstart[kGroups];
            int32_t gs[kGroups];
#pragma unroll
            for (int q = 0; q < kGroups; ++q) gs[q] = s_gstart[q];
            for (int32_t f = part * T + tid; f < cnt; f += parts * T) {
                const unsigned long long* sp[1] = {spike_ptr(X, sl, gs, f)};
                unsigned long long x[1] = {ld_dev64(sp[0])};
                await_tags(sp, tag, x, A.err);
                A.hist[(int64_t)sl * n + f] = (int32_t)(uint32_t)x[0];
            }
            if (part == 0 && tid == 0) A.hist_cnt[sl] = cnt;
        }
        // the next step's updates start counting from zero (the slot held step t_end - D)
        if (b == 0 && tid == 0) A.hist_cnt[(X.t0 + X.nsteps) % D] = 0;
    }
    __syncthreads();
    if (tid == 0 && s_ev) atomicAdd(A.events, s_ev);
}

// Replay delivery on the post-blocked layout.
//
// Workgroup (g, q) owns the neurons of m consecutive layout blocks [q*m, q*m + m) and the
// output steps [g*Kg, g*Kg + Kg) of a K-step launch. Its LDS holds a circular ring of
// R = W + D - 1 output steps for its neurons. It walks the input steps from D - 2 steps
// before its first output (contributions landing earlier belong to the previous time
// group) in windows of W steps; after the input steps [k0, k0 + W) the outputs
// [k0, k0 + W) are complete (delays are >= 1), so they are written to HBM and their slots
// cleared. Every synapse segment is read once per launch (plus a D - 2 step halo).
//
// The kernel is bound by memory requests in flight, so the per-window work is a three-
// stage software pipeline: while the segments of window w are delivered, the segment
// bounds (o0, o1) of window w+1 and the spike ids of window w+2 are already in flight.
// Delivery itself keeps two batches of kRR spikes per lane group in flight; lanes work in
// groups of LANES on one spike so the consecutive entries of its segment coalesce, and
// arrivals are int32 fixed point added with LDS atomics, so the sum is order independent.
constexpr int kMaxWin = 128;
constexpr int kTpCap = 1024;  // input steps whose tick pointers one pass keeps in LDS
constexpr int kRR = 4;        // spikes per lane group per batch
constexpr int kRU = 2;        // entries per lane loaded ahead (LANES * kRU per segment)
constexpr int kStageU = 4;    // spikes staged per thread per window (C <= kStageU * 1024)

struct ReplayArgs {
    const int32_t* off;
    const uint2* seg;
    int32_t B, P, rb, gs, os;   // layout
    int32_t m, Bq, G, W, R, C;  // merge factor, post blocks, time groups, window, ring slots, staged spikes
    int32_t n, nch, D, K;
    const int32_t* ids;
    const int32_t* tick_ptr;
    int32_t* out;
    unsigned long long* events;
    const uint4* slots;  // fixed-slot layout: 8 uint4 (16 entries) per (pre, post block)
    const uint2* ovf;
    const int64_t* row_ptr;  // CSR row starts (row lengths count the synaptic events)
    int32_t* err;            // device error flags (a bounded wait that gave up)
    int32_t wbits;           // packed slots: bits of the LDS word in an entry's meta
};

template <int LANES>
__global__ __launch_bounds__(1024) void nbx_replay_ring(ReplayArgs X) {
    extern __shared__ int32_t s_mem[];
    __shared__ int32_t s_tp[kTpCap + 1];
    const int32_t T = blockDim.x, tid = threadIdx.x;
    const int32_t gl = tid % LANES, grp = tid / LANES, ngrp = T / LANES;
    const int32_t Pr = X.m * X.P, words = X.nch * Pr, R = X.R, C = X.C, W = X.W;
    int32_t* s_ring = s_mem;                                           // [R][nch][Pr]
    int2* s_span = reinterpret_cast<int2*>(s_mem + (int64_t)R * words);  // [2][C]: {o0, len | step << 24}

    // consecutive post blocks of one time group share an XCD (and its L2)
    const int32_t nw = gridDim.x;
    const int32_t lin = (nw % kXcds == 0) ? (int32_t)(blockIdx.x % kXcds) * (nw / kXcds) + (int32_t)blockIdx.x / kXcds
                                          : (int32_t)blockIdx.x;
    const int32_t g = lin / X.Bq, q = lin % X.Bq;
    const int32_t b0 = q * X.m, b1 = min(b0 + X.m, X.B);
    const int32_t p0 = b0 * X.P;
    const int32_t np = max(0, min(Pr, X.n - p0));
    const int32_t Kg = (X.K + X.G - 1) / X.G;
    const int32_t o_begin = g * Kg, o_end = min(X.K, o_begin + Kg);
    if (np == 0 || o_begin >= o_end || b0 >= X.B) return;  // uniform over the workgroup
    const int32_t k_begin = max(0, o_begin - (X.D - 2));
    // offsets of [b0, b1) within a row: the pair never straddles a region (rb is even)
    const int32_t i0 = (b0 / X.rb) * X.gs + b0 % X.rb;
    const int32_t i1 = i0 + (b1 - b0);
    const uint32_t b0lo = (uint32_t)b0 & 0x7Fu;

    for (int32_t j = tid; j < R * words; j += T) s_ring[j] = 0;
    unsigned long long ev = 0;
    int32_t phys = 0;  // ring slot of the first output step of the current window

    for (int32_t s0 = k_begin; s0 < o_end; s0 += kTpCap) {
        const int32_t sl = min(kTpCap, o_end - s0);
        const int32_t nwin = (sl + W - 1) / W;
        __syncthreads();
        for (int32_t j = tid; j <= sl; j += T) s_tp[j] = X.tick_ptr[s0 + j];
        __syncthreads();
        auto win_lo = [&](int32_t w) { return s_tp[w * W]; };
        auto win_cnt = [&](int32_t w) { return s_tp[min(w * W + W, sl)] - s_tp[w * W]; };
        auto load_ids = [&](int32_t w, int32_t (&dst)[kStageU]) {
            const int32_t lo = (w < nwin) ? win_lo(w) : 0;
            const int32_t cnt = (w < nwin) ? min(win_cnt(w), C) : 0;
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                const int32_t f = tid + u * T;
                dst[u] = (f < cnt) ? X.ids[lo + f] : -1;
            }
        };
        auto load_bounds = [&](const int32_t (&id)[kStageU], int32_t (&o0)[kStageU], int32_t (&o1)[kStageU]) {
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                o0[u] = o1[u] = 0;
                if (id[u] >= 0) {
                    const int32_t* op = X.off + (int64_t)id[u] * X.os;
                    o0[u] = op[i0];
                    o1[u] = op[i1];
                }
            }
        };
        // window step of spike fa (absolute index) within window w, by binary search
        auto step_of = [&](int32_t w, int32_t fa) {
            const int32_t kw0 = w * W, wl = min(W, sl - kw0);
            int32_t lo = 0, hi = wl;
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if (s_tp[kw0 + mid] <= fa) lo = mid;
                else hi = mid;
            }
            return lo;
        };
        auto commit = [&](int32_t w, const int32_t (&id)[kStageU], const int32_t (&o0)[kStageU],
                          const int32_t (&o1)[kStageU], int32_t buf) {
            if (w >= nwin) return;
            const int32_t lo = win_lo(w);
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                const int32_t f = tid + u * T;
                if (id[u] >= 0) s_span[buf * C + f] = make_int2(o0[u], (o1[u] - o0[u]) | (step_of(w, lo + f) << 24));
            }
        };

        // prologue: window 0 staged, ids of window 1 in flight
        int32_t idn[kStageU], o0v[kStageU], o1v[kStageU];
        {
            int32_t id0[kStageU];
            load_ids(0, id0);
            load_bounds(id0, o0v, o1v);
            load_ids(1, idn);
            commit(0, id0, o0v, o1v, 0);
        }
        __syncthreads();

        for (int32_t w = 0; w < nwin; ++w) {
            const int32_t buf = w & 1;
            const int32_t kw0 = w * W, wl = min(W, sl - kw0), k0 = s0 + kw0;
            // prefetch: bounds of window w+1, ids of window w+2
            load_bounds(idn, o0v, o1v);
            int32_t idn2[kStageU];
            load_ids(w + 2, idn2);

            auto deposit = [&](const uint2 x, int32_t tkr) {
                const int32_t rel = tkr + (int32_t)((x.x >> 16) & 0x3Fu) - 1;  // output step - k0
                const int32_t ko = k0 + rel;
                if (ko >= o_begin && ko < o_end) {
                    int32_t slot = phys + rel;
                    if (slot >= R) slot -= R;
                    const int32_t sub = (int32_t)(((x.x >> 25) - b0lo) & 0x7Fu);
                    atomicAdd(&s_ring[slot * words + (int32_t)((x.x >> 22) & 3u) * Pr + sub * X.P +
                                      (int32_t)(x.x & 0xFFFFu)],
                              (int32_t)x.y);
                }
            };
            const int32_t lo = win_lo(w), cntw = win_cnt(w);
            for (int32_t c0 = 0; c0 < cntw; c0 += C) {
                const int32_t cc = min(C, cntw - c0);
                if (c0 > 0) {  // more spikes than one staging buffer holds: stage this chunk now
                    __syncthreads();
                    for (int32_t f = tid; f < cc; f += T) {
                        const int32_t fa = lo + c0 + f;
                        const int32_t* op = X.off + (int64_t)X.ids[fa] * X.os;
                        const int32_t a0 = op[i0], a1 = op[i1];
                        s_span[buf * C + f] = make_int2(a0, (a1 - a0) | (step_of(w, fa) << 24));
                    }
                    __syncthreads();
                }
                const int32_t step = ngrp * kRR;
                auto fetch = [&](int32_t base, int2 (&sp)[kRR], uint2 (&e)[kRR][kRU]) {
#pragma unroll
                    for (int r = 0; r < kRR; ++r) {
                        const int32_t f = base + r * ngrp;
                        sp[r] = (f < cc) ? s_span[buf * C + f] : make_int2(0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < kRR; ++r)
#pragma unroll
                        for (int u = 0; u < kRU; ++u) {
                            const int32_t j = gl + u * LANES;
                            e[r][u] = (j < (sp[r].y & 0xFFFFFF)) ? X.seg[sp[r].x + j] : make_uint2(0u, 0u);
                        }
                };
                auto consume = [&](const int2 (&sp)[kRR], const uint2 (&e)[kRR][kRU]) {
#pragma unroll
                    for (int r = 0; r < kRR; ++r) {
                        const int32_t len = sp[r].y & 0xFFFFFF, tk = (int32_t)((uint32_t)sp[r].y >> 24);
#pragma unroll
                        for (int u = 0; u < kRU; ++u)
                            if (gl + u * LANES < len) deposit(e[r][u], tk);
                        for (int32_t j = gl + kRU * LANES; j < len; j += LANES) deposit(X.seg[sp[r].x + j], tk);
                        // each spike is counted by the time group that owns its input step
                        if (gl == 0 && k0 + tk >= o_begin) ev += (unsigned long long)len;
                    }
                };
                int2 spA[kRR], spB[kRR];
                uint2 eA[kRR][kRU], eB[kRR][kRU];
                int32_t base = grp;
                if (base < cc) fetch(base, spA, eA);
                while (base < cc) {
                    const int32_t b2 = base + step;
                    if (b2 < cc) fetch(b2, spB, eB);
                    consume(spA, eA);
                    if (b2 >= cc) break;
                    const int32_t b3 = b2 + step;
                    if (b3 < cc) fetch(b3, spA, eA);
                    consume(spB, eB);
                    base = b3;
                }
            }
            // stage window w+1 into the other buffer (its bounds were loaded during delivery)
            commit(w + 1, idn, o0v, o1v, buf ^ 1);
            __syncthreads();
            // flush the window's complete output steps and clear their slots
            for (int32_t kk = 0; kk < wl; ++kk) {
                int32_t slot = phys + kk;
                if (slot >= R) slot -= R;
                int32_t* row = s_ring + slot * words;
                const bool own = k0 + kk >= o_begin;
                for (int32_t j = tid; j < words; j += T) {
                    const int32_t ch = j / Pr, pl = j - ch * Pr;
                    if (own && pl < np) X.out[((int64_t)(k0 + kk) * X.nch + ch) * X.n + p0 + pl] = row[j];
                    row[j] = 0;
                }
            }
            phys += wl;
            if (phys >= R) phys -= R;
#pragma unroll
            for (int u = 0; u < kStageU; ++u) idn[u] = idn2[u];
            __syncthreads();
        }
    }
    if (ev) atomicAdd(X.events, ev);
}

// Replay delivery on the fixed-slot layout: the same workgroup/time-group/window structure as
// nbx_replay_ring, but a spike's synapses onto this workgroup's neurons are one aligned
// 128-byte slot at a computed address, so each (spike, workgroup) pair costs one memory
// request and no offset lookup. Eight lanes share a slot (16 bytes, two entries each). A slot
// entry carries its LDS word within a ring step (channel * m*P + neuron) and delay - 1, and
// the ring has a power-of-two step count, so a deposit is one add, mask, multiply-add and LDS
// atomic; range checks run only in windows near the ends of the time group. The window's
// spike ids (with their window step) are staged in LDS, those of the next window are in
// flight while the current one is delivered; synaptic events are counted once per spike from
// its row length by the workgroup of post block 0.
constexpr int kSR = 4;  // spikes per lane group per batch (more spills registers with the prefetched batch)

__global__ __launch_bounds__(1024) void nbx_replay_slots(ReplayArgs X) {
    constexpr int LANES = 8;
    extern __shared__ int32_t s_mem[];
    __shared__ int32_t s_tp[kTpCap + 1];
    const int32_t T = blockDim.x, tid = threadIdx.x;
    const int32_t gl = tid % LANES, grp = tid / LANES, ngrp = T / LANES;
    const int32_t Pr = X.m * X.P, words = X.nch * Pr, R = X.R, C = X.C, W = X.W;
    const uint32_t rmask = (uint32_t)R - 1u;                 // R is a power of two
    int32_t* s_ring = s_mem;                                 // [R][nch][Pr]
    int32_t* s_sid = s_mem + (int64_t)R * words;             // [2][C]: pre | window step << 24

    const int32_t nw = gridDim.x;
    const int32_t lin = (nw % kXcds == 0) ? (int32_t)(blockIdx.x % kXcds) * (nw / kXcds) + (int32_t)blockIdx.x / kXcds
                                          : (int32_t)blockIdx.x;
    const int32_t g = lin / X.Bq, q = lin % X.Bq;
    const int32_t p0 = q * Pr;
    const int32_t np = max(0, min(Pr, X.n - p0));
    const int32_t Kg = (X.K + X.G - 1) / X.G;
    const int32_t o_begin = g * Kg, o_end = min(X.K, o_begin + Kg);
    if (np == 0 || o_begin >= o_end) return;  // uniform over the workgroup
    const int32_t k_begin = max(0, o_begin - (X.D - 2));
    // synaptic events are counted once per time group: the staged spike at window index f is
    // counted by the workgroup with q == f % Ba (Ba = blocks that hold neurons; the others have
    // returned above), which spreads the extra row reads evenly
    const int32_t Ba = min(X.Bq, (X.n + Pr - 1) / Pr);
    const __amdgpu_buffer_rsrc_t slot_rsrc =
        buf_rsrc(X.slots + (int64_t)q * X.n * (kSlotCap / 2), (uint32_t)X.n * 128u);

    for (int32_t j = tid; j < R * words; j += T) s_ring[j] = 0;
    unsigned long long ev = 0;
    int32_t phys = 0;

    for (int32_t s0 = k_begin; s0 < o_end; s0 += kTpCap) {
        const int32_t sl = min(kTpCap, o_end - s0);
        const int32_t nwin = (sl + W - 1) / W;
        __syncthreads();
        for (int32_t j = tid; j <= sl; j += T) s_tp[j] = X.tick_ptr[s0 + j];
        __syncthreads();
        auto win_lo = [&](int32_t w) { return s_tp[w * W]; };
        auto win_cnt = [&](int32_t w) { return s_tp[min(w * W + W, sl)] - s_tp[w * W]; };
        auto load_ids = [&](int32_t w, int32_t (&dst)[kStageU]) {
            const int32_t lo = (w < nwin) ? win_lo(w) : 0;
            const int32_t cnt = (w < nwin) ? min(win_cnt(w), C) : 0;
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                const int32_t f = tid + u * T;
                dst[u] = (f < cnt) ? X.ids[lo + f] : -1;
            }
        };
        auto load_deg = [&](const int32_t (&id)[kStageU], int32_t (&deg)[kStageU]) {
#pragma unroll
            for (int u = 0; u < kStageU; ++u)
                deg[u] = (id[u] >= 0 && (tid + u * T) % Ba == q)
                             ? (int32_t)(X.row_ptr[id[u] + 1] - X.row_ptr[id[u]]) : 0;
        };
        auto step_of = [&](int32_t w, int32_t fa) {
            const int32_t kw0 = w * W, wl = min(W, sl - kw0);
            int32_t lo = 0, hi = wl;
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if (s_tp[kw0 + mid] <= fa) lo = mid;
                else hi = mid;
            }
            return lo;
        };
        auto commit = [&](int32_t w, const int32_t (&id)[kStageU], const int32_t (&deg)[kStageU], int32_t buf) {
            if (w >= nwin) return;
            const int32_t lo = win_lo(w), kw = s0 + w * W;
#pragma unroll
            for (int u = 0; u < kStageU; ++u) {
                const int32_t f = tid + u * T;
                if (id[u] >= 0) {
                    const int32_t tk = step_of(w, lo + f);
                    s_sid[buf * C + f] = id[u] | (tk << 24);
                    if (kw + tk >= o_begin) ev += (unsigned long long)deg[u];  // 0 unless this workgroup counts f
                }
            }
        };

        // one slot batch of the staged spikes [base, base + kSR * ngrp) of staging buffer sb
        // (a lane group without a spike reads zeros: an entry of weight 0)
        auto fetch = [&](int32_t sb, int32_t cc, int32_t base, int32_t (&sid)[kSR], uint4 (&v)[kSR]) {
#pragma unroll
            for (int r = 0; r < kSR; ++r) {
                const int32_t f = base + r * ngrp;
                sid[r] = (f < cc) ? s_sid[sb * C + f] : -1;
            }
#pragma unroll
            for (int r = 0; r < kSR; ++r)
                v[r] = buf_load16(slot_rsrc, sid[r] >= 0 ? (uint32_t)(sid[r] & 0xFFFFFF) * 128u + (uint32_t)gl * 16u
                                                         : kBufOff);
        };
        // the first batch of the next window, fetched before this window's flush so its loads
        // are in flight while the completed steps are written out
        int32_t sidP[kSR];
        uint4 vP[kSR];
        bool pre = false;

        int32_t idn[kStageU], degn[kStageU];
        {
            int32_t id0[kStageU], deg0[kStageU];
            load_ids(0, id0);
            load_ids(1, idn);
            load_deg(id0, deg0);
            commit(0, id0, deg0, 0);
        }
        __syncthreads();

        for (int32_t w = 0; w < nwin; ++w) {
            const int32_t buf = w & 1;
            const int32_t kw0 = w * W, wl = min(W, sl - kw0), k0 = s0 + kw0;
            int32_t idn2[kStageU];
            load_deg(idn, degn);
            load_ids(w + 2, idn2);
            // every arrival of this window lands in [k0, k0 + wl + D - 3]: check only near the ends
            const bool checked = (k0 < o_begin) || (k0 + wl + X.D - 3 >= o_end);

            const int32_t lo = win_lo(w), cntw = win_cnt(w);
            for (int32_t c0 = 0; c0 < cntw; c0 += C) {
                const int32_t cc = min(C, cntw - c0);
                if (c0 > 0) {  // more spikes than one staging buffer holds: stage this chunk now
                    __syncthreads();
                    for (int32_t f = tid; f < cc; f += T) {
                        const int32_t fa = lo + c0 + f;
                        const int32_t pre = X.ids[fa], tk = step_of(w, fa);
                        s_sid[buf * C + f] = pre | (tk << 24);
                        if ((c0 + f) % Ba == q && k0 + tk >= o_begin)
                            ev += (unsigned long long)(X.row_ptr[pre + 1] - X.row_ptr[pre]);
                    }
                    __syncthreads();
                }
                const int32_t step = ngrp * kSR;
                auto consume = [&](const int32_t (&sid)[kSR], const uint4 (&v)[kSR], auto chk) {
                    constexpr bool CHK = decltype(chk)::value;
#pragma unroll
                    for (int r = 0; r < kSR; ++r) {
                        const uint32_t tk = (uint32_t)sid[r] >> 24;
                        const uint32_t tkp = (uint32_t)phys + tk;
                        auto deposit = [&](uint32_t meta, uint32_t qv) {
                            if (CHK) {
                                const int32_t ko = k0 + (int32_t)(tk + (meta >> 16));
                                if (ko < o_begin || ko >= o_end) return;
                            }
                            atomicAdd(&s_ring[((tkp + (meta >> 16)) & rmask) * (uint32_t)words + (meta & 0xFFFFu)],
                                      (int32_t)qv);
                        };
                        if (sid[r] < 0) continue;
                        if (v[r].x != kSlotEmpty) deposit(v[r].x, v[r].y);
                        if ((v[r].z & 0xFF000000u) == kSlotLink) {  // the rest of a long segment
                            const uint32_t cnt = v[r].z & 0xFFFFFFu;
                            for (uint32_t k = 0; k < cnt; ++k) {
                                const uint2 e = X.ovf[(int64_t)v[r].w + k];
                                deposit(e.x, e.y);
                            }
                        } else if (v[r].z != kSlotEmpty) {
                            deposit(v[r].z, v[r].w);
                        }
                    }
                };
                const bool use_pre = pre && c0 == 0;
                auto run = [&](auto chk) {
                    int32_t sidA[kSR], sidB[kSR];
                    uint4 vA[kSR], vB[kSR];
                    int32_t base = grp;
                    if (base < cc) {
                        if (use_pre) {
#pragma unroll
                            for (int r = 0; r < kSR; ++r) {
                                sidA[r] = sidP[r];
                                vA[r] = vP[r];
                            }
                        } else {
                            fetch(buf, cc, base, sidA, vA);
                        }
                    }
                    while (base < cc) {
                        const int32_t b2 = base + step;
                        if (b2 < cc) fetch(buf, cc, b2, sidB, vB);
                        consume(sidA, vA, chk);
                        if (b2 >= cc) break;
                        const int32_t b3 = b2 + step;
                        if (b3 < cc) fetch(buf, cc, b3, sidA, vA);
                        consume(sidB, vB, chk);
                        base = b3;
                    }
                };
                if (checked) run(std::true_type{});
                else run(std::false_type{});
            }
            commit(w + 1, idn, degn, buf ^ 1);
            __syncthreads();
            pre = w + 1 < nwin;
            if (pre) {
                const int32_t cc1 = min(C, win_cnt(w + 1));
                if (grp < cc1) fetch(buf ^ 1, cc1, grp, sidP, vP);
            }
            for (int32_t kk = 0; kk < wl; ++kk) {
                int32_t* row = s_ring + (((uint32_t)phys + (uint32_t)kk) & rmask) * (uint32_t)words;
                const bool own = k0 + kk >= o_begin;
                for (int32_t j = tid; j < words; j += T) {
                    const int32_t ch = j / Pr, pl = j - ch * Pr;
                    if (own && pl < np) X.out[((int64_t)(k0 + kk) * X.nch + ch) * X.n + p0 + pl] = row[j];
                    row[j] = 0;
                }
            }
            phys = (int32_t)(((uint32_t)phys + (uint32_t)wl) & rmask);
#pragma unroll
            for (int u = 0; u < kStageU; ++u) idn[u] = idn2[u];
            __syncthreads();
        }
    }
    if (ev) atomicAdd(X.events, ev);
}

// Replay delivery on the fixed-slot layout without workgroup barriers.
//
// Same ownership as nbx_replay_slots (workgroup = post block q x time group g, arrivals summed in
// an LDS ring of R output steps), but the 16 waves never wait for one another at a window
// boundary, so every wave keeps slot loads in flight all the time:
//  - input steps are taken in super steps of kss steps (about 1024 spikes each); in super step i
//    lane group h of wave v delivers the spikes lo_i + 8v + h + 128r, r = 0, 1, ..., in chunks
//    of 8 rounds: one load per lane fetches a chunk's 64 spike ids, and the input step of each
//    (a search over the super step's kss tick pointers in LDS) rides in the id's top byte;
//  - a wave pipelines its own work: the ids of the next chunk, the slots of the next half-chunk
//    (4 rounds: one 128-byte slot per lane group and round) and the LDS deposits of the current
//    one are in flight together, across super-step boundaries;
//  - wave v owns slice v of the ring's words: it writes out (and clears) that slice of every
//    output step whose inputs all waves have delivered, read from per-wave progress counters;
//  - before depositing super step i a wave waits until every slice of the ring slots that the
//    super step lands in has been cleared (per-wave flush counters). A waiting wave keeps
//    writing out its slice, and R >= kss + D - 1 leaves room for one super step past the
//    completed outputs, so the waits always end.
// Progress and flush counters are LDS words published with workgroup-scope release stores and
// read with acquire loads, so a slice is read after every deposit into it and cleared before
// the next deposit into the same ring slot.
constexpr int kStreamWaves = 16;    // waves per workgroup (1024 threads)
constexpr int kStreamSpikes = 1024; // spikes per super step the super-step length aims for

__device__ inline int32_t wave_min16(const int32_t* a, int32_t lane) {
    int32_t x = __hip_atomic_load(a + (lane & 15), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = min(x, __shfl_xor(x, 8));
    x = min(x, __shfl_xor(x, 4));
    x = min(x, __shfl_xor(x, 2));
    x = min(x, __shfl_xor(x, 1));
    return x;
}

// PACKED: 6-byte entries read by 4 lanes of 32 bytes per slot (build_replay_slots_packed);
// otherwise 8-byte entries read by 8 lanes of 16 bytes.
template <bool PACKED>
__global__ __launch_bounds__(1024) void nbx_replay_stream(ReplayArgs X) {
    constexpr int LPS = PACKED ? 4 : 8;  // lanes per slot
    constexpr int GW = 64 / LPS;         // lane groups per wave
    constexpr int NG = kStreamWaves * GW; // lane groups per workgroup
    constexpr int CH = 64 / GW;          // rounds per chunk (one id load per lane)
    constexpr int SR = CH / 2;           // rounds per half-chunk
    extern __shared__ int32_t s_mem[];
    __shared__ int32_t s_tp[kTpCap + 1];
    __shared__ int32_t s_prog[kStreamWaves], s_fl[kStreamWaves];
    const int32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t gl = lane % LPS, h = lane / LPS;
    const int32_t Pr = X.m * X.P, words = X.nch * Pr, R = X.R;
    int32_t* s_ring = s_mem;  // [R][nch][Pr]

    const int32_t nw = gridDim.x;
    const int32_t lin = (nw % kXcds == 0) ? (int32_t)(blockIdx.x % kXcds) * (nw / kXcds) + (int32_t)blockIdx.x / kXcds
                                          : (int32_t)blockIdx.x;
    const int32_t g = lin / X.Bq, q = lin % X.Bq;
    const int32_t p0 = q * Pr;
    const int32_t np = max(0, min(Pr, X.n - p0));
    const int32_t Kg = (X.K + X.G - 1) / X.G;
    const int32_t o_begin = g * Kg, o_end = min(X.K, o_begin + Kg);
    if (np == 0 || o_begin >= o_end) return;  // uniform over the workgroup
    const int32_t span = X.D - 2;             // output step - input step lies in [0, D - 2]
    const int32_t k_begin = max(0, o_begin - span);
    // super-step cap: W from the host (>= 1), at most R - D + 1 so the waits always end
    const int32_t kss_max = max(1, min(R - span - 1, X.W));
    // spike f is counted (its row length) by the workgroup of its time group with q == f % Ba
    const int32_t Ba = min(X.Bq, (X.n + Pr - 1) / Pr);
    // slices of whole 128-byte lines (32 words), so no output line is written in two pieces
    const int32_t per = 32 * (((words + 31) / 32 + kStreamWaves - 1) / kStreamWaves);
    const int32_t j0 = min(words, wv * per), j1 = min(words, j0 + per);
    const __amdgpu_buffer_rsrc_t slot_rsrc =
        buf_rsrc(X.slots + (int64_t)q * X.n * (kSlotCap / 2), (uint32_t)X.n * 128u);
    const __amdgpu_buffer_rsrc_t ids_rsrc = buf_rsrc(X.ids, 0x7FFFFFFFu);
    const __amdgpu_buffer_rsrc_t rowp_rsrc = buf_rsrc(X.row_ptr, (uint32_t)(X.n + 1) * 8u);

    for (int32_t j = tid; j < R * words; j += blockDim.x) s_ring[j] = 0;
    if (tid < kStreamWaves) s_fl[tid] = k_begin;
    unsigned long long ev = 0;
    int32_t fl = k_begin, fslot = 0;  // this wave's first output step not yet written, its ring slot
    bool bad = false;

    // write out and clear this wave's slice of the output steps [fl, upto)
    auto flush_to = [&](int32_t upto) {
        if (fl >= upto) return;
        for (; fl < upto; ++fl) {
            int32_t* row = s_ring + fslot * words;
            const bool own = fl >= o_begin;
            int64_t obase = (int64_t)fl * X.nch * X.n + p0;
            for (int32_t j = j0 + lane; j < j1; j += 64) {
                int32_t ch = 0, pl = j;
                while (pl >= Pr) {
                    pl -= Pr;
                    ++ch;
                }
                if (own && pl < np) X.out[obase + (int64_t)ch * X.n + pl] = row[j];
                row[j] = 0;
            }
            if (++fslot == R) fslot = 0;
        }
        __hip_atomic_store(&s_fl[wv], fl, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    };

    for (int32_t s0 = k_begin; s0 < o_end; s0 += kTpCap) {
        const int32_t sl = min(kTpCap, o_end - s0), s1 = s0 + sl;
        __syncthreads();  // every wave has finished the previous pass
        for (int32_t j = tid; j <= sl; j += blockDim.x) s_tp[j] = X.tick_ptr[s0 + j];
        if (tid < kStreamWaves) s_prog[tid] = 0;
        __syncthreads();
        const int32_t tot = s_tp[sl] - s_tp[0];
        const int32_t kss =
            max(1, min(kss_max, tot > 0 ? (int32_t)min<int64_t>(kss_max, (int64_t)kStreamSpikes * sl / tot) : kss_max));
        const int32_t nss = (sl + kss - 1) / kss;
        auto ss_lo = [&](int32_t i) { return s_tp[min(i * kss, sl)]; };
        auto ss_hi = [&](int32_t i) { return s_tp[min(i * kss + kss, sl)]; };
        // rounds of this wave in super step i (its lane group 0 has the most spikes)
        auto rounds = [&](int32_t i) {
            const int32_t c = ss_hi(i) - ss_lo(i) - GW * wv;
            return c > 0 ? (c + NG - 1) / NG : 0;
        };
        auto next_chunk = [&](int32_t& i, int32_t& c) {
            if (CH * (c + 1) < rounds(i)) {
                ++c;
                return;
            }
            c = 0;
            for (++i; i < nss && rounds(i) == 0; ++i) {
            }
        };
        // lane's spike of chunk (i, c): its id (a buffer load left in flight; zero when the lane
        // has no spike), its input step within the super step (tk, -1 when none) and whether this
        // workgroup counts its synaptic events
        struct Spk {
            int32_t id, tk;
            bool cnt;
        };
        auto load_ids = [&](int32_t i, int32_t c) {
            Spk r{0, -1, false};
            int32_t f = -1;
            if (i < nss) {
                f = ss_lo(i) + GW * wv + (lane % GW) + NG * (CH * c + lane / GW);
                if (f < ss_hi(i)) {
                    const int32_t t0 = i * kss;
                    int32_t a = 0, b = min(kss, sl - t0);
                    while (b - a > 1) {
                        const int32_t mid = (a + b) >> 1;
                        if (s_tp[t0 + mid] <= f) a = mid;
                        else b = mid;
                    }
                    r.tk = a;
                    r.cnt = (f % Ba == q) && (s0 + t0 + a >= o_begin);
                } else {
                    f = -1;
                }
            }
            r.id = buf_load_i32(ids_rsrc, f >= 0 ? (uint32_t)f * 4u : kBufOff);
            return r;
        };
        // slots of half `half` (rounds SR half .. SR half + SR - 1) of a chunk: group h's spike of
        // each round comes from the lane that loaded it; a round without a spike reads zeros
        constexpr int VL = PACKED ? 2 : 1;  // 16-byte loads per lane and slot
        auto fetch = [&](const Spk& c, int32_t half, int32_t (&sid)[SR], uint4 (&v)[SR][VL]) {
            const int32_t pk = c.tk >= 0 ? (c.id | (c.tk << 24)) : -1;
#pragma unroll
            for (int r = 0; r < SR; ++r) sid[r] = __shfl(pk, (half * SR + r) * GW + h);
#pragma unroll
            for (int r = 0; r < SR; ++r)
#pragma unroll
                for (int u = 0; u < VL; ++u)
                    v[r][u] = buf_load16(slot_rsrc, sid[r] >= 0 ? (uint32_t)(sid[r] & 0xFFFFFF) * 128u +
                                                                      (uint32_t)(gl * VL + u) * 16u
                                                                : kBufOff);
        };
        const uint32_t wmask = (1u << X.wbits) - 1u;
        auto consume = [&](const int32_t (&sid)[SR], const uint4 (&v)[SR][VL], int32_t base) {
#pragma unroll
            for (int r = 0; r < SR; ++r) {
                if (sid[r] < 0) continue;
                const int32_t b2 = base + ((uint32_t)sid[r] >> 24);
                auto deposit = [&](uint32_t meta, uint32_t qv) {
                    int32_t s = b2 + (int32_t)(meta >> 16);
                    if (s >= R) s -= R;
                    atomicAdd(&s_ring[s * words + (int32_t)(meta & 0xFFFFu)], (int32_t)qv);
                };
                if constexpr (PACKED) {
                    auto deposit_p = [&](uint32_t meta, uint32_t qv) {
                        int32_t s = b2 + (int32_t)(meta >> X.wbits);
                        if (s >= R) s -= R;
                        atomicAdd(&s_ring[s * words + (int32_t)(meta & wmask)], (int32_t)qv);
                    };
                    const uint4 a = v[r][0], b = v[r][1];
                    const uint32_t lw = b.w >> 16, c = lw & 7u;
                    if (c > 0) deposit_p(b.y & 0xFFFFu, a.x);
                    if (c > 1) deposit_p(b.y >> 16, a.y);
                    if (c > 2) deposit_p(b.z & 0xFFFFu, a.z);
                    if (c > 3) deposit_p(b.z >> 16, a.w);
                    if (c > 4) deposit_p(b.w & 0xFFFFu, b.x);
                    if (lw & 0x8000u) {  // the rest of a long segment (8-byte entries)
                        const uint32_t n = (lw >> 3) & 0xFFFu;
                        for (uint32_t k = 0; k < n; ++k) {
                            const uint2 e = X.ovf[(int64_t)b.x + k];
                            deposit(e.x, e.y);
                        }
                    }
                    continue;
                }
                const uint4 w4 = v[r][0];
                if (w4.x != kSlotEmpty) deposit(w4.x, w4.y);
                if ((w4.z & 0xFF000000u) == kSlotLink) {  // the rest of a long segment
                    const uint32_t n = w4.z & 0xFFFFFFu;
                    for (uint32_t k = 0; k < n; ++k) {
                        const uint2 e = X.ovf[(int64_t)w4.w + k];
                        deposit(e.x, e.y);
                    }
                } else if (w4.z != kSlotEmpty) {
                    deposit(w4.z, w4.w);
                }
            }
        };
        auto try_flush = [&]() { flush_to(min(s1, s0 + kss * wave_min16(s_prog, lane))); };
        auto publish = [&](int32_t done) {
            __hip_atomic_store(&s_prog[wv], done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            try_flush();
        };
        // every ring slot super step i lands in is cleared in every slice
        auto wait_ring = [&](int32_t i) {
            const int32_t need = s0 + kss * (i + 1) + span - R;
            uint32_t spins = 0;
            while (!bad && wave_min16(s_fl, lane) < need) {
                try_flush();
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) bad = true;
            }
        };

        int32_t ci = 0, cc = 0;
        while (ci < nss && rounds(ci) == 0) ++ci;
        if (ci > 0) publish(min(ci, nss));
        int32_t ni = ci, nc = cc;
        if (ni < nss) next_chunk(ni, nc);
        Spk P = load_ids(ci, cc), Q = load_ids(ni, nc);
        int32_t sidA[SR], sidB[SR];
        uint4 vA[SR][VL], vB[SR][VL];
        // Fetches are unconditional (a lane without a spike reads zeros) and the chunk records P, Q
        // swap roles by unrolling rather than by copies: a load under a branch, or a copy of a load
        // still in flight, makes the compiler wait for it.
        fetch(P, 0, sidA, vA);
        // one chunk: cur's second half and the first half of nxt are fetched while cur is deposited;
        // cur's record is then reloaded with the chunk after nxt
        auto step = [&](Spk& cur, const Spk& nxt) {
            fetch(cur, 1, sidB, vB);
            // row bounds of the counted spikes, summed after the deposits (no wait here)
            const uint32_t ro = cur.cnt ? (uint32_t)cur.id * 8u : kBufOff;
            const int64_t r0 = buf_load_i64(rowp_rsrc, ro), r1 = buf_load_i64(rowp_rsrc, cur.cnt ? ro + 8u : kBufOff);
            const int32_t base = (s0 - k_begin + ci * kss) % R;
            if (cc == 0) wait_ring(ci);
            consume(sidA, vA, base);
            fetch(nxt, 0, sidA, vA);
            consume(sidB, vB, base);
            ev += (unsigned long long)(r1 - r0);
            if (ni != ci) publish(min(ni, nss));
            ci = ni;
            cc = nc;
            if (ni < nss) next_chunk(ni, nc);
            cur = load_ids(ni, nc);
        };
        while (ci < nss) {
            step(P, Q);
            if (ci >= nss) break;
            step(Q, P);
        }
        // drain: write out this wave's slice of every output step of the pass once all waves
        // are done (others may still wait on this wave's slice)
        uint32_t spins = 0;
        while (!bad && fl < s1) {
            try_flush();
            if (fl < s1) __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) bad = true;
        }
    }
    // one atomic per wave for the event count
    for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off);
    if (lane == 0 && ev) atomicAdd(X.events, ev);
    if (bad && lane == 0) atomicOr(X.err, kErrBarrier);
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
            int32_t q = quantize(A.w[j], A.scale);
            const int32_t ch = arrival_channel(A.P.model, syn_type(en));
            if (arrival_negated(A.P.model, syn_type(en))) q = -q;
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
        case NBX_MODEL_LIF_EXP2: return persistent_kernel<NBX_MODEL_LIF_EXP2>(npt);
        case NBX_MODEL_LIF_EXP3: return persistent_kernel<NBX_MODEL_LIF_EXP3>(npt);
        case NBX_MODEL_POISSON: return persistent_kernel<NBX_MODEL_POISSON>(npt);
        case NBX_MODEL_TRIGGER_RELAY: return persistent_kernel<NBX_MODEL_TRIGGER_RELAY>(npt);
        default: return persistent_kernel<NBX_MODEL_IZH_CUR>(npt);
    }
}

struct Geometry {
    int32_t B = 0, P = 0, threads = 0, npt = 0, lds = 0;
    std::string why;
};

struct ReplayGeometry {
    bool ok = false;
    bool slots = false;
    bool stream = false;  // slots delivered by nbx_replay_stream (no window barriers)
    bool packed = false;  // streamed slots with 6-byte entries
    int32_t wbits = 0;
    int32_t P = 0, B = 0;  // layout block size and count
    int32_t m = 1, Bq = 0, G = 1, W = 0, R = 0, C = 0, lanes = 4;
    int64_t lds = 0;
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
    // the delay ring slice, then two spike-total tables of the step counters ([D][kGroups] each)
    g.lds = (net->D * net->nch * P + 2 * net->D * kGroups) * (int32_t)sizeof(int32_t);
    if (g.npt > kMaxNpt) g.why = "more than 4 neurons per thread";
    else if (g.lds > max_lds - 4096) g.why = "delay ring slice does not fit in LDS";
    if (P > 65535) g.why = "block too large for 16-bit local indices";
    return g;
}

}  // namespace

std::string persistent_reason(const nbx_net* net) {
    if (net->has_dense && !net->has_sparse) {  // small dense networks: the single-workgroup engine
        const std::string why = small_dense_reason(net);
        return why.empty() ? why : "dense delivery: " + why;
    }
    if (!net->has_sparse) return "needs sparse synapses";
    if (net->multi()) return "networks of several populations run on the graph engine";
    if (net->md.model == NBX_MODEL_IZH && (net->P.cond & kCondNmda))
        return "Izhikevich NMDA conductances run on the graph engine";
    if (net->has_dense) return "dense delivery runs on the graph engine";
    if (net->stdp) return "STDP networks use arrival-time delivery";
    if (!net->rules.empty()) return "networks with learning rules use arrival-time delivery";
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
    // regions of rb blocks (an even count), each with its own 128-byte aligned group of
    // rb + 1 offsets per row and a 128-byte aligned start in seg
    const int32_t per = B / kXcds;
    const int32_t rb = (per % 2 == 0) ? per : 2 * per;
    const int32_t nr = B / rb;
    const int32_t gs = (rb + 1 + 31) / 32 * 32;
    const int32_t os = nr * gs;
    constexpr int64_t kAlign = 16;  // uint2 entries per 128 bytes
    std::vector<int32_t> off((size_t)n * os, 0);
    std::vector<int32_t> cnt((size_t)B);
    int64_t pos = 0;
    for (int32_t p = 0; p < n; ++p) {  // pass 1: offsets
        const int64_t rbeg = net->h_row_ptr[(size_t)p], rend = net->h_row_ptr[(size_t)p + 1];
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int64_t j = rbeg; j < rend; ++j) cnt[(size_t)(syn_post(net->h_syn[(size_t)j]) / P)]++;
        int32_t* o = off.data() + (size_t)p * os;
        for (int32_t r = 0; r < nr; ++r) {
            pos = (pos + kAlign - 1) / kAlign * kAlign;
            if (pos >= ((int64_t)1 << 31) - 1) {
                set_error("persistent layout: more than 2^31 entries");
                return NBX_ERR_UNSUPPORTED;
            }
            o[r * gs] = (int32_t)pos;
            for (int32_t j = 0; j < rb; ++j) {
                pos += cnt[(size_t)(r * rb + j)];
                o[r * gs + j + 1] = (int32_t)pos;
            }
        }
    }
    std::vector<uint2> seg((size_t)pos, make_uint2(0u, 0u));
    for (int32_t p = 0; p < n; ++p) {  // pass 2: entries, in row order within each block
        const int64_t rbeg = net->h_row_ptr[(size_t)p], rend = net->h_row_ptr[(size_t)p + 1];
        const int32_t* o = off.data() + (size_t)p * os;
        for (int32_t bb = 0; bb < B; ++bb) cnt[(size_t)bb] = o[(bb / rb) * gs + bb % rb];
        for (int64_t j = rbeg; j < rend; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            const int32_t post = syn_post(e), ty = syn_type(e), blk = post / P;
            int32_t qv = (int32_t)std::lrint(net->h_w[(size_t)j] * scale);
            const int32_t ch = arrival_channel(net->P.model, ty);
            if (arrival_negated(net->P.model, ty)) qv = -qv;
            const uint32_t meta = (uint32_t)(post % P) | ((uint32_t)syn_delay(e) << 16) | ((uint32_t)ch << 22) |
                                  (((uint32_t)blk & 0x7Fu) << 25);
            seg[(size_t)cnt[(size_t)blk]++] = make_uint2(meta, (uint32_t)qv);
        }
    }
    L.rb = rb;
    L.nr = nr;
    L.gs = gs;
    L.seg_count = pos;
    L.max_row = 0;
    for (int32_t p = 0; p < n; ++p)
        L.max_row = std::max<int64_t>(L.max_row, net->h_row_ptr[(size_t)p + 1] - net->h_row_ptr[(size_t)p]);
    int32_t rc = net->upload(&L.d_off, off.data(), off.size());
    if (rc) return rc;
    rc = net->upload(&L.d_seg, seg.data(), seg.size());
    if (rc) return rc;
    L.gcap = (B + kGroups - 1) / kGroups * P;  // the neurons of one group
    rc = net->alloc(&L.d_glist, (size_t)net->D * kGroups * L.gcap);
    if (rc) return rc;
    rc = net->alloc(&L.d_gctr, (size_t)2 * (net->D * kGroups + 1) * kCtrLine);  // two banks, each with its start line
    if (rc) return rc;
    L.t_next = -1;  // the counters are cleared before the first launch
    L.epoch = 0;
    L.B = B;
    L.P = P;
    L.threads = g.threads;
    L.npt = g.npt;
    L.lds_bytes = g.lds;
    L.built = true;
    return NBX_OK;
}

static int32_t free_owned(nbx_net* net, void* p) {
    if (!p) return NBX_OK;
    auto it = std::find(net->allocs.begin(), net->allocs.end(), p);
    if (it != net->allocs.end()) net->allocs.erase(it);
    NBX_HIP_TRY(hipFree(p));
    return NBX_OK;
}

// Both blocked layouts carry pre-quantized weights: drop them when the weights change.
int32_t refresh_persistent_weights(nbx_net* net) {
    PersistentLayout& L = net->pl;
    if (L.built) {
        for (void* p : {(void*)L.d_off, (void*)L.d_seg, (void*)L.d_glist, (void*)L.d_gctr})
            NBX_PASS(free_owned(net, p));
        L = PersistentLayout();
    }
    ReplaySlots& RS = net->rs;
    if (RS.built) {
        for (void* p : {(void*)RS.d_slots, (void*)RS.d_ovf}) NBX_PASS(free_owned(net, p));
        RS = ReplaySlots();
    }
    return NBX_OK;
}

static int32_t build_replay_slots(nbx_net* net, int32_t m) {
    ReplaySlots& RS = net->rs;
    if (RS.built && !RS.packed && RS.m == m) return NBX_OK;
    if (RS.built) {
        for (void* p : {(void*)RS.d_slots, (void*)RS.d_ovf}) NBX_PASS(free_owned(net, p));
        RS = ReplaySlots();
    }
    const Geometry L = geometry(net);
    const int32_t n = net->n, Pr = m * L.P, Bq = (L.B + m - 1) / m;
    const float scale = std::ldexp(1.0f, net->shift);
    std::vector<uint2> slots((size_t)n * Bq * kSlotCap, make_uint2(kSlotEmpty, 0u));
    std::vector<uint2> ovf;
    std::vector<int32_t> cnt((size_t)Bq), fill((size_t)Bq);
    for (int32_t p = 0; p < n; ++p) {
        const int64_t rbeg = net->h_row_ptr[(size_t)p], rend = net->h_row_ptr[(size_t)p + 1];
        std::fill(cnt.begin(), cnt.end(), 0);
        std::fill(fill.begin(), fill.end(), 0);
        for (int64_t j = rbeg; j < rend; ++j) cnt[(size_t)(syn_post(net->h_syn[(size_t)j]) / Pr)]++;
        // block-major: the slots one workgroup reads (one post block) are contiguous
        auto slot = [&](int32_t qb) { return slots.data() + ((size_t)qb * n + (size_t)p) * kSlotCap; };
        for (int32_t qb = 0; qb < Bq; ++qb)
            if (cnt[(size_t)qb] > kSlotCap) {  // 15 entries in the slot, the rest behind a link
                slot(qb)[kSlotCap - 1] =
                    make_uint2(kSlotLink | (uint32_t)(cnt[(size_t)qb] - (kSlotCap - 1)), (uint32_t)ovf.size());
                ovf.resize(ovf.size() + (size_t)(cnt[(size_t)qb] - (kSlotCap - 1)));
            }
        for (int64_t j = rbeg; j < rend; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            const int32_t post = syn_post(e), ty = syn_type(e), qb = post / Pr;
            int32_t qv = (int32_t)std::lrint(net->h_w[(size_t)j] * scale);
            const int32_t ch = arrival_channel(net->P.model, ty);
            if (arrival_negated(net->P.model, ty)) qv = -qv;
            const uint2 en = make_uint2((uint32_t)(ch * Pr + post - qb * Pr) | ((uint32_t)(syn_delay(e) - 1) << 16),
                                        (uint32_t)qv);
            const int32_t k = fill[(size_t)qb]++;
            const bool linked = cnt[(size_t)qb] > kSlotCap;
            if (!linked || k < kSlotCap - 1) slot(qb)[k] = en;
            else ovf[(size_t)slot(qb)[kSlotCap - 1].y + (size_t)(k - (kSlotCap - 1))] = en;
        }
    }
    if ((int64_t)ovf.size() >= ((int64_t)1 << 31)) {
        set_error("replay slots: overflow list too long");
        return NBX_ERR_UNSUPPORTED;
    }
    NBX_PASS(net->upload(&RS.d_slots, slots.data(), slots.size()));
    NBX_PASS(net->upload(&RS.d_ovf, ovf.data(), ovf.size()));
    RS.ovf_count = (int64_t)ovf.size();
    RS.m = m;
    RS.Bq = Bq;
    RS.built = true;
    return NBX_OK;
}

// Packed slots: 128 bytes per (presynaptic neuron, post block of m*P neurons), read by four lanes
// of 32 bytes. Lane l holds entries 5l .. 5l+4 as q[0..4] (words 0-4, the quantized weights) and
// metas m[0..4] (16-bit halves of words 5-7: (delay - 1) << wbits | LDS word), and in the last
// half a lane word: entry count (bits 0-2); the last lane of a segment longer than 20 holds 4
// entries, the overflow list offset in q[4], the number of further entries in bits 3-14 and
// the link flag in bit 15. Overflow entries keep the 8-byte format.
static int32_t build_replay_slots_packed(nbx_net* net, int32_t m, int32_t wbits) {
    ReplaySlots& RS = net->rs;
    if (RS.built && RS.packed && RS.m == m && RS.wbits == wbits) return NBX_OK;
    if (RS.built) {
        for (void* p : {(void*)RS.d_slots, (void*)RS.d_ovf}) NBX_PASS(free_owned(net, p));
        RS = ReplaySlots();
    }
    const Geometry L = geometry(net);
    const int32_t n = net->n, Pr = m * L.P, Bq = (L.B + m - 1) / m;
    const float scale = std::ldexp(1.0f, net->shift);
    constexpr int kPer = 5, kLanes = 4, kCap = kPer * kLanes;
    std::vector<uint32_t> slots((size_t)Bq * n * 32, 0u);
    std::vector<uint2> ovf;
    std::vector<int32_t> cnt((size_t)Bq), fill((size_t)Bq);
    std::vector<int64_t> ovf_at((size_t)Bq);
    for (int32_t p = 0; p < n; ++p) {
        const int64_t rbeg = net->h_row_ptr[(size_t)p], rend = net->h_row_ptr[(size_t)p + 1];
        std::fill(cnt.begin(), cnt.end(), 0);
        std::fill(fill.begin(), fill.end(), 0);
        for (int64_t j = rbeg; j < rend; ++j) cnt[(size_t)(syn_post(net->h_syn[(size_t)j]) / Pr)]++;
        auto slot = [&](int32_t qb) { return slots.data() + ((size_t)qb * n + (size_t)p) * 32; };
        for (int32_t qb = 0; qb < Bq; ++qb) {
            const int32_t c = cnt[(size_t)qb];
            uint32_t* sp = slot(qb);
            const int32_t inslot = c > kCap ? kCap - 1 : c;
            for (int32_t l = 0; l < kLanes; ++l) {
                const int32_t lc = std::max(0, std::min(kPer, inslot - kPer * l));
                sp[l * 8 + 7] = (uint32_t)lc << 16;
            }
            if (c > kCap) {
                const int64_t excess = c - (kCap - 1);
                if (excess > 4095) {
                    set_error("replay slots: a segment too long for the packed format");
                    return NBX_ERR_UNSUPPORTED;
                }
                sp[3 * 8 + 4] = (uint32_t)ovf.size();  // lane 3, q[4]
                sp[3 * 8 + 7] |= (0x8000u | (uint32_t)excess << 3) << 16;
                ovf_at[(size_t)qb] = (int64_t)ovf.size();
                ovf.resize(ovf.size() + (size_t)excess);
            }
        }
        for (int64_t j = rbeg; j < rend; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            const int32_t post = syn_post(e), ty = syn_type(e), qb = post / Pr;
            int32_t qv = (int32_t)std::lrint(net->h_w[(size_t)j] * scale);
            const int32_t ch = arrival_channel(net->P.model, ty);
            if (arrival_negated(net->P.model, ty)) qv = -qv;
            const uint32_t word = (uint32_t)(ch * Pr + post - qb * Pr), dly = (uint32_t)(syn_delay(e) - 1);
            const int32_t k = fill[(size_t)qb]++;
            const int32_t c = cnt[(size_t)qb];
            if (c > kCap && k >= kCap - 1) {  // overflow list, 8-byte entries
                ovf[(size_t)(ovf_at[(size_t)qb] + (k - (kCap - 1)))] = make_uint2(word | dly << 16, (uint32_t)qv);
                continue;
            }
            uint32_t* lw = slot(qb) + (k / kPer) * 8;
            const int32_t i = k % kPer;
            lw[i] = (uint32_t)qv;
            const uint32_t meta = dly << wbits | word;
            uint16_t* half = reinterpret_cast<uint16_t*>(lw + 5);  // little endian: meta i is half i
            half[i] = (uint16_t)meta;
        }
    }
    if ((int64_t)ovf.size() >= ((int64_t)1 << 31)) {
        set_error("replay slots: overflow list too long");
        return NBX_ERR_UNSUPPORTED;
    }
    NBX_PASS(net->upload(&RS.d_slots, reinterpret_cast<const uint2*>(slots.data()), slots.size() / 2));
    NBX_PASS(net->upload(&RS.d_ovf, ovf.data(), ovf.size()));
    RS.ovf_count = (int64_t)ovf.size();
    RS.m = m;
    RS.Bq = Bq;
    RS.packed = true;
    RS.wbits = wbits;
    RS.built = true;
    return NBX_OK;
}

static int64_t longest_row(const nbx_net* net) {
    int64_t r = 0;
    for (int32_t p = 0; p < net->n; ++p)
        r = std::max<int64_t>(r, net->h_row_ptr[(size_t)p + 1] - net->h_row_ptr[(size_t)p]);
    return r;
}

hipError_t launch_coresident(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t stream) {
    static const bool profiled = std::getenv("ROCP_TOOL_LIBRARIES") != nullptr;
    if (profiled) return hipLaunchKernel(fn, grid, block, args, lds, stream);
    return hipLaunchCooperativeKernel(fn, grid, block, args, (unsigned)lds, stream);
}

static int32_t env_int(const char* name, int32_t def) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : def;
}

int32_t run_persistent(nbx_net* net, int32_t nsteps) {
    if (net->has_dense && !net->has_sparse) return run_small_dense(net, nsteps);
    int32_t rc = build_persistent_layout(net);
    if (rc) return rc;
    if (nsteps == 0) return NBX_OK;
    PersistentLayout& L = net->pl;
    void* fn = persistent_kernel_for(net->md.model, L.npt);
    if (!L.launch_checked) {  // the kernel, its block size and LDS are fixed with the layout
        int per_cu = 0, cus = 0;
        if (L.lds_bytes > 64 * 1024)
            NBX_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, L.lds_bytes));
        NBX_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, L.threads, (size_t)L.lds_bytes));
        NBX_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device));
        if ((int64_t)per_cu * cus < L.B) {
            set_error("persistent engine: " + std::to_string(L.B) + " workgroups cannot be co-resident");
            return NBX_ERR_UNSUPPORTED;
        }
        L.launch_checked = true;
    }
    PersistArgs X;
    X.A = net->args();
    X.off = L.d_off;
    X.seg = L.d_seg;
    X.B = L.B;
    X.P = L.P;
    X.rb = L.rb;
    X.gs = L.gs;
    X.os = L.nr * L.gs;
    X.t0 = net->t;
    X.nsteps = nsteps;
    // this launch's counter bank must be zero: each launch clears the other bank for the next one,
    // so it holds after a persistent launch; the first launch clears both. List entries carry
    // unique tags, so only a tag wrap needs them cleared
    const bool wrap = (uint64_t)L.epoch + (uint64_t)nsteps + 2u >= ((uint64_t)1 << 32);
    // uint64 per bank: the step counters, then one line holding the start counter
    const int64_t bank_words = ((int64_t)net->D * kGroups + 1) * kCtrLine;
    if (L.t_next < 0 || wrap)
        NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(L.d_gctr), 0, 2 * 2 * bank_words, net->stream));
    if (wrap) {
        NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(L.d_glist), 0, (int64_t)2 * net->D * kGroups * L.gcap,
                                    net->stream));
        L.epoch = 0;
    }
    X.gctr = L.d_gctr + L.bank * bank_words;
    X.gctr_next = L.d_gctr + (1 - L.bank) * bank_words;
    X.gstart = X.gctr + (int64_t)net->D * kGroups * kCtrLine;
    // NBX_PERSISTENT_TEST_ABSENT=k (tests of the failure path only) makes the launch wait for k
    // workgroups that do not exist, with a short wait: the agreement fails as it would if k
    // workgroups could not be resident
    const int32_t absent = std::max(0, env_int("NBX_PERSISTENT_TEST_ABSENT", 0));
    X.start_want = (unsigned long long)L.B + (unsigned long long)absent;
    X.start_spins = absent ? 256u : (1u << 16);
    X.glist = L.d_glist;
    X.gcap = L.gcap;
    X.epoch = L.epoch;
    void* kargs[] = {&X};
    // An ordinary dispatch: co-residency of all L.B workgroups is possible (occupancy x CUs,
    // checked above) and the workgroups agree on it before taking a step (residency agreement).
    NBX_HIP_TRY(hipLaunchKernel(fn, dim3(L.B), dim3(L.threads), kargs, (size_t)L.lds_bytes, net->stream));
    net->t += nsteps;
    L.t_next = net->t;
    L.bank = 1 - L.bank;
    L.epoch += (uint32_t)nsteps + 1u;  // the launch used tags epoch + 1 .. epoch + nsteps
    return NBX_OK;
}

static ReplayGeometry replay_geometry(const nbx_net* net) {
    struct {
        int32_t P, B, rb;
        int64_t max_row;
    } L;
    {
        const Geometry g = geometry(net);
        const int32_t per = g.B / kXcds;
        L.P = g.P;
        L.B = g.B;
        L.rb = (per % 2 == 0) ? per : 2 * per;  // the offset layout's region size
        L.max_row = longest_row(net);
    }
    ReplayGeometry r;
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, net->device) != hipSuccess ||
        max_lds <= 0)
        max_lds = 64 * 1024;
    const int64_t budget = (int64_t)max_lds - (int64_t)(kTpCap + 1) * 4 - 256;  // minus the static LDS
    auto env = [](const char* name, int32_t dflt) {
        const char* v = std::getenv(name);
        return (v && *v) ? (int32_t)std::atoi(v) : dflt;
    };
    // launch-shape overrides (merge factor, window, time groups, staged spikes, slot path) for
    // tuning and tests
    const int32_t m_req = env("NBX_REPLAY_M", 0), w_req = env("NBX_REPLAY_W", 0), g_req = env("NBX_REPLAY_G", 0);
    const int32_t c_req = env("NBX_REPLAY_C", 0), slots_req = env("NBX_REPLAY_SLOTS", -1);
    const int32_t stream_req = env("NBX_REPLAY_STREAM", 1);
    const int32_t c_min = 2048, c_max = kStageU * 1024;
    if (L.max_row >= (1 << 24)) return r;  // segment lengths share a word with the window step
    if (net->n > (1 << 24)) return r;        // staged ids share a word with the window step
    const double outdeg = (double)net->S / std::max<int32_t>(1, net->n);
    int cus = 0;  // the streamed kernel runs one workgroup per CU: as many time groups as CUs hold
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, net->device) != hipSuccess || cus <= 0)
        cus = 256;
    const int64_t sbudget = (int64_t)max_lds - (int64_t)(kTpCap + 1) * 4 - 2 * kStreamWaves * 4 - 256;
    auto stream_shape = [&](int32_t m, int64_t words, bool packed, int32_t wbits) {
        const int64_t Rs = std::min<int64_t>(sbudget / (words * 4), 64);
        if (Rs < net->D) return false;  // one super step of at least one input step past D - 2
        r.slots = r.stream = true;
        r.packed = packed;
        r.wbits = wbits;
        r.P = L.P;
        r.B = L.B;
        r.m = m;
        r.R = (int32_t)Rs;
        // super steps may fill the ring's free steps but two (measured on the 1e5-neuron network
        // with 20-step delays: R = 30, super steps of 6 steps 2.11 ms, of 5 steps 2.17 ms)
        r.W = w_req > 0 ? w_req : std::max<int32_t>(1, r.R - net->D - 1);
        r.lds = Rs * words * 4;
        r.Bq = (L.B + m - 1) / m;
        r.G = g_req > 0 ? g_req : std::max<int32_t>(1, cus / r.Bq);
        r.ok = true;
        return true;
    };
    // Packed slots (20 six-byte entries per 128 bytes) when an entry's LDS word and delay fit
    // 16 bits: the largest block merge whose mean segment stays below ~13.5 entries (about 1 %
    // of segments continue in the overflow list), i.e. the fewest slot reads per spike.
    if (stream_req && env("NBX_REPLAY_PACKED", 1) && slots_req != 0) {
        auto bits = [](int64_t maxval) {
            int32_t b = 0;
            while (((int64_t)1 << b) <= maxval) ++b;
            return b;
        };
        for (int32_t m : {3, 2, 1}) {
            if ((m_req && m != m_req) || m > L.B) continue;
            const double mean_seg = outdeg * m * L.P / std::max<int32_t>(1, net->n);
            if (slots_req < 0 && (mean_seg < 3.0 || mean_seg > 13.5)) continue;
            const int64_t words = (int64_t)net->nch * m * L.P;
            const int32_t wbits = bits(words - 1);
            if (wbits + bits(net->D - 2) > 16) continue;
            if (stream_shape(m, words, true, wbits)) return r;
        }
    }
    for (int32_t m : {3, 2, 1}) {
        if (m_req && m != m_req) continue;
        if (m == 3 && m_req != 3) continue;  // three blocks per workgroup only on request
        if (m > L.B) continue;
        const double mean_seg = outdeg * m * L.P / std::max<int32_t>(1, net->n);
        // a 16-entry slot suits mean segments of ~3..10 entries (longer ones overflow often,
        // shorter ones waste most of the 128 bytes); otherwise the offset layout is used
        // (the streamed kernel on three blocks per workgroup takes up to ~13: its deeper pipeline
        // absorbs the ~12 % of slots that continue in the overflow list)
        const double seg_hi = (m == 3 && stream_req) ? 13.5 : 10.5;
        const bool slots = slots_req >= 0 ? slots_req != 0 : (mean_seg >= 3.0 && mean_seg <= seg_hi);
        if (!slots && L.rb % m != 0) continue;
        const int64_t words = (int64_t)net->nch * m * L.P;
        if (slots && words > 65536) continue;  // a slot entry holds its LDS word in 16 bits
        if (slots && stream_req) {  // ring of as many steps as fit, no staging buffers
            if (stream_shape(m, words, false, 0)) return r;
            continue;
        }
        const int64_t stage_bytes = slots ? 8 : 16;  // per staged spike, two buffers
        const int32_t steps_fit = (int32_t)((budget - c_min * stage_bytes) / (words * 4));
        int32_t R = 0, W = 0;
        if (slots) {  // power-of-two ring
            R = 1;
            while (2 * R <= steps_fit) R *= 2;
            W = std::min(R - (net->D - 1), kMaxWin) / 4 * 4;
        } else {
            W = std::min(steps_fit - (net->D - 1), kMaxWin) / 4 * 4;
        }
        if (w_req) W = std::min(W, w_req);
        if (W < 4) continue;
        if (m == 2 && !m_req && W < 8) continue;  // a short window costs more than the larger segments gain
        if (!slots) R = W + net->D - 1;
        r.slots = slots;
        r.P = L.P;
        r.B = L.B;
        r.m = m;
        r.W = W;
        r.R = R;
        const int64_t ring = (int64_t)r.R * words * 4;
        r.C = (int32_t)std::min<int64_t>(c_max, (budget - ring) / stage_bytes / 64 * 64);
        if (c_req >= 64) r.C = std::min(r.C, c_req / 64 * 64);
        r.lds = ring + (int64_t)r.C * stage_bytes;
        r.Bq = (L.B + m - 1) / m;
        r.G = g_req > 0 ? g_req : m;
        r.lanes = mean_seg >= 12 ? 16 : (mean_seg >= 6 ? 8 : 4);
        r.ok = true;
        return r;
    }
    return r;
}

int32_t replay_path(const nbx_net* net) {
    if (net->stdp || !net->rules.empty() || net->multi()) return NBX_REPLAY_GLOBAL;
    const ReplayGeometry r = replay_geometry(net);
    if (!r.ok) return NBX_REPLAY_GLOBAL;
    if (r.slots) return r.stream ? NBX_REPLAY_STREAM : NBX_REPLAY_SLOTS;
    return persistent_reason(net).empty() || net->pl.built ? NBX_REPLAY_OFFSETS : NBX_REPLAY_GLOBAL;
}

int32_t launch_replay(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t nsteps, int32_t* d_out) {
    const StepArgs A = net->args();
    // the blocked layouts carry pre-quantized weights, so they are not used while STDP changes them
    const ReplayGeometry r = (net->stdp || !net->rules.empty()) ? ReplayGeometry() : replay_geometry(net);
    if (r.ok && (r.slots || build_persistent_layout(net) == NBX_OK)) {
        const PersistentLayout& L = net->pl;
        {
            ReplayArgs X;
            X.off = L.d_off;  // null on the slot path (the offset layout is not built for it)
            X.seg = L.d_seg;
            X.B = r.B;
            X.P = r.P;
            X.rb = L.rb;
            X.gs = L.gs;
            X.os = L.nr * L.gs;
            X.m = r.m;
            X.Bq = r.Bq;
            X.G = r.G;
            X.W = r.W;
            X.R = r.R;
            X.C = r.C;
            X.n = net->n;
            X.nch = net->nch;
            X.D = net->D;
            X.K = nsteps;
            X.ids = d_ids;
            X.tick_ptr = d_tick_ptr;
            X.out = d_out;
            X.events = A.events;
            X.slots = nullptr;
            X.ovf = nullptr;
            X.row_ptr = net->d_row_ptr;
            X.err = net->d_err;
            X.wbits = r.wbits;
            const void* fn = r.lanes == 16 ? (const void*)nbx_replay_ring<16>
                             : r.lanes == 8 ? (const void*)nbx_replay_ring<8>
                                            : (const void*)nbx_replay_ring<4>;
            if (r.packed) {
                NBX_PASS(build_replay_slots_packed(net, r.m, r.wbits));
                X.slots = reinterpret_cast<const uint4*>(net->rs.d_slots);
                X.ovf = net->rs.d_ovf;
                X.Bq = net->rs.Bq;
                fn = (const void*)nbx_replay_stream<true>;
            } else if (r.slots) {
                NBX_PASS(build_replay_slots(net, r.m));
                X.slots = reinterpret_cast<const uint4*>(net->rs.d_slots);
                X.ovf = net->rs.d_ovf;
                X.Bq = net->rs.Bq;
                fn = r.stream ? (const void*)nbx_replay_stream<false> : (const void*)nbx_replay_slots;
            }
            if (r.lds > 64 * 1024)
                NBX_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)r.lds));
            void* kargs[] = {&X};
            NBX_HIP_TRY(hipLaunchKernel(fn, dim3((unsigned)(r.Bq * r.G)), dim3(1024), kargs, (size_t)r.lds,
                                        net->stream));
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
