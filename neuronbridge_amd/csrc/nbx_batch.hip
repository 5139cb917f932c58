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
// captured hipGraph.
//
// Such a launch lasts as long as its busiest workgroup: an 8192-neuron instance puts 8192
// RK4 + Philox updates and the conversion of every fired row's 8192 weights on one CU while
// CUs holding small instances idle. The split schedule (nbx_batch_parts) therefore runs an
// instance as several 256-thread workgroups of at most 1024 (or 2048) columns, each owning a
// range of columns and exchanging spike words with the instance's other parts every step.
// The autotuner times both schedules (and the workgroup sizes of the first) and picks the
// split from the measured per-instance synaptic events. Between launches the last step's
// spikes are kept as the same stamped exchange words under both schedules.
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
#include <cstdlib>
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
constexpr uint32_t kSpinLimit = 1u << 24;  // ~0.5 s of s_sleep(1) polls before giving up

struct InstDev {
    int32_t n;
    int32_t ld;          // weight row stride (elements), n rounded up to kRowPad
    int32_t words;       // raster words per step
    int64_t w_off;       // element offset of W[pre][post] (bf16), row stride ld
    int64_t s_off;       // offset into the state arrays
    int64_t r_off;       // offset of this instance's raster rows (words)
    int64_t b_off;       // offset of this instance's spike exchange words
    int32_t fast;        // every |fixed-point weight| below 2^22: single-fma quantization
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
    // spike exchange words, [2][total words]: {step stamp << 32 | 32 spike bits}; step t's spikes
    // are in half t & 1 with stamp t + 1 (both halves start as stamp 0, no spikes)
    uint64_t* xb;
    int64_t xb_half;
    uint32_t* raster;
    int32_t rec_t0, rec_n;
    unsigned long long* events;
    unsigned long long* inst_events;  // [instances] per-instance event counts (autotuner)
    int32_t* err;
    int32_t t0, nsteps;
    int32_t max_n;  // spike-list length in LDS
};

// One part of an instance: columns [lo, lo + w) (lo a multiple of 32), gathered by cg column
// groups of NPT neurons times rg row groups that split the fired rows between them.
struct PartDev {
    int32_t inst, lo, w, cg, rg, parts;
};

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// A value the compiler must treat as changed here: keeps the step-invariant part of the Philox
// counter (the neuron index) from being hoisted out of the step loop, where the products it
// would keep alive for every neuron of the thread cost more registers than they save.
__device__ inline int32_t opaque(int32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

__device__ inline uint64_t ld_x(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_x(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- weight rows: NPT consecutive elements of a fired row per thread
template <int NPT>
__device__ inline void load_piece(const uint16_t* p, bool ok, uint32_t (&q)[NPT / 2]) {
    if constexpr (NPT == 4) {
        const uint2 v = ok ? *reinterpret_cast<const uint2*>(p) : make_uint2(0u, 0u);
        q[0] = v.x;
        q[1] = v.y;
    } else {
#pragma unroll
        for (int h = 0; h < NPT / 8; ++h) {
            const uint4 v = ok ? reinterpret_cast<const uint4*>(p)[h] : make_uint4(0u, 0u, 0u, 0u);
            q[4 * h] = v.x;
            q[4 * h + 1] = v.y;
            q[4 * h + 2] = v.z;
            q[4 * h + 3] = v.w;
        }
    }
}

// Adds one piece to the int32 fixed-point sums quantize(w, scale) = round-half-even(w * scale).
// FAST (every |w * scale| of the instance below 2^22): one fma(w, scale, 1.5 * 2^23) rounds
// w * scale to an integer exactly as quantize() does (the sum lies in [2^23, 2^24), where floats
// are the integers, and the multiply by a power of two is exact), and its bit pattern is that
// integer plus kMagicBits.
constexpr float kMagic = 12582912.0f;  // 1.5 * 2^23
constexpr int32_t kMagicBits = 0x4B400000;
constexpr int32_t kFastLimit = 1 << 22;

template <int NPT, bool FAST>
__device__ inline void add_piece(int32_t (&acc)[NPT], const uint32_t (&q)[NPT / 2], float scale) {
#pragma unroll
    for (int e = 0; e < NPT / 2; ++e) {
        const float lo = __uint_as_float(q[e] << 16), hi = __uint_as_float(q[e] & 0xFFFF0000u);
        if constexpr (FAST) {
            acc[2 * e] += (int32_t)__float_as_uint(__fmaf_rn(lo, scale, kMagic)) - kMagicBits;
            acc[2 * e + 1] += (int32_t)__float_as_uint(__fmaf_rn(hi, scale, kMagic)) - kMagicBits;
        } else {
            acc[2 * e] += quantize(lo, scale);
            acc[2 * e + 1] += quantize(hi, scale);
        }
    }
}

// Sums this thread's pieces of the rows list[s] over batches of kRows consecutive entries
// starting at s = first, first + step, ... (kRows loads in flight per batch).
template <int NPT, bool FAST, int kRows>
__device__ inline void gather_rows(int32_t (&acc)[NPT], const uint16_t* wbase, int32_t ld, const uint16_t* list,
                                   int32_t cnt, int32_t first, int32_t step, float scale) {
    for (int32_t s0 = first; s0 < cnt; s0 += step) {
        uint32_t q[kRows][NPT / 2];
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            const bool ok = s0 + r < cnt;
            load_piece<NPT>(wbase + (int64_t)(ok ? list[s0 + r] : 0) * ld, ok, q[r]);
        }
#pragma unroll
        for (int r = 0; r < kRows; ++r) add_piece<NPT, FAST>(acc, q[r], scale);
    }
}

// Largest |fixed-point weight| of one instance (it selects the FAST gather).
__global__ void nbx_batch_qmax(const uint16_t* __restrict__ w, int64_t count, float scale, int32_t* qmax) {
    int64_t m = 0;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = quantize(bf16_to_f32(w[j]), scale);
        m = max(m, q < 0 ? -q : q);
    }
    if (m) atomicMax(qmax, (int32_t)min(m, (int64_t)0x7FFFFFFF));
}

// Waits until exchange word p carries `stamp` and returns its spike bits. The spin is bounded:
// on time-out (or when another workgroup has raised the error flag) it returns false.
__device__ inline bool poll_word(const uint64_t* p, uint32_t stamp, uint32_t& bits, int32_t* err) {
    for (uint32_t spins = 0;; ++spins) {
        const uint64_t x = ld_x(p);
        if ((uint32_t)(x >> 32) == stamp) {
            bits = (uint32_t)x;
            return true;
        }
        if ((spins & 63u) == 63u && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
        if (spins > kSpinLimit) {
            atomicExch(err, kErrExchange);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Writes the neuron ids of the set bits of this thread's words b0 (word tid) and b1 (word
// tid + blockDim) to list[] in one workgroup-wide exclusive scan; returns the list length.
// Ids are listed per thread (word tid, then word tid + blockDim): the order is fixed and the
// arrival sums are integer, so it does not affect any result.
__device__ inline int32_t compact_words(uint32_t b0, uint32_t b1, uint16_t* list, int32_t* s_wsum) {
    const int32_t tid = threadIdx.x, T = blockDim.x, lane = tid & 63, wv = tid >> 6, nwv = T >> 6;
    const int32_t c = __popc(b0) + __popc(b1);
    int32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wsum[wv] = x;
    __syncthreads();
    int32_t pos = x - c, total = 0;
    for (int32_t q = 0; q < nwv; ++q) {
        const int32_t s = s_wsum[q];
        if (q < wv) pos += s;
        total += s;
    }
    for (uint32_t m = b0; m; m &= m - 1u) list[pos++] = (uint16_t)(tid * 32 + __builtin_ctz(m));
    for (uint32_t m = b1; m; m &= m - 1u) list[pos++] = (uint16_t)((tid + T) * 32 + __builtin_ctz(m));
    return total;
}

template <int NPT>
__global__ __launch_bounds__(1024) void nbx_batch_kernel(BatchArgs B) {
    static_assert(NPT % 8 == 0, "a thread owns whole 16-byte pieces of a row");
    extern __shared__ uint16_t s_ids[];  // [2][max_n] spike lists, then [blockDim] spike bits
    __shared__ int32_t s_cnt[2];
    __shared__ int32_t s_wsum[16];
    const int32_t inst_id = B.order[blockIdx.x];
    const InstDev& I = B.inst[inst_id];
    const int32_t n = I.n;
    const int32_t T = blockDim.x, tid = threadIdx.x;
    const int32_t c0 = tid * NPT;  // first neuron of this thread
    const bool own = c0 < n;
    uint16_t* s_rb = s_ids + 2 * B.max_n;
    constexpr int kPer = 32 / NPT;  // threads per spike word
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
    {  // the previous step's spikes, as published at the end of the previous launch
        const uint64_t* xw = B.xb + (int64_t)((B.t0 - 1) & 1) * B.xb_half + I.b_off;
        uint32_t b0 = 0, b1 = 0;
        bool bad = false;
        if (tid < I.words) {
            const uint64_t x = ld_x(xw + tid);
            bad |= (uint32_t)(x >> 32) != (uint32_t)B.t0;
            b0 = (uint32_t)x;
        }
        if (tid + T < I.words) {
            const uint64_t x = ld_x(xw + tid + T);
            bad |= (uint32_t)(x >> 32) != (uint32_t)B.t0;
            b1 = (uint32_t)x;
        }
        if (bad) atomicExch(B.err, kErrExchange);
        const int32_t c = compact_words(b0, b1, s_ids, s_wsum);
        if (tid == 0) s_cnt[0] = c;
    }
    __syncthreads();
    unsigned long long ev = 0;
    int cur = 0;
    uint32_t fired_bits = 0;
    const uint16_t* wbase = B.w + I.w_off + c0;
    constexpr int kRows = (NPT == 8) ? 16 : 4;
    constexpr int kRowsFast = kRows / 2;  // the fma path holds both halves of a piece: fewer rows in flight
    for (int32_t k = 0; k < B.nsteps; ++k) {
        const int32_t t = B.t0 + k;
        const int32_t cnt = s_cnt[cur];
        const uint16_t* ids = s_ids + cur * B.max_n;
        // gather the fired rows over this thread's neurons, kRows rows in flight
        int32_t acc[NPT];
#pragma unroll
        for (int j = 0; j < NPT; ++j) acc[j] = 0;
        if (own) {
            if (I.fast) gather_rows<NPT, true, kRowsFast>(acc, wbase, I.ld, ids, cnt, 0, kRowsFast, I.scale);
            else gather_rows<NPT, false, kRows>(acc, wbase, I.ld, ids, cnt, 0, kRows, I.scale);
        }
        if (tid == 0) s_cnt[cur ^ 1] = 0;
        __syncthreads();
        fired_bits = 0;
#pragma unroll
        for (int j = 0; j < NPT; ++j) {
            const int32_t i = c0 + j;
            if (i < n) {
                const int32_t a0 = acc[j] + poisson_count(I.seed_lo, I.seed_hi, opaque(i), t, I.tab, I.kmax) * I.q_bg;
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
    // persist state and publish the last step's spikes for the next launch
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        const int32_t i = c0 + j;
        if (i < n) {
            B.v[I.s_off + i] = V[j];
            B.last[I.s_off + i] = last[j];
        }
    }
    if (B.nsteps > 0) {
        const int32_t tl = B.t0 + B.nsteps - 1;
        __syncthreads();
        s_rb[tid] = (uint16_t)fired_bits;
        __syncthreads();
        uint64_t* xw = B.xb + (int64_t)(tl & 1) * B.xb_half + I.b_off;
        for (int32_t w = tid; w < I.words; w += T) {
            uint32_t word = 0;
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const int32_t th = w * kPer + q;
                if (th * NPT < n) word |= (uint32_t)s_rb[th] << (q * NPT);
            }
            st_x(xw + w, ((uint64_t)(uint32_t)(tl + 1) << 32) | word);
        }
    }
    if (tid == 0 && ev) {
        atomicAdd(B.events, ev);
        atomicAdd(B.inst_events + inst_id, ev);
    }
}

void* batch_kernel(int npt) { return npt == 8 ? (void*)nbx_batch_kernel<8> : (void*)nbx_batch_kernel<16>; }

// Split instances. Workgroup b runs part parts[b] of an instance: its column groups own NPT
// neurons each (state in registers), its row groups split the previous step's fired rows and
// add their partial sums through LDS in a fixed order. After the update each part publishes
// its spike words (stamp t + 1) with device-scope stores; then every part of the instance
// polls the instance's words until all carry the stamp and rebuilds the full spike list.
// There is no barrier: a word is written once per step and carries its own step, and the
// half-per-parity layout means a word is rewritten (two steps later) only after every part
// has read it (a part cannot publish step t + 1 before it has read all of step t, which every
// part publishes only after it has read all of step t - 1). All parts of a launch must be
// co-resident (cooperative launch); every wait is bounded by kSpinLimit and the error flag.
// An instance of one part keeps its exchange in LDS and publishes only its last step.
constexpr int kPartThreads = 256;

template <int NPT>
__global__ __launch_bounds__(kPartThreads) __attribute__((amdgpu_waves_per_eu(NPT == 4 ? 5 : 4)))
void nbx_batch_parts(BatchArgs B, const PartDev* __restrict__ parts) {
    static_assert(NPT == 4 || NPT == 8, "pieces of 8 or 16 bytes");
    constexpr int kGroupsPerWord = 32 / NPT;
    constexpr int kRows = NPT == 4 ? 16 : 12;  // loads in flight per thread (more spill registers at NPT 8)
    constexpr int kRowsFast = NPT == 4 ? 12 : 6;  // the fma path holds both halves of a piece
    extern __shared__ __align__(16) uint16_t s_list[];                          // [max_n] spike ids
    int32_t* s_red = reinterpret_cast<int32_t*>(s_list + ((B.max_n + 7) & ~7));  // [rg][NPT cg] partial sums
    __shared__ int32_t s_cnt, s_abort;
    __shared__ int32_t s_wsum[kPartThreads / 64];
    __shared__ uint8_t s_rb[kPartThreads];  // this step's spike bits per column group
    const PartDev pd = parts[blockIdx.x];
    const InstDev& I = B.inst[pd.inst];
    const int32_t n = I.n, tid = threadIdx.x;
    const int32_t cgi = tid % pd.cg, rgi = tid / pd.cg;
    const int32_t c0 = pd.lo + cgi * NPT;
    const bool act = rgi < pd.rg && cgi * NPT < pd.w;  // gathers columns c0 .. c0 + NPT - 1 (c0 < n)
    const bool owner = act && rgi == 0;                // and keeps their state
    const int32_t wl = (pd.w + 31) / 32;               // spike words of this part
    const int32_t wg0 = pd.lo / 32;                    // its first word in the instance
    const int32_t pitch = NPT * pd.cg;                 // s_red row-group stride
    ModelParams P = B.P;
    P.tau = I.tau;
    float V[NPT];
    int32_t last[NPT];
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        const int32_t i = c0 + j;
        V[j] = 0.0f;
        last[j] = 0;
        if (owner && i < n) {
            V[j] = B.v[I.s_off + i];
            last[j] = B.last[I.s_off + i];
        }
    }
    if (tid == 0) s_abort = 0;
    {  // the previous step's spikes, as published at the end of the previous launch
        const uint64_t* xw = B.xb + (int64_t)((B.t0 - 1) & 1) * B.xb_half + I.b_off;
        uint32_t b0 = 0, b1 = 0;
        bool bad = false;
        if (tid < I.words) {
            const uint64_t x = ld_x(xw + tid);
            bad |= (uint32_t)(x >> 32) != (uint32_t)B.t0;
            b0 = (uint32_t)x;
        }
        if (tid + kPartThreads < I.words) {
            const uint64_t x = ld_x(xw + tid + kPartThreads);
            bad |= (uint32_t)(x >> 32) != (uint32_t)B.t0;
            b1 = (uint32_t)x;
        }
        if (bad) atomicExch(B.err, kErrExchange);
        const int32_t c = compact_words(b0, b1, s_list, s_wsum);
        if (tid == 0) s_cnt = c;
    }
    __syncthreads();
    uint32_t nspk = 0;  // spikes of this thread's words
    const uint16_t* wbase = B.w + I.w_off + c0;
    // Poisson background counts of the coming step: they depend on no other part, so each step's
    // are drawn while the part waits for the instance's spike words of the step before
    int32_t pc[NPT];
    auto draw = [&](int32_t t) {
#pragma unroll
        for (int j = 0; j < NPT; ++j)
            pc[j] = (owner && c0 + j < n) ? poisson_count(I.seed_lo, I.seed_hi, opaque(c0 + j), t, I.tab, I.kmax) : 0;
    };
    draw(B.t0);
    for (int32_t k = 0; k < B.nsteps; ++k) {
        const int32_t t = B.t0 + k;
        const int32_t cnt = s_cnt;
        int32_t acc[NPT];
#pragma unroll
        for (int j = 0; j < NPT; ++j) acc[j] = 0;
        // row group g takes the batches of kRows consecutive list entries g, g + rg, g + 2 rg, ...
        if (act) {
            if (I.fast)
                gather_rows<NPT, true, kRowsFast>(acc, wbase, I.ld, s_list, cnt, rgi * kRowsFast, kRowsFast * pd.rg, I.scale);
            else gather_rows<NPT, false, kRows>(acc, wbase, I.ld, s_list, cnt, rgi * kRows, kRows * pd.rg, I.scale);
        }
        if (pd.rg > 1) {  // row groups 1.. hand their partial sums to row group 0 through LDS
            if (act && rgi > 0) {
                int4* dst = reinterpret_cast<int4*>(s_red + rgi * pitch + cgi * NPT);
#pragma unroll
                for (int h = 0; h < NPT / 4; ++h)
                    dst[h] = make_int4(acc[4 * h], acc[4 * h + 1], acc[4 * h + 2], acc[4 * h + 3]);
            }
            __syncthreads();
            if (owner)
                for (int32_t g = 1; g < pd.rg; ++g) {
                    const int4* src = reinterpret_cast<const int4*>(s_red + g * pitch + cgi * NPT);
#pragma unroll
                    for (int h = 0; h < NPT / 4; ++h) {
                        const int4 a = src[h];
                        acc[4 * h] += a.x;
                        acc[4 * h + 1] += a.y;
                        acc[4 * h + 2] += a.z;
                        acc[4 * h + 3] += a.w;
                    }
                }
        }
        uint32_t fired = 0;
        if (owner) {
#pragma unroll
            for (int j = 0; j < NPT; ++j) {
                const int32_t i = c0 + j;
                if (i < n) {
                    const int32_t a0 = acc[j] + pc[j] * I.q_bg;
                    if (last[j] > P.t_ref) V[j] += (float)a0 * I.inv_scale;
                    float g = 0.0f;
                    if (lif_step<false>(P, V[j], g, 0.0f, last[j])) fired |= 1u << j;
                }
            }
        }
        if (rgi == 0) s_rb[cgi] = (uint8_t)fired;
        __syncthreads();  // every gather read of s_list and s_red is done
        // this part's spike words: publish, record, count
        uint32_t word = 0;
        if (tid < wl) {
#pragma unroll
            for (int q = 0; q < kGroupsPerWord; ++q) word |= (uint32_t)s_rb[kGroupsPerWord * tid + q] << (q * NPT);
            if (pd.parts > 1 || k == B.nsteps - 1)
                st_x(B.xb + (int64_t)(t & 1) * B.xb_half + I.b_off + wg0 + tid, ((uint64_t)(uint32_t)(t + 1) << 32) | word);
            if (B.raster && t >= B.rec_t0 && t < B.rec_t0 + B.rec_n)
                B.raster[I.r_off + (int64_t)(t - B.rec_t0) * I.words + wg0 + tid] = word;
            nspk += (uint32_t)__popc(word);
        }
        draw(t + 1);
        // the instance's spike words of step t
        uint32_t b0 = 0, b1 = 0;
        if (pd.parts > 1) {
            const uint64_t* xw = B.xb + (int64_t)(t & 1) * B.xb_half + I.b_off;
            bool ok = true;
            if (tid < I.words) ok &= poll_word(xw + tid, (uint32_t)(t + 1), b0, B.err);
            if (tid + kPartThreads < I.words) ok &= poll_word(xw + tid + kPartThreads, (uint32_t)(t + 1), b1, B.err);
            if (!ok) s_abort = 1;
        } else {
            b0 = tid < wl ? word : 0u;  // one part: tid < wl covers every word of the instance
        }
        const int32_t c = compact_words(b0, b1, s_list, s_wsum);  // its barrier also publishes s_abort
        if (s_abort) break;
        if (tid == 0) s_cnt = c;
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NPT; ++j) {
        const int32_t i = c0 + j;
        if (owner && i < n) {
            B.v[I.s_off + i] = V[j];
            B.last[I.s_off + i] = last[j];
        }
    }
    if (nspk) {
        const unsigned long long ev = (unsigned long long)nspk * (unsigned long long)n;
        atomicAdd(B.events, ev);
        atomicAdd(B.inst_events + pd.inst, ev);
    }
}

const void* parts_kernel(int npt) { return npt == 4 ? (const void*)nbx_batch_parts<4> : (const void*)nbx_batch_parts<8>; }

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
    uint64_t* d_xb = nullptr;
    uint32_t* d_raster = nullptr;
    unsigned long long* d_events = nullptr;
    unsigned long long* d_inst_ev = nullptr;
    int32_t* d_err = nullptr;
    int32_t rec_t0 = 0, rec_n = 0;
    int32_t t = 0;
    int32_t threads = 0;  // 0 = not tuned yet
    int32_t chunk = 500;
    hipGraphExec_t gexec = nullptr;
    // split schedule: parts per instance (empty = one workgroup per instance, nbx_batch_kernel)
    std::vector<int32_t> h_nparts;
    std::vector<PartDev> h_parts;
    PartDev* d_parts = nullptr;
    size_t parts_cap = 0;
    int32_t part_npt = 8;        // neurons per thread of the split schedule's kernel (4 or 8)
    std::vector<double> h_work;  // per-instance work estimate (synaptic events), heaviest parts first

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
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }
    // neurons per thread: 8 or 16 (whole 16-byte row pieces); 32 means the size does not fit
    int npt_for(int32_t thr) const {
        const int need = (max_n + thr - 1) / thr;
        return need <= 8 ? 8 : (need <= 16 ? 16 : 32);
    }
    size_t part_lds(int32_t npt) const {
        return (size_t)((max_n + 7) & ~7) * sizeof(uint16_t) + (size_t)kPartThreads * npt * 4;
    }
    BatchArgs args(int32_t t0, int32_t nsteps) const {
        BatchArgs A{};
        A.P = P;
        A.inst = d_inst;
        A.order = d_order;
        A.w = d_w;
        A.v = d_v;
        A.last = d_last;
        A.xb = d_xb;
        A.xb_half = total_words;
        A.raster = d_raster;
        A.rec_t0 = rec_t0;
        A.rec_n = rec_n;
        A.events = d_events;
        A.inst_events = d_inst_ev;
        A.err = d_err;
        A.t0 = t0;
        A.nsteps = nsteps;
        A.max_n = max_n;
        return A;
    }
    hipError_t launch(int32_t t0, int32_t nsteps, int32_t thr) {
        BatchArgs A = args(t0, nsteps);
        void* fn = batch_kernel(npt_for(thr));
        const size_t lds = ((size_t)2 * max_n + (size_t)thr) * sizeof(uint16_t);
        void* kargs[] = {&A};
        return hipLaunchKernel(fn, dim3(count), dim3(thr), kargs, lds, stream);
    }
    // Every part of a launch must be co-resident: the cooperative launch has the runtime guarantee
    // it (capacity is also checked when the split is set).
    hipError_t launch_parts(int32_t t0, int32_t nsteps) {
        BatchArgs A = args(t0, nsteps);
        const PartDev* parts = d_parts;
        void* kargs[] = {&A, &parts};
        const void* fn = parts_kernel(part_npt);
        return launch_coresident(fn, dim3((unsigned)h_parts.size()), dim3(kPartThreads), kargs, part_lds(part_npt),
                                 stream);
    }
    int32_t part_capacity(int32_t npt, int32_t* cap) const {
        int per_cu = 0, cus = 0;
        NBX_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, parts_kernel(npt), kPartThreads,
                                                                 part_lds(npt)));
        NBX_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        *cap = per_cu * cus;
        return NBX_OK;
    }
    // Builds the parts of a split schedule (parts[i] requested for instance i; the part width is
    // rounded up to 32 columns, so an instance may get fewer) and uploads them, heaviest first.
    // Parts of at most 1024 columns run 4 neurons per thread, wider ones (up to 2048) 8.
    int32_t set_parts(const int32_t* req) {
        std::vector<int32_t> np((size_t)count), pwv((size_t)count);
        std::vector<PartDev> parts;
        std::vector<double> load;
        int32_t max_pw = 0;
        for (int32_t i = 0; i < count; ++i) {
            const int32_t n = h_inst[(size_t)i].n;
            const int32_t S = std::max(1, std::min(req[i], (n + 31) / 32));
            const int32_t pw = ((n + S - 1) / S + 31) / 32 * 32;
            if (pw > kPartThreads * 8) {
                set_error("nbx_batch_set_parts: instance " + std::to_string(i) + " (n = " + std::to_string(n) +
                          ") needs at least " + std::to_string((n + kPartThreads * 8 - 1) / (kPartThreads * 8)) +
                          " parts");
                return NBX_ERR_INVALID;
            }
            pwv[(size_t)i] = pw;
            max_pw = std::max(max_pw, pw);
        }
        const int32_t npt = max_pw <= kPartThreads * 4 ? 4 : 8;
        for (int32_t i = 0; i < count; ++i) {
            const int32_t n = h_inst[(size_t)i].n, pw = pwv[(size_t)i];
            np[(size_t)i] = (n + pw - 1) / pw;
            const int32_t cg = pw / npt;
            for (int32_t j = 0; j < np[(size_t)i]; ++j) {
                PartDev d{};
                d.inst = i;
                d.lo = j * pw;
                d.w = std::min(pw, n - d.lo);
                d.cg = cg;
                d.rg = kPartThreads / cg;
                d.parts = np[(size_t)i];
                parts.push_back(d);
                load.push_back(h_work[(size_t)i] / np[(size_t)i]);
            }
        }
        int32_t cap = 0;
        NBX_PASS(part_capacity(npt, &cap));
        if ((int64_t)parts.size() > cap) {
            set_error("nbx_batch_set_parts: " + std::to_string(parts.size()) + " parts exceed the " +
                      std::to_string(cap) + " co-resident workgroups");
            return NBX_ERR_UNSUPPORTED;
        }
        std::vector<int32_t> idx(parts.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return load[(size_t)x] > load[(size_t)y]; });
        h_parts.clear();
        for (int32_t k : idx) h_parts.push_back(parts[(size_t)k]);
        if (h_parts.size() > parts_cap) {
            NBX_PASS(alloc(&d_parts, h_parts.size()));
            parts_cap = h_parts.size();
        }
        NBX_HIP_TRY(hipStreamSynchronize(stream));
        NBX_HIP_TRY(hipMemcpy(d_parts, h_parts.data(), sizeof(PartDev) * h_parts.size(), hipMemcpyHostToDevice));
        h_nparts = np;
        part_npt = npt;
        drop_graph();
        return NBX_OK;
    }
    // Split plan for the autotuner: the fewest parts of at most 256 * npt columns, then one more
    // part at a time for the instance with the most work per part while that exceeds twice the
    // mean work per part, parts keep at least 256 columns and the co-resident capacity allows.
    // Parts of similar work spread evenly over the CUs whatever the dispatcher's placement.
    std::vector<int32_t> plan_parts(int32_t npt, int32_t cap) const {
        const int32_t maxw = kPartThreads * npt;
        std::vector<int32_t> S((size_t)count);
        double total = 0.0;
        int64_t used = 0;
        for (int32_t i = 0; i < count; ++i) {
            S[(size_t)i] = (h_inst[(size_t)i].n + maxw - 1) / maxw;
            total += h_work[(size_t)i];
            used += S[(size_t)i];
        }
        std::vector<bool> done((size_t)count, false);
        while (used < cap) {
            int32_t best = -1;
            double bw = 0.0;
            for (int32_t i = 0; i < count; ++i) {
                const double w = h_work[(size_t)i] / S[(size_t)i];
                if (!done[(size_t)i] && w > bw) {
                    bw = w;
                    best = i;
                }
            }
            if (best < 0 || bw <= 2.0 * total / (double)used) break;
            const int32_t n = h_inst[(size_t)best].n, s2 = S[(size_t)best] + 1;
            if ((n + s2 - 1) / s2 < 256) {
                done[(size_t)best] = true;
                continue;
            }
            S[(size_t)best] = s2;
            ++used;
        }
        return S;
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

// Snapshot of the mutable state, for timing runs that must leave no trace.
struct BatchSnapshot {
    std::vector<float> v;
    std::vector<int32_t> last;
    std::vector<uint64_t> xb;
    std::vector<unsigned long long> inst_ev;
    unsigned long long ev = 0;
    int32_t save(nbx_batch* b) {
        const size_t n = (size_t)b->total_n, nw = 2 * (size_t)b->total_words;
        v.resize(n);
        last.resize(n);
        xb.resize(nw);
        inst_ev.resize((size_t)b->count);
        NBX_HIP_TRY(hipStreamSynchronize(b->stream));
        NBX_HIP_TRY(hipMemcpy(v.data(), b->d_v, sizeof(float) * n, hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(last.data(), b->d_last, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(xb.data(), b->d_xb, sizeof(uint64_t) * nw, hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(inst_ev.data(), b->d_inst_ev, sizeof(unsigned long long) * inst_ev.size(),
                              hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(&ev, b->d_events, sizeof(ev), hipMemcpyDeviceToHost));
        return NBX_OK;
    }
    int32_t restore(nbx_batch* b) const {
        NBX_HIP_TRY(hipStreamSynchronize(b->stream));
        NBX_HIP_TRY(hipMemcpy(b->d_v, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
        NBX_HIP_TRY(hipMemcpy(b->d_last, last.data(), sizeof(int32_t) * last.size(), hipMemcpyHostToDevice));
        NBX_HIP_TRY(hipMemcpy(b->d_xb, xb.data(), sizeof(uint64_t) * xb.size(), hipMemcpyHostToDevice));
        NBX_HIP_TRY(hipMemcpy(b->d_inst_ev, inst_ev.data(), sizeof(unsigned long long) * inst_ev.size(),
                              hipMemcpyHostToDevice));
        NBX_HIP_TRY(hipMemcpy(b->d_events, &ev, sizeof(ev), hipMemcpyHostToDevice));
        return NBX_OK;
    }
};

int32_t check_batch_error(nbx_batch* b) {
    int32_t e = 0;
    NBX_HIP_TRY(hipMemcpy(&e, b->d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (e == 0) return NBX_OK;
    const int32_t z = 0;
    NBX_HIP_TRY(hipMemcpy(b->d_err, &z, sizeof(int32_t), hipMemcpyHostToDevice));
    set_error(e == kErrExchange ? "batched engine: spike exchange out of step or timed out (parts not co-resident?)"
                                : "batched engine: device error flag " + std::to_string(e));
    return NBX_ERR_STATE;
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
        I.b_off = b->total_words;
        I.tau = s.tau;
        I.scale = std::ldexp(1.0f, s.fixed_shift);
        I.inv_scale = 1.0f / I.scale;
        I.q_bg = (int32_t)std::lrint(d->w_bg * I.scale);
        I.seed_lo = s.seed_lo;
        I.seed_hi = s.seed_hi;
        I.kmax = s.pois_lambda > 0.0 ? poisson_tab(s.pois_lambda, I.tab, kMaxPoisTab) : 0;
        b->h_inst.push_back(I);
        b->h_work.push_back((double)s.n * (double)s.n);  // until the autotuner has measured
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
    // instances whose fixed-point weights all stay below 2^22 take the single-fma quantization;
    // NBX_BATCH_EXACT=1 keeps every instance on the scale-and-round path (for tests)
    {
        const char* keep = std::getenv("NBX_BATCH_EXACT");
        int32_t* d_qmax = nullptr;
        if ((rc = b->alloc(&d_qmax, (size_t)b->count))) return rc;
        NBX_HIP_TRY(hipMemsetAsync(d_qmax, 0, sizeof(int32_t) * (size_t)b->count, b->stream));
        auto grid = [](int64_t cnt) { return (unsigned)std::min<int64_t>((cnt + 255) / 256, 2048); };
        for (int32_t k = 0; k < d->count; ++k) {
            const InstDev& I = b->h_inst[(size_t)k];
            const int64_t cnt = (int64_t)I.n * I.ld;
            nbx_batch_qmax<<<grid(cnt), 256, 0, b->stream>>>(b->d_w + I.w_off, cnt, I.scale, d_qmax + k);
        }
        NBX_HIP_TRY(hipGetLastError());
        std::vector<int32_t> qmax((size_t)b->count);
        NBX_HIP_TRY(hipStreamSynchronize(b->stream));
        NBX_HIP_TRY(hipMemcpy(qmax.data(), d_qmax, sizeof(int32_t) * qmax.size(), hipMemcpyDeviceToHost));
        for (int32_t k = 0; k < d->count; ++k) {
            InstDev& I = b->h_inst[(size_t)k];
            I.fast = !(keep && keep[0] == '1') && qmax[(size_t)k] < kFastLimit;
        }
    }
    std::vector<int32_t> last((size_t)b->total_n, 10000);
    if ((rc = b->alloc(&b->d_v, (size_t)b->total_n))) return rc;
    if ((rc = b->alloc(&b->d_last, (size_t)b->total_n))) return rc;
    if ((rc = b->alloc(&b->d_xb, 2 * (size_t)b->total_words))) return rc;
    if ((rc = b->alloc(&b->d_events, 1))) return rc;
    if ((rc = b->alloc(&b->d_inst_ev, (size_t)b->count))) return rc;
    if ((rc = b->alloc(&b->d_err, 1))) return rc;
    if ((rc = b->alloc(&b->d_inst, b->h_inst.size()))) return rc;
    if ((rc = b->alloc(&b->d_order, (size_t)b->count))) return rc;
    NBX_HIP_TRY(hipMemcpy(b->d_v, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(b->d_last, last.data(), sizeof(int32_t) * last.size(), hipMemcpyHostToDevice));
    // clears ordered with the batch's (non-blocking) stream, finished before the call returns
    NBX_HIP_TRY(hipMemsetAsync(b->d_xb, 0, sizeof(uint64_t) * 2 * (size_t)b->total_words, b->stream));  // stamp 0
    NBX_HIP_TRY(hipMemsetAsync(b->d_events, 0, sizeof(unsigned long long), b->stream));
    NBX_HIP_TRY(hipMemsetAsync(b->d_inst_ev, 0, sizeof(unsigned long long) * (size_t)b->count, b->stream));
    NBX_HIP_TRY(hipMemsetAsync(b->d_err, 0, sizeof(int32_t), b->stream));
    NBX_HIP_TRY(hipStreamSynchronize(b->stream));
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
        NBX_HIP_TRY(hipMemsetAsync(b->d_raster, 0, sizeof(uint32_t) * (size_t)off, b->stream));
        NBX_HIP_TRY(hipStreamSynchronize(b->stream));  // the stream is non-blocking
    }
    b->drop_graph();
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
    b->h_nparts.clear();  // one workgroup per instance
    b->drop_graph();
    return NBX_OK;
}

int32_t nbx_batch_get_schedule(const nbx_batch* b, int32_t* threads, int32_t* chunk) {
    if (!b || !threads || !chunk) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    *threads = b->h_nparts.empty() ? b->threads : kPartThreads;
    *chunk = b->chunk;
    return NBX_OK;
}

int32_t nbx_batch_set_parts(nbx_batch* b, const int32_t* parts, int32_t count) {
    if (!b || !parts || count != b->count) {
        set_error("nbx_batch_set_parts: one entry per instance required");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    return b->set_parts(parts);
}

int32_t nbx_batch_get_parts(const nbx_batch* b, int32_t* parts, int32_t count) {
    if (!b || !parts || count != b->count) {
        set_error("nbx_batch_get_parts: one entry per instance required");
        return NBX_ERR_INVALID;
    }
    for (int32_t i = 0; i < count; ++i) parts[i] = b->h_nparts.empty() ? 0 : b->h_nparts[(size_t)i];
    return NBX_OK;
}

int32_t nbx_batch_run_async(nbx_batch* b, int32_t nsteps);

// Autotuner: times `steps` steps for each workgroup size of the one-workgroup-per-instance
// kernel, measures each instance's synaptic events, then times the split schedule planned
// from them, all on a snapshot of the state (restored afterwards), and keeps the fastest.
int32_t nbx_batch_tune(nbx_batch* b, int32_t steps, int32_t* chosen) {
    if (!b || steps <= 0) {
        set_error("nbx_batch_tune: bad arguments");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(b->device));
    BatchSnapshot snap;
    NBX_PASS(snap.save(b));
    const int32_t saved_rec = b->rec_n;
    b->rec_n = 0;  // no raster writes while tuning
    float best = 1e30f;
    int32_t best_thr = 0;
    hipEvent_t e0, e1;
    NBX_HIP_TRY(hipEventCreate(&e0));
    NBX_HIP_TRY(hipEventCreate(&e1));
    auto timed = [&](auto&& go, float* ms) -> int32_t {
        NBX_PASS(snap.restore(b));
        NBX_HIP_TRY(go(b->t, std::min(steps, 20)));  // warm
        NBX_HIP_TRY(hipEventRecord(e0, b->stream));
        NBX_HIP_TRY(go(b->t + std::min(steps, 20), steps));
        NBX_HIP_TRY(hipEventRecord(e1, b->stream));
        NBX_HIP_TRY(hipEventSynchronize(e1));
        NBX_HIP_TRY(hipEventElapsedTime(ms, e0, e1));
        return check_batch_error(b);
    };
    int32_t rc = NBX_OK;
    for (int32_t thr : {256, 512, 1024}) {
        if (b->npt_for(thr) > 16 || rc) continue;
        float ms = 0.0f;
        rc = timed([&](int32_t t0, int32_t k) { return b->launch(t0, k, thr); }, &ms);
        if (!rc && ms < best) {
            best = ms;
            best_thr = thr;
        }
    }
    // per-instance work of the last timed run, then the split schedule planned from it
    std::vector<unsigned long long> iev((size_t)b->count);
    bool split = false;
    if (!rc) {
        rc = [&]() -> int32_t {
            NBX_HIP_TRY(hipMemcpy(iev.data(), b->d_inst_ev, sizeof(unsigned long long) * iev.size(),
                                  hipMemcpyDeviceToHost));
            return NBX_OK;
        }();
    }
    if (!rc) {
        for (int32_t i = 0; i < b->count; ++i) {
            const double ev = (double)(iev[(size_t)i] - snap.inst_ev[(size_t)i]);
            // a neuron update costs about as much as gathering 8 synaptic events
            b->h_work[(size_t)i] = ev + 8.0 * (double)b->h_inst[(size_t)i].n * (steps + std::min(steps, 20));
        }
        int32_t cap4 = 0, cap8 = 0;
        rc = b->part_capacity(4, &cap4);
        if (!rc) rc = b->part_capacity(8, &cap8);
        if (!rc) {
            int64_t min4 = 0;
            for (const InstDev& I : b->h_inst) min4 += (I.n + kPartThreads * 4 - 1) / (kPartThreads * 4);
            const int32_t npt = min4 <= cap4 ? 4 : 8;
            const int32_t cap = npt == 4 ? cap4 : cap8;
            const std::vector<int32_t> S = b->plan_parts(npt, cap);
            int64_t total = 0;
            for (int32_t s : S) total += s;
            if (total <= cap && (rc = b->set_parts(S.data())) == NBX_OK) {
                float ms = 0.0f;
                rc = timed([&](int32_t t0, int32_t k) { return b->launch_parts(t0, k); }, &ms);
                if (!rc && ms < best) {
                    best = ms;
                    split = true;
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->rec_n = saved_rec;
    const int32_t rs = snap.restore(b);
    if (rc) return rc;
    if (rs) return rs;
    if (best_thr == 0 && !split) {
        set_error("nbx_batch_tune: no workgroup size fits the largest instance");
        return NBX_ERR_UNSUPPORTED;
    }
    if (!split) {
        b->h_nparts.clear();
        b->threads = best_thr;
    }
    b->drop_graph();
    if (chosen) *chosen = split ? kPartThreads : best_thr;
    return NBX_OK;
}

// Runs nsteps: one workgroup per instance as a captured graph of `chunk`-step launches (rebuilt
// for every call, as the step numbers are baked into the launches), or the split schedule as
// direct cooperative launches of `chunk` steps.
int32_t nbx_batch_run_async(nbx_batch* b, int32_t nsteps) {
    if (!b || nsteps < 0) {
        set_error("nbx_batch_run: bad arguments");
        return NBX_ERR_INVALID;
    }
    if (nsteps == 0) return NBX_OK;
    NBX_HIP_TRY(hipSetDevice(b->device));
    if (!b->h_nparts.empty()) {
        for (int32_t done = 0; done < nsteps; done += b->chunk)
            NBX_HIP_TRY(b->launch_parts(b->t + done, std::min(b->chunk, nsteps - done)));
        b->t += nsteps;
        return NBX_OK;
    }
    if (b->threads == 0) {  // untuned: the smallest workgroup that gives each thread at most 8 neurons
        int32_t thr = 1024;
        while (thr > 256 && b->max_n <= thr / 2 * 8) thr /= 2;
        b->threads = thr;
    }
    b->drop_graph();
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
    return check_batch_error(b);
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
