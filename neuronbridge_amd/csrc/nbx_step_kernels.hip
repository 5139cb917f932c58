// Per-step kernels of the launch-per-phase and graph engines.
//
//   nbx_update        neuron update of one step over SoA state; reads and clears the
//                     step's delay-ring slot, adds Poisson background and external
//                     current, integrates (Euler/RK4), and compacts spikes with a
//                     wavefront ballot + LDS prefix into the firing-history ring
//                     (reference: UpdateDenseUnifiedKernel, DenseUnifiedNeuronRuntime.cu:45-149,
//                     and the legacy *_GPU_Interface_Update kernels)
//   nbx_deliver_push  CSR fan-out of the step's spikes into ring[(t+d) % D]
//                     (reference: TimeDrivenInternalSpike.cpp:37-69 + PropogatedSpikeGroup.cpp:33-71)
//   nbx_deliver_dense dense W·s column gather into the ring (all-to-all networks)
//   nbx_pull          arrival-time delivery from the firing-history ring with pre-side
//                     STDP (reference: Mark/Drain + ApplyDenseLearningPreKernel,
//                     GpuPropagationRuntime.cu:150-316)
//   nbx_ltp           post-side STDP over each fired neuron's plastic inputs
//                     (reference: STDP::ApplyPostSynaticSpike, STDP.cpp:37-61)
#include <algorithm>

#include "nbx_internal.h"

namespace nbx {

namespace {

constexpr int kUpdateThreads = 256;
constexpr int kWaves = kUpdateThreads / 64;

__device__ inline int32_t step_of(const int32_t* tbase, int32_t k) { return tbase ? (*tbase + k) : k; }

// Signed fixed-point contribution of one synapse and its ring channel.
__device__ inline void syn_effect(const StepArgs& A, uint32_t e, float w, int32_t& ch, int32_t& q) {
    q = quantize(w, A.scale);
    const int32_t type = syn_type(e);
    ch = arrival_channel(A.P.model, type);
    if (arrival_negated(A.P.model, type)) q = -q;
}

// Fan out one presynaptic row [b, e) into the delay ring of step t. Loads are issued
// UNROLL at a time ahead of the atomics so several row segments are in flight per thread.
template <int THREADS>
__device__ inline void push_row(const StepArgs& A, int32_t t, int64_t b, int64_t e) {
    constexpr int UNROLL = 4;
    for (int64_t j0 = b + threadIdx.x; j0 < e; j0 += (int64_t)THREADS * UNROLL) {
        uint32_t en[UNROLL];
        float w[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t j = j0 + (int64_t)u * THREADS;
            en[u] = j < e ? A.syn[j] : 0u;
            w[u] = j < e ? A.w[j] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            if (j0 + (int64_t)u * THREADS < e) {
                int32_t ch, q;
                syn_effect(A, en[u], w[u], ch, q);
                const int32_t dslot = (t + syn_delay(en[u])) % A.D;
                atomicAdd(&A.ring[((int64_t)dslot * A.nch + ch) * A.n + syn_post(en[u])], q);
            }
        }
    }
}

// Reads and clears neuron i's ring entries of step t, advances it with the parameters P and
// stores its state. TYPED: the ring holds typed arrivals (several populations), folded first.
template <int MODEL, bool TYPED>
__device__ inline bool update_neuron(const StepArgs& A, const ModelParams& P, int32_t i, int32_t t, int32_t slot) {
    int32_t* r = A.ring + (int64_t)slot * A.nch * A.n;
    int32_t a0 = r[i];
    r[i] = 0;
    int32_t a1 = 0, a2 = 0;
    if (A.nch > 1) {
        a1 = r[A.n + i];
        r[A.n + i] = 0;
    }
    if (A.nch > 2) {
        a2 = r[2 * A.n + i];
        r[2 * A.n + i] = 0;
    }
    if (TYPED) fold_typed(MODEL, a0, a1, a2);
    constexpr bool izh_nmda = MODEL == NBX_MODEL_IZH;
    float V = A.v[i];
    float S1 = model_uses_s1(MODEL) ? A.s1[i] : 0.0f;
    float GA = model_uses_g2(MODEL) ? A.ga[i] : 0.0f;
    float GG = model_uses_g2(MODEL) ? A.gg[i] : 0.0f;
    float GN = (izh_nmda && A.g3) ? A.g3[i] : 0.0f;
    float I = A.iext[i];
    int32_t last = A.last[i];
    const bool fired = neuron_tick_p<MODEL>(A, P, i, t, a0, a1, a2, V, S1, GA, GG, GN, I, last);
    A.v[i] = V;
    if (model_uses_s1(MODEL)) A.s1[i] = S1;
    if (model_uses_g2(MODEL)) {
        A.ga[i] = GA;
        A.gg[i] = GG;
    }
    if (izh_nmda && A.g3) A.g3[i] = GN;
    A.iext[i] = I;
    A.last[i] = last;
    return fired;
}

// MODEL = kModelTyped: a network of several populations; each neuron takes its population's
// model and parameters from the population table.
template <int MODEL>
__global__ __launch_bounds__(kUpdateThreads) void nbx_update(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    const int32_t i = blockIdx.x * kUpdateThreads + threadIdx.x;
    const int32_t slot = t % A.D;
    bool fired = false;
    // the neuron's out-degree for the event count, loaded with its state rather than after the
    // spike decision (one dependent load less)
    unsigned long long deg = 0;
    if (i < A.n) {
        if (A.row_ptr) deg = (unsigned long long)(A.row_ptr[i + 1] - A.row_ptr[i]);
        if (A.wt) deg += (unsigned long long)A.n;
        if constexpr (MODEL == kModelTyped) {
            const ModelParams P = A.pops[A.pop_of[i]];
            switch (P.model) {
                case NBX_MODEL_LIF_VJ: fired = update_neuron<NBX_MODEL_LIF_VJ, true>(A, P, i, t, slot); break;
                case NBX_MODEL_LIF_EXP: fired = update_neuron<NBX_MODEL_LIF_EXP, true>(A, P, i, t, slot); break;
                case NBX_MODEL_IZH: fired = update_neuron<NBX_MODEL_IZH, true>(A, P, i, t, slot); break;
                case NBX_MODEL_IZH_CUR: fired = update_neuron<NBX_MODEL_IZH_CUR, true>(A, P, i, t, slot); break;
                case NBX_MODEL_LIF_EXP2: fired = update_neuron<NBX_MODEL_LIF_EXP2, true>(A, P, i, t, slot); break;
                case NBX_MODEL_LIF_EXP3: fired = update_neuron<NBX_MODEL_LIF_EXP3, true>(A, P, i, t, slot); break;
                case NBX_MODEL_POISSON: fired = update_neuron<NBX_MODEL_POISSON, true>(A, P, i, t, slot); break;
                default: fired = update_neuron<NBX_MODEL_TRIGGER_RELAY, true>(A, P, i, t, slot); break;
            }
        } else {
            fired = update_neuron<MODEL, false>(A, A.P, i, t, slot);
        }
        if (A.fmask) A.fmask[i] = (A.fmask[i] << 1) | (fired ? 1ull : 0ull);
    }
    // ---- spike compaction: ballot per wavefront, prefix over the block's wavefronts
    const uint64_t mask = __ballot(fired);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    __shared__ int32_t s_wcnt[kWaves];
    __shared__ int32_t s_base;
    __shared__ unsigned long long s_ev;
    if (threadIdx.x == 0) s_ev = 0;
    if (lane == 0) s_wcnt[wave] = __popcll(mask);
    if (A.raster && lane == 0 && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n) {
        // 64 consecutive neurons per wavefront -> two aligned raster words
        const int32_t i0 = blockIdx.x * kUpdateThreads + wave * 64;
        uint32_t* row = A.raster + (int64_t)(t - A.rec_t0) * A.words;
        if (i0 < A.n) row[i0 >> 5] = (uint32_t)mask;
        if (i0 + 32 < A.n) row[(i0 >> 5) + 1] = (uint32_t)(mask >> 32);
    }
    __syncthreads();
    // synaptic events are counted at the spike: out-degree (+ n for dense delivery)
    if (fired && deg) atomicAdd(&s_ev, deg);
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
        for (int w = 0; w < kWaves; ++w) {
            const int32_t c = s_wcnt[w];
            s_wcnt[w] = tot;
            tot += c;
        }
        if (s_ev) atomicAdd(A.events, s_ev);
        s_base = tot ? atomicAdd(&A.hist_cnt[slot], tot) : 0;
        if (tot && A.counts && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n) atomicAdd(&A.counts[t - A.rec_t0], tot);
        // the slot of step t+1 held step t+1-D, which no later phase reads any more
        if (blockIdx.x == 0) A.hist_cnt[(t + 1) % A.D] = 0;
    }
    __syncthreads();
    if (fired) A.hist[(int64_t)slot * A.n + s_base + s_wcnt[wave] + __popcll(mask & ((1ull << lane) - 1ull))] = i;
}

constexpr int kDeliverThreads = 256;

__global__ __launch_bounds__(kDeliverThreads) void nbx_deliver_push(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    const int32_t slot = t % A.D;
    const int32_t cnt = A.hist_cnt[slot];
    const int32_t* ids = A.hist + (int64_t)slot * A.n;
    for (int32_t s = blockIdx.x; s < cnt; s += gridDim.x) {
        const int32_t pre = ids[s];
        push_row<kDeliverThreads>(A, t, A.row_ptr[pre], A.row_ptr[pre + 1]);
    }
}

// Dense delivery: thread per postsynaptic neuron, ascending-id loop over the
// step's spikes reading column wt[pre][*] coalesced across the block.
__global__ __launch_bounds__(kDeliverThreads) void nbx_deliver_dense(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    const int32_t slot = t % A.D;
    const int32_t cnt = A.hist_cnt[slot];
    const int32_t* ids = A.hist + (int64_t)slot * A.n;
    const int32_t i = blockIdx.x * kDeliverThreads + threadIdx.x;
    if (i >= A.n || cnt == 0) return;
    int32_t pos = 0, neg = 0;
    for (int32_t s = 0; s < cnt; ++s) {
        const float w = A.wt[(int64_t)ids[s] * A.n + i];
        // a negative dense weight is an inhibitory synapse of magnitude -w; models with one signed
        // channel take the signed value directly (rint is symmetric, so quantize(w) == -quantize(-w))
        if (A.nch > 1 && w < 0.0f) neg += quantize(-w, A.scale);
        else pos += quantize(w, A.scale);
    }
    const int32_t dslot = (t + A.dense_delay) % A.D;
    int32_t* r = A.ring + (int64_t)dslot * A.nch * A.n;
    atomicAdd(&r[i], pos);
    if (A.nch > 1 && neg) atomicAdd(&r[A.n + i], neg);
}

// STDP_State::SetUpdate on the network's trace tables.
__device__ inline void stdp_decay(const StepArgs& A, TraceHead& r, int32_t t) {
    stdp_decay_tab(A.ltp_tab, A.ltd_tab, A.tab_len, r, t);
}

// ---- trigger-driven rules
// STDP_State::SetUpdate for rule r (STDP_State.cpp:17-26): both traces decay over the integer
// step delta, looked up in the rule's tables.
__device__ inline void rule_decay(const StepArgs& A, const RuleDev& R, SynTrace& r, int32_t t) {
    int32_t dt = t - r.slast;
    if (dt > R.tab_len - 1) dt = R.tab_len - 1;
    r.apre = r.apre * A.rtab[R.tab_off + dt];
    r.apost = r.apost * A.rtab[R.tab_off + R.tab_len + dt];
    r.slast = t;
}

// Interconnections::UpdateWeight (Interconnections.h:120-128)
__device__ inline float update_weight(float w, float inc, float wmax) {
    w += inc;
    if (w > wmax) w = wmax;
    else if (w < 0.0f) w = 0.0f;
    return w;
}

// ApplyBoundedWeightIncrement (CerebullarLearningRule.cpp:27-38)
__device__ inline float bounded_increment(float w, float inc, float wmax, float wmin) {
    w += inc;
    if (w > wmax) w = wmax;
    else if (w < wmin) w = wmin;
    return w;
}

// StableUnitRandom (CerebullarLearningRule.cpp:11-25): 24-bit uniform keyed by
// (seed, rule, connection, step)
__device__ inline float stable_unit_random(uint32_t seed, int32_t rule_id, int32_t synapse_id, int32_t step) {
    uint32_t v = seed;
    v ^= (uint32_t)rule_id + 0x9e3779b9u;
    v *= 1664525u;
    v += 1013904223u;
    v ^= (uint32_t)synapse_id + 0x85ebca6bu;
    v *= 2246822519u;
    v ^= (uint32_t)step + 0xc2b2ae35u;
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return (float)(v & 0x00ffffffu) / (float)0x01000000u;
}

// Pre-side update of a rule-plastic synapse at its arrival (after its weight was delivered).
__device__ inline void rule_pre(const StepArgs& A, int64_t j, int32_t t, int32_t r) {
    const RuleDev R = A.rules[r];
    if (R.kind == NBX_RULE_CEREBELLAR) {  // CerebullarLearningRule.cpp:89-103
        const float u = stable_unit_random(R.seed, r, A.conn_index[j], t);
        A.w[j] = bounded_increment(A.w[j], R.a1pre * u, A.trace[j].rmax, R.min_weight);
        // keep the arrival step; dropping one still inside the kernel window is an error
        const int64_t c = A.cidx[j];
        const int32_t h = A.chead[c];
        int32_t* ring = A.chist + c * A.ccap;
        if (h >= A.ccap && (float)ring[h % A.ccap] * A.P.dt > (float)t * A.P.dt - R.maxtime)
            atomicOr(A.err, kErrHistory);
        ring[h % A.ccap] = t;
        A.chead[c] = h + 1;
        return;
    }
    SynTrace tr = A.trace[j];
    if (R.kind == NBX_RULE_RSTDP) {  // R_STDP.cpp:40-48
        rule_decay(A, R, tr, t);
        tr.apre += 1.0f;
        tr.elig += -R.max_ltd * tr.apost;
    } else {  // AdditiveKernalChange.cpp:65-70: weight first, then the trace
        A.w[j] = update_weight(A.w[j], R.a1pre, tr.rmax);
        rule_decay(A, R, tr, t);
        tr.apre += 1.0f;
    }
    A.trace[j] = tr;
}

// A trigger synapse arrived: stamp it, and list its target once per step.
__device__ inline void note_trigger(const StepArgs& A, int64_t j, int32_t post, int32_t t) {
    A.tstamp[j] = t;
    if (atomicExch(&A.tpost_stamp[post], t) != t) {
        const int32_t k = atomicAdd(&A.tcount[t & 1], 1);
        A.tposts[(int64_t)(t & 1) * A.n + k] = post;
    }
}

// Arrival-time delivery: work items are (delay d, spike of step t-d) pairs; each wave walks
// the slice of synapses with that delay in the spiking neuron's row (slices are short, so a
// wave per item keeps four items per workgroup in flight).
//
// MERGED (plain STDP, no rules): the post-side updates of step t - 1 run in the same launch
// (ltp_items below) instead of a kernel of their own after the update of step t - 1. A synapse
// whose target fired at t - 1 and whose presynaptic spike arrives at t gets both updates here,
// in the reference's order (post-side of t - 1, delivery of the new weight, pre-side of t); the
// post-side pass skips it. fmask tells both sides which case applies.
template <bool MERGED>
__device__ inline void pull_items(const StepArgs& A, int32_t t, const int32_t* s_pref, int32_t bi, int32_t nb) {
    const int32_t total = s_pref[A.D];
    int32_t* racc = A.ring + (int64_t)(t % A.D) * A.nch * A.n;
    constexpr int32_t kWaves = kDeliverThreads / 64;
    const int32_t lane = threadIdx.x & 63;
    for (int32_t p = bi * kWaves + (int32_t)(threadIdx.x >> 6); p < total; p += nb * kWaves) {
        int d = 1;
        while (d + 1 < A.D && s_pref[d + 1] <= p) ++d;
        const int32_t pre = A.hist[(int64_t)((t - d) % A.D) * A.n + (p - s_pref[d])];
        const int64_t b = A.dptr[(int64_t)pre * (A.D + 1) + d];
        const int64_t e = A.dptr[(int64_t)pre * (A.D + 1) + d + 1];
        for (int64_t j = b + lane; j < e; j += 64) {
            const uint32_t en = A.syn[j];
            float w = A.w[j];
            // every synapse has a trace record: load it with the flag rather than after it
            TraceHead tr = A.stdp ? load_head(A.heads, A.head_stride, j) : TraceHead{};
            const bool plastic = A.stdp && A.plastic[j];
            if (MERGED && plastic && (A.fmask[syn_post(en)] & 1ull)) {
                // STDP.cpp:37-59 for step t - 1 first: the target fired then
                stdp_decay(A, tr, t - 1);
                tr.apost += 1.0f;
                w = stdp_clip(w + A.max_ltp * tr.apre, tr.wmax);  // stored below with the pre-side result
            }
            int32_t ch, q;
            syn_effect(A, en, w, ch, q);
            atomicAdd(&racc[(int64_t)ch * A.n + syn_post(en)], q);
            if (plastic) {  // STDP.cpp:19-35: decay, Apre += 1, w -= A- * Apost, clip
                stdp_decay(A, tr, t);
                tr.apre += 1.0f;
                store_head(A.heads, A.head_stride, j, tr);
                A.w[j] = stdp_clip(w + -A.max_ltd * tr.apost, tr.wmax);
            }
            if (A.nrules) {
                const int32_t r = A.syn_rule[j];
                if (r >= 0) rule_pre(A, j, t, r);
                if (A.syn_trig[j] >= 0) note_trigger(A, j, syn_post(en), t);
            }
        }
    }
}

// s_pref[d] = spikes of the steps t - 1 .. t - d + 1 (d = 1 .. D - 1), s_pref[D] = all of them: lane d
// of wave 0 loads the count of step t - d (all loads in flight together), then a wave prefix sum
__device__ inline void pull_prefix(const StepArgs& A, int32_t t, int32_t* s_pref) {
    if (threadIdx.x < 64) {
        const int32_t d = (int32_t)threadIdx.x;
        const int32_t c = (d >= 1 && d < A.D && t - d >= 0) ? A.hist_cnt[(t - d) % A.D] : 0;
        int32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t z = __shfl_up(incl, o, 64);
            if (d >= o) incl += z;
        }
        if (d < A.D) s_pref[d] = incl - c;
        if (d == A.D - 1) s_pref[A.D] = incl;
    }
    __syncthreads();
}

__device__ inline void pull_reset(const StepArgs& A, int32_t t) {
    // the trigger list of step t + 1 (same parity as t - 1, consumed by then) starts empty
    if (A.nrules && blockIdx.x == 0 && threadIdx.x == 0) A.tcount[(t + 1) & 1] = 0;
    // hist_cnt[t % D] belongs to step t-D, no longer read; the update of step t reuses it
    if (blockIdx.x == 0 && threadIdx.x == 0) A.hist_cnt[t % A.D] = 0;
}

__global__ __launch_bounds__(kDeliverThreads) void nbx_pull(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    __shared__ int32_t s_pref[kMaxDelay + 2];
    pull_prefix(A, t, s_pref);
    pull_items<false>(A, t, s_pref, (int32_t)blockIdx.x, (int32_t)gridDim.x);
    pull_reset(A, t);
}

// Post-side updates of the neurons that fired at step t.
template <bool MERGED>
__device__ inline void ltp_items(const StepArgs& A, int32_t t, int32_t bi, int32_t nb) {
    const int32_t slot = t % A.D;
    const int32_t cnt = A.hist_cnt[slot];
    const int32_t* ids = A.hist + (int64_t)slot * A.n;
    for (int32_t s = bi; s < cnt; s += nb) {
        const int32_t p = ids[s];
        if (A.stdp) {  // STDP.cpp:37-59: decay, Apost += 1, w += A+ * Apre, clip
            // two synapses per thread, every load issued before the dependent arithmetic (the
            // synapses of one target are distinct, so the stores cannot alias the other's loads)
            const int64_t b = A.in_ptr[p], e = A.in_ptr[p + 1];
            for (int64_t j0 = b + threadIdx.x; j0 < e; j0 += 2 * kDeliverThreads) {
                const int64_t j1 = j0 + kDeliverThreads;
                bool two = j1 < e, one = true;
                const int64_t s0 = A.in_syn[j0];
                const int64_t s1 = two ? A.in_syn[j1] : s0;
                if (MERGED) {  // skip a synapse whose presynaptic spike arrives at t + 1: pull_items has it
                    const int32_t pd0 = A.in_pd[j0], pd1 = two ? A.in_pd[j1] : pd0;
                    one = !((A.fmask[pd0 & 0xFFFFFF] >> ((pd0 >> 24) - 1)) & 1ull);
                    two = two && !((A.fmask[pd1 & 0xFFFFFF] >> ((pd1 >> 24) - 1)) & 1ull);
                }
                TraceHead r0 = load_head(A.heads, A.head_stride, s0), r1 = load_head(A.heads, A.head_stride, s1);
                const float w0 = A.w[s0], w1 = A.w[s1];
                stdp_decay(A, r0, t);
                stdp_decay(A, r1, t);
                r0.apost += 1.0f;
                r1.apost += 1.0f;
                if (one) {
                    store_head(A.heads, A.head_stride, s0, r0);
                    A.w[s0] = stdp_clip(w0 + A.max_ltp * r0.apre, r0.wmax);
                }
                if (two) {
                    store_head(A.heads, A.head_stride, s1, r1);
                    A.w[s1] = stdp_clip(w1 + A.max_ltp * r1.apre, r1.wmax);
                }
            }
        }
        if (A.rstdp) {  // R_STDP.cpp:19-37: decay, Apost += 1, e += Apre * max_ltp
            for (int32_t r = 0; r < A.nrules; ++r) {
                const RuleDev R = A.rules[r];
                if (R.kind != NBX_RULE_RSTDP) continue;
                const int64_t b = A.rin_ptr[(int64_t)p * A.nrules + r], e = A.rin_ptr[(int64_t)p * A.nrules + r + 1];
                for (int64_t j = b + threadIdx.x; j < e; j += kDeliverThreads) {
                    const int64_t sid = A.rin_syn[j];
                    SynTrace tr = A.trace[sid];
                    rule_decay(A, R, tr, t);
                    tr.apost += 1.0f;
                    tr.elig += tr.apre * R.max_ltp;
                    A.trace[sid] = tr;
                }
            }
        }
    }
}

__global__ __launch_bounds__(kDeliverThreads) void nbx_ltp(StepArgs A, const int32_t* tbase, int32_t k) {
    ltp_items<false>(A, step_of(tbase, k), (int32_t)blockIdx.x, (int32_t)gridDim.x);
}

// Plain STDP: the pull of step t and the post-side updates of step t - 1 in one launch, on
// disjoint workgroups (the first kLtpBlocks take the fired neurons of t - 1) so that the two
// chains of dependent loads overlap instead of adding up.
constexpr int32_t kLtpBlocks = 256;

__global__ __launch_bounds__(kDeliverThreads) void nbx_pull_ltp(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    __shared__ int32_t s_pref[kMaxDelay + 2];
    if ((int32_t)blockIdx.x < kLtpBlocks) {
        ltp_items<true>(A, t - 1, (int32_t)blockIdx.x, kLtpBlocks);
    } else {
        pull_prefix(A, t, s_pref);
        pull_items<true>(A, t, s_pref, (int32_t)blockIdx.x - kLtpBlocks, (int32_t)gridDim.x - kLtpBlocks);
    }
    pull_reset(A, t);
}

// Trigger arrivals of step t: one wave per target that had any; its trigger synapses are
// applied one after another in ascending synapse order, each to all plastic inputs of its rule
// (R_STDP.cpp:51-69, AdditiveKernalChange.cpp:71-82).
__global__ __launch_bounds__(kDeliverThreads) void nbx_trigger(StepArgs A, const int32_t* tbase, int32_t k) {
    const int32_t t = step_of(tbase, k);
    const int32_t cnt = A.tcount[t & 1];
    const int32_t* posts = A.tposts + (int64_t)(t & 1) * A.n;
    constexpr int32_t kWaves = kDeliverThreads / 64;
    const int32_t lane = threadIdx.x & 63;
    // one wave per target. Its trigger inputs are tested 64 at a time (stamp loads in parallel)
    // and the arrived ones applied in ascending synapse order (ballot bits, lowest first). A
    // synapse sits at a fixed position of its rule's input list and is always handled by the
    // same lane, so each lane's program order is the order between the target's triggers.
    for (int32_t i = blockIdx.x * kWaves + (int32_t)(threadIdx.x >> 6); i < cnt; i += gridDim.x * kWaves) {
        const int32_t p = posts[i];
        const int64_t qb = A.tin_ptr[p], qe = A.tin_ptr[p + 1];
        for (int64_t c0 = qb; c0 < qe; c0 += 64) {
            const int64_t q = c0 + lane;
            const int64_t tq = q < qe ? A.tin_syn[q] : -1;
            uint64_t m = __ballot(tq >= 0 && A.tstamp[tq] == t);
            while (m) {
                const int32_t src = __builtin_ctzll(m);
                m &= m - 1ull;
                const int64_t tj = __shfl(tq, src, 64);
                const int32_t r = A.syn_trig[tj];
                const RuleDev R = A.rules[r];
                const float factor = syn_type(A.syn[tj]) == 1 ? R.punishment : R.reward;
                const int64_t b = A.rin_ptr[(int64_t)p * A.nrules + r], e = A.rin_ptr[(int64_t)p * A.nrules + r + 1];
                if (R.kind == NBX_RULE_CEREBELLAR) {  // CerebullarLearningRule.cpp:105-129
                    const float realtime = (float)t * A.P.dt;
                    const float thr = realtime - R.maxtime;
                    for (int64_t j = b + lane; j < e; j += 64) {
                        const int64_t sid = A.rin_syn[j];
                        const int64_t c = A.cidx[sid];
                        const int32_t h = A.chead[c];
                        const int32_t* ring = A.chist + c * A.ccap;
                        float w = A.w[sid];
                        const float wmax = A.trace[sid].rmax;
                        for (int32_t q2 = h - 1; q2 >= 0 && q2 >= h - A.ccap; --q2) {  // newest first
                            const float et = (float)ring[q2 % A.ccap] * A.P.dt;
                            if (!(et > thr)) break;
                            int32_t ti = (int32_t)((realtime - et) * R.inv_step);
                            if (ti < 0) continue;
                            if (ti >= R.kern_len) ti = R.kern_len - 1;
                            w = bounded_increment(w, R.a2prepre * A.rtab[R.kern_off + ti], wmax, R.min_weight);
                        }
                        A.w[sid] = w;
                    }
                } else {
                    for (int64_t j = b + lane; j < e; j += 64) {
                        const int64_t sid = A.rin_syn[j];
                        SynTrace tr = A.trace[sid];
                        rule_decay(A, R, tr, t);
                        if (R.kind == NBX_RULE_RSTDP) {
                            A.w[sid] = update_weight(A.w[sid], factor * tr.elig, tr.rmax);
                            if (R.clear) tr.elig = 0.0f;
                        } else {
                            A.w[sid] = update_weight(A.w[sid], R.a2prepre * tr.apre, tr.rmax);
                        }
                        A.trace[sid] = tr;
                    }
                }
            }
        }
    }
}

__global__ void nbx_advance(int32_t* tbase, int32_t by) {
    if (threadIdx.x == 0) *tbase += by;
}

// Injected spikes of step t: appended to the step's firing list (the host entry point has
// checked that no id is repeated or already listed).
__global__ void nbx_inject(StepArgs A, const int32_t* ids, int32_t count, int32_t t) {
    const int32_t slot = t % A.D;
    __shared__ int32_t s_base;
    if (threadIdx.x == 0) s_base = atomicAdd(&A.hist_cnt[slot], count);
    __syncthreads();
    for (int32_t j = threadIdx.x; j < count; j += blockDim.x) {
        const int32_t id = ids[j];
        A.hist[(int64_t)slot * A.n + s_base + j] = id;
        if (A.fmask) atomicOr(&A.fmask[id], 1ull);  // after this step's update: bit 0 is step t
        if (A.raster && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n)
            atomicOr(&A.raster[(int64_t)(t - A.rec_t0) * A.words + (id >> 5)], 1u << (id & 31));
    }
    if (threadIdx.x == 0 && A.counts && t >= A.rec_t0 && t < A.rec_t0 + A.rec_n)
        atomicAdd(&A.counts[t - A.rec_t0], count);
    // events of injected spikes, counted at the spike like those of the update kernel
    unsigned long long ev = 0;
    for (int32_t j = threadIdx.x; j < count; j += blockDim.x) {
        const int32_t id = ids[j];
        if (A.row_ptr) ev += (unsigned long long)(A.row_ptr[id + 1] - A.row_ptr[id]);
        if (A.wt) ev += (unsigned long long)A.n;
    }
    if (ev) atomicAdd(A.events, ev);
}

// External inputs of step t (interface channels, arrivals from outside the synapses), resolved on
// the host: an arrival of synapse type r.kind (0..2) adds rint(value * 2^shift) to the target's ring
// entry exactly as a synapse's arrival does; kind kInputCurrent replaces I_ext. A record with
// src >= 0 takes its value from the device array at src times r.value (device-to-device copies;
// zero values skipped unless r.overwrite). |value * 2^shift| above qmax raises kErrInput.
__global__ void nbx_inputs(StepArgs A, const InputRec* __restrict__ recs, int32_t count,
                           const float* __restrict__ vals, int32_t t, float qmax) {
    int32_t* r0 = A.ring + (int64_t)(t % A.D) * A.nch * A.n;
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count; j += gridDim.x * blockDim.x) {
        const InputRec r = recs[j];
        float applied = r.value;
        if (r.src >= 0) {
            applied = vals[r.src] * r.value;
            if (applied == 0.0f && !r.overwrite) continue;
        }
        if (r.kind == kInputCurrent) {
            A.iext[r.target] = applied;
            continue;
        }
        const float qs = applied * A.scale;
        if (!(fabsf(qs) <= qmax)) {
            atomicOr(A.err, kErrInput);
            continue;
        }
        int32_t q = __float2int_rn(qs);
        if (arrival_negated(A.P.model, r.kind)) q = -q;
        atomicAdd(&r0[(int64_t)arrival_channel(A.P.model, r.kind) * A.n + r.target], q);
    }
}

// Pending arrivals of step t as floats, channel-major [ch][n] (a snapshot of the ring slot).
__global__ void nbx_pending_float(StepArgs A, int32_t t, float* out) {
    const int32_t* r0 = A.ring + (int64_t)(t % A.D) * A.nch * A.n;
    const int64_t total = (int64_t)A.nch * A.n;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x)
        out[j] = (float)r0[j] * A.inv_scale;
}

// Did-fire flags of step t from its firing list (the flags were cleared before).
__global__ void nbx_fire_flags(StepArgs A, int32_t t, uint8_t* flags) {
    const int32_t slot = t % A.D;
    const int32_t cnt = A.hist_cnt[slot];
    for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x)
        flags[A.hist[(int64_t)slot * A.n + j]] = 1;
}

__global__ void nbx_fill_u8(uint8_t* p, int64_t count) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = 0;
}

__global__ void nbx_fill_i32(int32_t* p, int32_t value, int64_t count) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x)
        p[j] = value;
}

int deliver_grid() { return 1024; }

}  // namespace

hipError_t launch_update(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    const int grid = (A.n + kUpdateThreads - 1) / kUpdateThreads;
    switch (A.P.model) {
        case kModelTyped: nbx_update<kModelTyped><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
        case NBX_MODEL_LIF_VJ: nbx_update<NBX_MODEL_LIF_VJ><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
        case NBX_MODEL_LIF_EXP: nbx_update<NBX_MODEL_LIF_EXP><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
        case NBX_MODEL_IZH: nbx_update<NBX_MODEL_IZH><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
        case NBX_MODEL_LIF_EXP2:
            nbx_update<NBX_MODEL_LIF_EXP2><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k);
            break;
        case NBX_MODEL_LIF_EXP3:
            nbx_update<NBX_MODEL_LIF_EXP3><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k);
            break;
        case NBX_MODEL_POISSON: nbx_update<NBX_MODEL_POISSON><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
        case NBX_MODEL_TRIGGER_RELAY:
            nbx_update<NBX_MODEL_TRIGGER_RELAY><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k);
            break;
        default: nbx_update<NBX_MODEL_IZH_CUR><<<grid, kUpdateThreads, 0, s>>>(A, tbase, k); break;
    }
    return hipGetLastError();
}

hipError_t launch_deliver_push(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_deliver_push<<<deliver_grid(), kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_deliver_dense(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_deliver_dense<<<(A.n + kDeliverThreads - 1) / kDeliverThreads, kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_pull(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_pull<<<deliver_grid(), kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_ltp(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_ltp<<<deliver_grid(), kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_pull_ltp(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_pull_ltp<<<kLtpBlocks + deliver_grid(), kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_trigger(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s) {
    nbx_trigger<<<deliver_grid(), kDeliverThreads, 0, s>>>(A, tbase, k);
    return hipGetLastError();
}

hipError_t launch_advance(int32_t* tbase, int32_t by, hipStream_t s) {
    nbx_advance<<<1, 64, 0, s>>>(tbase, by);
    return hipGetLastError();
}

hipError_t launch_inject(const StepArgs& A, const int32_t* d_ids, int32_t count, int32_t t, hipStream_t s) {
    nbx_inject<<<1, 256, 0, s>>>(A, d_ids, count, t);
    return hipGetLastError();
}

hipError_t launch_inputs(const StepArgs& A, const InputRec* d_recs, int32_t count, const float* d_vals, int32_t t,
                         float qmax, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    nbx_inputs<<<std::min((count + 255) / 256, 1024), 256, 0, s>>>(A, d_recs, count, d_vals, t, qmax);
    return hipGetLastError();
}

hipError_t launch_pending_float(const StepArgs& A, int32_t t, float* d_out, hipStream_t s) {
    const int64_t total = (int64_t)A.nch * A.n;
    nbx_pending_float<<<(int)std::min<int64_t>((total + 255) / 256, 4096), 256, 0, s>>>(A, t, d_out);
    return hipGetLastError();
}

hipError_t launch_fire_flags(const StepArgs& A, int32_t t, uint8_t* d_flags, hipStream_t s) {
    nbx_fill_u8<<<(int)std::min<int64_t>(((int64_t)A.n + 255) / 256, 4096), 256, 0, s>>>(d_flags, A.n);
    nbx_fire_flags<<<std::max(1, std::min((A.n + 255) / 256, 1024)), 256, 0, s>>>(A, t, d_flags);
    return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    nbx_fill_i32<<<(int)blocks, 256, 0, s>>>(p, value, count);
    return hipGetLastError();
}

}  // namespace nbx
