// Internal declarations shared by the HIP translation units of libnbx.
//
// Numerics contract (what makes device results bit-identical to the CPU
// restatement of the reference): every file is compiled with -ffp-contract=off
// and correctly rounded fp32 division, no transcendental runs on the device
// (decay factors and STDP exponentials are host-computed tables), and synaptic
// arrivals accumulate as int32 fixed point so their sum is independent of the
// order in which atomics land.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "nbx_core.h"

namespace nbx {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);

// Launches a kernel whose workgroups wait for one another (the persistent and batched engines).
// A cooperative launch has the runtime guarantee that all of them are resident at once. Under
// rocprofv3 (which sets ROCP_TOOL_LIBRARIES) the launch is a regular one: the profiler loses its
// results when it exits after tracing a cooperative dispatch, also for a 40-line program with one
// empty cooperative kernel and no part of this library, while the same program with regular
// launches profiles cleanly. The caller has checked that occupancy x CUs covers the grid, and a
// profiled process runs alone on the GPU, so the workgroups are co-resident there too.
hipError_t launch_coresident(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t stream);
#define NBX_HIP_TRY(expr)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            ::nbx::set_error(std::string(#expr " failed: ") + hipGetErrorString(e_));            \
            return NBX_ERR_HIP;                                                                  \
        }                                                                                        \
    } while (0)

// return a non-zero nbx status to the caller
#define NBX_PASS(expr)                  \
    do {                                \
        const int32_t rc_ = (expr);     \
        if (rc_ != 0) return rc_;       \
    } while (0)

// ----------------------------------------------------------------- packed synapse entry
// bits 0..23 post neuron, 24..29 delay (1..63), 30..31 type (0 exc, 1 inh)
constexpr uint32_t kPostMask = 0x00FFFFFFu;
constexpr int kMaxNeurons = 1 << 24;
constexpr int kMaxDelay = 63;
__host__ __device__ inline uint32_t pack_syn(int32_t post, int32_t delay, int32_t type) {
    return (uint32_t)post | ((uint32_t)delay << 24) | ((uint32_t)type << 30);
}
__host__ __device__ inline int32_t syn_post(uint32_t e) { return (int32_t)(e & kPostMask); }
__host__ __device__ inline int32_t syn_delay(uint32_t e) { return (int32_t)((e >> 24) & 63u); }
__host__ __device__ inline int32_t syn_type(uint32_t e) { return (int32_t)(e >> 30); }

// ----------------------------------------------------------------- model parameters (by value)
struct ModelParams {
    int32_t model, integrator, n, t_ref;
    int32_t begin;  // first neuron of the population (0 for a single-population network)
    float dt, hdt, dt6;
    float v_rest, tau, v_th, r, v_reset, e_rev;
    float izh_v_th, izh_r, e_ampa, e_gaba;
    float g_decay, ampa_decay, gaba_decay, nmda_decay;
    float rate_bias, rate_gain;
    int32_t cond;  // kCond* bits: which conductance inputs the network has (the reference's CheckType flags)
};

// Conductance inputs present in a network (set at finalize from synapse types, dense signs and the
// Poisson background); the double/triple-exponential LIF models gate terms and decay on them.
constexpr int32_t kCondExc = 1, kCondInh = 2, kCondNmda = 4;

// ----------------------------------------------------------------- arrival channels
// Ring channel of an arrival through a synapse of `type` (0 exc, 1 inh, 2 NMDA) and whether its
// fixed-point value enters negated, per model (the reference models' ProcessSpike):
//   LIF_VJ, LIF_EXP, POISSON: one signed channel (V += w / g += w / rate += w, minus for inhibitory)
//   IZH, LIF_EXP2, TRIGGER_RELAY: channel 0 excitatory, channel 1 inhibitory, both positive
//   LIF_EXP3: channel = type (AMPA, GABA, NMDA), positive
//   IZH_CUR: one channel, weights already signed
//   IZH: channel 2 carries NMDA arrivals (type 2) when the network has them
// Networks of several populations deliver by type instead (kModelTyped below): channel = type,
// magnitudes as stored, and each population's update folds the channels into its model's inputs.
constexpr int32_t kModelTyped = 15;
__host__ __device__ inline int32_t model_channels(int32_t model) {
    return (model == NBX_MODEL_LIF_EXP3 || model == kModelTyped) ? 3
         : (model == NBX_MODEL_IZH || model == NBX_MODEL_LIF_EXP2 || model == NBX_MODEL_TRIGGER_RELAY) ? 2 : 1;
}
__host__ __device__ inline int32_t arrival_channel(int32_t model, int32_t type) {
    if (model == NBX_MODEL_LIF_EXP3 || model == kModelTyped) return type;
    if (model == NBX_MODEL_IZH) return type;
    if (model_channels(model) == 2) return type == 1 ? 1 : 0;
    return 0;
}
__host__ __device__ inline bool arrival_negated(int32_t model, int32_t type) {
    return type == 1 && (model == NBX_MODEL_LIF_VJ || model == NBX_MODEL_LIF_EXP || model == NBX_MODEL_POISSON);
}
// Largest synapse type a model accepts (NMDA for the triple-exponential LIF and Izhikevich).
__host__ __device__ inline int32_t model_max_type(int32_t model) {
    return (model == NBX_MODEL_LIF_EXP3 || model == NBX_MODEL_IZH) ? 2 : 1;
}

// Which SoA state arrays a model keeps besides v, I_ext and LastSpike.
__host__ __device__ constexpr bool model_uses_s1(int32_t m) {
    return m == NBX_MODEL_LIF_EXP || m == NBX_MODEL_IZH || m == NBX_MODEL_IZH_CUR || m == NBX_MODEL_LIF_EXP3;
}
__host__ __device__ constexpr bool model_uses_g2(int32_t m) {
    return m == NBX_MODEL_IZH || m == NBX_MODEL_LIF_EXP2 || m == NBX_MODEL_LIF_EXP3;
}

// One trigger-driven learning rule on the device; its trace decay tables sit at
// rtab[tab_off .. tab_off + tab_len) (Apre, tau ltp) and the next tab_len entries (Apost, tau ltd).
struct RuleDev {
    int32_t kind, clear;
    float max_ltp, max_ltd, reward, punishment, a1pre, a2prepre;
    int32_t tab_off, tab_len;
    // cerebellar: kernel table at rtab[kern_off .. kern_off + kern_len), its window (ms) and 1/step
    float min_weight, maxtime, inv_step;
    int32_t kern_off, kern_len;
    uint32_t seed;
};
// Device error codes (nbx_net::d_err; the batched engine raises kErrExchange, replay STDP kErrRaster)
constexpr int32_t kErrBarrier = 1, kErrHistory = 2, kErrExchange = 4, kErrRaster = 8, kErrInput = 16;
constexpr int32_t kErrResidency = 32;  // a persistent launch's workgroups were not all resident: no step taken

// One external input of a step: an arrival of synapse type kind (0..2) or (kInputCurrent) a new
// I_ext; value is the applied value, or the scale of d_values[src] when src >= 0.
constexpr int32_t kInputCurrent = 3;
struct InputRec {
    int32_t target, kind, src, overwrite;
    float value;
};
constexpr int32_t kMaxRules = 64;
constexpr int32_t kMaxPops = 64;  // populations per network

// Device pointers and scalars the step kernels need; passed by value.
// Per-synapse learning state in one 32-byte record, so an update touches this line and the
// weight's: the pair traces (STDP_State: Apre, Apost, time of the last update), the STDP weight
// bound, the R-STDP eligibility and the learning rules' weight bound (the connection's maximum).
struct alignas(32) SynTrace {
    float apre, apost;
    int32_t slast;
    float wmax;  // STDP bound
    float elig;  // R-STDP eligibility
    float rmax;  // learning-rule bound
    int32_t pad0, pad1;
};
static_assert(sizeof(SynTrace) == 32, "one 32-byte record per synapse");

// All that pair STDP reads and writes of a synapse: a 16-byte record of its own when the network
// has no learning rules, else the first 16 bytes of its SynTrace (heads at a stride of 2 records).
struct alignas(16) TraceHead {
    float apre, apost;
    int32_t slast;
    float wmax;
};
static_assert(sizeof(TraceHead) == 16, "one 16-byte head per synapse");
__device__ inline TraceHead load_head(const TraceHead* h, int32_t stride, int64_t j) { return h[j * stride]; }
__device__ inline void store_head(TraceHead* h, int32_t stride, int64_t j, const TraceHead& v) { h[j * stride] = v; }

// STDP_State::SetUpdate (STDP_State.cpp:17-26) with the exponentials looked up by the integer
// step delta (tab[k] = expf(-k dt / tau), k < len; the last entry covers every longer delta);
// the record's last-update step becomes t. A negative delta (a record already updated past t)
// decays nothing.
__device__ inline void stdp_decay_tab(const float* __restrict__ ltp_tab, const float* __restrict__ ltd_tab,
                                      int32_t len, TraceHead& r, int32_t t) {
    int32_t dt = t - r.slast;
    if (dt > len - 1) dt = len - 1;
    if (dt < 0) dt = 0;
    r.apre = r.apre * ltp_tab[dt];
    r.apost = r.apost * ltd_tab[dt];
    r.slast = t;
}

// STDP.cpp:29-34: below 0 -> 0, above the synapse's maximum weight -> that maximum (selects,
// no branches)
__device__ inline float stdp_clip(float w, float wm) { return w < 0.0f ? 0.0f : (w > wm ? wm : w); }

struct StepArgs {
    ModelParams P;
    int32_t n, nch, D, words;
    // state, SoA
    float* v;
    float* s1;    // LIF_EXP: g ; IZH: u
    float* ga;    // IZH: g_ampa
    float* gg;    // IZH: g_gaba
    float* iext;  // I_ext
    float* g3;    // IZH: g_nmda (networks with Izhikevich NMDA synapses), else null
    int32_t* last;
    const float *a, *b, *c, *d;
    // delivery
    int32_t* ring;      // [D][nch][n]
    int32_t* hist;      // [D][n]
    int32_t* hist_cnt;  // [D]
    uint64_t* fmask;    // [n] firing history, bit k: fired k steps before the last update (STDP), or null
    const int64_t* row_ptr;
    const uint32_t* syn;
    float* w;
    const int64_t* dptr;  // [n][D+1] start of each delay inside a row
    const uint8_t* plastic;
    const float* wt;  // dense [n][n]
    int32_t dense_delay;
    float scale, inv_scale;
    unsigned long long* events;
    // Poisson background
    int32_t pois;
    uint32_t seed_lo, seed_hi;
    int32_t kmax;
    const uint32_t* pois_tab;
    int32_t q_bg;
    // external current
    const float* cur;
    int32_t cur_t0, cur_n;
    // STDP
    int32_t stdp;
    float max_ltp, max_ltd, w_max;
    SynTrace* trace;  // [S] learning records of networks with rules (pair traces, eligibility, bounds), else null
    TraceHead* heads; // pair-STDP records: heads[j * head_stride]
    int32_t head_stride;
    const float* ltp_tab;
    const float* ltd_tab;
    int32_t tab_len;
    const int64_t* in_ptr;
    const int64_t* in_syn;
    const int32_t* in_pd;  // [P] presynaptic neuron | delay << 24 of each in_syn entry
    // trigger-driven rules
    int32_t nrules, rstdp;        // rule count; any R-STDP rule (post-side updates)
    const RuleDev* rules;
    const int8_t* syn_rule;       // [S] rule of a rule-plastic synapse, else -1
    const int8_t* syn_trig;       // [S] rule of a trigger synapse, else -1
    const float* rtab;
    const int64_t* rin_ptr;       // [n * nrules + 1] rule-plastic inputs of (post, rule)
    const int64_t* rin_syn;
    const int64_t* tin_ptr;       // [n + 1] trigger synapses into each post, ascending
    const int64_t* tin_syn;
    int32_t* tstamp;              // [S] step of a trigger synapse's last arrival
    int32_t* tpost_stamp;         // [n] step a post last had a trigger arrival
    int32_t* tposts;              // [2][n] posts with trigger arrivals, by step parity
    int32_t* tcount;              // [2]
    const int32_t* conn_index;    // [S] connection index of each synapse (cerebellar random stream)
    const int32_t* cidx;          // [S] row of a cerebellar synapse in chist / chead, else -1
    int32_t* chist;               // [rows][ccap] arrival steps, circular
    int32_t* chead;               // [rows] arrivals so far
    int32_t ccap;
    // recording
    int32_t* counts;
    uint32_t* raster;
    int32_t rec_t0, rec_n;
    // error flag (device-side failures such as a barrier time-out)
    int32_t* err;
    // populations of a network with several (P.model == kModelTyped): per-neuron population and
    // the populations' parameters; null for one population
    const uint8_t* pop_of;
    const ModelParams* pops;
};

// ----------------------------------------------------------------- device helpers
#define NBX_POIS_STREAM 0x504F4953u

__device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                     uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // each 32x32 -> 64-bit product is one v_mad_u64_u32 (not a mul_lo / mul_hi pair)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__device__ inline int32_t poisson_count(uint32_t seed_lo, uint32_t seed_hi, int32_t i, int32_t t,
                                        const uint32_t* __restrict__ tab, int32_t kmax) {
    uint32_t o[4];
    philox4x32_10((uint32_t)i, (uint32_t)t, NBX_POIS_STREAM, 0u, seed_lo, seed_hi, o);
    int32_t k = 0;
    while (k < kmax && o[0] >= tab[k]) ++k;
    return k;
}

__device__ inline int32_t quantize(float w, float scale) { return __float2int_rn(w * scale); }

// ---- LIF right-hand side: R*I/tau [+ g*(E-V)/tau] + (V_rest - V)/tau
template <bool WITH_G>
__device__ inline float lif_rhs(const ModelParams& P, float V, float I, float g) {
    float cur = P.r * I / P.tau;
    if (WITH_G) cur += g * (P.e_rev - V) / P.tau;
    return cur + (P.v_rest - V) / P.tau;
}

// ---- NMDA magnesium block 1 / (1 + exp(-0.062 V) * (1.2 / 3.57)). The exponential is evaluated in
// double and rounded to float (the oracle does the same); the reference calls float std::exp.
__device__ inline float nmda_block(float V) {
    const float e = (float)exp((double)(-0.062f * V));
    return 1.0f / (1.0f + e * (1.2f / 3.57f));
}

// TimeDrivenIzhikevic_Exponential_Decay.cpp:62-75: I_tot = g_ampa (E_ampa - v) + g_gaba (E_gaba - v)
// [+ g_nmda B(v) (E_ampa - v) when the network has NMDA synapses] + I, left to right
template <bool COND>
__device__ inline void izh_rhs(const ModelParams& P, float a, float b, float V, float U, float ga, float gg, float gn,
                               float I, float& dv, float& du) {
    float tot = 0.0f;
    if (COND) {
        tot += ga * (P.e_ampa - V);
        tot += gg * (P.e_gaba - V);
        if (P.cond & kCondNmda) tot += gn * nmda_block(V) * (P.e_ampa - V);
    }
    tot += I;
    dv = 0.04f * V * V + 5.0f * V + 140.0f - U + P.izh_r * tot;
    du = a * (b * V - U);
}

// One LIF step after arrivals are applied; returns true on a spike.
template <bool WITH_G>
__device__ inline bool lif_step(const ModelParams& P, float& V, float& g, float I, int32_t& last) {
    if (last > P.t_ref) {
        if (P.integrator == NBX_EULER) {
            V += lif_rhs<WITH_G>(P, V, I, g) * P.dt;
        } else {
            const float k1 = lif_rhs<WITH_G>(P, V, I, g);
            const float k2 = lif_rhs<WITH_G>(P, V + k1 * P.hdt, I, g);
            const float k3 = lif_rhs<WITH_G>(P, V + k2 * P.hdt, I, g);
            const float k4 = lif_rhs<WITH_G>(P, V + k3 * P.dt, I, g);
            float acc = k1 + 2.0f * k2;
            acc = acc + 2.0f * k3;
            acc = acc + k4;
            V += acc * P.dt6;
        }
    }
    if (WITH_G) g *= P.g_decay;
    last += 1;
    if (V > P.v_th) {
        V = P.v_reset;
        last = 0;
        return true;
    }
    return false;
}

template <bool COND>
__device__ inline bool izh_step(const ModelParams& P, float a, float b, float c, float d, float& V, float& U,
                                float& ga, float& gg, float& gn, float I, int32_t& last) {
    if (last > 0) {
        if (P.integrator == NBX_EULER) {
            float dv, du;
            izh_rhs<COND>(P, a, b, V, U, ga, gg, gn, I, dv, du);
            V += dv * P.dt;
            U += du * P.dt;
        } else {
            float v1, u1, v2, u2, v3, u3, v4, u4;
            izh_rhs<COND>(P, a, b, V, U, ga, gg, gn, I, v1, u1);
            izh_rhs<COND>(P, a, b, V + v1 * P.hdt, U + u1 * P.hdt, ga, gg, gn, I, v2, u2);
            izh_rhs<COND>(P, a, b, V + v2 * P.hdt, U + u2 * P.hdt, ga, gg, gn, I, v3, u3);
            izh_rhs<COND>(P, a, b, V + v3 * P.dt, U + u3 * P.dt, ga, gg, gn, I, v4, u4);
            float av = v1 + 2.0f * v2;
            av = av + 2.0f * v3;
            av = av + v4;
            float au = u1 + 2.0f * u2;
            au = au + 2.0f * u3;
            au = au + u4;
            V += av * P.dt6;
            U += au * P.dt6;
        }
    }
    if (COND) {
        ga *= P.ampa_decay;
        gg *= P.gaba_decay;
        if (P.cond & kCondNmda) gn *= P.nmda_decay;  // :96-97
    }
    last += 1;
    if (V >= P.izh_v_th) {
        V = c;
        U += d;
        last = 0;
        return true;
    }
    return false;
}

// ---- LIF with excitatory and inhibitory conductances (TimeDrivenLIF_Exponential_double.cpp:61-74):
// R*I/tau + gexc*(Eexc-V)/tau + ginh*(Einh-V)/tau + (V_rest-V)/tau, left to right
__device__ inline float lif2_rhs(const ModelParams& P, float V, float I, float ge, float gi) {
    float cur = P.r * I / P.tau;
    cur += ge * (P.e_ampa - V) / P.tau;
    cur += gi * (P.e_gaba - V) / P.tau;
    return cur + (P.v_rest - V) / P.tau;
}

// ---- triple-exponential LIF (TimeDrivenLIF_Exponential_triple.cpp:79-105):
// (R * (gampa(Eampa-V) + ggaba(Egaba-V) + gnmda*B(V)*(Eampa-V) + I) + V_rest - V) / tau
__device__ inline float lif3_rhs(const ModelParams& P, float V, float I, float ga, float gg, float gn) {
    float cur = 0.0f;
    if (P.cond & kCondExc) cur += ga * (P.e_ampa - V);
    if (P.cond & kCondInh) cur += gg * (P.e_gaba - V);
    if ((P.cond & kCondNmda) && gn != 0.0f) cur += gn * nmda_block(V) * (P.e_ampa - V);  // gn == 0 adds a zero
    cur += I;
    return (P.r * cur + P.v_rest - V) / P.tau;
}

__device__ inline float decay_or_zero(float g, float k) { return g < 1.0e-9f ? 0.0f : g * k; }

template <int NG>
__device__ inline bool lif_cond_step(const ModelParams& P, float& V, float& ga, float& gg, float& gn, float I,
                                     int32_t& last) {
    auto rhs = [&](float v) { return NG == 2 ? lif2_rhs(P, v, I, ga, gg) : lif3_rhs(P, v, I, ga, gg, gn); };
    if (last > P.t_ref) {
        if (P.integrator == NBX_EULER) {
            V += rhs(V) * P.dt;
        } else {
            const float k1 = rhs(V);
            const float k2 = rhs(V + k1 * P.hdt);
            const float k3 = rhs(V + k2 * P.hdt);
            const float k4 = rhs(V + k3 * P.dt);
            float acc = k1 + 2.0f * k2;
            acc = acc + 2.0f * k3;
            acc = acc + k4;
            V += acc * P.dt6;
        }
    }
    if (NG == 2) {  // TimeDrivenLIF_Exponential_double.cpp:76-86: both decay once either input exists
        if (P.cond & (kCondExc | kCondInh)) {
            ga *= P.ampa_decay;
            gg *= P.gaba_decay;
        }
    } else {  // TimeDrivenLIF_Exponential_triple.cpp:107-128
        ga = decay_or_zero(ga, P.ampa_decay);
        gg = decay_or_zero(gg, P.gaba_decay);
        gn = decay_or_zero(gn, P.nmda_decay);
    }
    last += 1;
    if (V > P.v_th) {
        V = P.v_reset;
        last = 0;
        return true;
    }
    return false;
}

// ---- PoissonRate hash (PoissonRate.cpp:11-22)
__host__ __device__ inline float hash_uniform01(int32_t neuron, int32_t t) {
    uint32_t x = (uint32_t)neuron * 747796405u + (uint32_t)(t + 1) * 2891336453u + 0x9e3779b9u;
    x ^= x >> 16;
    x *= 2246822519u;
    x ^= x >> 13;
    x *= 3266489917u;
    x ^= x >> 16;
    return (float)(x & 0x00FFFFFFu) / (float)0x01000000u;
}

// Applies one neuron's arrivals (a0/a1/a2 fixed point) and external inputs for step t,
// then advances it one step with the parameters P of its population. Shared by every engine so
// their arithmetic is identical. GN is the Izhikevich NMDA conductance (unused otherwise).
// POIS_GIVEN: the step's Poisson background count was drawn ahead by the caller (pc).
template <int MODEL, bool LOAD_CUR = true, bool POIS_GIVEN = false>
__device__ inline bool neuron_tick_p(const StepArgs& A, const ModelParams& P, int32_t i, int32_t t, int32_t a0,
                                     int32_t a1, int32_t a2, float& V, float& S1, float& GA, float& GG, float& GN,
                                     float& I, int32_t& last, int32_t pc = 0) {
    if (LOAD_CUR && A.cur && t >= A.cur_t0 && t < A.cur_t0 + A.cur_n) I = A.cur[(int64_t)(t - A.cur_t0) * A.n + i];
    if (A.pois) a0 += (POIS_GIVEN ? pc : poisson_count(A.seed_lo, A.seed_hi, i, t, A.pois_tab, A.kmax)) * A.q_bg;
    const float f0 = (float)a0 * A.inv_scale;
    if (MODEL == NBX_MODEL_LIF_EXP2) {
        GA += f0;
        GG += (float)a1 * A.inv_scale;
        float unused = 0.0f;
        return lif_cond_step<2>(P, V, GA, GG, unused, I, last);
    } else if (MODEL == NBX_MODEL_LIF_EXP3) {
        GA += f0;
        GG += (float)a1 * A.inv_scale;
        S1 += (float)a2 * A.inv_scale;
        return lif_cond_step<3>(P, V, GA, GG, S1, I, last);
    } else if (MODEL == NBX_MODEL_POISSON) {  // PoissonRate.cpp:55-77; V holds the transient rate input
        V += f0;
        float rate = P.rate_bias + V + P.rate_gain * I;
        rate = (0.0f < rate) ? rate : 0.0f;
        float p = rate * P.dt * 0.001f;
        p = (p < 1.0f) ? p : 1.0f;
        V = 0.0f;
        if (hash_uniform01(i - P.begin, t) < p) {  // the hash takes the neuron's index within its model
            last = 0;
            return true;
        }
        last += 1;
        return false;
    } else if (MODEL == NBX_MODEL_TRIGGER_RELAY) {  // V holds last_exc_input
        V = f0;
        if (f0 > P.v_th) {
            last = 0;
            return true;
        }
        last += 1;
        return false;
    } else if (MODEL == NBX_MODEL_LIF_VJ) {
        if (last > P.t_ref) V += f0;
        float g = 0.0f;
        return lif_step<false>(P, V, g, I, last);
    } else if (MODEL == NBX_MODEL_LIF_EXP) {
        S1 += f0;
        return lif_step<true>(P, V, S1, I, last);
    } else if (MODEL == NBX_MODEL_IZH) {
        GA += f0;
        GG += (float)a1 * A.inv_scale;
        GN += (float)a2 * A.inv_scale;
        return izh_step<true>(P, A.a[i], A.b[i], A.c[i], A.d[i], V, S1, GA, GG, GN, I, last);
    } else {
        const float Itot = I + f0;
        float gn = 0.0f;
        return izh_step<false>(P, A.a[i], A.b[i], A.c[i], A.d[i], V, S1, GA, GG, gn, Itot, last);
    }
}

// The single-population form (the network's own parameters, no NMDA conductance of Izhikevich).
template <int MODEL, bool LOAD_CUR = true, bool POIS_GIVEN = false>
__device__ inline bool neuron_tick(const StepArgs& A, int32_t i, int32_t t, int32_t a0, int32_t a1, int32_t a2,
                                   float& V, float& S1, float& GA, float& GG, float& I, int32_t& last,
                                   int32_t pc = 0) {
    float gn = 0.0f;
    return neuron_tick_p<MODEL, LOAD_CUR, POIS_GIVEN>(A, A.P, i, t, a0, a1, a2, V, S1, GA, GG, gn, I, last, pc);
}

// Typed arrivals (channel = synapse type, magnitudes as stored) folded into a model's inputs, the
// same integers a network of that model alone accumulates: one signed channel (inhibitory
// subtracted) for LIF_VJ / LIF_EXP / POISSON, the signed sum for current-based Izhikevich,
// channels as they are for the conductance models.
__device__ inline void fold_typed(int32_t model, int32_t& a0, int32_t& a1, int32_t& a2) {
    if (model == NBX_MODEL_LIF_VJ || model == NBX_MODEL_LIF_EXP || model == NBX_MODEL_POISSON) {
        a0 = a0 - a1;
        a1 = 0;
    } else if (model == NBX_MODEL_IZH_CUR) {
        a0 = a0 + a1;
        a1 = 0;
    }
}

// ----------------------------------------------------------------- launchers (host)
hipError_t launch_update(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_deliver_push(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_deliver_dense(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_pull(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_ltp(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_pull_ltp(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_trigger(const StepArgs& A, const int32_t* tbase, int32_t k, hipStream_t s);
hipError_t launch_advance(int32_t* tbase, int32_t by, hipStream_t s);
hipError_t launch_inject(const StepArgs& A, const int32_t* d_ids, int32_t count, int32_t t, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int32_t value, int64_t count, hipStream_t s);
hipError_t launch_inputs(const StepArgs& A, const InputRec* d_recs, int32_t count, const float* d_vals, int32_t t,
                         float qmax, hipStream_t s);
hipError_t launch_pending_float(const StepArgs& A, int32_t t, float* d_out, hipStream_t s);
hipError_t launch_fire_flags(const StepArgs& A, int32_t t, uint8_t* d_flags, hipStream_t s);

}  // namespace nbx
