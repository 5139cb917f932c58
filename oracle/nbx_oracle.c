/*
 * CPU oracle for the NeuronBridge per-timestep hot path (see nbx_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY — never linked into or called by the product library.
 * Built with -ffp-contract=off so every float operation rounds exactly as written,
 * matching the reference's x86-64 build (no FMA contraction, SSE float arithmetic).
 *
 * Parity pins, per function (what checks it, so a dropped term or wrong index fails):
 *   orc_philox4x32_10      Random123 known-answer vectors (tests/test_oracle.py)
 *   orc_hash_uniform01     reference poisson_hash restatement in numerical_accuracy_tests.py:118-128
 *                          and its expected sequence rule (:139); brute force vs Python
 *   orc_poisson_*          chi-square against the Poisson pmf; thresholds vs closed form
 *   orc_update_f32 (Euler) reference KATs: numerical_accuracy_report.md (Izhikevich first run
 *                          V=-65.03999328613281, U=-13.01200008392334; voltage jump -5 -> -4.5 -> -3.55)
 *                          and the LIF closed form of numerical_accuracy_tests.py:54-64
 *   orc_update_f32 (RK4)   closed form for the linear LIF ODE (growth factor 1-h+h^2/2-h^3/6+h^4/24),
 *                          4th-order convergence, FP64 instantiation
 *   orc_run delivery       brute-force push-ring image checked against the pull delivery every tick
 *                          (returns -2 on mismatch); delay KAT of tests/native/dense_cuda_smoke.cpp:109-118
 *   STDP                   closed-form pair updates; replay scenario of
 *                          tests/model_codegen/pair_stdp_gpu_trajectory.cpp:52-66
 *   orc_replay_stdp        orc_run's own final weights and traces when fed orc_run's raster
 *                          (single and chained calls); per synapse, orc_stdp_replay on that
 *                          synapse's arrival / spike schedule
 *   conductance terms      closed forms with the conductances held (tests/test_oracle_pins.py):
 *                          LIF_EXP / LIF_EXP2 / LIF_EXP3 contract to V_inf = (sum g E + R I +
 *                          V_rest) / (1 + sum g) by 1 - h (Euler) or the RK4 polynomial in h;
 *                          Izhikevich AMPA / GABA / NMDA currents vanish at their reversal
 *                          potential, change sign across it and add linearly (scaled by R);
 *                          decays equal exp(-n dt / tau)
 *   orc_stable_unit_random bit-exact against the reference's own StableUnitRandom compiled from
 *                          its source (oracle/_ref, tests/golden/stable_unit_random.npz)
 */
#define _POSIX_C_SOURCE 200809L /* clock_gettime */
#include "nbx_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

int orc_nstate(int model) {
    switch (model) {
        case ORC_LIF_VJ: return 2;        /* V, I_ext            (TimeDrivenLIF_Voltage_jump.h:34) */
        case ORC_LIF_EXP: return 3;       /* V, g, I_ext         (TimeDrivenLIF_Exponential_Decay.h) */
        case ORC_LIF_EXP2: return 4;      /* V, gexc, ginh, I_ext (TimeDrivenLIF_Exponential_double.h:27-33) */
        case ORC_LIF_EXP3: return 5;      /* V, ampa, gaba, nmda, I_ext (..._triple.h:19-25) */
        case ORC_POISSON: return 2;       /* rate_hz, I_ext      (PoissonRate.h:14-16) */
        case ORC_TRIGGER_RELAY: return 1; /* last excitatory input (DenseTriggerRelayModel.h) */
        default: return 6;                /* v, u, g_ampa, g_gaba, g_nmda, I_ext (Izhikevic .h) */
    }
}

/* column of I_ext in the AoS state, -1 when the model has none */
static int orc_iext_index(int model) {
    switch (model) {
        case ORC_LIF_VJ: return 1;
        case ORC_LIF_EXP: return 2;
        case ORC_LIF_EXP2: return 3;
        case ORC_LIF_EXP3: return 4;
        case ORC_POISSON: return 1;
        case ORC_TRIGGER_RELAY: return -1;
        default: return 5;
    }
}

/* exp(-dt * (1/tau)) the way ResetConductanceDecayLookup computes it
 * (TimeDrivenLIF_Exponential_Decay.cpp:37-39, TimeDrivenIzhikevic_Exponential_Decay.cpp:32-39). */
static float ref_decay(float dt, float tau) {
    const float inv = (tau != 0.0f) ? (1.0f / tau) : 0.0f;
    return expf(-dt * inv);
}

void orc_model_prepare(orc_model* m) {
    m->g_decay = ref_decay(m->dt, m->g_tau);
    m->ampa_decay = ref_decay(m->dt, m->ampa_tau);
    m->gaba_decay = ref_decay(m->dt, m->gaba_tau);
    /* Izhikevich NMDA: the same single-entry lookup (TimeDrivenIzhikevic_Exponential_Decay.cpp:32-39) */
    m->nmda_decay = m->model == ORC_IZH ? ref_decay(m->dt, m->nmda_tau) : 1.0f;
    if (m->model == ORC_LIF_EXP3) { /* TimeDrivenLIF_Exponential_triple.cpp:56-60: exp(-dt / tau) */
        m->ampa_decay = expf(-m->dt / m->ampa_tau);
        m->gaba_decay = expf(-m->dt / m->gaba_tau);
        m->nmda_decay = expf(-m->dt / m->nmda_tau);
    }
}

/* Ring channel of an arrival and whether it enters negated (each model's ProcessSpike):
 *   LIF_VJ V += w / -= w; LIF_EXP g += w / -= w; PoissonRate rate += w / -= w (PoissonRate.cpp:80-88)
 *   IZH g_ampa / g_gaba; LIF_EXP2 gexc / ginh (..._double.cpp:105-113); TRIGGER_RELAY exc / (unused) inh
 *   LIF_EXP3 ampa / gaba / nmda by type (..._triple.cpp:142-155); IZH_CUR signed current */
static int orc_channel(int model, int type) {
    if (model == ORC_LIF_EXP3 || model == ORC_IZH) return type; /* Izhikevich NMDA: channel 2 (:119-131) */
    if (model == ORC_IZH || model == ORC_LIF_EXP2 || model == ORC_TRIGGER_RELAY) return type == 1 ? 1 : 0;
    return 0;
}
static int orc_negated(int model, int type) {
    return type == 1 && (model == ORC_LIF_VJ || model == ORC_LIF_EXP || model == ORC_POISSON);
}

#define REAL float
#define FN(name) name##_f32
#include "orc_model.inc"
#undef REAL
#undef FN
#define REAL double
#define FN(name) name##_f64
#include "orc_model.inc"
#undef REAL
#undef FN

int32_t orc_update_f32(const orc_model* m, float* x, int32_t* last, const float* isyn, int32_t* spikes) {
    return update_f32(m, x, last, isyn, spikes);
}
int32_t orc_update_f64(const orc_model* m, double* x, int32_t* last, const double* isyn, int32_t* spikes) {
    return update_f64(m, x, last, isyn, spikes);
}

/* ---------------------------------------------------------------- random numbers */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
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

/* PoissonRate.cpp:7-21 */
float orc_hash_uniform01(int32_t neuron, int32_t t) {
    uint32_t x = (uint32_t)neuron * 747796405u + (uint32_t)(t + 1) * 2891336453u + 0x9e3779b9u;
    x ^= x >> 16;
    x *= 2246822519u;
    x ^= x >> 13;
    x *= 3266489917u;
    x ^= x >> 16;
    return (float)(x & 0x00FFFFFFu) / (float)0x01000000u;
}

int32_t orc_poisson_thresholds(double lambda, uint32_t* tab, int32_t cap) {
    double p = exp(-lambda), cdf = 0.0;
    int32_t k = 0;
    for (; k < cap; ++k) {
        cdf += p;
        double v = floor(cdf * 4294967296.0);
        uint32_t t = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
        tab[k] = t;
        if (t == 0xFFFFFFFFu) return k;
        p = p * lambda / (double)(k + 1);
    }
    return cap - 1;
}

#define ORC_POIS_STREAM 0x504F4953u /* "POIS" */

int32_t orc_poisson_count(uint32_t seed_lo, uint32_t seed_hi, int32_t neuron, int32_t t, const uint32_t* tab,
                          int32_t kmax) {
    const uint32_t ctr[4] = {(uint32_t)neuron, (uint32_t)t, ORC_POIS_STREAM, 0u};
    const uint32_t key[2] = {seed_lo, seed_hi};
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    int32_t k = 0;
    while (k < kmax && out[0] >= tab[k]) ++k;
    return k;
}

int32_t orc_quantize(float w, int32_t shift) { return (int32_t)lrintf(w * ldexpf(1.0f, shift)); }

void orc_init_state(const orc_model* m, float* x, int32_t* last) {
    const int ns = orc_nstate(m->model);
    for (int i = 0; i < m->n; ++i) {
        float* s = x + (size_t)i * ns;
        memset(s, 0, sizeof(float) * (size_t)ns);
        if (m->model == ORC_LIF_VJ) {
            s[0] = m->v_reset; /* TimeDrivenLIF_Voltage_jump.cpp:37 */
        } else if (m->model == ORC_LIF_EXP) {
            s[0] = m->v_rest; /* TimeDrivenLIF_Exponential_Decay.cpp:51 */
        } else if (m->model == ORC_LIF_EXP2) {
            s[0] = m->v_rest; /* TimeDrivenLIF_Exponential_double.cpp:51 */
        } else if (m->model == ORC_LIF_EXP3) {
            s[0] = m->v_reset; /* TimeDrivenLIF_Exponential_triple.cpp:71 */
        } else if (m->model == ORC_POISSON || m->model == ORC_TRIGGER_RELAY) {
            /* all zero (PoissonRate.cpp:45) */
        } else {
            s[0] = m->v_rest;            /* V_rest + init[0], Izhikevic .cpp:49 */
            s[1] = m->b[i] * m->v_rest;  /* b * V_rest + init[1] */
        }
        last[i] = 10000; /* Neuron_State_Vector.cpp:43-67 */
    }
}

/* ---------------------------------------------------------------- network run */

typedef struct {
    orc_net* net;
    int32_t acc_mode;
    float* x;
    int32_t* last;
    int32_t* acc0;
    int32_t* acc1;
    int32_t* acc2;
    float* isyn;
    float scale;
    int ns;
    int typed;              /* several populations: fixed-point arrivals kept per synapse type */
    const uint8_t* pop_of;  /* population of each neuron (typed) */
} run_ctx;

/* the model (and parameters) of neuron i */
static const orc_model* model_of(const run_ctx* c, int32_t i) {
    return c->typed ? &c->net->pops[c->pop_of[i]] : &c->net->m;
}

/* One synaptic arrival. ProcessSpike semantics:
 *   voltage jump: V += w (exc) / V -= w (inh) only outside refractory (Voltage_jump.cpp:100-113)
 *   exp decay:    g += w / g -= w, ungated (Exponential_Decay.cpp:108-119)
 *   Izhikevich:   g_ampa += w (exc), g_gaba += w (inh) (Izhikevic .cpp:119-131)
 *   current-based Izhikevich: signed w into this tick's synaptic current. */
static void arrive(run_ctx* c, int32_t post, int32_t type, float w) {
    const orc_model* m = model_of(c, post);
    float* s = c->x + (size_t)post * c->ns;
    if (c->acc_mode == ORC_ACC_FIXED) {
        int32_t q = (int32_t)lrintf(w * c->scale);
        int32_t* acc[3] = {c->acc0, c->acc1, c->acc2};
        if (c->typed) { /* by synapse type; apply_fixed folds them per model */
            acc[type][post] += q;
            return;
        }
        if (orc_negated(m->model, type)) q = -q;
        acc[orc_channel(m->model, type)][post] += q;
        return;
    }
    switch (m->model) {
        case ORC_LIF_VJ:
            if (c->last[post] > m->t_ref) s[0] += (type == 1) ? -w : w;
            break;
        case ORC_LIF_EXP: s[1] += (type == 1) ? -w : w; break;
        case ORC_IZH: s[2 + type] += w; break; /* g_ampa / g_gaba / g_nmda by type (Izhikevic .cpp:119-131) */
        case ORC_LIF_EXP2: s[1 + (type == 1)] += w; break;  /* ..._double.cpp:105-113 */
        case ORC_LIF_EXP3: s[1 + type] += w; break;         /* ..._triple.cpp:142-155 */
        case ORC_POISSON: s[0] += (type == 1) ? -w : w; break;
        case ORC_TRIGGER_RELAY:
            if (type != 1) c->isyn[post] += w; /* this step's excitatory input */
            break;
        default: c->isyn[post] += w; break;
    }
}

/* Fixed-point arrivals are applied once per tick, before the update. */
static void apply_fixed(run_ctx* c, int32_t n) {
    const float inv = 1.0f / c->scale;
    for (int i = 0; i < n; ++i) {
        const orc_model* m = model_of(c, i);
        float* s = c->x + (size_t)i * c->ns;
        int32_t q0 = c->acc0[i];
        if (c->typed) { /* the integers a network of this model alone accumulates */
            if (m->model == ORC_LIF_VJ || m->model == ORC_LIF_EXP || m->model == ORC_POISSON) q0 -= c->acc1[i];
            else if (m->model == ORC_IZH_CUR) q0 += c->acc1[i];
        }
        const float a0 = (float)q0 * inv;
        switch (m->model) {
            case ORC_LIF_VJ:
                if (c->last[i] > m->t_ref) s[0] += a0;
                break;
            case ORC_LIF_EXP: s[1] += a0; break;
            case ORC_IZH:
                s[2] += a0;
                s[3] += (float)c->acc1[i] * inv;
                s[4] += (float)c->acc2[i] * inv;
                break;
            case ORC_LIF_EXP2:
                s[1] += a0;
                s[2] += (float)c->acc1[i] * inv;
                break;
            case ORC_LIF_EXP3:
                s[1] += a0;
                s[2] += (float)c->acc1[i] * inv;
                s[3] += (float)c->acc2[i] * inv;
                break;
            case ORC_POISSON: s[0] += a0; break;
            default: c->isyn[i] = a0; break;
        }
    }
}

/* STDP_State::SetUpdate (STDP_State.cpp:17-26) */
static void stdp_set_update(orc_net* net, int64_t s, int32_t t, float inv_ltp, float inv_ltd) {
    const float dT = (float)(t - net->slast[s]) * net->m.dt;
    net->apre[s] *= expf(-dT * inv_ltp);
    net->apost[s] *= expf(-dT * inv_ltd);
    net->slast[s] = t;
}

static void stdp_clip(orc_net* net, int64_t s) {
    const float wm = net->stdp_maxw ? net->stdp_maxw[s] : net->w_max;
    if (net->w[s] < 0.0f) net->w[s] = 0.0f;
    else if (net->w[s] > wm) net->w[s] = wm;
}

/* ---------------------------------------------------------------- trigger-driven rules */

float orc_stable_unit_random(uint32_t seed, int32_t rule_id, int32_t synapse_id, int32_t spike_time) {
    uint32_t value = seed;
    value ^= (uint32_t)rule_id + 0x9e3779b9u;
    value *= 1664525u;
    value += 1013904223u;
    value ^= (uint32_t)synapse_id + 0x85ebca6bu;
    value *= 2246822519u;
    value ^= (uint32_t)spike_time + 0xc2b2ae35u;
    value ^= value >> 16;
    value *= 0x7feb352du;
    value ^= value >> 15;
    value *= 0x846ca68bu;
    value ^= value >> 16;
    return (float)(value & 0x00ffffffu) / (float)0x01000000u;
}

int32_t orc_cerebellar_kernel(const orc_rule* r, float* tab, int32_t cap, float* maxtime, float* inv_step) {
    float maxpos = r->maxpos;
    if (maxpos <= r->initpos) maxpos = r->initpos + 1e-6f;
    const float width = maxpos - r->initpos;
    float mt = maxpos;
    for (;;) {
        const float value = (1.0f / width) * mt * expf(-(mt / width) + 1.0f);
        if (value < 1e-2f) break;
        mt += r->kernel_step;
    }
    mt += r->initpos;
    const float inv = 1.0f / r->kernel_step;
    const int32_t n = (int32_t)(mt * inv) + 1;
    if (maxtime) *maxtime = mt;
    if (inv_step) *inv_step = inv;
    for (int32_t i = 0; tab && i < n && i < cap; ++i) {
        const float time = (float)i * r->kernel_step;
        if (time < r->initpos) {
            tab[i] = 0.0f;
        } else {
            const float shifted = time - r->initpos;
            tab[i] = (1.0f / width) * shifted * expf(-(shifted / width) + 1.0f);
        }
    }
    return n;
}

/* Interconnections::UpdateWeight (Interconnections.h:120-128) */
static float update_weight(float w, float inc, float wmax) {
    w += inc;
    if (w > wmax) w = wmax;
    else if (w < 0.0f) w = 0.0f;
    return w;
}

/* STDP_State::SetUpdate with the rule's time constants (the additive rule's state uses ltp_tau
 * for both traces, AdditiveKernalChange.cpp:16-20) */
static void rule_set_update(orc_net* net, const orc_rule* R, int64_t s, int32_t t) {
    const float inv_a = (float)(1.0 / (double)R->ltp_tau);
    const float inv_b = (float)(1.0 / (double)(R->kind == ORC_RULE_ADDITIVE ? R->ltp_tau : R->ltd_tau));
    const float dT = (float)(t - net->slast[s]) * net->m.dt;
    net->apre[s] *= expf(-dT * inv_a);
    net->apost[s] *= expf(-dT * inv_b);
    net->slast[s] = t;
}

typedef struct {
    float time;      /* arrival time in ms */
    int64_t synapse;
} cb_entry;

/* per (target, rule) buffer of plastic-input arrival times, oldest first (BufferedActivityTime) */
typedef struct {
    cb_entry* e;
    int64_t first, count, cap;
} cb_buffer;

typedef struct {
    float* tab;
    int32_t len;
    float maxtime, inv_step;
} cb_kernel;

typedef struct {
    int64_t *rin_ptr, *rin_syn, *tin_ptr, *tin_syn; /* ascending synapse order in every list */
    int32_t *tstamp, *post_mark, *posts;
    int32_t nposts;
    cb_buffer* cb;   /* [n * nrules] */
    cb_kernel* ck;   /* [nrules] */
    size_t ncb;
    int nck;
} rule_lists;

/* ApplyBoundedWeightIncrement (CerebullarLearningRule.cpp:27-38) */
static float bounded_increment(float w, float inc, float wmax, float wmin) {
    w += inc;
    if (w > wmax) w = wmax;
    else if (w < wmin) w = wmin;
    return w;
}

/* Pre-side update at the synapse's own arrival (after its weight was delivered). */
static void rule_pre(orc_net* net, rule_lists* L, int64_t s, int32_t t) {
    const int r = net->syn_rule[s];
    const orc_rule* R = &net->rules[r];
    if (R->kind == ORC_RULE_CEREBELLAR) { /* CerebullarLearningRule.cpp:89-103 */
        const float realtime = (float)t * net->m.dt;
        const int32_t cid = net->conn_index ? net->conn_index[s] : (int32_t)s;
        const float u = orc_stable_unit_random(R->seed, r, cid, t);
        net->w[s] = bounded_increment(net->w[s], R->a1pre * u, net->maxw[s], R->min_weight);
        cb_buffer* b = &L->cb[(size_t)net->post[s] * net->nrules + r];
        const float thr = realtime - L->ck[r].maxtime;
        while (b->count > 0 && !(b->e[b->first].time > thr)) { /* expired: never read again */
            b->first++;
            b->count--;
        }
        if (b->first + b->count == b->cap) {
            if (b->first > 0) {
                memmove(b->e, b->e + b->first, sizeof(cb_entry) * (size_t)b->count);
                b->first = 0;
            } else {
                b->cap = b->cap ? 2 * b->cap : 16;
                b->e = (cb_entry*)realloc(b->e, sizeof(cb_entry) * (size_t)b->cap);
            }
        }
        b->e[b->first + b->count].time = realtime;
        b->e[b->first + b->count].synapse = s;
        b->count++;
        return;
    }
    if (R->kind == ORC_RULE_RSTDP) { /* R_STDP.cpp:40-48 */
        rule_set_update(net, R, s, t);
        net->apre[s] += 1.0f;
        net->elig[s] += -R->max_ltd * net->apost[s];
    } else { /* AdditiveKernalChange.cpp:65-70 */
        net->w[s] = update_weight(net->w[s], R->a1pre, net->maxw[s]);
        rule_set_update(net, R, s, t);
        net->apre[s] += 1.0f;
    }
}

static void rule_lists_build(const orc_net* net, rule_lists* L) {
    const int n = net->m.n, nr = net->nrules;
    const int64_t S = net->row_ptr[n];
    L->rin_ptr = (int64_t*)calloc((size_t)n * nr + 1, sizeof(int64_t));
    L->tin_ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t s = 0; s < S; ++s) {
        if (net->syn_rule[s] >= 0) L->rin_ptr[(size_t)net->post[s] * nr + net->syn_rule[s] + 1]++;
        if (net->syn_trig[s] >= 0) L->tin_ptr[net->post[s] + 1]++;
    }
    for (size_t k = 1; k <= (size_t)n * nr; ++k) L->rin_ptr[k] += L->rin_ptr[k - 1];
    for (int k = 1; k <= n; ++k) L->tin_ptr[k] += L->tin_ptr[k - 1];
    L->rin_syn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L->rin_ptr[(size_t)n * nr] + 1));
    L->tin_syn = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L->tin_ptr[n] + 1));
    int64_t* rf = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n * nr + 1));
    int64_t* tf = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(rf, L->rin_ptr, sizeof(int64_t) * ((size_t)n * nr + 1));
    memcpy(tf, L->tin_ptr, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t s = 0; s < S; ++s) {
        if (net->syn_rule[s] >= 0) L->rin_syn[rf[(size_t)net->post[s] * nr + net->syn_rule[s]]++] = s;
        if (net->syn_trig[s] >= 0) L->tin_syn[tf[net->post[s]]++] = s;
    }
    free(rf);
    free(tf);
    L->tstamp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(S > 0 ? S : 1));
    for (int64_t s = 0; s < S; ++s) L->tstamp[s] = -1;
    L->post_mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int i = 0; i < n; ++i) L->post_mark[i] = -1;
    L->posts = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    L->nposts = 0;
    L->cb = (cb_buffer*)calloc((size_t)n * nr, sizeof(cb_buffer));
    L->ck = (cb_kernel*)calloc((size_t)nr, sizeof(cb_kernel));
    L->ncb = (size_t)n * nr;
    L->nck = nr;
    for (int r = 0; r < nr; ++r) {
        if (net->rules[r].kind != ORC_RULE_CEREBELLAR) continue;
        cb_kernel* k = &L->ck[r];
        k->len = orc_cerebellar_kernel(&net->rules[r], NULL, 0, &k->maxtime, &k->inv_step);
        k->tab = (float*)malloc(sizeof(float) * (size_t)k->len);
        orc_cerebellar_kernel(&net->rules[r], k->tab, k->len, NULL, NULL);
    }
}

static void rule_lists_free(rule_lists* L) {
    free(L->rin_ptr);
    free(L->rin_syn);
    free(L->tin_ptr);
    free(L->tin_syn);
    free(L->tstamp);
    free(L->post_mark);
    free(L->posts);
    for (size_t k = 0; k < L->ncb; ++k) free(L->cb[k].e);
    free(L->cb);
    for (int r = 0; r < L->nck; ++r) free(L->ck[r].tab);
    free(L->ck);
}

/* Trigger arrivals of tick t: each target applies its arrived trigger synapses in ascending
 * synapse order to every plastic input of the trigger's rule (R_STDP.cpp:51-69,
 * AdditiveKernalChange.cpp:71-82). Targets are independent (a synapse has one target). */
static void rule_triggers(orc_net* net, rule_lists* L, int32_t t) {
    const int nr = net->nrules;
    for (int k = 0; k < L->nposts; ++k) {
        const int32_t p = L->posts[k];
        for (int64_t q = L->tin_ptr[p]; q < L->tin_ptr[p + 1]; ++q) {
            const int64_t tj = L->tin_syn[q];
            if (L->tstamp[tj] != t) continue;
            const int r = net->syn_trig[tj];
            const orc_rule* R = &net->rules[r];
            if (R->kind == ORC_RULE_CEREBELLAR) { /* CerebullarLearningRule.cpp:105-129 */
                const float realtime = (float)t * net->m.dt;
                const cb_kernel* k = &L->ck[r];
                cb_buffer* b = &L->cb[(size_t)p * nr + r];
                const float thr = realtime - k->maxtime;
                int64_t keep = 0;  /* newest first; stop at the first arrival outside the window */
                for (int64_t q2 = b->count - 1; q2 >= 0; --q2) {
                    const cb_entry* e = &b->e[b->first + q2];
                    if (!(e->time > thr)) break;
                    keep++;
                    const float elapse = realtime - e->time;
                    int ti = (int)(elapse * k->inv_step);
                    if (ti < 0) continue;
                    if (ti >= k->len) ti = k->len - 1;
                    const int64_t sj = e->synapse;
                    net->w[sj] = bounded_increment(net->w[sj], R->a2prepre * k->tab[ti], net->maxw[sj], R->min_weight);
                }
                b->first += b->count - keep; /* older arrivals are dropped */
                b->count = keep;
                continue;
            }
            const float factor = net->type[tj] == 1 ? R->punishment : R->reward;
            for (int64_t j = L->rin_ptr[(size_t)p * nr + r]; j < L->rin_ptr[(size_t)p * nr + r + 1]; ++j) {
                const int64_t s = L->rin_syn[j];
                rule_set_update(net, R, s, t);
                if (R->kind == ORC_RULE_RSTDP) {
                    net->w[s] = update_weight(net->w[s], factor * net->elig[s], net->maxw[s]);
                    if (R->clear) net->elig[s] = 0.0f;
                } else {
                    net->w[s] = update_weight(net->w[s], R->a2prepre * net->apre[s], net->maxw[s]);
                }
            }
        }
    }
    L->nposts = 0;
}

/* Post-side R-STDP at the spike tick (R_STDP.cpp:19-37). */
static void rule_post(orc_net* net, rule_lists* L, int32_t p, int32_t t) {
    const int nr = net->nrules;
    for (int r = 0; r < nr; ++r) {
        const orc_rule* R = &net->rules[r];
        if (R->kind != ORC_RULE_RSTDP) continue;
        for (int64_t j = L->rin_ptr[(size_t)p * nr + r]; j < L->rin_ptr[(size_t)p * nr + r + 1]; ++j) {
            const int64_t s = L->rin_syn[j];
            rule_set_update(net, R, s, t);
            net->apost[s] += 1.0f;
            net->elig[s] += net->apre[s] * R->max_ltp;
        }
    }
}

int32_t orc_run(orc_net* net, int32_t acc_mode, float* x, int32_t* last, int32_t* hist, int32_t* hist_cnt,
                int32_t* ring, int32_t t0, int32_t nticks, int32_t* spike_counts, uint32_t* raster,
                int64_t* events) {
    orc_model* m = &net->m;
    const int typed = net->npops > 0;
    if (typed) { /* network-wide step and size for the learning code */
        m->dt = net->pops[0].dt;
        m->n = net->pop_begin[net->npops];
        for (int k = 0; k < net->npops; ++k) orc_model_prepare(&net->pops[k]);
    } else {
        orc_model_prepare(m);
    }
    const int n = m->n;
    const int D = net->d_slots;
    const int words = (n + 31) / 32;
    run_ctx c;
    c.net = net;
    c.acc_mode = acc_mode;
    c.x = x;
    c.last = last;
    c.typed = typed;
    uint8_t* pop_of = NULL;
    c.ns = 0;
    if (typed) {
        pop_of = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));
        for (int k = 0; k < net->npops; ++k) {
            for (int32_t i = net->pop_begin[k]; i < net->pop_begin[k + 1]; ++i) pop_of[i] = (uint8_t)k;
            if (orc_nstate(net->pops[k].model) > c.ns) c.ns = orc_nstate(net->pops[k].model);
        }
    } else {
        c.ns = orc_nstate(m->model);
    }
    c.pop_of = pop_of;
    c.scale = ldexpf(1.0f, net->fixed_shift);
    c.acc0 = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    c.acc1 = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    c.acc2 = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    c.isyn = (float*)calloc((size_t)n, sizeof(float));
    int32_t* spikes = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t tab[64];
    int32_t kmax = 0;
    if (net->pois_on) kmax = orc_poisson_thresholds(net->pois_lambda, tab, 64);
    const float inv_ltp = (float)(1.0 / (double)net->ltp_tau); /* STDP_State.cpp:6-7 */
    const float inv_ltd = (float)(1.0 / (double)net->ltd_tau);
    int64_t ev = 0;
    int32_t status = 0;
    /* conductance inputs present in each model (CheckType): types of the synapses into it, dense
     * weight signs, background */
    if (typed) {
        for (int k = 0; k < net->npops; ++k) net->pops[k].cond = net->pois_on ? ORC_COND_EXC : 0;
        if (net->row_ptr)
            for (int64_t s = 0; s < net->row_ptr[n]; ++s) net->pops[pop_of[net->post[s]]].cond |= 1 << net->type[s];
        for (int k = 0; k < net->npops; ++k) net->pops[k].cond &= ORC_COND_EXC | ORC_COND_INH | ORC_COND_NMDA;
    } else {
        m->cond = net->pois_on ? ORC_COND_EXC : 0;
        if (net->row_ptr)
            for (int64_t s = 0; s < net->row_ptr[n]; ++s) m->cond |= 1 << net->type[s];
        if (net->dense_wt)
            for (size_t k = 0; k < (size_t)n * n; ++k) m->cond |= net->dense_wt[k] < 0.0f ? ORC_COND_INH : ORC_COND_EXC;
        m->cond &= ORC_COND_EXC | ORC_COND_INH | ORC_COND_NMDA;
    }
    rule_lists rl;
    memset(&rl, 0, sizeof(rl));
    const int rules_on = net->nrules > 0 && net->row_ptr;
    if (rules_on) rule_lists_build(net, &rl);

    for (int32_t t = t0; t < t0 + nticks; ++t) {
        memset(c.acc0, 0, sizeof(int32_t) * (size_t)n);
        memset(c.acc1, 0, sizeof(int32_t) * (size_t)n);
        memset(c.acc2, 0, sizeof(int32_t) * (size_t)n);
        memset(c.isyn, 0, sizeof(float) * (size_t)n);
        /* inputs from outside the synapses: new external currents (stateful, ProcessCurrent); their
         * arrivals follow the synaptic ones below */
        if (net->inputs && t >= net->input_t0 && t < net->input_t0 + net->input_nsteps) {
            for (int32_t k = net->input_ptr[t - net->input_t0]; k < net->input_ptr[t - net->input_t0 + 1]; ++k) {
                const orc_input* in = &net->inputs[k];
                if (in->kind == ORC_INPUT_CURRENT) {
                    const int ix = orc_iext_index(model_of(&c, in->target)->model);
                    if (ix >= 0) x[(size_t)in->target * c.ns + ix] = in->value;
                }
            }
        }
        /* external current replaces I_ext for the tick (stateful current input) */
        if (net->cur) {
            const float* cr = net->cur + (size_t)(t - net->cur_t_begin) * n;
            for (int i = 0; i < n; ++i) {
                const int ix = orc_iext_index(model_of(&c, i)->model);
                if (ix >= 0) x[(size_t)i * c.ns + ix] = cr[i];
            }
        }
        /* sparse arrivals: spikes fired at t-d through synapses of delay d
         * (group time = t_spike + delay, TimeDrivenInternalSpike.cpp:17; delivery
         * then pre-learning per synapse, PropogatedSpikeGroup.cpp:42-58) */
        if (net->row_ptr) {
            for (int d = 1; d < D; ++d) {
                const int32_t ts = t - d;
                if (ts < 0) continue;
                const int slot = ts % D;
                const int32_t* hl = hist + (size_t)slot * n;
                for (int k = 0; k < hist_cnt[slot]; ++k) {
                    const int32_t pre = hl[k];
                    /* rows are sorted by delay: binary search the start of the delay-d run */
                    int64_t lo = net->row_ptr[pre], hi = net->row_ptr[pre + 1];
                    while (lo < hi) {
                        const int64_t mid = lo + (hi - lo) / 2;
                        if (net->delay[mid] < d) lo = mid + 1;
                        else hi = mid;
                    }
                    for (int64_t s = lo; s < net->row_ptr[pre + 1] && net->delay[s] == d; ++s) {
                        arrive(&c, net->post[s], net->type[s], net->w[s]);
                        if (net->stdp_on && net->plastic[s]) { /* STDP.cpp:19-35 */
                            stdp_set_update(net, s, t, inv_ltp, inv_ltd);
                            net->apre[s] += 1.0f;
                            net->w[s] += -net->max_ltd * net->apost[s];
                            stdp_clip(net, s);
                        }
                        if (rules_on) {
                            if (net->syn_rule[s] >= 0) rule_pre(net, &rl, s, t);
                            if (net->syn_trig[s] >= 0) {
                                rl.tstamp[s] = t;
                                if (rl.post_mark[net->post[s]] != t) {
                                    rl.post_mark[net->post[s]] = t;
                                    rl.posts[rl.nposts++] = net->post[s];
                                }
                            }
                        }
                    }
                }
            }
        }
        if (rules_on) rule_triggers(net, &rl, t);
        /* dense all-to-all arrivals (column of the fired presynaptic neuron) */
        if (net->dense_wt && t - net->dense_delay >= 0) {
            const int slot = (t - net->dense_delay) % D;
            const int32_t* hl = hist + (size_t)slot * n;
            for (int k = 0; k < hist_cnt[slot]; ++k) {
                const float* col = net->dense_wt + (size_t)hl[k] * n;
                for (int i = 0; i < n; ++i) {
                    float w = col[i];
                    int32_t type = 0;
                    if (m->model != ORC_IZH_CUR && w < 0.0f) { /* a negative weight is inhibitory, magnitude -w */
                        type = 1;
                        w = -w;
                    }
                    arrive(&c, i, type, w);
                }
            }
        }
        /* push-ring image: must hold exactly the pull result for this tick */
        if (ring && acc_mode == ORC_ACC_FIXED) {
            int32_t* r = ring + (size_t)(t % D) * 3 * n;
            for (int i = 0; i < n; ++i) {
                if (r[i] != c.acc0[i] || r[n + i] != c.acc1[i] || r[2 * n + i] != c.acc2[i]) status = -2;
                r[i] = 0;
                r[n + i] = 0;
                r[2 * n + i] = 0;
            }
        }
        /* arrivals from outside the synapses (not part of the push-ring image) */
        if (net->inputs && t >= net->input_t0 && t < net->input_t0 + net->input_nsteps) {
            for (int32_t k = net->input_ptr[t - net->input_t0]; k < net->input_ptr[t - net->input_t0 + 1]; ++k) {
                const orc_input* in = &net->inputs[k];
                if (in->kind != ORC_INPUT_CURRENT) arrive(&c, in->target, in->kind, in->value);
            }
        }
        /* Poisson background: count excitatory arrivals of weight w_bg */
        if (net->pois_on) {
            for (int i = 0; i < n; ++i) {
                const int32_t cnt = orc_poisson_count(net->seed_lo, net->seed_hi, i, t, tab, kmax);
                for (int k = 0; k < cnt; ++k) arrive(&c, i, 0, net->w_bg);
            }
        }
        if (acc_mode == ORC_ACC_FIXED) apply_fixed(&c, n);

        int32_t nspk = 0;
        if (typed) { /* each model over its own neurons, in population order (ascending ids) */
            for (int k = 0; k < net->npops; ++k) {
                orc_model* pm = &net->pops[k];
                const int32_t b = net->pop_begin[k];
                pm->step = t;
                const int cur_in = pm->model == ORC_IZH_CUR || pm->model == ORC_TRIGGER_RELAY;
                nspk += update_range_f32(pm, x + (size_t)b * c.ns, last + b, cur_in ? c.isyn + b : NULL, spikes + nspk,
                                         c.ns, b);
            }
        } else {
            m->step = t;
            nspk = update_f32(m, x, last, (m->model == ORC_IZH_CUR || m->model == ORC_TRIGGER_RELAY) ? c.isyn : NULL,
                              spikes);
        }

        /* post-synaptic learning at the spike tick (TimeDrivenInternalSpike.cpp:70-82, STDP.cpp:37-61) */
        if (net->stdp_on && net->in_ptr) {
            for (int k = 0; k < nspk; ++k) {
                const int32_t p = spikes[k];
                for (int64_t j = net->in_ptr[p]; j < net->in_ptr[p + 1]; ++j) {
                    const int64_t s = net->in_syn[j];
                    stdp_set_update(net, s, t, inv_ltp, inv_ltd);
                    net->apost[s] += 1.0f;
                    net->w[s] += net->max_ltp * net->apre[s];
                    stdp_clip(net, s);
                }
            }
        }
        if (rules_on)
            for (int k = 0; k < nspk; ++k) rule_post(net, &rl, spikes[k], t);
        /* push-ring image of this tick's spikes */
        if (ring && acc_mode == ORC_ACC_FIXED) {
            for (int k = 0; k < nspk; ++k) {
                const int32_t pre = spikes[k];
                if (net->row_ptr) {
                    for (int64_t s = net->row_ptr[pre]; s < net->row_ptr[pre + 1]; ++s) {
                        int32_t* r = ring + (size_t)((t + net->delay[s]) % D) * 3 * n;
                        int32_t q = (int32_t)lrintf(net->w[s] * c.scale);
                        if (typed) {
                            r[(size_t)net->type[s] * n + net->post[s]] += q;
                            continue;
                        }
                        if (orc_negated(m->model, net->type[s])) q = -q;
                        r[(size_t)orc_channel(m->model, net->type[s]) * n + net->post[s]] += q;
                    }
                }
                if (net->dense_wt) {
                    int32_t* r = ring + (size_t)((t + net->dense_delay) % D) * 3 * n;
                    const float* col = net->dense_wt + (size_t)pre * n;
                    for (int i = 0; i < n; ++i) {
                        const float w = col[i];
                        if (m->model != ORC_IZH_CUR && w < 0.0f) {
                            int32_t q = (int32_t)lrintf(-w * c.scale);
                            if (orc_negated(m->model, 1)) q = -q;
                            r[(size_t)orc_channel(m->model, 1) * n + i] += q;
                        } else {
                            r[i] += (int32_t)lrintf(w * c.scale);
                        }
                    }
                }
            }
        }
        /* synaptic events: one per outgoing synapse of every spike, counted at the spike */
        for (int k = 0; k < nspk; ++k) {
            if (net->row_ptr) ev += net->row_ptr[spikes[k] + 1] - net->row_ptr[spikes[k]];
            if (net->dense_wt) ev += n;
        }
        /* record history, counts and raster */
        const int slot = t % D;
        memcpy(hist + (size_t)slot * n, spikes, sizeof(int32_t) * (size_t)nspk);
        hist_cnt[slot] = nspk;
        if (spike_counts) spike_counts[t - t0] = nspk;
        if (raster) {
            uint32_t* row = raster + (size_t)(t - t0) * words;
            for (int k = 0; k < nspk; ++k) row[spikes[k] >> 5] |= 1u << (spikes[k] & 31);
        }
    }
    if (events) *events += ev;
    if (rules_on) rule_lists_free(&rl);
    free(pop_of);
    free(c.acc0);
    free(c.acc1);
    free(c.acc2);
    free(c.isyn);
    free(spikes);
    return status;
}

/* ---------------------------------------------------------------- queue-parallel run
 * The legacy engine's OpenMP structure (the CPU baseline): neurons are split into nq queues, one
 * thread each (Simulation.cu:635-639, NumberOfOpenMPQueues); each presynaptic neuron keeps its
 * synapses per target queue (NeuronPropogationStructure.h:16-24, built per queue by
 * Network.cu:483-560), so a queue's thread delivers every spike's arrivals into its own neurons
 * as per-event float adds (ProcessSpike, PropogatedSpikeGroup.cpp:33-71), draws their background,
 * updates them, and the queues meet once per tick. Per target the events arrive in the order of
 * orc_run's float-event mode, so the results are identical to it. Supports what the sparse
 * configurations use (sparse synapses, Poisson background, external current table; no learning,
 * dense weights, populations or external inputs): returns -3 otherwise. */
static double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int32_t orc_run_queues(orc_net* net, int32_t nq, float* x, int32_t* last, int32_t* hist, int32_t* hist_cnt,
                       int32_t t0, int32_t nticks, int32_t* spike_counts, int64_t* events, double* loop_seconds) {
    if (nq < 1 || net->stdp_on || net->nrules > 0 || net->dense_wt || net->npops > 0 || net->inputs || !net->row_ptr)
        return -3;
    orc_model* m = &net->m;
    orc_model_prepare(m);
    const int n = m->n, D = net->d_slots, ns = orc_nstate(m->model);
    if (nq > n) nq = n;
    m->cond = net->pois_on ? ORC_COND_EXC : 0;
    for (int64_t s = 0; s < net->row_ptr[n]; ++s) m->cond |= 1 << net->type[s];
    m->cond &= ORC_COND_EXC | ORC_COND_INH | ORC_COND_NMDA;
    int32_t* qbeg = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nq + 1));
    for (int q = 0; q <= nq; ++q) qbeg[q] = (int32_t)((int64_t)n * q / nq);
    /* per queue: CSR by presynaptic neuron of the synapses into the queue, in CSR order */
    int64_t** qptr = (int64_t**)malloc(sizeof(int64_t*) * (size_t)nq);
    int64_t** qsyn = (int64_t**)malloc(sizeof(int64_t*) * (size_t)nq);
    uint32_t tab[64];
    const int32_t kmax = net->pois_on ? orc_poisson_thresholds(net->pois_lambda, tab, 64) : 0;
    float* isyn = (float*)calloc((size_t)n, sizeof(float));
    int32_t* spk = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* qcnt = (int32_t*)calloc((size_t)nq, sizeof(int32_t));
    const int iidx = orc_iext_index(m->model);
    int64_t ev = 0;
#pragma omp parallel num_threads(nq)
    {
#ifdef _OPENMP
        const int q = omp_get_thread_num();
#else
        const int q = 0;
#endif
        int64_t cnt = 0;
        for (int64_t s = 0; s < net->row_ptr[n]; ++s)
            if (net->post[s] >= qbeg[q] && net->post[s] < qbeg[q + 1]) ++cnt;
        qptr[q] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
        qsyn[q] = (int64_t*)malloc(sizeof(int64_t) * (size_t)(cnt > 0 ? cnt : 1));
        int64_t k = 0;
        for (int p = 0; p < n; ++p) {
            qptr[q][p] = k;
            for (int64_t s = net->row_ptr[p]; s < net->row_ptr[p + 1]; ++s)
                if (net->post[s] >= qbeg[q] && net->post[s] < qbeg[q + 1]) qsyn[q][k++] = s;
        }
        qptr[q][n] = k;
        run_ctx c;
        memset(&c, 0, sizeof(c));
        c.net = net;
        c.acc_mode = ORC_ACC_FLOAT_EVENTS;
        c.x = x;
        c.last = last;
        c.isyn = isyn;
        c.ns = ns;
        c.scale = 1.0f;
        const int32_t b = qbeg[q], e = qbeg[q + 1];
        double t_begin = 0.0;
#pragma omp barrier
        if (q == 0) t_begin = orc_now(); /* the stepping loop only, after the per-queue lists */
        for (int32_t t = t0; t < t0 + nticks; ++t) {
            for (int32_t i = b; i < e; ++i) isyn[i] = 0.0f;
            if (net->cur && iidx >= 0) {
                const float* cr = net->cur + (size_t)(t - net->cur_t_begin) * n;
                for (int32_t i = b; i < e; ++i) x[(size_t)i * ns + iidx] = cr[i];
            }
            for (int d = 1; d < D; ++d) { /* spikes of t - d through this queue's delay-d synapses */
                const int32_t ts = t - d;
                if (ts < 0) continue;
                const int slot = ts % D;
                const int32_t* hl = hist + (size_t)slot * n;
                for (int kk = 0; kk < hist_cnt[slot]; ++kk) {
                    const int32_t pre = hl[kk];
                    int64_t lo = qptr[q][pre], hi = qptr[q][pre + 1];
                    while (lo < hi) {
                        const int64_t mid = lo + (hi - lo) / 2;
                        if (net->delay[qsyn[q][mid]] < d) lo = mid + 1;
                        else hi = mid;
                    }
                    for (int64_t j = lo; j < qptr[q][pre + 1] && net->delay[qsyn[q][j]] == d; ++j) {
                        const int64_t s = qsyn[q][j];
                        arrive(&c, net->post[s], net->type[s], net->w[s]);
                    }
                }
            }
            if (net->pois_on)
                for (int32_t i = b; i < e; ++i) {
                    const int32_t pc = orc_poisson_count(net->seed_lo, net->seed_hi, i, t, tab, kmax);
                    for (int kk = 0; kk < pc; ++kk) arrive(&c, i, 0, net->w_bg);
                }
            orc_model mq = *m;
            mq.n = e - b;
            mq.step = t;
            if (mq.a) {
                mq.a = m->a + b;
                mq.b = m->b + b;
                mq.c = m->c + b;
                mq.d = m->d + b;
            }
            const int cur_in = m->model == ORC_IZH_CUR || m->model == ORC_TRIGGER_RELAY;
            qcnt[q] = update_range_f32(&mq, x + (size_t)b * ns, last + b, cur_in ? isyn + b : NULL, spk + b, ns, b);
#pragma omp barrier
#pragma omp single
            { /* the tick's spike list in ascending order, its events and counts */
                const int slot = t % D;
                int32_t tot = 0;
                for (int qq = 0; qq < nq; ++qq) {
                    memmove(hist + (size_t)slot * n + tot, spk + qbeg[qq], sizeof(int32_t) * (size_t)qcnt[qq]);
                    tot += qcnt[qq];
                }
                hist_cnt[slot] = tot;
                for (int32_t kk = 0; kk < tot; ++kk) {
                    const int32_t p = hist[(size_t)slot * n + kk];
                    ev += net->row_ptr[p + 1] - net->row_ptr[p];
                }
                if (spike_counts) spike_counts[t - t0] = tot;
            }
        }
#pragma omp barrier
        if (q == 0 && loop_seconds) *loop_seconds = orc_now() - t_begin;
        free(qptr[q]);
        free(qsyn[q]);
    }
    if (events) *events += ev;
    free(qptr);
    free(qsyn);
    free(qbeg);
    free(isyn);
    free(spk);
    free(qcnt);
    return 0;
}

/* Replay STDP over a recorded raster: the STDP part of orc_run (STDP.cpp:19-61 at the delivery
 * of each arrival, then at each spike of the step) with the neuron update replaced by the raster.
 * ids[tick_ptr[k] .. tick_ptr[k+1]) fired at step t0 - lead + k; updates of steps [t0, t0+nsteps). */
int64_t orc_replay_stdp(orc_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t lead,
                        int32_t nsteps) {
    const int32_t D = net->d_slots, base = t0 - lead, K = lead + nsteps;
    const float inv_ltp = (float)(1.0 / (double)net->ltp_tau);
    const float inv_ltd = (float)(1.0 / (double)net->ltd_tau);
    int64_t updates = 0;
    for (int32_t t = t0; t < t0 + nsteps; ++t) {
        /* pre-side: spikes of t - d through their synapses of delay d */
        for (int d = 1; d < D; ++d) {
            const int32_t k = t - d - base;
            if (k < 0 || k >= K) continue;
            for (int32_t f = tick_ptr[k]; f < tick_ptr[k + 1]; ++f) {
                const int32_t pre = ids[f];
                for (int64_t s = net->row_ptr[pre]; s < net->row_ptr[pre + 1]; ++s) {
                    if (net->delay[s] != d || !net->plastic[s]) continue;
                    stdp_set_update(net, s, t, inv_ltp, inv_ltd);
                    net->apre[s] += 1.0f;
                    net->w[s] += -net->max_ltd * net->apost[s];
                    stdp_clip(net, s);
                    ++updates;
                }
            }
        }
        /* post-side: spikes of t through the plastic inputs of each target */
        const int32_t k = t - base;
        for (int32_t f = tick_ptr[k]; f < tick_ptr[k + 1]; ++f) {
            const int32_t p = ids[f];
            for (int64_t j = net->in_ptr[p]; j < net->in_ptr[p + 1]; ++j) {
                const int64_t s = net->in_syn[j];
                stdp_set_update(net, s, t, inv_ltp, inv_ltd);
                net->apost[s] += 1.0f;
                net->w[s] += net->max_ltp * net->apre[s];
                stdp_clip(net, s);
                ++updates;
            }
        }
    }
    return updates;
}

void orc_stdp_replay(float max_ltp, float ltp_tau, float max_ltd, float ltd_tau, float dt, float w0, float w_max,
                     const uint8_t* pre, const uint8_t* post, int32_t nticks, float* w_out) {
    float w = w0, apre = 0.0f, apost = 0.0f;
    int32_t lastu = 0;
    const float inv_ltp = (float)(1.0 / (double)ltp_tau);
    const float inv_ltd = (float)(1.0 / (double)ltd_tau);
    w_out[0] = w;
    for (int32_t t = 0; t < nticks; ++t) {
        if (pre[t] || post[t]) {
            for (int side = 0; side < 2; ++side) {
                if (!(side == 0 ? pre[t] : post[t])) continue;
                const float dT = (float)(t - lastu) * dt;
                apre *= expf(-dT * inv_ltp);
                apost *= expf(-dT * inv_ltd);
                lastu = t;
                if (side == 0) {
                    apre += 1.0f;
                    w += -max_ltd * apost;
                } else {
                    apost += 1.0f;
                    w += max_ltp * apre;
                }
                if (w < 0.0f) w = 0.0f;
                else if (w > w_max) w = w_max;
            }
        }
        w_out[t + 1] = w;
    }
}
