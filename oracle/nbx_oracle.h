/*
 * CPU oracle for the NeuronBridge per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * legacy CPU/OpenMP semantics (neuron update, spike delivery through integer
 * delays, lazy pair STDP) used to check the HIP library. Only tests, the smoke
 * check and the CPU-baseline leg of the benchmark load it; the product library
 * never links it and never falls back to it.
 *
 * Reference files restated (paths relative to the reference root;
 * L/ = src/native/legacy/source_file_realtime_v1_async/):
 *   L/Intergration/inc/FixStep/ForwardEulerMethod.h:29-55
 *   L/NeuralModel/src/TimeDriven/TimeDrivenLIF_Voltage_jump.cpp:49-115
 *   L/NeuralModel/src/TimeDriven/TimeDrivenLIF_Exponential_Decay.cpp:37-121
 *   L/NeuralModel/src/TimeDriven/TimeDrivenIzhikevic_Exponential_Decay.cpp:32-131
 *   L/NeuralModel/src/TimeDriven/TimeDrivenLIF_Exponential_double.cpp:36-123
 *   L/NeuralModel/src/TimeDriven/TimeDrivenLIF_Exponential_triple.cpp:56-155
 *   L/NeuralModel/src/TimeDriven/PoissonRate.cpp:7-88
 *   src/native/gpu/include/dense_subnetwork/model/DenseTriggerRelayDeviceUpdate.cuh:10-32
 *   L/NeuralModel/src/NeuralStateVector/Neuron_State_Vector.cpp:43-67
 *   L/Event/inc/Event.h:10 (deliveries before updates within a tick)
 *   L/Event/src/Spike/PropogatedSpikeGroup.cpp:33-71 (ProcessSpike, then pre-learning)
 *   L/Event/src/Spike/TimeDrivenInternalSpike.cpp:11-144 (group time = t + delay; post-learning)
 *   L/LearningRule/src/Rule/STDP.cpp:19-61, L/LearningRule/src/State/STDP_State.cpp:17-34
 *   L/LearningRule/src/Rule/R_STDP.cpp:19-70, L/LearningRule/src/Rule/AdditiveKernalChange.cpp:64-82
 *   L/LearningRule/src/Rule/CerebullarLearningRule.cpp:9-134, L/LearningRule/inc/SpikeBuffer/BufferedActivityTime.h
 *   L/connections/inc/Interconnections.h:120-128 (UpdateWeight)
 *
 * Parity status of each function is listed in the header comment of nbx_oracle.c.
 */
#ifndef NBX_ORACLE_H
#define NBX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_LIF_VJ = 0,
    ORC_LIF_EXP = 1,
    ORC_IZH = 2,
    ORC_IZH_CUR = 3,
    ORC_LIF_EXP2 = 4,     /* excitatory + inhibitory conductance LIF (state V, gexc, ginh, I) */
    ORC_LIF_EXP3 = 5,     /* AMPA + GABA + NMDA conductance LIF (state V, ampa, gaba, nmda, I) */
    ORC_POISSON = 6,      /* PoissonRate (state rate, I) */
    ORC_TRIGGER_RELAY = 7 /* fires when the step's excitatory input exceeds v_th (state last input) */
};
/* conductance inputs present in a network: the reference models' CheckType flags */
enum { ORC_COND_EXC = 1, ORC_COND_INH = 2, ORC_COND_NMDA = 4 };
enum { ORC_EULER = 0, ORC_RK4 = 1 };
enum { ORC_ACC_FIXED = 0, ORC_ACC_FLOAT_EVENTS = 1 };

typedef struct orc_model {
    int32_t model;       /* ORC_* model id */
    int32_t integrator;  /* ORC_EULER | ORC_RK4 */
    int32_t n;           /* neurons */
    int32_t t_ref;       /* LIF refractory period in steps */
    float dt;            /* ms */
    /* LIF (model-wide) */
    float v_rest, tau, v_th, r, v_reset, g_tau, e_rev;
    /* Izhikevich (model-wide) */
    float izh_v_th, izh_r, e_ampa, e_gaba, ampa_tau, gaba_tau;
    /* Izhikevich per-neuron a, b, c, d */
    const float *a, *b, *c, *d;
    /* filled by orc_model_prepare: exp(-dt/tau) factors computed the reference way */
    float g_decay, ampa_decay, gaba_decay;
    /* ORC_LIF_EXP2 uses e_ampa / ampa_tau (Eexc, gexc_tau) and e_gaba / gaba_tau (Einh, ginh_tau) */
    float nmda_tau;               /* ORC_LIF_EXP3 */
    float rate_bias, rate_gain;   /* ORC_POISSON */
    float nmda_decay;             /* filled by orc_model_prepare */
    int32_t cond;                 /* ORC_COND_* bits (orc_run derives them from the network) */
    int32_t step;                 /* tick being updated (the PoissonRate hash input; orc_run sets it) */
} orc_model;

typedef struct orc_net {
    orc_model m;
    int32_t d_slots;         /* delay ring slots = max_delay + 1 */
    int32_t fixed_shift;     /* arrivals accumulate as int32 in units of 2^-fixed_shift */
    /* sparse synapses, CSR by presynaptic neuron, rows sorted by (delay, post) */
    const int64_t* row_ptr;  /* n + 1 (NULL: no sparse synapses) */
    const int32_t* post;
    const uint8_t* delay;    /* 1 .. d_slots-1 */
    const uint8_t* type;     /* 0 excitatory, 1 inhibitory */
    float* w;                /* magnitudes (signed for ORC_IZH_CUR); STDP updates in place */
    /* dense all-to-all delivery: wt[pre*n + post] (NULL: none), fixed delay */
    const float* dense_wt;
    int32_t dense_delay;
    /* Poisson background: count ~ Poisson(lambda) per (neuron, tick), w_bg excitatory */
    int32_t pois_on;
    uint32_t seed_lo, seed_hi;
    double pois_lambda;
    float w_bg;
    /* external current per tick: cur[(t - t_begin) * n + i] replaces I_ext (NULL: none) */
    const float* cur;
    int32_t cur_t_begin;
    /* pair STDP (reference STDP::Apply*SynaticSpike) */
    int32_t stdp_on;
    const uint8_t* plastic;  /* per synapse */
    float max_ltp, ltp_tau, max_ltd, ltd_tau, w_max;
    float* apre;             /* per synapse */
    float* apost;
    int32_t* slast;
    const int64_t* in_ptr;   /* n + 1: plastic incoming synapses of each post */
    const int64_t* in_syn;   /* synapse ids (CSR positions) */
    /* trigger-driven rules (LearningRule/src/Rule/R_STDP.cpp, AdditiveKernalChange.cpp); traces in
     * apre / apost / slast like STDP (a synapse is STDP-plastic, rule-plastic or a trigger) */
    int32_t nrules;
    const struct orc_rule* rules;
    const int8_t* syn_rule;  /* per synapse: rule of a rule-plastic synapse, else -1 */
    const int8_t* syn_trig;  /* per synapse: rule of a trigger synapse, else -1 */
    const float* maxw;       /* per synapse weight bound (Interconnections::maximum_weight) */
    float* elig;             /* per synapse R-STDP eligibility */
    const float* stdp_maxw;  /* per synapse STDP weight bound (STDP.cpp:29-34), NULL: w_max for all */
    const int32_t* conn_index; /* per synapse connection index (Interconnections::Index), NULL: CSR position */
    /* several neuron populations (npops > 0; m is then unused): population k owns neurons
     * [pop_begin[k], pop_begin[k+1]) with its own model pops[k] (the legacy engine's several
     * NeuronModel objects joined by Interconnections). Each synapse's arrival goes to its target's
     * model; fixed-point arrivals accumulate per synapse type and each model folds them. The AoS
     * state x[i * ns + j] uses ns = the largest orc_nstate of the populations. */
    int32_t npops;
    orc_model* pops;
    const int32_t* pop_begin;
    /* inputs from outside the synapses for ticks [input_t0, input_t0 + input_nsteps): those of tick t
     * are inputs[input_ptr[t - input_t0] .. input_ptr[t - input_t0 + 1]) */
    const struct orc_input* inputs;
    const int32_t* input_ptr;
    int32_t input_t0, input_nsteps;
} orc_net;

/* One external input of a tick: kind 0 / 1 / 2 an arrival of that synapse type (value the weight,
 * as a synapse's arrival), kind 3 a new I_ext (value replaces the target's external current, before
 * the tick's current table). */
enum { ORC_INPUT_CURRENT = 3 };
typedef struct orc_input {
    int32_t target, kind;
    float value;
} orc_input;

enum { ORC_RULE_RSTDP = 1, ORC_RULE_ADDITIVE = 2, ORC_RULE_CEREBELLAR = 3 };
typedef struct orc_rule {
    int32_t kind, clear;     /* clear: R-STDP ClearEligibilityAfterTrigger */
    float max_ltp, ltp_tau, max_ltd, ltd_tau, reward, punishment, a1pre, a2prepre;
    float initpos, maxpos, kernel_step, min_weight; /* cerebellar (CerebullarLearningRule.h:9-21) */
    uint32_t seed;
} orc_rule;

int orc_nstate(int model);
void orc_model_prepare(orc_model* m);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* PoissonRate hash (PoissonRate.cpp:7-21): 24-bit uniform in [0,1). */
float orc_hash_uniform01(int32_t neuron, int32_t t);
/* CerebullarLearningRule.cpp:11-25: 24-bit uniform keyed by (seed, rule, connection, step). */
float orc_stable_unit_random(uint32_t seed, int32_t rule_id, int32_t synapse_id, int32_t spike_time);
/* Cerebellar kernel table (CerebullarLearningRule.cpp:56-88); returns its length, writes the
 * window (maxtimemeasured, ms) and 1/kernel_step; tab may be NULL for a size query. */
int32_t orc_cerebellar_kernel(const orc_rule* r, float* tab, int32_t cap, float* maxtime, float* inv_step);
/* Inverse-CDF thresholds of Poisson(lambda) in units of 2^-32; returns kmax. */
int32_t orc_poisson_thresholds(double lambda, uint32_t* tab, int32_t cap);
int32_t orc_poisson_count(uint32_t seed_lo, uint32_t seed_hi, int32_t neuron, int32_t t,
                          const uint32_t* tab, int32_t kmax);
/* fixed-point image of a weight: rint(w * 2^shift) */
int32_t orc_quantize(float w, int32_t shift);

/* initial state like Neuron_State_Vector::InitNeuronState with sigma = 0 */
void orc_init_state(const orc_model* m, float* x, int32_t* last);

/* one neuron-update pass (FP32 reference order / FP64); returns spike count */
int32_t orc_update_f32(const orc_model* m, float* x, int32_t* last, const float* isyn, int32_t* spikes);
int32_t orc_update_f64(const orc_model* m, double* x, int32_t* last, const double* isyn, int32_t* spikes);

/*
 * Closed-loop network run for ticks [t0, t0 + nticks).
 *   acc_mode ORC_ACC_FIXED:        arrivals accumulate in int32 fixed point, applied once per tick
 *   acc_mode ORC_ACC_FLOAT_EVENTS: every arrival is applied as its own float add (reference order)
 * hist/hist_cnt carry the last d_slots ticks of spike lists between calls:
 *   hist[(t % d_slots) * n + k], hist_cnt[t % d_slots].
 * ring (optional, fixed mode, non-plastic): push-delivery image ring[slot * 2n + ch * n + i]
 *   maintained alongside the pull delivery so tests can compare device ring contents.
 * Outputs: spike_counts[nticks], raster (optional) bits [nticks][ceil(n/32)], events delivered.
 */
int32_t orc_run(orc_net* net, int32_t acc_mode, float* x, int32_t* last, int32_t* hist, int32_t* hist_cnt,
                int32_t* ring, int32_t t0, int32_t nticks, int32_t* spike_counts, uint32_t* raster,
                int64_t* events);

/* The legacy engine's queue-parallel OpenMP run (one thread per queue of neurons, per-queue
 * synapse lists, per-event float delivery): the CPU baseline. Results equal orc_run in
 * ORC_ACC_FLOAT_EVENTS mode. Sparse networks without learning, dense weights, populations or
 * external inputs (else returns -3). loop_seconds (may be NULL): wall time of the stepping loop,
 * without building the per-queue synapse lists. */
int32_t orc_run_queues(orc_net* net, int32_t nq, float* x, int32_t* last, int32_t* hist, int32_t* hist_cnt,
                       int32_t t0, int32_t nticks, int32_t* spike_counts, int64_t* events, double* loop_seconds);

/* Replay STDP over a recorded raster: ids[tick_ptr[k] .. tick_ptr[k+1]) fired at step
 * t0 - lead + k (k < lead + nsteps); applies the pair-STDP updates of steps [t0, t0 + nsteps)
 * in orc_run's order (arrivals of a step, then its spikes). Returns the number of updates. */
int64_t orc_replay_stdp(orc_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t lead,
                        int32_t nsteps);

/* STDP replay on one synapse for a given event schedule (pair_stdp_gpu_trajectory.cpp:52-66). */
void orc_stdp_replay(float max_ltp, float ltp_tau, float max_ltd, float ltd_tau, float dt, float w0, float w_max,
                     const uint8_t* pre, const uint8_t* post, int32_t nticks, float* w_out);

#ifdef __cplusplus
}
#endif
#endif
