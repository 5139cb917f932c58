/*
 * nbx_core.h — C ABI of the MI355X (gfx950) simulation core for NeuronBridge's
 * per-timestep hot path: neuron update (LIF voltage-jump, LIF exponential-decay,
 * Izhikevich; forward Euler or RK4), spike delivery (CSR fan-out into an integer
 * delay ring, pull-at-arrival delivery from a firing-history ring, dense W·s),
 * and lazy pair STDP.
 *
 * Plain C types only: host pointers for uploads/downloads, opaque handles, int
 * status codes, and a thread-local last-error string. Every call is synchronous
 * on return unless its comment says otherwise, which is the contract the
 * reference's callers rely on (every dense-runtime call ends with a device sync,
 * src/native/gpu/src/gpu_runtime/GpuPropagationRuntime.cu:1687-1827).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   G/ = src/native/gpu/include/, L/ = src/native/legacy/source_file_realtime_v1_async/
 *   npgr::GpuPropagationRuntime           G/gpu_runtime/GpuPropagationRuntime.h:79-144
 *   npgr::GpuPropagationLayout            G/gpu_runtime/GpuPropagationLayout.h:20-70
 *   npgr::DenseUnifiedNeuronRuntime       G/dense_subnetwork/DenseUnifiedNeuronRuntime.h:15-41
 *   npgr::DeviceCommonBuffersView         G/dense_subnetwork/DenseNeuronDeviceViews.h:6-13
 *   legacy GPU neuron models (the *_GPU_Interface classes in L/NeuralModel/inc/TimeDrivenGPU/)
 *                                         (UpdateState/ProcessSpike, NeuronModel.h:66-149)
 */
#ifndef NBX_CORE_H
#define NBX_CORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define NBX_OK 0
#define NBX_ERR_INVALID (-1)     /* bad argument or layout (reference: false + *reason) */
#define NBX_ERR_HIP (-2)         /* HIP runtime error */
#define NBX_ERR_UNSUPPORTED (-3) /* valid request this build does not implement */
#define NBX_ERR_STATE (-4)       /* call out of order, or a device-side error flag */

/* Message of the last failing call on this thread (reference: std::string* reason). */
const char* nbx_last_error(void);
/* ABI version (major*10000 + minor*100 + patch). */
int32_t nbx_version(void);
/* Number of visible HIP devices (0 when none). */
int32_t nbx_device_count(void);

/* ------------------------------------------------------------------ models */
enum {
    NBX_MODEL_LIF_VJ = 0,  /* TimeDrivenLIF_Voltage_jump (TimeDrivenLIF_Voltage_jump.cpp:49-115) */
    NBX_MODEL_LIF_EXP = 1, /* TimeDrivenLIF_Exponential_Decay (.cpp:62-121) */
    NBX_MODEL_IZH = 2,     /* TimeDrivenIzhikevic_Exponential_Decay, AMPA/GABA conductances (.cpp:59-131) */
    NBX_MODEL_IZH_CUR = 3, /* Izhikevich with current-based synapses (signed weights, one-tick current) */
    NBX_MODEL_LIF_EXP2 = 4, /* TimeDrivenLIF_Exponential_double: excitatory + inhibitory conductances (.cpp:61-123) */
    NBX_MODEL_LIF_EXP3 = 5, /* TimeDrivenLIF_Exponential_triple: AMPA, GABA, NMDA (Mg block) conductances (.cpp:79-155) */
    NBX_MODEL_POISSON = 6,  /* PoissonRate: rate-driven hash spikes (PoissonRate.cpp:7-88) */
    NBX_MODEL_TRIGGER_RELAY = 7 /* DenseTriggerRelay: fires when the step's excitatory input exceeds v_th
                                   (dense_subnetwork/model/DenseTriggerRelayDeviceUpdate.cuh:10-32) */
};
enum { NBX_EULER = 0, NBX_RK4 = 1 };

/* Neuron population description (reference: DenseNeuronModelSpec / legacy model parameters). */
typedef struct nbx_model_desc {
    int32_t model;       /* NBX_MODEL_* */
    int32_t integrator;  /* NBX_EULER (the reference's only integrator) or NBX_RK4 */
    int32_t n;           /* neurons */
    int32_t t_ref;       /* LIF refractory period in steps (reference t_ref, int) */
    float dt;            /* step in ms */
    float v_rest, tau, v_th, r, v_reset, g_tau, e_rev;             /* LIF, model-wide */
    float izh_v_th, izh_r, e_ampa, e_gaba, ampa_tau, gaba_tau;     /* Izhikevich, model-wide */
    const float *izh_a, *izh_b, *izh_c, *izh_d;                    /* Izhikevich, per neuron (host) */
    /* LIF_EXP2 uses e_ampa / ampa_tau for its excitatory and e_gaba / gaba_tau for its inhibitory
     * conductance (the reference's Eexc, gexc_tau, Einh, ginh_tau); LIF_EXP3 uses all three. */
    float nmda_tau;                        /* LIF_EXP3 */
    float rate_bias_hz, rate_gain_hz;      /* POISSON: rate = bias + arrivals + gain * I_ext */
} nbx_model_desc;

/* Synapses as CSR by presynaptic neuron, each row sorted by (delay, post) —
 * the reference's SynapseArrays + PreDelaySlices (GpuPropagationLayout.h:20-49)
 * folded into one ordering. Host pointers; copied at nbx_net_set_sparse. */
typedef struct nbx_sparse_desc {
    int64_t count;           /* synapses */
    int32_t max_delay;       /* steps; delay ring has max_delay + 1 slots */
    const int64_t* row_ptr;  /* n + 1 */
    const int32_t* post;     /* count */
    const uint8_t* delay;    /* count, 1 .. max_delay */
    const uint8_t* type;     /* count, 0 excitatory, 1 inhibitory (reference Interconnections::type) */
    const float* weight;     /* count; magnitudes (signed for NBX_MODEL_IZH_CUR) */
    const uint8_t* plastic;  /* count or NULL; 1 = STDP synapse */
} nbx_sparse_desc;

/* Pair STDP parameters (reference STDP.h:14-20, STDP.cpp:19-61). */
typedef struct nbx_stdp_desc {
    float max_ltp, ltp_tau, max_ltd, ltd_tau, w_max;
} nbx_stdp_desc;

/* Trigger-driven learning rules of the legacy engine (LearningRule/src/Rule/):
 *   NBX_RULE_RSTDP     R_STDP.cpp:19-70 — pair traces feed an eligibility (pre arrival:
 *                      e += -max_ltd * Apost; post spike: e += Apre * max_ltp); a trigger arrival
 *                      at the target adds factor * e to each plastic input's weight (factor =
 *                      reward for excitatory, punishment for inhibitory trigger synapses) and
 *                      clears e when clear_eligibility is set
 *   NBX_RULE_ADDITIVE  AdditiveKernalChange.cpp:64-82 — pre arrival: w += a1pre, Apre decays with
 *                      ltp_tau and jumps by 1; a trigger arrival adds a2prepre * Apre to each
 *                      plastic input of its target
 *   NBX_RULE_CEREBELLAR CerebullarLearningRule.cpp:56-134 — pre arrival: w += a1pre * U(seed, rule,
 *                      connection index, step); a trigger arrival adds a2prepre * kernel(elapsed)
 *                      for every arrival of each plastic input within the kernel's window, newest
 *                      first, weights bounded below by min_weight; the kernel is
 *                      (t - initpos) / width * exp(1 - (t - initpos) / width) tabulated every
 *                      kernel_step ms (width = maxpos - initpos). Each synapse keeps its last
 *                      `history` arrival steps (0: 16); a run that would drop one still inside the
 *                      window stops with NBX_ERR_STATE.
 * Weights stay in [0, max_weight] of their synapse (Interconnections::UpdateWeight). */
enum { NBX_RULE_STDP = 0, NBX_RULE_RSTDP = 1, NBX_RULE_ADDITIVE = 2, NBX_RULE_CEREBELLAR = 3 };
/* (NBX_RULE_STDP: nbx_net_assign_nbnet_rules only) */
typedef struct nbx_rule_desc {
    int32_t kind;                         /* NBX_RULE_* */
    float max_ltp, ltp_tau, max_ltd, ltd_tau; /* R-STDP; ADDITIVE uses ltp_tau */
    float reward, punishment;             /* R-STDP trigger factors (R_STDP.h:16-18: 1, -1) */
    int32_t clear_eligibility;            /* R-STDP (R_STDP.h:20: true) */
    float a1pre, a2prepre;                /* ADDITIVE (AdditiveKernalChange.h:8-10: 0.002, -0.08), CEREBELLAR */
    float initpos, maxpos, kernel_step, min_weight; /* CEREBELLAR (CerebullarLearningRule.h:9-19: 120, 150,
                                                       0.1 ms, 0) */
    uint32_t seed;                        /* CEREBELLAR random_seed (17) */
    int32_t history;                      /* CEREBELLAR arrivals kept per synapse (0: 16) */
} nbx_rule_desc;

/* Engines for nbx_net_run. */
enum {
    NBX_ENGINE_KERNELS = 0,    /* one launch per phase per step */
    NBX_ENGINE_GRAPH = 1,      /* the same launches captured in a hipGraph of many steps */
    NBX_ENGINE_PERSISTENT = 2, /* one cooperative launch; post-blocked delivery into an LDS ring
                                  (dense networks of up to 1024 neurons: one workgroup for all steps) */
    NBX_ENGINE_AUTO = 3        /* persistent when the network fits it, else graph */
};

typedef struct nbx_net nbx_net;

/* Create a network on `device` (reference: DenseUnifiedNeuronRuntime::Initialize +
 * AllocateDeviceBuffers, legacy InitStateVector). State starts like
 * Neuron_State_Vector::InitNeuronState with sigma 0 (LastSpike = 10000). */
int32_t nbx_net_create(const nbx_model_desc* model, int32_t device, nbx_net** out);
/* A network of `count` (1..64) neuron populations (reference: DenseUnifiedNeuronRuntime::Initialize
 * with one DenseNeuronModelSpec per range, DenseUnifiedNeuronRuntime.h:20-23; the legacy engine's
 * several NeuronModel objects joined by Interconnections). Population k owns the neurons
 * [pops[0].n + ... + pops[k-1].n, + pops[k].n) and keeps its own model, integrator and
 * parameters; every population steps with the same dt, and synapses of the network join any two
 * neurons (each synapse's type must be an input of its target's model). Arrivals are kept by
 * synapse type (ring channel = type: 0 excitatory, 1 inhibitory, 2 NMDA, magnitudes as stored)
 * and each population folds them into its model's inputs, so a population computes exactly what a
 * network of its model alone computes from the same arrivals. State is x[i * nstate + j] with
 * nstate the largest of the populations' (columns a model does not have read 0). Such networks run
 * on the kernel and graph engines and the per-phase calls; dense all-to-all weights need one
 * population. nbx_net_create(md, ...) is the one-population case. */
int32_t nbx_net_create_populations(const nbx_model_desc* pops, int32_t count, int32_t device, nbx_net** out);
int32_t nbx_net_destroy(nbx_net* net);

/* Sparse synapses (reference: GpuPropagationRuntime::Initialize(layout)). */
int32_t nbx_net_set_sparse(nbx_net* net, const nbx_sparse_desc* syn);
/* Dense all-to-all weights wt[pre * n + post] with one delay (FP32, host). */
int32_t nbx_net_set_dense(nbx_net* net, const float* wt, int32_t delay);
/* Poisson background: per (neuron, step) count ~ Poisson(lambda) of excitatory
 * arrivals of weight w_bg, drawn from Philox4x32-10 keyed by the seed. */
int32_t nbx_net_set_poisson(nbx_net* net, uint32_t seed_lo, uint32_t seed_hi, double lambda, float w_bg);
/* Enable pair STDP on the synapses flagged plastic. */
int32_t nbx_net_set_stdp(nbx_net* net, const nbx_stdp_desc* stdp);
/* Trigger-driven rules (at most 64). In CSR order: rule[j] >= 0 makes synapse j plastic under
 * rules[rule[j]], trigger[j] >= 0 makes it a trigger of rules[trigger[j]] (the .nbnet record's
 * synapse_rule / trigger_rule; NbnetReader.cpp:505-509), max_weight[j] bounds its weight. A
 * synapse is at most one of: STDP-plastic, rule-plastic, trigger. Within a step, every arrival's
 * own update comes first, then each target applies its trigger arrivals in ascending synapse
 * order, then the neurons update, then post-side updates run (Event.h:10 orders deliveries
 * before updates). Before nbx_net_finalize. */
int32_t nbx_net_set_rules(nbx_net* net, const nbx_rule_desc* rules, int32_t count, const int32_t* rule,
                          const int32_t* trigger, const float* max_weight);
/* Learning rules of a network whose synapses came from nbx_net_load_nbnet, by the file's rule
 * index (the order of its metadata's "learning_rules"): a record's synapse_rule / trigger_rule
 * select rules[...]; NBX_RULE_STDP rules become pair STDP on their synapses (one parameter set per
 * network), the others as in nbx_net_set_rules. Every rule-driven weight is bounded by its record's
 * max_weight (STDP.cpp:29-34, Interconnections.h:120-128); a trigger id naming an STDP rule is
 * ignored, as the reference ignores triggers of rules without a trigger side (Network.cu:438-445).
 * Replaces earlier nbx_net_set_stdp / nbx_net_set_rules settings. Before nbx_net_finalize. */
int32_t nbx_net_assign_nbnet_rules(nbx_net* net, const nbx_rule_desc* rules, int32_t count);
/* Per-synapse rule state (CSR order; entries of other synapses read 0). */
int32_t nbx_net_get_rule_state(nbx_net* net, float* apre, float* apost, float* eligibility,
                               int32_t* last_update, int64_t count);
/* External current per step: cur[(t - t_begin) * n + i] replaces I_ext at step t
 * (reference: ProcessCurrent / InterfaceChannelValue with overwrite). Host array, copied. */
int32_t nbx_net_set_current(nbx_net* net, const float* cur, int32_t t_begin, int32_t nsteps);
/* Arrivals accumulate as int32 in units of 2^-shift; 0 picks the largest shift
 * that cannot overflow for this network. Must precede nbx_net_finalize, which rejects a
 * requested shift whose int32 sums could overflow (NBX_ERR_INVALID). */
int32_t nbx_net_set_fixed_shift(nbx_net* net, int32_t shift);
int32_t nbx_net_get_fixed_shift(const nbx_net* net, int32_t* shift);
/* Upload and build device layouts; required before stepping. */
int32_t nbx_net_finalize(nbx_net* net);

/* State in the reference's AoS order x[i * nstate + j] (nstate per model: LIF_VJ 2, LIF_EXP 3,
 * IZH / IZH_CUR 6, LIF_EXP2 4, LIF_EXP3 5, POISSON 2, TRIGGER_RELAY 1)
 * plus LastSpike. */
int32_t nbx_net_nstate(const nbx_net* net);
int32_t nbx_net_set_state(nbx_net* net, const float* x, const int32_t* last);
int32_t nbx_net_get_state(nbx_net* net, float* x, int32_t* last);

/* Record per-step spike counts and/or a spike raster (bit i of word
 * [(t - t_begin) * ceil(n/32) + i/32]) for steps [t_begin, t_begin + nsteps). */
enum { NBX_RECORD_COUNTS = 1, NBX_RECORD_RASTER = 2 };
int32_t nbx_net_record(nbx_net* net, int32_t t_begin, int32_t nsteps, int32_t flags);
int32_t nbx_net_get_counts(nbx_net* net, int32_t* counts, int32_t nsteps);
int32_t nbx_net_get_raster(nbx_net* net, uint32_t* raster, int64_t nwords);

/* Advance the network by nsteps closed-loop steps (reference: Simulation::RunSimulationStep
 * driving DenseSubnetworkRuntimeGpu::RunStep once per step). */
int32_t nbx_net_run(nbx_net* net, int32_t nsteps, int32_t engine);
/* Same, but returns once the work is queued on the network's stream (graph-capturable
 * engines only); nbx_net_sync waits for it. */
int32_t nbx_net_run_async(nbx_net* net, int32_t nsteps, int32_t engine);
int32_t nbx_net_sync(nbx_net* net);
/* Engine nbx_net_run would use for NBX_ENGINE_AUTO. */
int32_t nbx_net_pick_engine(const nbx_net* net, int32_t* engine);
/* Current step (the next step to be simulated). */
int32_t nbx_net_time(const nbx_net* net, int32_t* t);
/* Synaptic events delivered so far (one per synapse per presynaptic spike). */
int32_t nbx_net_events(nbx_net* net, int64_t* events);

/* ---------------- per-phase calls (the reference runtime's step sequence) ----------------
 * GpuPropagationRuntime::BeginStep / RunPropagationOnly / CommitCurrentDeviceFiringsFromExternal /
 * ApplyPostLearningFromDeviceFirings and DenseUnifiedNeuronRuntime::Step, in the order
 * DenseSubnetworkRuntimeGpu::RunStep calls them (DenseSubnetworkRuntimeGpu.cpp:625-663). */
int32_t nbx_net_begin_step(nbx_net* net, int32_t t);
/* Deliver arrivals due at the current step (pull delivery and pre-side STDP). */
int32_t nbx_net_run_propagation(nbx_net* net);
/* Neuron update of the current step; optionally copies out the firing ids (ascending)
 * and count (reference: DenseUnifiedNeuronRuntime::Step output_firing_ids). */
int32_t nbx_net_neuron_step(nbx_net* net, int32_t* ids_out, int32_t ids_cap, int32_t* count_out);
/* Output neurons (reference: DenseUnifiedNeuronRuntime::ConfigureOutputNeuronMask,
 * DenseUnifiedNeuronRuntime.h:26-27): mask[i] != 0 marks neuron i as an output of the network;
 * NULL marks every neuron. The mask selects the output list of nbx_net_neuron_step_outputs; every
 * firing is still delivered, recorded and committed. */
int32_t nbx_net_set_output_mask(nbx_net* net, const uint8_t* mask);
/* The neuron update of nbx_net_neuron_step, copying out the step's firings among the output
 * neurons and (full_ids non-NULL) all of them, each ascending (reference: Step(common, t,
 * output_firing_ids, full_firing_ids, export_full_firings_to_host)). */
int32_t nbx_net_neuron_step_outputs(nbx_net* net, int32_t* out_ids, int32_t out_cap, int32_t* out_count,
                                    int32_t* full_ids, int32_t full_cap, int32_t* full_count);
/* Post-side STDP for the neurons that fired this step. */
int32_t nbx_net_apply_post_learning(nbx_net* net);
/* Push this step's firings into the delay ring and close the step. */
int32_t nbx_net_commit_firings(nbx_net* net);
/* Commit with the caller's full firing list of the current step (reference:
 * CommitCurrentDeviceFiringsFromExternal(t, d_external_firing_ids, d_external_firing_count),
 * GpuPropagationRuntime.h:96-99; device pointers). The list replaces the step's list: ids beyond the
 * neurons that fired in the update are injected as firings of the step (as nbx_net_inject_spikes),
 * and then the step is committed (nbx_net_commit_firings). A list that omits a neuron that fired,
 * repeats an id or holds one out of range is rejected (NBX_ERR_INVALID, nothing committed). The
 * pointers of nbx_net_firing_device_view commit the step's own list. */
int32_t nbx_net_commit_external(nbx_net* net, int32_t t, const int32_t* d_ids, const int32_t* d_count);
/* Device views of the current step's firings after its neuron update (reference:
 * DenseUnifiedNeuronRuntime::device_full_firing_ids / device_full_firing_count /
 * device_current_did_fire): the step's id list (unordered) and count inside the firing-history
 * ring, and per-neuron did-fire flags (a uint8 array of the network, rebuilt by this call). */
int32_t nbx_net_firing_device_view(nbx_net* net, int32_t** d_ids, int32_t** d_count, uint8_t** d_did_fire);
/* Inject spikes as if `ids` fired at the current step (reference: QueueInterfaceSpike /
 * AddExternalSpikeActivity); call between nbx_net_neuron_step and nbx_net_commit_firings.
 * A neuron fires at most once per step: an id repeated in the call or already in the step's
 * firing list (fired in the update, or injected earlier) is rejected (NBX_ERR_INVALID, nothing
 * injected). */
int32_t nbx_net_inject_spikes(nbx_net* net, const int32_t* ids, int32_t count);

/* ---------------- interface channels (inputs from outside the network) ----------------
 * GpuPropagationRuntime::SetInterfaceTargetNeurons / UploadInterfaceChannelValues /
 * UploadInterfaceChannelValuesDevice / ApplyInterfaceStateVectors (GpuPropagationRuntime.h:103-111,
 * GpuPropagationRuntime.cu:68-140,1431-1560). A record names its target neuron directly
 * (target_neuron >= 0) or through an interface slot (target_neuron < 0: the neuron mapped to
 * interface_slot); records whose target resolves to no neuron are skipped, as the reference does.
 * Channels: NBX_PENDING_EXC / _INH / _NMDA add value * scale as an arrival of synapse type 0 / 1 / 2
 * for the current step (int32 fixed point like every arrival; an inhibitory value is a magnitude);
 * NBX_PENDING_CURRENT with overwrite set replaces the target's I_ext by value * scale (the
 * reference's stateful current slots; a per-step current table of nbx_net_set_current still wins
 * for the steps it covers); NBX_PENDING_MODULATORY and channels the target's model does not take
 * are skipped. The current step is the one being run (after nbx_net_begin_step, before the neuron
 * update) or, between steps, the next one. Records sharing a (target, channel) accumulate; an
 * overwrite record of a conductance channel adds its value, since the reference's pending buffer
 * holds nothing of the step when the interface is applied. Values whose fixed-point image could
 * overflow the arrival sums are rejected (host records: NBX_ERR_INVALID, nothing applied; device
 * copies: the value is dropped and the next call reports NBX_ERR_STATE). */
enum { NBX_PENDING_EXC = 0, NBX_PENDING_INH = 1, NBX_PENDING_CURRENT = 2, NBX_PENDING_NMDA = 3,
       NBX_PENDING_MODULATORY = 4 };
typedef struct nbx_interface_value {   /* reference InterfaceChannelValue */
    int32_t interface_slot, target_neuron;
    uint8_t pending_channel, overwrite;
    float value, scale;
} nbx_interface_value;
typedef struct nbx_interface_copy {    /* reference DeviceToDeviceInterfaceCopySpec */
    int32_t interface_slot, target_neuron;
    uint8_t pending_channel, overwrite;
    int32_t source_index;
    float scale;
} nbx_interface_copy;
/* Slot -> target neuron map (ids in [0, n), or -1 for an unmapped slot). */
int32_t nbx_net_set_interface_targets(nbx_net* net, const int32_t* neuron_ids, int32_t count);
/* Stages host records for nbx_net_apply_interface (the record count must equal the slot count
 * when a slot map is set; 0 records clears the stage). A zero value is skipped, as the reference's
 * kernel skips it; a current record must have overwrite set. */
int32_t nbx_net_upload_interface_values(nbx_net* net, const nbx_interface_value* values, int32_t count);
/* Applies the staged host records to the current step (they stay staged until replaced). */
int32_t nbx_net_apply_interface(nbx_net* net);
/* Applies device-to-device records at once: value = d_values[source_index] * scale, read on the
 * device (d_values: device pointer); a zero value is skipped unless overwrite is set, and a current
 * record must have overwrite set. */
int32_t nbx_net_upload_interface_values_device(nbx_net* net, const float* d_values, const nbx_interface_copy* specs,
                                               int32_t count);

/* Arrivals from outside the network's synapses (reference: ProcessSpike calls of other models'
 * and queues' spike groups on this model, which the legacy GPU models stage in AuxStateCPU,
 * TimeDrivenLIF_Exponential_Decay_GPU_Interface.cu:50-58): adds weights[k] to neuron targets[k] as
 * an arrival of synapse type types[k] (0 exc, 1 inh, 2 NMDA; magnitudes, signed for current-based
 * Izhikevich) delivered at step t, exactly like a synapse's arrival at that step. t must lie in
 * [current step, current step + slots - 1] (the delay ring's reach; the current step only before its
 * neuron update). Weights whose fixed-point sum could overflow are rejected (NBX_ERR_INVALID). */
int32_t nbx_net_add_arrivals(nbx_net* net, const int32_t* targets, const uint8_t* types, const float* weights,
                             int32_t count, int32_t t);

/* Current synapses (synapse type 3; reference: type-3 Interconnections registered as input-current
 * slots of their target by the model's CheckType, TimeDrivenLIF_Voltage_jump.cpp:127-147, and
 * CurrentSynapse, Current/src/CurrentSynapes.cpp; filled by PropogatedCurrent events,
 * Event/src/Current/PropogatedCurrent.cpp:15-25). Declares `count` current synapses src[k] -> dst[k]
 * with delay[k] (1..63) steps; each target's slots are numbered in declaration order and start at 0.
 * A current injected at source s for step t (nbx_net_inject_currents; the reference's InputCurrent,
 * Event/src/Current/InputCurrent.cpp:13-24) reaches each current synapse of s at step t + delay, sets
 * that synapse's slot to the current (no weight is applied, as in the reference), and the target's
 * I_ext becomes the float sum of its slots in slot order (TimeDrivenLIF_Voltage_jump.cpp:117-125).
 * One step's arrivals are applied before its neuron update, in injection order (a later current for
 * the same synapse wins); a per-step current table (nbx_net_set_current) still wins for the steps it
 * covers. Runs stop before every step with arrivals. Call after nbx_net_finalize; a new call replaces
 * the set and clears slots and pending currents; ResetState clears both. */
int32_t nbx_net_set_current_synapses(nbx_net* net, const int32_t* src, const int32_t* dst, const uint8_t* delay,
                                     int32_t count);
/* Source neuron src[k] emits current[k] at step t (t >= the current step). */
int32_t nbx_net_inject_currents(nbx_net* net, const int32_t* src, const float* current, int32_t count, int32_t t);

/* Back to the state after nbx_net_finalize with the synapses kept (reference:
 * GpuPropagationRuntime::ResetState + DenseUnifiedNeuronRuntime::ResetState): neuron state and
 * LastSpike to their initial values, delay ring, firing history and event count cleared, learning
 * traces (STDP and rule records, eligibilities, cerebellar arrival histories) cleared, the step
 * counter to 0. Weights are kept, as the reference keeps d_syn_weight_. */
int32_t nbx_net_reset_state(nbx_net* net);

/* Pending inputs of the current step as floats, channel-major pending[ch * n + i] =
 * arrivals * 2^-shift (reference: DownloadPendingChannels, GpuPropagationRuntime.h:131-134). */
int32_t nbx_net_download_pending(nbx_net* net, float* pending, int64_t count, int32_t* channel_count,
                                 int32_t* channel_stride);

/* Device float snapshot of the current step's pending arrivals, [ch][n] (the float pending pool of
 * GetDeviceCommonBuffersView, DenseNeuronDeviceViews.h:6-13); the buffer belongs to the network
 * and is overwritten by the next call. */
int32_t nbx_net_pending_device_float(nbx_net* net, float** d_pending);

/* Delay ring contents, int32 fixed point ring[slot][channel][n] (slot = step % slots). */
int32_t nbx_net_ring_shape(const nbx_net* net, int32_t* slots, int32_t* channels);
int32_t nbx_net_get_ring(nbx_net* net, int32_t* ring, int64_t count);
/* Spike list of step t if still held in the firing-history ring
 * (reference: ExportFiringHistoryRing). Returns count in *count_out. */
int32_t nbx_net_get_history(nbx_net* net, int32_t t, int32_t* ids, int32_t cap, int32_t* count_out);

/* Pending-input view for a foreign neuron kernel (reference: GetDeviceCommonBuffersView).
 * d_pending is int32 fixed point, channel-major d_pending[ch * stride + i] for the current step. */
typedef struct nbx_pending_view {
    int32_t* d_pending;
    int32_t channel_count;
    int32_t channel_stride;
    int32_t neuron_count;
    float inv_scale;
} nbx_pending_view;
int32_t nbx_net_pending_view(nbx_net* net, nbx_pending_view* view);

/* Weights and STDP state (reference: Download/UploadSynapticWeights, Get/SetSynapseWeight). */
int32_t nbx_net_get_weights(nbx_net* net, float* w, int64_t count);
/* Replaces every sparse weight (CSR order), between steps. Arrivals of spikes already fired but
 * not yet delivered take the new weights, as the reference reads a synapse's weight at delivery
 * (PropogatedSpikeGroup.cpp:33-71). Rejected (NBX_ERR_INVALID, weights unchanged) when a weight is
 * not finite or when the new weights could overflow the int32 arrival sums at the fixed shift
 * chosen at nbx_net_finalize (sum of |w| into one neuron, plastic synapses at their bound);
 * NBX_ERR_STATE inside an open per-phase step. */
int32_t nbx_net_set_weights(nbx_net* net, const float* w, int64_t count);
/* One synapse by CSR index (reference: GpuPropagationRuntime::Get/SetSynapseWeight,
 * GpuPropagationRuntime.h:139-141), with the checks and delivery rule of nbx_net_set_weights. */
int32_t nbx_net_get_synapse_weight(nbx_net* net, int64_t synapse, float* w);
int32_t nbx_net_set_synapse_weight(nbx_net* net, int64_t synapse, float w);
int32_t nbx_net_get_stdp_state(nbx_net* net, float* apre, float* apost, int32_t* last_update, int64_t count);

/* ---------------- replay delivery (multi-step, injected spikes) ----------------
 * Delivers the spikes of steps [t0, t0 + nsteps) given as CSR lists
 * (ids[tick_ptr[k] .. tick_ptr[k+1]) fired at step t0 + k) through the sparse
 * synapses, and writes per-step arrivals out[k][channel][n] (int32 fixed point)
 * for the steps [t0 + 1, t0 + nsteps] — what a closed-loop run would read from the
 * ring, given only these spikes. Host arrays. */
int32_t nbx_replay_deliver(nbx_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t nsteps,
                           int32_t* out);

/* ---------------- replay STDP (multi-step, recorded raster) ----------------
 * Replaces the per-step STDP::ApplyPreSynaticSpike / ApplyPostSynaticSpike calls
 * (LearningRule/src/Rule/STDP.cpp:19-61) of a replayed spike schedule
 * (tests/model_codegen/pair_stdp_gpu_trajectory.cpp:52-66: per tick the presynaptic side,
 * then the postsynaptic side) for every plastic synapse at once.
 * ids[tick_ptr[k] .. tick_ptr[k+1]) fired at step t0 - lead + k, k < lead + nsteps (a neuron
 * at most once per step). Applies, in each synapse's time order, the pre-side update of every
 * arrival (spike step + delay) at a step in [t0, t0 + nsteps) and the post-side update of every
 * spike of a step in [t0, t0 + nsteps): what nsteps closed-loop steps whose neurons fire
 * exactly these spikes leave in the weights and trace records, bit for bit. The lead steps
 * (lead <= t0; D - 1 of them continue an earlier call exactly) only supply arrivals. Neuron
 * state, delay ring, firing history and the step counter are not touched. *updates (may be
 * NULL) receives the number of plastic updates applied. lead + nsteps <= 65535 (longer
 * rasters: chained calls). The network keeps a workspace of about 12 * n * ceil((lead + nsteps)
 * / 64) + 4 * spikes + 48 * n bytes for later calls. Networks with trigger rules are rejected.
 * Host arrays. */
int32_t nbx_replay_stdp(nbx_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t lead,
                        int32_t nsteps, int64_t* updates);
/* Device-resident variant for timing: device pointers, enqueued on the network's stream;
 * nspikes = tick_ptr[lead + nsteps] sizes the workspace. *d_updates (device, may be NULL) is
 * incremented. An invalid raster changes no synapse and is reported (NBX_ERR_INVALID) by the
 * next nbx_net_sync (or nbx_net_run); the next replay-STDP call starts with the report cleared. */
int32_t nbx_replay_stdp_device(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0, int32_t lead,
                               int32_t nsteps, int64_t nspikes, uint64_t* d_updates);
/* Allocation-free variant of nbx_replay_stdp_device. It has no counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, semantics and results; the call's
 * workspace (per-neuron spike lists, spike bit words, scan storage, decay table) lives in one
 * static device array of NBX_REPLAY_STDP_SCRATCH_BYTES per GPU, held until the process ends, and
 * the call makes no hipMalloc / hipFree or other allocation. A call whose workspace would exceed
 * the array (about 12 * n * ceil((lead + nsteps) / 64) + 4 * nspikes + 48 * n bytes) is rejected
 * with NBX_ERR_INVALID, as is a network whose blocked delivery layouts are built (freeing them
 * would deallocate). Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from any network or thread) race and must be serialised by the caller. */
#define NBX_REPLAY_STDP_SCRATCH_BYTES ((int64_t)32 << 20)
int32_t nbx_replay_stdp_device_noalloc(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0,
                                       int32_t lead, int32_t nsteps, int64_t nspikes, uint64_t* d_updates);
/* ---------------- .nbnet connection tables ----------------
 * The reference's binary network description (python/src/neuronbridge/nbnet.py:34-38,
 * read by src/native/shared/src/streaming_build/NbnetReader.cpp): header, metadata JSON,
 * 32-byte connection records. Validation and error messages follow NbnetReader
 * (magic, major version 1, record size 32, metadata <= 64 MiB, exact file size, optional
 * SHA-256 over metadata + records, neuron ids in range, finite weights). */
typedef struct {
    int64_t file_bytes, metadata_bytes, connection_count;
    int32_t version_major, version_minor, flags;
    int32_t checksum_verified;
} nbx_nbnet_header;
/* Validates a file (and its checksum when verify_checksum != 0) and reports its header. */
int32_t nbx_nbnet_info(const char* path, int32_t verify_checksum, nbx_nbnet_header* info);
/* Copies the metadata JSON into buf (cap bytes); buf = NULL only reports its length. */
int32_t nbx_nbnet_metadata(const char* path, char* buf, int64_t cap, int64_t* len);
/* Replaces NbnetReader::ForEachConnectionBatch feeding the network builder: loads the
 * connection table as the network's sparse synapses (before nbx_net_finalize). Records are
 * streamed to the device and sorted there by (presynaptic, delay, postsynaptic), equal keys in
 * file order. Supported subset: delays 1..63 steps, synapse type 0/1, no trigger rules;
 * synapse_rule >= 0 marks a synapse plastic (STDP via nbx_net_set_stdp, whose w_max applies). */
int32_t nbx_net_load_nbnet(nbx_net* net, const char* path, int32_t verify_checksum);

/* ---------------- monitor export (DebugMonitor directory schema) ----------------
 * Writes meta.json and the eight CSV tables of the reference's DebugMonitorWriter
 * (src/native/gpu/src/debug_monitor/DebugMonitorWriter.cpp:193-235) into output_dir:
 * spikes.csv from the recorded raster (steps that are multiples of sample_interval_steps),
 * and, when requested, one snapshot of neuron state ("v", "state_1", ...), pending arrivals
 * of the next step and weights at export time. neuron_ids = NULL selects every neuron. */
typedef struct {
    const char* output_dir;
    const char* component_name; /* NULL = "main" */
    int32_t sample_interval_steps, flush_interval_steps;
    int32_t record_spikes, record_state, record_weights, record_pending_channels;
    const int32_t* neuron_ids;
    int32_t neuron_count;
} nbx_monitor_desc;
int32_t nbx_net_export_monitor(nbx_net* net, const nbx_monitor_desc* desc);
/* Time-series monitor (reference: SimulationDebugMonitor::Initialize / CaptureStep / Flush,
 * SimulationDebugMonitor.cpp:27-96, and DebugMonitorWriter's tables): writes meta.json and the
 * eight tables' headers into output_dir now; from then on every step t whose update completes
 * (nbx_net_run on any engine, or nbx_net_commit_firings) with t % sample_interval_steps == 0 is
 * sampled as one frame, all of whose rows carry time_step t: spikes of the selected neurons at t
 * (ascending ids), their state after t's update (as nbx_net_export_monitor writes it), the arrivals
 * already pending for step t + 1 (delay ring, float image; arrivals of plastic synapses are pulled at
 * their arrival step and are not pending), and every synapse's weight. Frames are captured on the
 * device and written every flush_interval_steps frames (<= 0: 100), earlier when the device frame
 * buffers (at most 512 MiB) are full, at nbx_net_monitor_flush and at nbx_net_monitor_stop (also
 * called by nbx_net_destroy). A run with the monitor on is cut after each sampled step. */
int32_t nbx_net_monitor_start(nbx_net* net, const nbx_monitor_desc* desc);
int32_t nbx_net_monitor_flush(nbx_net* net);
int32_t nbx_net_monitor_stop(nbx_net* net);
/* Frames written so far and frames captured but not yet written. */
int32_t nbx_net_monitor_frames(const nbx_net* net, int64_t* written, int32_t* held);

/* Which replay kernel nbx_replay_deliver uses for this network (launch-shape overrides
 * in the environment included): NBX_REPLAY_GLOBAL (global atomics; STDP networks and
 * layouts that do not fit), NBX_REPLAY_OFFSETS (post-blocked offset layout, LDS ring),
 * NBX_REPLAY_SLOTS (fixed 128-byte slot per presynaptic neuron and post block, windows of input
 * steps between workgroup barriers) or NBX_REPLAY_STREAM (the same slots, every wave of a
 * workgroup pipelining its own spikes with no barriers; the default where the slots fit). */
enum { NBX_REPLAY_GLOBAL = 0, NBX_REPLAY_OFFSETS = 1, NBX_REPLAY_SLOTS = 2, NBX_REPLAY_STREAM = 3 };
int32_t nbx_replay_path(nbx_net* net, int32_t* path);
/* Device-resident variant for timing: all pointers are device pointers and the
 * call only enqueues the launch on the network's stream. */
int32_t nbx_replay_deliver_device(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t nsteps,
                                  int32_t* d_out);

/* ---------------- host-side network builder ----------------
 * Fixed in-degree random topology: each post neuron j draws k_in distinct presynaptic
 * neurons (Floyd's sampling; Philox4x32-10 with counter (j, m, "TOPO", 0) and key
 * (seed_lo, seed_hi)) and a delay in [d_min, d_max] per synapse. Output is the CSR by
 * presynaptic neuron that nbx_net_set_sparse takes, rows sorted by (delay, post);
 * type = 1 for presynaptic ids >= n_exc. Caller allocates row_ptr[n+1] and
 * post/delay/type[n*k_in]. Host only (no device needed). Replaces the per-source,
 * per-delay connection runs of Network::CaculateOutputConnection (Network.cu:483-560). */
int32_t nbx_build_fixed_indegree(int32_t n, int32_t k_in, int32_t d_min, int32_t d_max, uint32_t seed_lo,
                                 uint32_t seed_hi, int32_t n_exc, int32_t nthreads, int64_t* row_ptr, int32_t* post,
                                 uint8_t* delay, uint8_t* type);

/* Dense N(0, sigma^2) weights rounded to BF16 (round to nearest even), w[pre * n + post]:
 * Philox4x32-10 counter (pre, post, "WGHT", 0) -> Box-Muller in double on the host.
 * Host only. Used for the batched multi-instance workload. */
int32_t nbx_build_dense_normal_bf16(int32_t n, float sigma, uint32_t seed_lo, uint32_t seed_hi, int32_t nthreads,
                                    uint16_t* out);

/* ---------------- batched multi-instance engine (no reference counterpart) ----------------
 * Hundreds of independent LIF voltage-jump networks of uneven size and parameters, each with
 * dense all-to-all BF16 weights (delay 1 step), its own tau, Poisson background and
 * fixed-point shift, advanced together, state in registers, the fired rows of W gathered per
 * step. Two schedules: one workgroup per instance (chunks of steps replayed from a captured
 * hipGraph), or a split schedule in which an instance runs as several column parts that
 * exchange their spike words every step (cooperative launches; the autotuner splits the
 * instances with the most synaptic events). Arrivals are the same int32 fixed-point sums the
 * single-network engines compute from the same (BF16-valued) weights, so each instance's
 * raster equals the per-instance FP32 run under either schedule. */
typedef struct nbx_batch nbx_batch;
typedef struct nbx_batch_instance {
    int32_t n;               /* neurons, 1 .. 16384 */
    float tau;               /* membrane time constant, ms */
    double pois_lambda;      /* Poisson background count per step */
    uint32_t seed_lo, seed_hi;
    int32_t fixed_shift;     /* arrivals in units of 2^-shift */
    const uint16_t* w_bf16;  /* host [n][n] pre-major */
    const float* v0;         /* host initial V or NULL (V_reset) */
} nbx_batch_instance;
typedef struct nbx_batch_desc {
    int32_t count;
    int32_t integrator;      /* NBX_EULER | NBX_RK4 */
    float dt;
    int32_t t_ref;
    float v_rest, v_th, r, v_reset, w_bg;
    const nbx_batch_instance* inst;
} nbx_batch_desc;
int32_t nbx_batch_create(const nbx_batch_desc* desc, int32_t device, nbx_batch** out);
int32_t nbx_batch_destroy(nbx_batch* b);
/* Autotuner: times `steps` steps per candidate workgroup size and the split schedule planned
 * from the measured per-instance events (state restored), keeps the fastest; threads_out is
 * 256 when the split schedule won. */
int32_t nbx_batch_tune(nbx_batch* b, int32_t steps, int32_t* threads_out);
/* One workgroup per instance with `threads` threads (256, 512, 1024); clears a split. */
int32_t nbx_batch_set_schedule(nbx_batch* b, int32_t threads, int32_t chunk);
int32_t nbx_batch_get_schedule(const nbx_batch* b, int32_t* threads, int32_t* chunk);
/* Split schedule: parts[i] >= 1 column parts for instance i (256-thread workgroups of at most
 * 2048 columns, widths rounded up to 32 columns, so an instance may get fewer parts than asked);
 * NBX_ERR_INVALID if an instance needs more parts, NBX_ERR_UNSUPPORTED if all parts cannot be
 * co-resident. get_parts reports the parts in use (0 for every instance when not split). */
int32_t nbx_batch_set_parts(nbx_batch* b, const int32_t* parts, int32_t count);
int32_t nbx_batch_get_parts(const nbx_batch* b, int32_t* parts, int32_t count);
int32_t nbx_batch_record(nbx_batch* b, int32_t t_begin, int32_t nsteps);
int32_t nbx_batch_run(nbx_batch* b, int32_t nsteps);
int32_t nbx_batch_run_async(nbx_batch* b, int32_t nsteps);
int32_t nbx_batch_sync(nbx_batch* b);
int32_t nbx_batch_get_raster(nbx_batch* b, int32_t instance, uint32_t* raster, int64_t nwords);
int32_t nbx_batch_get_state(nbx_batch* b, int32_t instance, float* v, int32_t* last);
int32_t nbx_batch_events(nbx_batch* b, int64_t* events);
void* nbx_batch_stream(nbx_batch* b);

/* ---------------- differentiable recurrent LIF (no reference counterpart) ----------------
 * Elementwise step kernels of a surrogate-gradient adjoint; the caller runs the products with
 * the weight matrices as GEMMs between them. Device pointers, [n][batch] float32 row-major,
 * enqueued on `stream` (a hipStream_t, NULL = default stream).
 *   forward:  Vm = g*V + (1-g)*(V_rest + R*I);  s = Vm > V_th;  V = Vm - delta_reset*s
 *   backward: ds = R_t + G_t - delta_reset*dV;  dVm = dV + sigma'(Vm - V_th)*ds;
 *             dI = (1-g)*R*dVm;  dV <- g*dVm
 * mode 1: Heaviside spike, surrogate sigma'(x) = 1/(1 + slope*|x|)^2 (fast sigmoid)
 * mode 0: Heaviside spike, sigma' = 0 (the exact gradient away from threshold crossings)
 * mode 2: smooth spike s = 1/2 + u/(2(1 + |u|)), u = slope*x, with its exact derivative. */
typedef struct nbx_lif_adj_params {
    int32_t n, batch;
    float g, one_minus_g;  /* RK4 growth factor of the linear LIF ODE over one step, and 1 - g */
    float r, v_rest, v_th, delta_reset, slope;
    int32_t mode;
} nbx_lif_adj_params;
int32_t nbx_lif_fwd_step(const nbx_lif_adj_params* p, const float* d_I, float* d_V, float* d_Vm, float* d_s,
                         void* stream);
int32_t nbx_lif_bwd_step(const nbx_lif_adj_params* p, const float* d_G, const float* d_R, const float* d_Vm,
                         float* d_dV, float* d_dI, void* stream);
/* Fused recurrent steps (the product with W_rec on FP32 matrix cores inside the step kernel), over
 * a segment of nsteps consecutive steps, all [n][batch] device arrays on `stream`:
 *   forward:  for k: s_k = step(V, IN_k + W s_{k-1}) with s_{-1} = s_init (null: zero); S and Vm
 *             (optional) get nsteps slices;
 *   backward: for k = nsteps-1 .. 0: dI_k = adjoint(dV, R_k (optional) + W^T dI_{k+1}, Vm_k) with
 *             dI_nsteps = dI_after (null: no recurrent term), WT = W_rec transposed. */
int32_t nbx_lif_fwd_segment(const nbx_lif_adj_params* p, const float* d_W, const float* d_IN, const float* d_s_init,
                            float* d_V, float* d_S, float* d_Vm, int32_t nsteps, void* stream);
int32_t nbx_lif_bwd_segment(const nbx_lif_adj_params* p, const float* d_WT, const float* d_dI_after, const float* d_R,
                            const float* d_Vm, float* d_dV, float* d_dI, int32_t nsteps, void* stream);
/* The forward segment for binary spikes (modes 0 and 1): W_rec s_{k-1} as the sum of the rows of
 * WT = W_rec^T of the neurons that fired at k-1, in ascending neuron order (deterministic, so a
 * recomputed segment is bit-identical), instead of a dense product. Same results as
 * nbx_lif_fwd_segment up to the rounding of the recurrent sum. d_bits: caller workspace of
 * 2 * batch * ceil(n / 32) uint32 (the spikes of consecutive steps as bit words). Mode 2 is
 * rejected (its spikes are not binary). */
int32_t nbx_lif_fwd_segment_sparse(const nbx_lif_adj_params* p, const float* d_WT, const float* d_IN,
                                   const float* d_s_init, float* d_V, float* d_S, float* d_Vm, uint32_t* d_bits,
                                   int32_t nsteps, void* stream);
/* Leaky readout y_t = beta*y_{t-1} + (1-beta)*Z_t over [T][m] (state carried in d_state if given),
 * and its adjoint dZ_t = (1-beta)*(gy_t + beta*(...)) as a reverse scan. */
int32_t nbx_readout_fwd(int32_t T, int64_t m, float beta, const float* d_Z, float* d_y, float* d_state, void* stream);
int32_t nbx_readout_bwd(int32_t T, int64_t m, float beta, const float* d_gy, float* d_dZ, float* d_carry,
                        void* stream);

/* HIP stream handle (hipStream_t) the network runs on, for event timing by callers. */
void* nbx_net_stream(nbx_net* net);

#ifdef __cplusplus
}
#endif
#endif
