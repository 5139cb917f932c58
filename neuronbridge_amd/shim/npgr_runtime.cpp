// Implementation of npgr::GpuPropagationRuntime and npgr::DenseUnifiedNeuronRuntime (the headers
// under shim/npgr/) over the libnbx C ABI (include/nbx_core.h). Host code only: every device
// operation is a libnbx call. A maintainer compiles this file into NeuronBridge::DenseRuntime in
// place of GpuPropagationRuntime.cu/.cpp and DenseUnifiedNeuronRuntime.cu, with shim/npgr ahead of
// src/native/gpu/include on the include path.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dense_subnetwork/DenseUnifiedNeuronRuntime.h"
#include "gpu_runtime/GpuPropagationRuntime.h"
#include "nbx_core.h"

namespace nbx_shim {

// The network both runtime objects of one dense subnetwork drive.
struct Binding {
    nbx_net* net = nullptr;
    int device = 0;
    // layout as libnbx's CSR by presynaptic neuron (rows ordered by delay slot)
    bool layout_ok = false;
    int n = 0;
    int max_delay = 1;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> post;
    std::vector<uint8_t> delay, type, plastic;
    std::vector<float> weight;
    std::vector<int64_t> csr_of_syn;  // reference synapse id -> CSR position (-1: in no slice)
    bool stdp = false;
    nbx_stdp_desc sd{};
    npgr::RuntimeConfig config;
    std::vector<int> iface_targets;
    std::vector<unsigned char> out_mask;
    // host mirror of the firing history: slot = step % steps_to_keep
    std::vector<npgr::FiringStepView> history;
    std::vector<int32_t> step_ids;  // the current step's firings (after its neuron update)
    int step = -1;                  // step of step_ids
    bool paired = false;

    ~Binding() {
        if (net) nbx_net_destroy(net);
    }
};

namespace {
thread_local std::weak_ptr<Binding> g_unpaired;

bool fail(std::string* reason, const std::string& msg) {
    if (reason) *reason = msg;
    return false;
}
bool check(int32_t rc, std::string* reason) {
    if (rc == NBX_OK) return true;
    return fail(reason, nbx_last_error());
}
bool ready(const std::shared_ptr<Binding>& b, std::string* reason) {
    if (!b || !b->net) return fail(reason, "dense runtime is not uploaded (DenseUnifiedNeuronRuntime::UploadHostToDevice)");
    return true;
}

// Learning field value of synapse s ("stdp.*" field, field-major pool)
bool field_value(const npgr::DenseLearningHostFieldTable& t, const char* name, int s, float* out) {
    auto it = t.field_id_by_name.find(name);
    if (it == t.field_id_by_name.end()) return false;
    for (const npgr::DenseFieldSpan& f : t.fields) {
        if (f.field_id != it->second) continue;
        if (f.storage != npgr::DenseFieldStorage::Float32 || s >= f.count) return false;
        *out = t.float_pool[(size_t)(f.offset + s)];
        return true;
    }
    return false;
}

// ---- model specs -> nbx populations
int model_of_spec(const npgr::DenseNeuronModelSpec& sp) {
    using F = npgr::DenseNeuronModelFactory;
    const std::string& nm = sp.legacy_model_name;
    auto is = [&](const char* a) { return nm == a || nm == std::string(a) + "_GPU"; };
    if (sp.factory_model_id == F::kLifVoltageJumpModelId || is("TimeDrivenLIF_Voltage_jump")) return NBX_MODEL_LIF_VJ;
    if (sp.factory_model_id == F::kLifExponentialDecayModelId || is("TimeDrivenLIF_Exponential_Decay"))
        return NBX_MODEL_LIF_EXP;
    if (sp.factory_model_id == F::kLifExponentialDoubleModelId || is("TimeDrivenLIF_Exponential_double"))
        return NBX_MODEL_LIF_EXP2;
    if (sp.factory_model_id == F::kLifExponentialTripleModelId || is("TimeDrivenLIF_Exponential_triple"))
        return NBX_MODEL_LIF_EXP3;
    if (sp.factory_model_id == F::kIzhikevichExponentialDecayModelId || is("TimeDrivenIzhikevic_Exponential_Decay") ||
        is("TimeDrivenIzhikevic_I_Exponential_Decay"))
        return NBX_MODEL_IZH;
    if (sp.factory_model_id == F::kPoissonRateModelId || is("PoissonRate")) return NBX_MODEL_POISSON;
    if (sp.factory_model_id == F::kTriggerRelayModelId || nm == "TriggerRelayNeuronModel") return NBX_MODEL_TRIGGER_RELAY;
    return -1;
}

// value of parameter `names` (first alias present) for local neuron k of the spec, else dflt
float param(const npgr::DenseNeuronModelSpec& sp, std::initializer_list<const char*> names, int k, float dflt) {
    for (const char* nm : names) {
        auto it = sp.per_neuron_params.find(nm);
        if (it == sp.per_neuron_params.end() || it->second.empty()) continue;
        const std::vector<float>& v = it->second;
        return v.size() == 1 ? v[0] : v[(size_t)std::min<int>(k, (int)v.size() - 1)];
    }
    return dflt;
}

// The model description of local neuron k (Izhikevich a, b, c, d excluded: they stay per neuron);
// the dense models' parameter names and defaults (dense_subnetwork/model/Dense*Model.h)
nbx_model_desc desc_at(const npgr::DenseNeuronModelSpec& sp, int model, int k, float dt) {
    nbx_model_desc m{};
    m.model = model;
    m.integrator = NBX_EULER;
    m.dt = dt;
    m.r = param(sp, {"R", "r"}, k, 1.0f);
    m.tau = param(sp, {"tau", "tau_m_ms"}, k, 20.0f);
    m.t_ref = (int32_t)param(sp, {"t_ref"}, k, 50.0f);
    switch (model) {
        case NBX_MODEL_LIF_VJ:
        case NBX_MODEL_LIF_EXP:
            m.v_rest = param(sp, {"V_rest", "v_rest"}, k, -65.0f);
            m.v_reset = param(sp, {"V_reset", "v_reset"}, k, -65.0f);
            m.v_th = param(sp, {"V_th", "v_threshold"}, k, model == NBX_MODEL_LIF_VJ ? -50.0f : -20.0f);
            m.g_tau = param(sp, {"g_tau"}, k, 12.0f);
            m.e_rev = param(sp, {"E"}, k, 0.0f);
            break;
        case NBX_MODEL_LIF_EXP2:
            m.v_rest = param(sp, {"V_rest", "v_rest"}, k, -65.0f);
            m.v_reset = param(sp, {"V_reset", "v_reset"}, k, -65.0f);
            m.v_th = param(sp, {"V_th", "v_threshold"}, k, -50.0f);
            m.e_ampa = param(sp, {"Eexc", "e_exc"}, k, 0.0f);
            m.e_gaba = param(sp, {"Einh", "Einhibitory", "e_inh"}, k, -80.0f);
            m.ampa_tau = param(sp, {"gexc_tau", "tau_exc_ms"}, k, 5.0f);
            m.gaba_tau = param(sp, {"ginh_tau", "tau_inh_ms"}, k, 10.0f);
            break;
        case NBX_MODEL_LIF_EXP3:
            m.v_rest = param(sp, {"V_rest", "v_rest"}, k, -60.0f);
            m.v_reset = param(sp, {"V_reset", "v_reset"}, k, -60.0f);
            m.v_th = param(sp, {"V_th", "v_threshold"}, k, -50.0f);
            m.e_ampa = param(sp, {"E_ampa", "Eexc", "e_exc"}, k, 0.0f);
            m.e_gaba = param(sp, {"E_gaba", "Einh", "e_inh"}, k, -80.0f);
            m.ampa_tau = param(sp, {"ampa_tau", "gexc_tau", "tau_exc_ms"}, k, 10.0f);
            m.gaba_tau = param(sp, {"gaba_tau", "ginh_tau", "tau_inh_ms"}, k, 10.0f);
            m.nmda_tau = param(sp, {"nmda_tau"}, k, 10.0f);
            break;
        case NBX_MODEL_IZH:
            m.v_rest = param(sp, {"V_rest", "v_rest"}, k, -65.0f);
            m.izh_v_th = param(sp, {"V_th", "v_threshold"}, k, -50.0f);
            m.izh_r = param(sp, {"R", "r"}, k, 1.0f);
            m.e_ampa = param(sp, {"E_ampa", "Eexc", "e_exc"}, k, 0.0f);
            m.e_gaba = param(sp, {"E_gaba", "Einh", "e_inh"}, k, -80.0f);
            m.ampa_tau = param(sp, {"ampa_tau", "tau_exc_ms"}, k, 5.0f);
            m.gaba_tau = param(sp, {"gaba_tau", "tau_inh_ms"}, k, 10.0f);
            m.nmda_tau = param(sp, {"nmda_tau"}, k, 100.0f);
            break;
        case NBX_MODEL_POISSON:
            m.rate_bias_hz = param(sp, {"poisson_rate_bias_hz", "rate_bias_hz"}, k, 0.0f);
            m.rate_gain_hz = param(sp, {"rate_gain_hz_per_current", "rate_gain_hz"}, k, 1.0f);
            break;
        default:  // trigger relay
            m.v_th = param(sp, {"threshold", "trigger_threshold"}, k, 0.0f);
    }
    return m;
}

bool same_desc(const nbx_model_desc& a, const nbx_model_desc& b) {
    return a.model == b.model && a.t_ref == b.t_ref && a.v_rest == b.v_rest && a.tau == b.tau && a.v_th == b.v_th &&
           a.r == b.r && a.v_reset == b.v_reset && a.g_tau == b.g_tau && a.e_rev == b.e_rev &&
           a.izh_v_th == b.izh_v_th && a.izh_r == b.izh_r && a.e_ampa == b.e_ampa && a.e_gaba == b.e_gaba &&
           a.ampa_tau == b.ampa_tau && a.gaba_tau == b.gaba_tau && a.nmda_tau == b.nmda_tau &&
           a.rate_bias_hz == b.rate_bias_hz && a.rate_gain_hz == b.rate_gain_hz;
}

}  // namespace
}  // namespace nbx_shim

namespace npgr {

using nbx_shim::Binding;
using nbx_shim::check;
using nbx_shim::fail;
using nbx_shim::ready;

// ------------------------------------------------------------------ GpuPropagationRuntime

GpuPropagationRuntime::GpuPropagationRuntime() : b_(std::make_shared<Binding>()) {}
GpuPropagationRuntime::~GpuPropagationRuntime() = default;

bool GpuPropagationRuntime::Initialize(const GpuPropagationLayout& layout, const RuntimeConfig& config,
                                       std::string* reason) {
    std::string why;
    if (!layout.IsValid(&why)) return fail(reason, why);
    if (layout.stats.delay_slot_count <= 0) return fail(reason, "layout delay_slot_count must be positive");
    if (config.steps_to_keep < layout.stats.delay_slot_count + 1)  // GpuPropagationRuntime.cpp:117-121
        return fail(reason, "steps_to_keep must be at least layout delay_slot_count + 1");
    if (b_->net) b_ = std::make_shared<Binding>();
    Binding& B = *b_;
    const SynapseArrays& sa = layout.synapses;
    const PreDelaySlices& sl = layout.pre_delay_slices;
    const int n = layout.stats.neuron_count, D = sl.delay_slot_count;
    const size_t S = sa.post_neuron.size();
    B.n = n;
    B.max_delay = std::max(1, D - 1);
    B.row_ptr.assign((size_t)n + 1, 0);
    B.post.clear();
    B.delay.clear();
    B.type.clear();
    B.plastic.clear();
    B.weight.clear();
    B.csr_of_syn.assign(S, -1);
    B.stdp = false;
    bool have_params = false;
    for (int p = 0; p < n; ++p) {
        for (int d = 1; d < D; ++d) {  // slot d holds the synapses of delay d
            const int f = sl.FlatIndex(p, d);
            for (int k = 0; k < sl.count[(size_t)f]; ++k) {
                const int s = sl.synapse_ids[(size_t)(sl.start[(size_t)f] + k)];
                // the spike-effect channel selects the arrival type (PendingChannel -> synapse type)
                int ty = sa.type[(size_t)s];
                if (!sa.effect_channel.empty()) {
                    switch ((PendingChannel)sa.effect_channel[(size_t)s]) {
                        case PendingChannel::ExcitatoryConductance: ty = 0; break;
                        case PendingChannel::InhibitoryConductance: ty = 1; break;
                        case PendingChannel::NmdaConductance: ty = 2; break;
                        default:
                            return fail(reason, "synapse effect channel " + std::to_string(sa.effect_channel[(size_t)s]) +
                                                    " (current / modulatory synapses) is driven through the interface "
                                                    "channels here");
                    }
                }
                float w = sa.weight[(size_t)s];
                if (!sa.effect_scale.empty() && sa.effect_scale[(size_t)s] != 1.0f)
                    return fail(reason, "synapse effect scales other than 1 are not supported");
                bool pl = false;
                if (!sa.plastic_rule_id.empty() && sa.plastic_rule_id[(size_t)s] >= 0) {
                    const int pm = sa.plastic_model_id.empty() ? kDenseLearningStdpModelId : sa.plastic_model_id[(size_t)s];
                    if (pm != kDenseLearningStdpModelId)
                        return fail(reason, "only pair-STDP plastic synapses are supported by this boundary");
                    nbx_stdp_desc q{};
                    float inv_ltp = 0, inv_ltd = 0;
                    const DenseLearningHostFieldTable& lf = layout.learning_fields;
                    if (!nbx_shim::field_value(lf, "stdp.max_ltp", s, &q.max_ltp) ||
                        !nbx_shim::field_value(lf, "stdp.max_ltd", s, &q.max_ltd) ||
                        !nbx_shim::field_value(lf, "stdp.inv_ltp_tau", s, &inv_ltp) ||
                        !nbx_shim::field_value(lf, "stdp.inv_ltd_tau", s, &inv_ltd) ||
                        !nbx_shim::field_value(lf, "stdp.max_weight", s, &q.w_max) || inv_ltp <= 0 || inv_ltd <= 0)
                        return fail(reason, "pair-STDP synapse without its stdp.* learning fields");
                    q.ltp_tau = 1.0f / inv_ltp;
                    q.ltd_tau = 1.0f / inv_ltd;
                    if (have_params && (q.max_ltp != B.sd.max_ltp || q.max_ltd != B.sd.max_ltd ||
                                        q.ltp_tau != B.sd.ltp_tau || q.ltd_tau != B.sd.ltd_tau || q.w_max != B.sd.w_max))
                        return fail(reason, "pair-STDP parameters must be the same for every plastic synapse");
                    B.sd = q;
                    have_params = true;
                    B.stdp = true;
                    pl = true;
                }
                B.csr_of_syn[(size_t)s] = (int64_t)B.post.size();
                B.post.push_back(sa.post_neuron[(size_t)s]);
                B.delay.push_back((uint8_t)d);
                B.type.push_back((uint8_t)ty);
                B.weight.push_back(w);
                B.plastic.push_back(pl ? 1 : 0);
            }
        }
        B.row_ptr[(size_t)p + 1] = (int64_t)B.post.size();
    }
    B.config = config;
    B.history.assign((size_t)config.steps_to_keep, FiringStepView{-1, {}});
    B.layout_ok = true;
    config_ = config;
    profiling_ = ProfilingStats{};
    pending_.pending_channels.assign((size_t)std::max(3, layout.stats.pending_channel_count) * n, 0.0f);
    learning_.syn_pending_dwt.assign(S, 0.0f);
    nbx_shim::g_unpaired = b_;  // the neuron runtime initialized next on this thread pairs with it
    return true;
}

bool GpuPropagationRuntime::IsInitialized() const { return b_ && b_->layout_ok; }

bool GpuPropagationRuntime::ResetState(std::string* reason) {
    if (!ready(b_, reason)) return false;
    for (FiringStepView& v : b_->history) v = FiringStepView{-1, {}};
    b_->step = -1;
    return check(nbx_net_reset_state(b_->net), reason);
}

bool GpuPropagationRuntime::BeginStep(int time_step, std::string* reason) {
    if (!ready(b_, reason)) return false;
    return check(nbx_net_begin_step(b_->net, time_step), reason);
}

bool GpuPropagationRuntime::RunPropagationOnly(std::string* reason) {
    if (!ready(b_, reason)) return false;
    return check(nbx_net_run_propagation(b_->net), reason);
}

bool GpuPropagationRuntime::CommitCurrentDeviceFiringsFromExternal(int time_step, const int* d_external_firing_ids,
                                                                   const int* d_external_firing_count,
                                                                   std::string* reason) {
    if (!ready(b_, reason)) return false;
    Binding& B = *b_;
    if (!check(nbx_net_commit_external(B.net, time_step, d_external_firing_ids, d_external_firing_count), reason))
        return false;
    // the committed list, kept for ExportFiringHistoryRing (slot = step % steps_to_keep)
    std::vector<int32_t> ids((size_t)B.n);
    int32_t cnt = 0;
    if (!check(nbx_net_get_history(B.net, time_step, ids.data(), B.n, &cnt), reason)) return false;
    FiringStepView& v = B.history[(size_t)(time_step % B.config.steps_to_keep)];
    v.time_step = time_step;
    v.firing_ids.assign(ids.begin(), ids.begin() + cnt);
    return true;
}

bool GpuPropagationRuntime::ApplyPostLearningFromDeviceFirings(const unsigned char*, std::string* reason) {
    if (!ready(b_, reason)) return false;
    return check(nbx_net_apply_post_learning(b_->net), reason);
}

bool GpuPropagationRuntime::SetInterfaceTargetNeurons(const std::vector<int>& ids, std::string* reason) {
    b_->iface_targets = ids;
    if (!b_->net) return true;  // applied when the network is uploaded
    return check(nbx_net_set_interface_targets(b_->net, ids.data(), (int32_t)ids.size()), reason);
}

bool GpuPropagationRuntime::UploadInterfaceChannelValues(const std::vector<InterfaceChannelValue>& values,
                                                         std::string* reason) {
    if (!ready(b_, reason)) return false;
    std::vector<nbx_interface_value> v(values.size());
    for (size_t k = 0; k < values.size(); ++k)
        v[k] = nbx_interface_value{values[k].interface_slot, values[k].target_neuron, values[k].pending_channel,
                                   values[k].overwrite, values[k].value, values[k].scale};
    return check(nbx_net_upload_interface_values(b_->net, v.data(), (int32_t)v.size()), reason);
}

bool GpuPropagationRuntime::UploadInterfaceChannelValuesDevice(const float* device_values,
                                                               const std::vector<DeviceToDeviceInterfaceCopySpec>& specs,
                                                               std::string* reason) {
    if (!ready(b_, reason)) return false;
    std::vector<nbx_interface_copy> c(specs.size());
    for (size_t k = 0; k < specs.size(); ++k)
        c[k] = nbx_interface_copy{specs[k].interface_slot, specs[k].target_neuron, specs[k].pending_channel,
                                  (uint8_t)(specs[k].overwrite ? 1 : 0), specs[k].source_index, specs[k].scale};
    return check(nbx_net_upload_interface_values_device(b_->net, device_values, c.data(), (int32_t)c.size()), reason);
}

bool GpuPropagationRuntime::ApplyInterfaceStateVectors(std::string* reason) {
    if (!ready(b_, reason)) return false;
    return check(nbx_net_apply_interface(b_->net), reason);
}

const RuntimeConfig& GpuPropagationRuntime::config() const { return config_; }
const ProfilingStats& GpuPropagationRuntime::profiling() const { return profiling_; }
const PendingBuffersHostView& GpuPropagationRuntime::pending() const {
    std::vector<float> p;
    int cc = 0, cs = 0;
    if (DownloadPendingChannels(&p, &cc, &cs, nullptr)) pending_.pending_channels.swap(p);
    return pending_;
}
const LearningBuffersHostView& GpuPropagationRuntime::learning() const { return learning_; }
bool GpuPropagationRuntime::IsCudaEnabledBuild() const { return true; }  // the HIP build
bool GpuPropagationRuntime::IsDeviceReady() const { return b_ && b_->layout_ok && nbx_device_count() > 0; }
bool GpuPropagationRuntime::IsCarlsimLikeGpuPathActive() const { return IsDeviceReady(); }

bool GpuPropagationRuntime::ExportFiringHistoryRing(std::vector<FiringStepView>* history, std::string* reason) const {
    if (!history) return fail(reason, "history output must not be null");
    if (!b_->layout_ok) return fail(reason, "propagation runtime is not initialized");
    *history = b_->history;
    return true;
}

bool GpuPropagationRuntime::DownloadPendingChannels(std::vector<float>* pending_channels, int* channel_count,
                                                    int* channel_stride, std::string* reason) const {
    if (!pending_channels) return fail(reason, "pending channel output must not be null");
    if (!ready(b_, reason)) return false;
    int32_t slots = 0, ch = 0;
    if (!check(nbx_net_ring_shape(b_->net, &slots, &ch), reason)) return false;
    const int n = b_->n;
    std::vector<float> own((size_t)ch * n);
    if (!check(nbx_net_download_pending(b_->net, own.data(), (int64_t)own.size(), nullptr, nullptr), reason))
        return false;
    // the reference's channel planes: excitatory, inhibitory, current, NMDA, modulatory; the
    // network's planes are arrivals by type (one-model networks of a single signed channel keep
    // the signed sum in the excitatory plane); the current plane reads 0
    const int count = 5;
    pending_channels->assign((size_t)count * n, 0.0f);
    const int to_ref[3] = {0, 1, 3};
    for (int c = 0; c < ch && c < 3; ++c)
        std::copy(own.begin() + (size_t)c * n, own.begin() + (size_t)(c + 1) * n,
                  pending_channels->begin() + (size_t)to_ref[c] * n);
    if (channel_count) *channel_count = count;
    if (channel_stride) *channel_stride = n;
    return true;
}

bool GpuPropagationRuntime::DownloadSynapticWeights(std::vector<float>* weights, std::string* reason) const {
    if (!weights) return fail(reason, "weight output vector must not be null");
    if (!ready(b_, reason)) return false;
    std::vector<float> w(b_->post.size());
    if (!check(nbx_net_get_weights(b_->net, w.data(), (int64_t)w.size()), reason)) return false;
    weights->assign(b_->csr_of_syn.size(), 0.0f);
    for (size_t s = 0; s < b_->csr_of_syn.size(); ++s)
        if (b_->csr_of_syn[s] >= 0) (*weights)[s] = w[(size_t)b_->csr_of_syn[s]];
    return true;
}

bool GpuPropagationRuntime::UploadSynapticWeights(const std::vector<float>& weights, std::string* reason) {
    if (!ready(b_, reason)) return false;
    if (weights.size() != b_->csr_of_syn.size()) return fail(reason, "weight vector size must equal the synapse count");
    std::vector<float> w(b_->post.size());
    if (!check(nbx_net_get_weights(b_->net, w.data(), (int64_t)w.size()), reason)) return false;
    for (size_t s = 0; s < weights.size(); ++s)
        if (b_->csr_of_syn[s] >= 0) w[(size_t)b_->csr_of_syn[s]] = weights[s];
    return check(nbx_net_set_weights(b_->net, w.data(), (int64_t)w.size()), reason);
}

bool GpuPropagationRuntime::GetSynapseWeight(int synapse_index, float* weight, std::string* reason) const {
    if (!weight) return fail(reason, "weight output must not be null");
    if (!ready(b_, reason)) return false;
    if (synapse_index < 0 || (size_t)synapse_index >= b_->csr_of_syn.size() || b_->csr_of_syn[(size_t)synapse_index] < 0)
        return fail(reason, "synapse index out of range");
    return check(nbx_net_get_synapse_weight(b_->net, b_->csr_of_syn[(size_t)synapse_index], weight), reason);
}

bool GpuPropagationRuntime::SetSynapseWeight(int synapse_index, float weight, std::string* reason) {
    if (!ready(b_, reason)) return false;
    if (synapse_index < 0 || (size_t)synapse_index >= b_->csr_of_syn.size() || b_->csr_of_syn[(size_t)synapse_index] < 0)
        return fail(reason, "synapse index out of range");
    return check(nbx_net_set_synapse_weight(b_->net, b_->csr_of_syn[(size_t)synapse_index], weight), reason);
}

nbx_net* GpuPropagationRuntime::nbx_handle() const { return b_ ? b_->net : nullptr; }

DeviceCommonBuffersView GpuPropagationRuntime::GetDeviceCommonBuffersView() const {
    DeviceCommonBuffersView v;
    if (!b_->net) return v;
    float* d = nullptr;
    int32_t slots = 0, ch = 0;
    if (nbx_net_pending_device_float(b_->net, &d) != NBX_OK || nbx_net_ring_shape(b_->net, &slots, &ch) != NBX_OK)
        return v;
    v.d_pending_channels = d;
    v.pending_channel_count = ch;
    v.pending_channel_stride = b_->n;
    v.neuron_count = b_->n;
    return v;
}

// ------------------------------------------------------------------ DenseUnifiedNeuronRuntime

DenseUnifiedNeuronRuntime::DenseUnifiedNeuronRuntime() = default;
DenseUnifiedNeuronRuntime::~DenseUnifiedNeuronRuntime() = default;

bool DenseUnifiedNeuronRuntime::Initialize(const std::vector<DenseNeuronModelSpec>& specs, int neuron_count, float dt_ms,
                                           std::string* reason) {
    if (neuron_count <= 0) return fail(reason, "neuron_count must be positive");
    if (!(dt_ms > 0.0f)) return fail(reason, "dt_ms must be positive");
    std::vector<DenseNeuronModelSpec> sorted = specs;
    std::sort(sorted.begin(), sorted.end(),
              [](const DenseNeuronModelSpec& a, const DenseNeuronModelSpec& b) { return a.range.begin < b.range.begin; });
    int next = 0;
    for (const DenseNeuronModelSpec& sp : sorted) {
        if (sp.range.begin != next || sp.range.count <= 0)
            return fail(reason, "model ranges must cover the neurons contiguously, each non-empty");
        if (nbx_shim::model_of_spec(sp) < 0) return fail(reason, "unsupported dense neuron model '" + sp.legacy_model_name + "'");
        next += sp.range.count;
    }
    if (next != neuron_count) return fail(reason, "model ranges do not cover neuron_count neurons");
    specs_ = sorted;
    neuron_count_ = neuron_count;
    dt_ms_ = dt_ms;
    std::shared_ptr<Binding> b = nbx_shim::g_unpaired.lock();
    if (b && !b->paired && b->n == neuron_count) {
        b->paired = true;
        nbx_shim::g_unpaired.reset();
        b_ = b;
    } else {
        b_ = std::make_shared<Binding>();  // no propagation runtime: a network without synapses
        b_->n = neuron_count;
        b_->row_ptr.assign((size_t)neuron_count + 1, 0);
        b_->config.steps_to_keep = 3;
        b_->history.assign(3, FiringStepView{-1, {}});
        b_->layout_ok = true;
    }
    return true;
}

bool DenseUnifiedNeuronRuntime::AllocateDeviceBuffers(std::string* reason) {
    if (!b_) return fail(reason, "dense unified neuron runtime is not initialized");
    return true;  // the network's buffers are allocated with it at UploadHostToDevice
}

bool DenseUnifiedNeuronRuntime::UploadHostToDevice(std::string* reason) {
    if (!b_) return fail(reason, "dense unified neuron runtime is not initialized");
    Binding& B = *b_;
    if (B.net) {
        nbx_net_destroy(B.net);
        B.net = nullptr;
    }
    // populations: runs of neurons whose model parameters are equal (Izhikevich a, b, c, d per neuron)
    std::vector<nbx_model_desc> pops;
    std::vector<std::vector<float>> keep;
    for (const DenseNeuronModelSpec& sp : specs_) {
        const int model = nbx_shim::model_of_spec(sp);
        int k0 = 0;
        while (k0 < sp.range.count) {
            const nbx_model_desc d0 = nbx_shim::desc_at(sp, model, k0, dt_ms_);
            int k1 = k0 + 1;
            while (k1 < sp.range.count && nbx_shim::same_desc(nbx_shim::desc_at(sp, model, k1, dt_ms_), d0)) ++k1;
            nbx_model_desc d = d0;
            d.n = k1 - k0;
            if (model == NBX_MODEL_IZH) {
                keep.emplace_back();
                std::vector<float>& abcd = keep.back();
                abcd.resize((size_t)4 * d.n);
                for (int k = k0; k < k1; ++k) {
                    abcd[(size_t)(k - k0)] = nbx_shim::param(sp, {"a"}, k, 0.02f);
                    abcd[(size_t)d.n + (k - k0)] = nbx_shim::param(sp, {"b"}, k, 0.2f);
                    abcd[(size_t)2 * d.n + (k - k0)] = nbx_shim::param(sp, {"c", "V_reset", "v_reset"}, k, -65.0f);
                    abcd[(size_t)3 * d.n + (k - k0)] = nbx_shim::param(sp, {"d"}, k, 8.0f);
                }
            }
            pops.push_back(d);
            k0 = k1;
        }
    }
    if (pops.size() > 64) return fail(reason, "more than 64 distinct parameter runs across the model ranges");
    size_t ki = 0;
    for (nbx_model_desc& d : pops)
        if (d.model == NBX_MODEL_IZH) {
            std::vector<float>& abcd = keep[ki++];
            d.izh_a = abcd.data();
            d.izh_b = abcd.data() + d.n;
            d.izh_c = abcd.data() + 2 * d.n;
            d.izh_d = abcd.data() + 3 * d.n;
        }
    if (!check(nbx_net_create_populations(pops.data(), (int32_t)pops.size(), B.device, &B.net), reason)) return false;
    if (!B.post.empty()) {
        nbx_sparse_desc sd{};
        sd.count = (int64_t)B.post.size();
        sd.max_delay = B.max_delay;
        sd.row_ptr = B.row_ptr.data();
        sd.post = B.post.data();
        sd.delay = B.delay.data();
        sd.type = B.type.data();
        sd.weight = B.weight.data();
        sd.plastic = B.stdp ? B.plastic.data() : nullptr;
        if (!check(nbx_net_set_sparse(B.net, &sd), reason)) return false;
        if (B.stdp && !check(nbx_net_set_stdp(B.net, &B.sd), reason)) return false;
    }
    if (!check(nbx_net_finalize(B.net), reason)) return false;
    if (!B.iface_targets.empty() &&
        !check(nbx_net_set_interface_targets(B.net, B.iface_targets.data(), (int32_t)B.iface_targets.size()), reason))
        return false;
    return check(nbx_net_set_output_mask(B.net, output_mask_.empty() ? nullptr : output_mask_.data()), reason);
}

bool DenseUnifiedNeuronRuntime::ConfigureOutputNeuronMask(const std::vector<unsigned char>& mask, std::string* reason) {
    if (!mask.empty() && (int)mask.size() != neuron_count_)
        return fail(reason, "output neuron mask size must equal the neuron count");
    output_mask_ = mask;
    if (b_ && b_->net)
        return check(nbx_net_set_output_mask(b_->net, output_mask_.empty() ? nullptr : output_mask_.data()), reason);
    return true;
}

bool DenseUnifiedNeuronRuntime::Step(const DeviceCommonBuffersView&, int current_time_step,
                                     std::vector<int>* output_firing_ids, std::vector<int>* full_firing_ids,
                                     bool export_full_firings_to_host, std::string* reason) {
    if (!ready(b_, reason)) return false;
    int32_t t = -1;
    nbx_net_time(b_->net, &t);
    if (t != current_time_step) return fail(reason, "Step time does not match the propagation step");
    std::vector<int32_t> out((size_t)neuron_count_), full((size_t)neuron_count_);
    int32_t oc = 0, fc = 0;
    if (!check(nbx_net_neuron_step_outputs(b_->net, out.data(), neuron_count_, &oc, full.data(), neuron_count_, &fc),
               reason))
        return false;
    if (output_firing_ids) output_firing_ids->assign(out.begin(), out.begin() + oc);
    if (full_firing_ids) {
        if (export_full_firings_to_host) full_firing_ids->assign(full.begin(), full.begin() + fc);
        else full_firing_ids->clear();
    }
    return true;
}

bool DenseUnifiedNeuronRuntime::ResetState(std::string* reason) {
    if (!ready(b_, reason)) return false;
    return check(nbx_net_reset_state(b_->net), reason);
}

bool DenseUnifiedNeuronRuntime::SyncHostStateFromDevice(std::string* reason) const { return ready(b_, reason); }

bool DenseUnifiedNeuronRuntime::BuildDebugSnapshots(std::vector<DenseNeuronDebugSnapshot>* out,
                                                    std::string* reason) const {
    if (!out) return fail(reason, "snapshot output must not be null");
    if (!ready(b_, reason)) return false;
    const int ns = nbx_net_nstate(b_->net);
    std::vector<float> x((size_t)neuron_count_ * ns);
    std::vector<int32_t> last((size_t)neuron_count_);
    if (!check(nbx_net_get_state(b_->net, x.data(), last.data()), reason)) return false;
    out->clear();
    for (const DenseNeuronModelSpec& sp : specs_) {
        DenseNeuronDebugSnapshot snap;
        snap.factory_model_id = sp.factory_model_id;
        snap.model_id = sp.model_id;
        snap.legacy_model_name = sp.legacy_model_name;
        std::vector<float> v((size_t)sp.range.count), steps((size_t)sp.range.count);
        for (int k = 0; k < sp.range.count; ++k) {
            v[(size_t)k] = x[(size_t)(sp.range.begin + k) * ns];
            steps[(size_t)k] = (float)last[(size_t)(sp.range.begin + k)];
        }
        snap.float_state_vectors["v"] = v;
        snap.float_state_vectors["steps_since_last_spike"] = steps;
        out->push_back(snap);
    }
    return true;
}

const int* DenseUnifiedNeuronRuntime::device_full_firing_ids() const {
    int32_t* d = nullptr;
    if (!b_ || !b_->net || nbx_net_firing_device_view(b_->net, &d, nullptr, nullptr) != NBX_OK) return nullptr;
    return d;
}

const int* DenseUnifiedNeuronRuntime::device_full_firing_count() const {
    int32_t* d = nullptr;
    if (!b_ || !b_->net || nbx_net_firing_device_view(b_->net, nullptr, &d, nullptr) != NBX_OK) return nullptr;
    return d;
}

const unsigned char* DenseUnifiedNeuronRuntime::device_current_did_fire() const {
    uint8_t* d = nullptr;
    if (!b_ || !b_->net || nbx_net_firing_device_view(b_->net, nullptr, nullptr, &d) != NBX_OK) return nullptr;
    return d;
}

}  // namespace npgr
