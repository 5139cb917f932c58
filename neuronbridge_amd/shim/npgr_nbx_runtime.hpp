// C++ host mirror of the reference's dense-runtime interface over the nbx C ABI.
//
// A maintainer swaps the CUDA translation units of NeuronBridge::DenseRuntime for this
// header + libnbx.so: the methods keep the reference's names, argument meaning and error
// convention (bool return, optional std::string* reason; GpuPropagationRuntime.cu:19-29),
// and the step sequence of DenseSubnetworkRuntimeGpu::RunStep
// (src/native/gpu/src/dense_subnetwork/DenseSubnetworkRuntimeGpu.cpp:625-663) drives it
// unchanged:
//   BeginStep(t) -> RunPropagationOnly() -> Step(...) -> ApplyPostLearningFromDeviceFirings()
//   -> CommitCurrentDeviceFiringsFromExternal(t)
// One object plays both GpuPropagationRuntime (src/native/gpu/include/gpu_runtime/
// GpuPropagationRuntime.h:79-144) and DenseUnifiedNeuronRuntime (src/native/gpu/include/
// dense_subnetwork/DenseUnifiedNeuronRuntime.h:15-41), because in nbx the ring buffer the
// propagation writes and the neuron kernel reads is one device allocation.
//
// Layout input follows the reference's GpuPropagationLayout (GpuPropagationLayout.h:20-49):
// SynapseArrays (post_neuron, weight, type) + PreDelaySlices (start/count/synapse_ids per
// (pre, delay slot)); it is converted once to nbx's CSR-by-(pre, delay).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "nbx_core.h"

namespace npgr_nbx {

struct RuntimeConfig {
    int steps_to_keep = 0;  // firing-history depth (nbx keeps max(max_delay, 2) committed steps)
    bool enable_learning_buffers = false;
    float dt_ms = 1.0f;
};

struct SynapseArrays {
    std::vector<int> post_neuron;
    std::vector<float> weight;
    std::vector<std::uint8_t> type;           // 0 excitatory, 1 inhibitory
    std::vector<unsigned char> plastic_flags;  // optional, 1 = STDP
};

struct PreDelaySlices {
    int neuron_count = 0;
    int delay_slot_count = 0;  // max delay + 1
    std::vector<int> start;    // [neuron * delay_slot_count + slot]
    std::vector<int> count;
    std::vector<int> synapse_ids;
};

struct GpuPropagationLayout {
    SynapseArrays synapses;
    PreDelaySlices pre_delay_slices;
};

struct FiringStepView {
    int time_step = 0;
    std::vector<int> firing_ids;
};

class DenseRuntime {
public:
    DenseRuntime() = default;
    DenseRuntime(const DenseRuntime&) = delete;
    DenseRuntime& operator=(const DenseRuntime&) = delete;
    ~DenseRuntime() { nbx_net_destroy(net_); }

    // DenseUnifiedNeuronRuntime::Initialize + GpuPropagationRuntime::Initialize in one call.
    bool Initialize(const nbx_model_desc& model, const GpuPropagationLayout& layout, const RuntimeConfig& config,
                    const nbx_stdp_desc* stdp = nullptr, std::string* reason = nullptr) {
        if (net_) return Fail(reason, "already initialized");
        if (nbx_net_create(&model, 0, &net_) != NBX_OK) return Fail(reason);
        const PreDelaySlices& sl = layout.pre_delay_slices;
        const int n = model.n;
        if (sl.neuron_count != n || sl.delay_slot_count < 2 ||
            (int)sl.start.size() != n * sl.delay_slot_count || sl.count.size() != sl.start.size())
            return Fail(reason, "pre_delay_slices do not match the neuron count");
        // CSR by presynaptic neuron, rows ordered by delay slot (slot d = delay d)
        row_ptr_.assign((size_t)n + 1, 0);
        post_.clear();
        delay_.clear();
        type_.clear();
        weight_.clear();
        plastic_.clear();
        syn_of_row_.clear();
        const bool has_plastic = !layout.synapses.plastic_flags.empty();
        for (int p = 0; p < n; ++p) {
            for (int d = 1; d < sl.delay_slot_count; ++d) {
                const int f = p * sl.delay_slot_count + d;
                for (int k = 0; k < sl.count[(size_t)f]; ++k) {
                    const int s = sl.synapse_ids[(size_t)(sl.start[(size_t)f] + k)];
                    post_.push_back(layout.synapses.post_neuron[(size_t)s]);
                    delay_.push_back((std::uint8_t)d);
                    type_.push_back(layout.synapses.type[(size_t)s]);
                    weight_.push_back(layout.synapses.weight[(size_t)s]);
                    if (has_plastic) plastic_.push_back(layout.synapses.plastic_flags[(size_t)s] ? 1 : 0);
                    syn_of_row_.push_back(s);
                }
            }
            row_ptr_[(size_t)p + 1] = (int64_t)post_.size();
        }
        row_of_syn_.assign(layout.synapses.post_neuron.size(), -1);
        for (size_t j = 0; j < syn_of_row_.size(); ++j) row_of_syn_[(size_t)syn_of_row_[j]] = (int64_t)j;
        nbx_sparse_desc sd{};
        sd.count = (int64_t)post_.size();
        sd.max_delay = sl.delay_slot_count - 1;
        sd.row_ptr = row_ptr_.data();
        sd.post = post_.data();
        sd.delay = delay_.data();
        sd.type = type_.data();
        sd.weight = weight_.data();
        sd.plastic = has_plastic ? plastic_.data() : nullptr;
        if (nbx_net_set_sparse(net_, &sd) != NBX_OK) return Fail(reason);
        if (stdp && nbx_net_set_stdp(net_, stdp) != NBX_OK) return Fail(reason);
        if (nbx_net_finalize(net_) != NBX_OK) return Fail(reason);
        (void)config;
        return true;
    }

    bool BeginStep(int time_step, std::string* reason = nullptr) {
        return Check(nbx_net_begin_step(net_, time_step), reason);
    }
    bool RunPropagationOnly(std::string* reason = nullptr) { return Check(nbx_net_run_propagation(net_), reason); }
    // DenseUnifiedNeuronRuntime::Step: neuron update of the current step; firing ids ascending.
    bool Step(std::vector<int>* output_firing_ids, std::string* reason = nullptr) {
        std::vector<int32_t> ids((size_t)Neurons());
        int32_t cnt = 0;
        if (!Check(nbx_net_neuron_step(net_, ids.data(), (int32_t)ids.size(), &cnt), reason)) return false;
        if (output_firing_ids) output_firing_ids->assign(ids.begin(), ids.begin() + cnt);
        return true;
    }
    // QueueInterfaceSpike for the current step (after Step, before the commit).
    bool QueueInterfaceSpikes(const std::vector<int>& ids, std::string* reason = nullptr) {
        return Check(nbx_net_inject_spikes(net_, ids.data(), (int32_t)ids.size()), reason);
    }
    bool ApplyPostLearningFromDeviceFirings(std::string* reason = nullptr) {
        return Check(nbx_net_apply_post_learning(net_), reason);
    }
    bool CommitCurrentDeviceFiringsFromExternal(int time_step, std::string* reason = nullptr) {
        int32_t t = -1;
        nbx_net_time(net_, &t);
        if (t != time_step) return Fail(reason, "commit for a step that is not the current one");
        return Check(nbx_net_commit_firings(net_), reason);
    }
    bool ExportFiringHistoryRing(std::vector<FiringStepView>* history, std::string* reason = nullptr) const {
        int32_t t = 0, slots = 0, ch = 0;
        nbx_net_time(net_, &t);
        if (nbx_net_ring_shape(net_, &slots, &ch) != NBX_OK) return Fail(reason);
        history->clear();
        std::vector<int32_t> ids((size_t)Neurons());
        // after a commit the ring holds steps t-1 down to t-1-(slots-2): slots - 1 steps
        for (int s = t - 1; s >= 0 && s >= t - slots + 1; --s) {
            int32_t cnt = 0;
            if (nbx_net_get_history(net_, s, ids.data(), (int32_t)ids.size(), &cnt) != NBX_OK) return Fail(reason);
            history->push_back({s, std::vector<int>(ids.begin(), ids.begin() + cnt)});
        }
        return true;
    }
    // Weights in the reference's synapse order (the layout's synapse ids).
    bool DownloadSynapticWeights(std::vector<float>* weights, std::string* reason = nullptr) const {
        std::vector<float> w(post_.size());
        if (nbx_net_get_weights(net_, w.data(), (int64_t)w.size()) != NBX_OK) return Fail(reason);
        weights->assign(row_of_syn_.size(), 0.0f);
        for (size_t s = 0; s < row_of_syn_.size(); ++s)
            if (row_of_syn_[s] >= 0) (*weights)[s] = w[(size_t)row_of_syn_[s]];
        return true;
    }
    bool UploadSynapticWeights(const std::vector<float>& weights, std::string* reason = nullptr) {
        if (weights.size() != row_of_syn_.size()) return Fail(reason, "weight count mismatch");
        std::vector<float> w(post_.size());
        for (size_t s = 0; s < row_of_syn_.size(); ++s)
            if (row_of_syn_[s] >= 0) w[(size_t)row_of_syn_[s]] = weights[s];
        return Check(nbx_net_set_weights(net_, w.data(), (int64_t)w.size()), reason);
    }
    bool GetSynapseWeight(int synapse_index, float* weight, std::string* reason = nullptr) const {
        std::vector<float> w;
        if (!DownloadSynapticWeights(&w, reason)) return false;
        if (synapse_index < 0 || (size_t)synapse_index >= w.size()) return Fail(reason, "synapse index out of range");
        *weight = w[(size_t)synapse_index];
        return true;
    }
    // Whole-run stepping for callers that do not need the per-phase hooks.
    bool RunSteps(int steps, std::string* reason = nullptr) {
        return Check(nbx_net_run(net_, steps, NBX_ENGINE_AUTO), reason);
    }
    bool SetState(const std::vector<float>& x, const std::vector<int32_t>& last, std::string* reason = nullptr) {
        return Check(nbx_net_set_state(net_, x.data(), last.data()), reason);
    }
    bool SetCurrent(const std::vector<float>& cur, int t_begin, int steps, std::string* reason = nullptr) {
        return Check(nbx_net_set_current(net_, cur.data(), t_begin, steps), reason);
    }
    nbx_net* handle() const { return net_; }

private:
    int Neurons() const { return (int)row_ptr_.size() - 1; }
    static bool Fail(std::string* reason, const char* msg = nullptr) {
        if (reason) *reason = msg ? msg : nbx_last_error();
        return false;
    }
    static bool Check(int32_t rc, std::string* reason) { return rc == NBX_OK ? true : Fail(reason); }

    nbx_net* net_ = nullptr;
    std::vector<int64_t> row_ptr_;
    std::vector<int32_t> post_;
    std::vector<std::uint8_t> delay_, type_, plastic_;
    std::vector<float> weight_;
    std::vector<int> syn_of_row_;
    std::vector<int64_t> row_of_syn_;
};

}  // namespace npgr_nbx
