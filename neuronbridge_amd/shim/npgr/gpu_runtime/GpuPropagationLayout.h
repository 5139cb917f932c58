// npgr::GpuPropagationLayout: the packed synapse layout the dense build finalizer hands the
// propagation runtime (reference: src/native/gpu/include/gpu_runtime/GpuPropagationLayout.h:12-70,
// with the learning tables of dense_subnetwork/learning/DenseLearningModelSpec.h:24-81 and
// DenseLearningFieldTable.h:14-20). Member names and meaning are the reference's, so the reference's
// builder fills this struct unchanged; libnbx reads the synapse arrays, the (pre, delay) slices and,
// for pair STDP, the learning fields "stdp.*".
#ifndef NPGR_GPU_PROPAGATION_LAYOUT_H
#define NPGR_GPU_PROPAGATION_LAYOUT_H

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace npgr {

enum class PendingChannel : std::uint8_t {
    ExcitatoryConductance = 0,
    InhibitoryConductance = 1,
    Current = 2,
    NmdaConductance = 3,
    Modulatory = 4,
};

enum class DenseFieldStorage : std::uint8_t { Float32 = 0, Int32 = 1, UInt8 = 2 };
enum class DenseFieldRole : std::uint8_t { Parameter = 0, State = 1, Derived = 2, Debug = 3 };

struct DenseFieldSpan {
    int field_id = -1;
    std::string name;
    DenseFieldStorage storage = DenseFieldStorage::Float32;
    DenseFieldRole role = DenseFieldRole::Parameter;
    int offset = 0;
    int count = 0;
};

// field-major pools: pool[span.offset + synapse_id]
struct DenseLearningHostFieldTable {
    std::vector<DenseFieldSpan> fields;
    std::unordered_map<std::string, int> field_id_by_name;
    std::vector<float> float_pool;
    std::vector<int> int_pool;
    std::vector<unsigned char> byte_pool;
};

struct DenseLearningModelSpan {
    int model_id = -1;
    int factory_model_id = -1;
    int field_index_offset = 0;
    int field_index_count = 0;
};

struct DenseLearningModelFieldIndexTable {
    std::vector<DenseLearningModelSpan> spans;
    std::vector<int> field_indices;
};

struct DenseLearningTriggerRouteTable {
    std::vector<int> start;
    std::vector<int> count;
    std::vector<int> synapse_ids;
};

struct DenseLearningSpikeBufferTable {
    std::vector<int> synapse_bucket_id;
    std::vector<int> bucket_start;
    std::vector<int> bucket_capacity;
    std::vector<int> bucket_head;
    std::vector<int> bucket_count;
    std::vector<float> spike_time;
    std::vector<int> spike_synapse_id;
    std::vector<unsigned char> valid;
};

// learning model ids of the reference registry (dense_subnetwork/learning/DenseLearningModelList.inc)
constexpr int kDenseLearningStdpModelId = 0;
constexpr int kDenseLearningRStdpModelId = 1;
constexpr int kDenseLearningAdditiveKernelModelId = 2;
constexpr int kDenseLearningCerebellarModelId = 3;

struct SynapseArrays {
    std::vector<int> post_neuron;
    std::vector<int> post_model;
    std::vector<float> weight;
    std::vector<float> max_weight;
    std::vector<std::uint8_t> type;
    // compiled spike-effect channel per synapse (PendingChannel) and its multiplier
    std::vector<std::uint8_t> effect_channel;
    std::vector<float> effect_scale;
    std::vector<int> plastic_rule_id;
    std::vector<int> plastic_model_id;
    std::vector<unsigned char> plastic_flags;
    std::vector<int> plastic_state_index;
    std::vector<int> trigger_rule_id;
    std::vector<int> trigger_model_id;
    std::vector<unsigned char> trigger_flags;
};

struct PreDelaySlices {
    int neuron_count = 0;
    int delay_slot_count = 0;
    std::vector<int> start;  // [pre_neuron * delay_slot_count + delay_slot]
    std::vector<int> count;
    std::vector<int> synapse_ids;

    int FlatIndex(int pre_neuron, int delay_slot) const { return pre_neuron * delay_slot_count + delay_slot; }
};

struct LayoutStats {
    int neuron_count = 0;
    int synapse_count = 0;
    int delay_slot_count = 0;
    int model_count = 0;
    int pending_channel_count = 0;
};

class GpuPropagationLayout {
public:
    LayoutStats stats;
    SynapseArrays synapses;
    PreDelaySlices pre_delay_slices;
    DenseLearningHostFieldTable learning_fields;
    DenseLearningModelFieldIndexTable learning_model_indices;
    DenseLearningTriggerRouteTable learning_trigger_routes;
    DenseLearningSpikeBufferTable learning_spike_buffers;

    // the structural checks of the reference's layout: sizes agree and ids are in range
    bool IsValid(std::string* reason = nullptr) const {
        auto fail = [&](const char* m) {
            if (reason) *reason = m;
            return false;
        };
        const size_t S = synapses.post_neuron.size();
        if (synapses.weight.size() != S || synapses.type.size() != S) return fail("synapse arrays differ in size");
        const PreDelaySlices& sl = pre_delay_slices;
        if (sl.neuron_count != stats.neuron_count) return fail("pre-delay slices do not match the neuron count");
        if (sl.delay_slot_count < 2) return fail("delay_slot_count must be at least 2");
        const size_t F = (size_t)sl.neuron_count * (size_t)sl.delay_slot_count;
        if (sl.start.size() != F || sl.count.size() != F) return fail("pre-delay slice tables have the wrong size");
        for (int post : synapses.post_neuron)
            if (post < 0 || post >= stats.neuron_count) return fail("synapse post neuron out of range");
        for (size_t f = 0; f < F; ++f)
            if (sl.count[f] < 0 || sl.start[f] < 0 || (size_t)(sl.start[f] + sl.count[f]) > sl.synapse_ids.size())
                return fail("pre-delay slice outside synapse_ids");
        for (int s : sl.synapse_ids)
            if (s < 0 || (size_t)s >= S) return fail("slice synapse id out of range");
        return true;
    }
};

}  // namespace npgr

#endif
