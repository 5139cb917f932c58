// npgr::GpuPropagationRuntime over libnbx: the reference class's public interface, method for
// method (reference: src/native/gpu/include/gpu_runtime/GpuPropagationRuntime.h:21-144; same
// names, arguments, bool + std::string* reason convention, and the structs it declares).
// Together with npgr::DenseUnifiedNeuronRuntime it replaces the CUDA translation units of
// NeuronBridge::DenseRuntime: DenseSubnetworkRuntimeGpu, which owns one object of each, drives
// them unchanged. The two objects share one nbx network: Initialize of the neuron runtime pairs it
// with the propagation runtime most recently initialized on the same thread (the order of
// DenseSubnetworkRuntimeGpu::Initialize), and its UploadHostToDevice builds the network.
#ifndef NPGR_GPU_PROPAGATION_RUNTIME_H
#define NPGR_GPU_PROPAGATION_RUNTIME_H

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "dense_subnetwork/DenseNeuronDeviceViews.h"
#include "gpu_runtime/GpuPropagationLayout.h"

struct nbx_net;
namespace nbx_shim {
struct Binding;
}

namespace npgr {

struct RuntimeConfig {
    // recent steps kept in the firing history (at least the layout's delay_slot_count + 1)
    int steps_to_keep = 0;
    bool enable_learning_buffers = false;
    float dt_ms = 1.0f;
};

struct ProfilingStats {
    double neuron_update_ms = 0.0;
    double slice_propagation_ms = 0.0;
    double pending_flush_ms = 0.0;
    double learning_apply_ms = 0.0;
};

struct FiringStepView {
    int time_step = 0;
    std::vector<int> firing_ids;
};

struct PendingBuffersHostView {
    // pending_channels[channel * runtime_stride + neuron]
    std::vector<float> pending_channels;
};

struct InterfaceChannelValue {
    int interface_slot = -1;
    int target_neuron = -1;
    std::uint8_t pending_channel = static_cast<std::uint8_t>(PendingChannel::ExcitatoryConductance);
    float value = 0.0f;
    float scale = 1.0f;
    unsigned char overwrite = 0;
};

struct DeviceToDeviceInterfaceCopySpec {
    int interface_slot = -1;
    int target_neuron = -1;
    std::uint8_t pending_channel = static_cast<std::uint8_t>(PendingChannel::Current);
    int source_index = -1;
    float scale = 1.0f;
    bool overwrite = false;
};

struct LearningBuffersHostView {
    std::vector<float> syn_pending_dwt;
};

class GpuPropagationRuntime {
public:
    GpuPropagationRuntime();
    ~GpuPropagationRuntime();

    bool Initialize(const GpuPropagationLayout& layout, const RuntimeConfig& config, std::string* reason = nullptr);
    bool IsInitialized() const;
    bool ResetState(std::string* reason = nullptr);

    bool BeginStep(int time_step, std::string* reason = nullptr);
    bool RunPropagationOnly(std::string* reason = nullptr);
    bool CommitCurrentDeviceFiringsFromExternal(int time_step,
                                                const int* d_external_firing_ids,
                                                const int* d_external_firing_count,
                                                std::string* reason = nullptr);
    bool ApplyPostLearningFromDeviceFirings(const unsigned char* d_current_did_fire,
                                            std::string* reason = nullptr);
    bool SetInterfaceTargetNeurons(const std::vector<int>& interface_target_neuron_ids, std::string* reason = nullptr);
    bool UploadInterfaceChannelValues(const std::vector<InterfaceChannelValue>& values,
                                      std::string* reason = nullptr);
    bool UploadInterfaceChannelValuesDevice(const float* device_values,
                                            const std::vector<DeviceToDeviceInterfaceCopySpec>& specs,
                                            std::string* reason = nullptr);
    bool ApplyInterfaceStateVectors(std::string* reason = nullptr);

    const RuntimeConfig& config() const;
    const ProfilingStats& profiling() const;
    const PendingBuffersHostView& pending() const;
    const LearningBuffersHostView& learning() const;
    bool IsCudaEnabledBuild() const;
    bool IsDeviceReady() const;
    bool IsCarlsimLikeGpuPathActive() const;
    bool ExportFiringHistoryRing(std::vector<FiringStepView>* history, std::string* reason = nullptr) const;
    bool DownloadPendingChannels(std::vector<float>* pending_channels,
                                 int* channel_count,
                                 int* channel_stride,
                                 std::string* reason = nullptr) const;
    bool DownloadSynapticWeights(std::vector<float>* weights, std::string* reason = nullptr) const;
    bool UploadSynapticWeights(const std::vector<float>& weights, std::string* reason = nullptr);
    bool GetSynapseWeight(int synapse_index, float* weight, std::string* reason = nullptr) const;
    bool SetSynapseWeight(int synapse_index, float weight, std::string* reason = nullptr);
    DeviceCommonBuffersView GetDeviceCommonBuffersView() const;

    // the shared network (libnbx handle; null before the neuron runtime's UploadHostToDevice), for
    // callers that use the nbx C ABI beside this class
    nbx_net* nbx_handle() const;

private:
    std::shared_ptr<nbx_shim::Binding> b_;
    RuntimeConfig config_;
    ProfilingStats profiling_;
    mutable PendingBuffersHostView pending_;
    LearningBuffersHostView learning_;
};

}  // namespace npgr

#endif
