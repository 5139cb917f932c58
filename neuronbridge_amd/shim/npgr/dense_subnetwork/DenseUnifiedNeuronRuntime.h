// npgr::DenseUnifiedNeuronRuntime over libnbx: the reference class's public interface, method for
// method (reference: src/native/gpu/include/dense_subnetwork/DenseUnifiedNeuronRuntime.h:15-41).
// Each DenseNeuronModelSpec range becomes nbx populations (a range whose per-neuron parameters
// differ is split into runs of equal parameters; Izhikevich a, b, c, d stay per neuron). The
// neuron update follows the legacy CPU models' arithmetic, the parity target of this core.
#ifndef NPGR_DENSE_UNIFIED_NEURON_RUNTIME_H
#define NPGR_DENSE_UNIFIED_NEURON_RUNTIME_H

#include <cstddef>
#include <memory>
#include <string>
#include <vector>

#include "dense_subnetwork/DenseNeuronDeviceViews.h"
#include "dense_subnetwork/DenseNeuronModelSpec.h"

namespace nbx_shim {
struct Binding;
}

namespace npgr {

class DenseUnifiedNeuronRuntime {
public:
    DenseUnifiedNeuronRuntime();
    ~DenseUnifiedNeuronRuntime();

    bool Initialize(const std::vector<DenseNeuronModelSpec>& specs,
                    int neuron_count,
                    float dt_ms,
                    std::string* reason = nullptr);
    bool AllocateDeviceBuffers(std::string* reason = nullptr);
    bool UploadHostToDevice(std::string* reason = nullptr);
    bool ConfigureOutputNeuronMask(const std::vector<unsigned char>& output_neuron_mask,
                                   std::string* reason = nullptr);
    bool Step(const DeviceCommonBuffersView& common,
              int current_time_step,
              std::vector<int>* output_firing_ids,
              std::vector<int>* full_firing_ids,
              bool export_full_firings_to_host,
              std::string* reason = nullptr);
    bool ResetState(std::string* reason = nullptr);
    bool SyncHostStateFromDevice(std::string* reason = nullptr) const;
    bool BuildDebugSnapshots(std::vector<DenseNeuronDebugSnapshot>* out,
                             std::string* reason = nullptr) const;

    const int* device_full_firing_ids() const;
    const int* device_full_firing_count() const;
    const unsigned char* device_current_did_fire() const;

private:
    std::shared_ptr<nbx_shim::Binding> b_;
    std::vector<DenseNeuronModelSpec> specs_;
    std::vector<unsigned char> output_mask_;
    int neuron_count_ = 0;
    float dt_ms_ = 1.0f;
};

}  // namespace npgr

#endif
