// npgr::DeviceCommonBuffersView, the view the propagation runtime lends the neuron runtime
// (reference: src/native/gpu/include/dense_subnetwork/DenseNeuronDeviceViews.h:6-13; same members).
// Over libnbx the pool is a float snapshot of the current step's pending arrivals.
#ifndef NPGR_DENSE_NEURON_DEVICE_VIEWS_H
#define NPGR_DENSE_NEURON_DEVICE_VIEWS_H

namespace npgr {

struct DeviceCommonBuffersView {
    // channel-major pending inputs: d_pending_channels[channel * pending_channel_stride + neuron]
    float* d_pending_channels = nullptr;
    int pending_channel_count = 0;
    int pending_channel_stride = 0;
    int neuron_count = 0;
};

}  // namespace npgr

#endif
