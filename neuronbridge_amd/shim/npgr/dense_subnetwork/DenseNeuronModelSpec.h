// npgr::DenseNeuronModelSpec and the debug snapshot (reference:
// src/native/gpu/include/dense_subnetwork/DenseNeuronModelSpec.h:11-31; same members). Model ids
// are the reference registry's (dense_subnetwork/model/DenseNeuronModelList.inc): 1 LIF double,
// 2 LIF exponential decay, 3 PoissonRate, 4 Izhikevich, 5 trigger relay, 6 LIF voltage jump,
// 7 LIF triple.
#ifndef NPGR_DENSE_NEURON_MODEL_SPEC_H
#define NPGR_DENSE_NEURON_MODEL_SPEC_H

#include <map>
#include <string>
#include <vector>

namespace npgr {

struct DenseNeuronModelFactory {
    static constexpr int kUnknownModelId = 0;
    static constexpr int kLifExponentialDoubleModelId = 1;
    static constexpr int kLifExponentialDecayModelId = 2;
    static constexpr int kPoissonRateModelId = 3;
    static constexpr int kIzhikevichExponentialDecayModelId = 4;
    static constexpr int kTriggerRelayModelId = 5;
    static constexpr int kLifVoltageJumpModelId = 6;
    static constexpr int kLifExponentialTripleModelId = 7;
};

struct DenseNeuronRange {
    int begin = 0;
    int count = 0;
};

struct DenseNeuronModelSpec {
    int factory_model_id = DenseNeuronModelFactory::kUnknownModelId;
    std::string legacy_model_name;
    int model_id = -1;
    DenseNeuronRange range;
    std::map<std::string, std::vector<float> > per_neuron_params;
};

struct DenseNeuronDebugSnapshot {
    int factory_model_id = DenseNeuronModelFactory::kUnknownModelId;
    int model_id = -1;
    std::string legacy_model_name;
    std::map<std::string, std::vector<float> > float_state_vectors;
    std::map<std::string, std::vector<unsigned char> > byte_state_vectors;
};

}  // namespace npgr

#endif
