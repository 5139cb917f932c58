// Native C++ checks of the reference-interface boundary (npgr::GpuPropagationRuntime +
// npgr::DenseUnifiedNeuronRuntime over libnbx), in the style of the reference's CTest executables:
// print result=PASS and exit 0.
//   1. delay known answer of tests/native/dense_cuda_smoke.cpp:109-118 (a spike committed at t=0
//      through a delay-1 synapse makes its target fire at t=1, not t=0). The spike enters through
//      CommitCurrentDeviceFiringsFromExternal with a caller-owned device list (the reference's
//      contract: the external list replaces the step's firings).
//   2. pair-STDP weight trajectory of tests/model_codegen/pair_stdp_gpu_trajectory.cpp: pre spikes
//      at ticks 0 and 7 (delay 1, dt 1 ms), reference STDP defaults carried in the layout's
//      "stdp.*" learning fields, weight 8 of 20; the GPU trajectory is replayed through the CPU pair
//      rule (STDP.cpp:19-61) and must agree within the reference's CPU/GPU tolerance of 5e-6.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "dense_subnetwork/DenseUnifiedNeuronRuntime.h"
#include "gpu_runtime/GpuPropagationRuntime.h"

namespace {

int fail(const std::string& what) {
    std::printf("result=FAIL reason=%s\n", what.c_str());
    return 1;
}

npgr::DenseNeuronModelSpec lif(int n, float v_th) {
    npgr::DenseNeuronModelSpec s;
    s.factory_model_id = npgr::DenseNeuronModelFactory::kLifVoltageJumpModelId;
    s.legacy_model_name = "TimeDrivenLIF_Voltage_jump";
    s.range = {0, n};
    s.per_neuron_params = {{"V_rest", {-60.0f}}, {"V_reset", {-60.0f}}, {"tau", {20.0f}}, {"V_th", {v_th}},
                           {"R", {1.0f}}, {"t_ref", {0.0f}}};
    return s;
}

// synapse s: s -> s + 1, delay slot 1; plastic synapses carry the reference's STDP fields
npgr::GpuPropagationLayout chain(int n, const std::vector<float>& w, bool plastic) {
    npgr::GpuPropagationLayout L;
    L.stats.neuron_count = n;
    L.stats.synapse_count = n - 1;
    L.stats.delay_slot_count = 2;
    L.stats.model_count = 1;
    L.stats.pending_channel_count = 3;
    for (int s = 0; s + 1 < n; ++s) {
        L.synapses.post_neuron.push_back(s + 1);
        L.synapses.post_model.push_back(0);
        L.synapses.weight.push_back(w[(size_t)s]);
        L.synapses.max_weight.push_back(20.0f);
        L.synapses.type.push_back(0);
        L.synapses.effect_channel.push_back((uint8_t)npgr::PendingChannel::ExcitatoryConductance);
        L.synapses.effect_scale.push_back(1.0f);
        L.synapses.plastic_rule_id.push_back(plastic ? 0 : -1);
        L.synapses.plastic_model_id.push_back(plastic ? npgr::kDenseLearningStdpModelId : -1);
    }
    L.pre_delay_slices.neuron_count = n;
    L.pre_delay_slices.delay_slot_count = 2;
    L.pre_delay_slices.start.assign((size_t)n * 2, 0);
    L.pre_delay_slices.count.assign((size_t)n * 2, 0);
    for (int s = 0; s + 1 < n; ++s) {
        L.pre_delay_slices.start[(size_t)s * 2 + 1] = s;
        L.pre_delay_slices.count[(size_t)s * 2 + 1] = 1;
        L.pre_delay_slices.synapse_ids.push_back(s);
    }
    if (plastic) {  // DenseStdpLearningModel synapse fields, field-major pools
        const char* names[] = {"stdp.max_ltp", "stdp.max_ltd", "stdp.inv_ltp_tau", "stdp.inv_ltd_tau", "stdp.max_weight"};
        const float vals[] = {0.92f, 0.53f, (float)(1.0 / 16.8), (float)(1.0 / 33.1), 20.0f};
        const int S = n - 1;
        for (int f = 0; f < 5; ++f) {
            npgr::DenseFieldSpan sp;
            sp.field_id = f;
            sp.name = names[f];
            sp.offset = f * S;
            sp.count = S;
            L.learning_fields.fields.push_back(sp);
            L.learning_fields.field_id_by_name[names[f]] = f;
            for (int s = 0; s < S; ++s) L.learning_fields.float_pool.push_back(vals[f]);
        }
    }
    return L;
}

struct Rt {
    npgr::GpuPropagationRuntime prop;
    npgr::DenseUnifiedNeuronRuntime neur;
    int* d_ext = nullptr;  // caller-owned firing list: [count, ids...]
    bool init(const npgr::DenseNeuronModelSpec& spec, const npgr::GpuPropagationLayout& L, int keep, std::string* why) {
        npgr::RuntimeConfig cfg;
        cfg.steps_to_keep = keep;
        cfg.dt_ms = 1.0f;
        if (hipMalloc((void**)&d_ext, sizeof(int) * 64) != hipSuccess) return false;
        return prop.Initialize(L, cfg, why) && neur.Initialize({spec}, spec.range.count, 1.0f, why) &&
               neur.AllocateDeviceBuffers(why) && neur.UploadHostToDevice(why);
    }
    ~Rt() { hipFree(d_ext); }
    // BeginStep / RunStep of DenseSubnetworkRuntimeGpu; `extra` fires in addition at step t
    bool step(int t, const std::vector<int>& extra, std::vector<int>* fired, std::string* why) {
        if (!prop.BeginStep(t, why) || !prop.RunPropagationOnly(why)) return false;
        std::vector<int> out, full;
        if (!neur.Step(prop.GetDeviceCommonBuffersView(), t, &out, &full, true, why)) return false;
        if (!prop.ApplyPostLearningFromDeviceFirings(neur.device_current_did_fire(), why)) return false;
        if (extra.empty()) {
            if (!prop.CommitCurrentDeviceFiringsFromExternal(t, neur.device_full_firing_ids(), neur.device_full_firing_count(), why))
                return false;
        } else {
            std::vector<int> all = full;
            all.insert(all.end(), extra.begin(), extra.end());
            const int cnt = (int)all.size();
            if (hipMemcpy(d_ext, &cnt, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_ext + 1, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice) != hipSuccess)
                return false;
            if (!prop.CommitCurrentDeviceFiringsFromExternal(t, d_ext + 1, d_ext, why)) return false;
            full = all;
        }
        *fired = full;
        return true;
    }
};

}  // namespace

int main() {
    std::string reason;
    {  // 1. delay known answer
        Rt rt;
        if (!rt.init(lif(3, -59.8f), chain(3, {12.0f, 8.0f}, false), 4, &reason)) return fail(reason);
        std::vector<int> at[3];
        for (int t = 0; t < 3; ++t)
            if (!rt.step(t, t == 0 ? std::vector<int>{0} : std::vector<int>{}, &at[t], &reason)) return fail(reason);
        if (at[0] != std::vector<int>{0}) return fail("t0 must hold only the committed spike");
        if (at[1] != std::vector<int>{1}) return fail("delay-1 target must fire at t1");
        if (at[2] != std::vector<int>{2}) return fail("second hop must fire at t2");
        std::vector<npgr::FiringStepView> ring;
        if (!rt.prop.ExportFiringHistoryRing(&ring, &reason) || ring.size() != 4) return fail("history ring: " + reason);
        for (int t = 0; t < 3; ++t)
            if (ring[(size_t)t].time_step != t || ring[(size_t)t].firing_ids != at[t]) return fail("history slot");
        if (ring[3].time_step != -1) return fail("an unwritten history slot must read step -1");
        std::string err;
        if (rt.prop.BeginStep(9, &err) || err.empty()) return fail("out-of-order BeginStep must be rejected");
        std::printf("delay_kat=PASS history t0=%zu t1=%zu t2=%zu\n", at[0].size(), at[1].size(), at[2].size());
    }
    {  // 2. pair-STDP trajectory
        const float ltp = 0.92f, ltd = 0.53f, wmax = 20.0f;
        const float inv_ltp = (float)(1.0 / 16.8), inv_ltd = (float)(1.0 / 33.1);
        Rt rt;
        if (!rt.init(lif(2, -59.8f), chain(2, {8.0f}, true), 3, &reason)) return fail(reason);
        std::vector<float> gpu_w;
        std::vector<int> pre_events, post_events;
        const int ticks = 16;
        for (int t = 0; t < ticks; ++t) {
            std::vector<int> fired;
            const bool inject = (t == 0 || t == 7);
            if (!rt.step(t, inject ? std::vector<int>{0} : std::vector<int>{}, &fired, &reason)) return fail(reason);
            float w = 0.0f;
            if (!rt.prop.GetSynapseWeight(0, &w, &reason)) return fail(reason);
            gpu_w.push_back(w);
            if (inject) pre_events.push_back(t + 1);  // arrival one step later
            for (int id : fired)
                if (id == 1) post_events.push_back(t);
        }
        // CPU replay of the pair rule on the GPU-observed events (pre at arrival before post)
        float w = 8.0f, apre = 0.0f, apost = 0.0f;
        int last = 0;
        float worst = 0.0f;
        int changed = 0;
        auto decay = [&](int t) {
            const float dT = (float)(t - last) * 1.0f;
            apre *= std::exp(-dT * inv_ltp);
            apost *= std::exp(-dT * inv_ltd);
            last = t;
        };
        for (int t = 0; t < ticks; ++t) {
            for (int a : pre_events)
                if (a == t) {
                    decay(t);
                    apre += 1.0f;
                    w += -ltd * apost;
                    w = w < 0.0f ? 0.0f : (w > wmax ? wmax : w);
                }
            for (int p : post_events)
                if (p == t) {
                    decay(t);
                    apost += 1.0f;
                    w += ltp * apre;
                    w = w < 0.0f ? 0.0f : (w > wmax ? wmax : w);
                }
            worst = std::fmax(worst, std::fabs(w - gpu_w[(size_t)t]));
            if (t > 0 && gpu_w[(size_t)t] != gpu_w[(size_t)t - 1]) ++changed;
        }
        if (post_events.empty()) return fail("post neuron never fired");
        if (worst > 5e-6f) return fail("pair-STDP CPU/GPU trajectory mismatch " + std::to_string(worst));
        if (changed == 0) return fail("pair-STDP trajectory did not change the weight");
        std::printf("pair_stdp=PASS post_spikes=%zu max_error=%g\n", post_events.size(), (double)worst);
    }
    std::printf("result=PASS\n");
    return 0;
}
