// Native C++ checks of the C++ runtime mirror (npgr_nbx::DenseRuntime) over libnbx, in the
// style of the reference's CTest executables: print result=PASS and exit 0.
//   1. delay known answer of tests/native/dense_cuda_smoke.cpp:109-118 (a spike injected at
//      t=0 through a delay-1 synapse makes its target fire at t=1, not t=0), driven through
//      the reference step sequence BeginStep/RunPropagationOnly/Step/.../Commit
//   2. pair-STDP weight trajectory of tests/model_codegen/pair_stdp_gpu_trajectory.cpp: pre
//      spikes at ticks 0 and 7 (delay 1, dt 1 ms), reference STDP defaults, weight 8 of 20; the
//      GPU trajectory is replayed through the CPU pair rule (STDP.cpp:19-61) and must agree
//      within the reference's CPU/GPU tolerance of 5e-6.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "npgr_nbx_runtime.hpp"

namespace {

int fail(const std::string& what) {
    std::printf("result=FAIL reason=%s\n", what.c_str());
    return 1;
}

nbx_model_desc lif(int n, float v_th) {
    nbx_model_desc m{};
    m.model = NBX_MODEL_LIF_VJ;
    m.integrator = NBX_EULER;
    m.n = n;
    m.t_ref = 0;
    m.dt = 1.0f;
    m.v_rest = -60.0f;
    m.tau = 20.0f;
    m.v_th = v_th;
    m.r = 1.0f;
    m.v_reset = -60.0f;
    return m;
}

npgr_nbx::GpuPropagationLayout chain(int n, const std::vector<float>& w, const std::vector<unsigned char>& plastic) {
    // synapse s: s -> s+1, delay slot 1
    npgr_nbx::GpuPropagationLayout L;
    for (int s = 0; s + 1 < n; ++s) {
        L.synapses.post_neuron.push_back(s + 1);
        L.synapses.weight.push_back(w[(size_t)s]);
        L.synapses.type.push_back(0);
    }
    L.synapses.plastic_flags = plastic;
    L.pre_delay_slices.neuron_count = n;
    L.pre_delay_slices.delay_slot_count = 2;
    L.pre_delay_slices.start.assign((size_t)n * 2, 0);
    L.pre_delay_slices.count.assign((size_t)n * 2, 0);
    for (int s = 0; s + 1 < n; ++s) {
        L.pre_delay_slices.start[(size_t)s * 2 + 1] = s;
        L.pre_delay_slices.count[(size_t)s * 2 + 1] = 1;
        L.pre_delay_slices.synapse_ids.push_back(s);
    }
    return L;
}

bool step(npgr_nbx::DenseRuntime& rt, int t, const std::vector<int>& inject, std::vector<int>* fired,
          std::string* reason) {
    return rt.BeginStep(t, reason) && rt.RunPropagationOnly(reason) && rt.Step(fired, reason) &&
           (inject.empty() || rt.QueueInterfaceSpikes(inject, reason)) && rt.ApplyPostLearningFromDeviceFirings(reason) &&
           rt.CommitCurrentDeviceFiringsFromExternal(t, reason);
}

}  // namespace

int main() {
    std::string reason;
    {  // 1. delay known answer
        npgr_nbx::DenseRuntime rt;
        npgr_nbx::RuntimeConfig cfg;
        cfg.steps_to_keep = 4;
        if (!rt.Initialize(lif(3, -59.8f), chain(3, {12.0f, 8.0f}, {}), cfg, nullptr, &reason)) return fail(reason);
        // the history ring holds the last max(max_delay, 2) committed steps: read each step's
        // firings right after its commit
        std::vector<int> at[3];
        for (int t = 0; t < 3; ++t) {
            std::vector<int> fired;
            if (!step(rt, t, t == 0 ? std::vector<int>{0} : std::vector<int>{}, &fired, &reason)) return fail(reason);
            std::vector<npgr_nbx::FiringStepView> ring;
            if (!rt.ExportFiringHistoryRing(&ring, &reason)) return fail(reason);
            if (ring.empty() || ring.front().time_step != t) return fail("history ring must start at the last commit");
            at[t] = ring.front().firing_ids;
        }
        if (at[0] != std::vector<int>{0}) return fail("t0 must hold only the injected spike");
        if (at[1] != std::vector<int>{1}) return fail("delay-1 target must fire at t1");
        if (at[2] != std::vector<int>{2}) return fail("second hop must fire at t2");
        std::string err;
        if (rt.BeginStep(9, &err) || err.empty()) return fail("out-of-order BeginStep must be rejected");
        std::printf("delay_kat=PASS history t0=%zu t1=%zu t2=%zu\n", at[0].size(), at[1].size(), at[2].size());
    }
    {  // 2. pair-STDP trajectory
        const float ltp = 0.92f, ltp_tau = 16.8f, ltd = 0.53f, ltd_tau = 33.1f, wmax = 20.0f;
        nbx_stdp_desc sd{ltp, ltp_tau, ltd, ltd_tau, wmax};
        npgr_nbx::DenseRuntime rt;
        npgr_nbx::RuntimeConfig cfg;
        if (!rt.Initialize(lif(2, -59.8f), chain(2, {8.0f}, {1}), cfg, &sd, &reason)) return fail(reason);
        std::vector<float> gpu_w;
        std::vector<int> pre_events, post_events;
        const int ticks = 16;
        for (int t = 0; t < ticks; ++t) {
            std::vector<int> fired;
            const bool inject = (t == 0 || t == 7);
            if (!step(rt, t, inject ? std::vector<int>{0} : std::vector<int>{}, &fired, &reason)) return fail(reason);
            float w = 0.0f;
            if (!rt.GetSynapseWeight(0, &w, &reason)) return fail(reason);
            gpu_w.push_back(w);
            if (inject) pre_events.push_back(t + 1);  // arrival one step later
            for (int id : fired)
                if (id == 1) post_events.push_back(t);
        }
        // CPU replay of the pair rule on the GPU-observed events (pre at arrival before post)
        float w = 8.0f, apre = 0.0f, apost = 0.0f;
        int last = 0;
        const float inv_ltp = (float)(1.0 / (double)ltp_tau), inv_ltd = (float)(1.0 / (double)ltd_tau);
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
