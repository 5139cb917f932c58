// Native C++ parity check of the reference-interface boundary (npgr::GpuPropagationRuntime +
// npgr::DenseUnifiedNeuronRuntime over libnbx) against the CPU oracle (oracle/liborc.so, the
// checker), in the style of the reference's CTest executables: print result=PASS and exit 0.
//
// A two-population network (LIF voltage jump + conductance Izhikevich with NMDA synapses) is
// described the way the reference's dense build finalizer describes it (GpuPropagationLayout:
// SynapseArrays with effect channels + PreDelaySlices; DenseNeuronModelSpec ranges with
// per-neuron parameters), driven only through interface channels, and stepped with the exact call
// sequence of DenseSubnetworkRuntimeGpu::BeginStep + RunStep (DenseSubnetworkRuntimeGpu.cpp:474-663):
//   BeginStep -> UploadInterfaceChannelValues -> ApplyInterfaceStateVectors ->
//   UploadInterfaceChannelValuesDevice -> RunPropagationOnly -> GetDeviceCommonBuffersView ->
//   Step -> ApplyPostLearningFromDeviceFirings(device_current_did_fire) ->
//   CommitCurrentDeviceFiringsFromExternal(t, device_full_firing_ids, device_full_firing_count)
// The oracle runs the same network with the inputs the reference's interface kernels would apply
// (GpuPropagationRuntime.cu:68-140). Checked bit for bit: every step's firings, the final state,
// ExportFiringHistoryRing over steps_to_keep steps, the output mask, then ResetState followed by a
// repeated run, SetSynapseWeight / GetSynapseWeight and DownloadPendingChannels.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dense_subnetwork/DenseUnifiedNeuronRuntime.h"
#include "gpu_runtime/GpuPropagationRuntime.h"
#include "nbx_core.h"
#include "nbx_oracle.h"

namespace {

int fail(const std::string& what) {
    std::printf("result=FAIL reason=%s\n", what.c_str());
    return 1;
}

constexpr int kLif = 450, kIzh = 150, kN = kLif + kIzh, kSlots = 6, kT = 160, kNSlots = 80, kNSrc = 16;

struct Net {
    npgr::GpuPropagationLayout layout;
    std::vector<npgr::DenseNeuronModelSpec> specs;
    std::vector<float> a, b, c, d;
    // CSR of the same synapses for the oracle
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> post;
    std::vector<uint8_t> delay, type;
    std::vector<float> w;
};

Net build(std::mt19937& rng) {
    Net N;
    std::uniform_int_distribution<int> pick(0, kN - 1), dly(1, kSlots - 1);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    // 40 synapses out of every neuron; inhibitory presynaptic neurons: the last 20% of each population
    struct S {
        int pre, post, d, ty;
        float w;
    };
    std::vector<S> syn;
    for (int p = 0; p < kN; ++p) {
        const bool inh = (p < kLif) ? (p >= kLif * 8 / 10) : (p - kLif >= kIzh * 8 / 10);
        for (int k = 0; k < 40; ++k) {
            const int q = pick(rng);
            const bool to_izh = q >= kLif;
            int ty = inh ? 1 : 0;
            if (!inh && to_izh && U(rng) < 0.3f) ty = 2;  // NMDA into Izhikevich
            const float w = to_izh ? (ty == 1 ? 0.08f : (ty == 2 ? 0.03f : 0.04f)) : (ty == 1 ? 0.6f : 0.15f);
            syn.push_back({p, q, dly(rng), ty, w});
        }
    }
    // reference layout: synapse ids in creation order, slices per (pre, delay slot)
    npgr::GpuPropagationLayout& L = N.layout;
    L.stats.neuron_count = kN;
    L.stats.synapse_count = (int)syn.size();
    L.stats.delay_slot_count = kSlots;
    L.stats.model_count = 2;
    L.stats.pending_channel_count = 4;
    const uint8_t chan_of_type[3] = {(uint8_t)npgr::PendingChannel::ExcitatoryConductance,
                                     (uint8_t)npgr::PendingChannel::InhibitoryConductance,
                                     (uint8_t)npgr::PendingChannel::NmdaConductance};
    for (const S& s : syn) {
        L.synapses.post_neuron.push_back(s.post);
        L.synapses.post_model.push_back(s.post >= kLif ? 1 : 0);
        L.synapses.weight.push_back(s.w);
        L.synapses.max_weight.push_back(1.0f);
        L.synapses.type.push_back((uint8_t)std::min(s.ty, 1));
        L.synapses.effect_channel.push_back(chan_of_type[s.ty]);
        L.synapses.effect_scale.push_back(1.0f);
        L.synapses.plastic_rule_id.push_back(-1);
    }
    L.pre_delay_slices.neuron_count = kN;
    L.pre_delay_slices.delay_slot_count = kSlots;
    L.pre_delay_slices.start.assign((size_t)kN * kSlots, 0);
    L.pre_delay_slices.count.assign((size_t)kN * kSlots, 0);
    N.row_ptr.assign(kN + 1, 0);
    for (int p = 0; p < kN; ++p) {
        for (int dd = 1; dd < kSlots; ++dd) {
            const int f = p * kSlots + dd;
            L.pre_delay_slices.start[(size_t)f] = (int)L.pre_delay_slices.synapse_ids.size();
            for (size_t s = 0; s < syn.size(); ++s)
                if (syn[s].pre == p && syn[s].d == dd) {
                    L.pre_delay_slices.synapse_ids.push_back((int)s);
                    N.post.push_back(syn[s].post);
                    N.delay.push_back((uint8_t)dd);
                    N.type.push_back((uint8_t)syn[s].ty);
                    N.w.push_back(syn[s].w);
                }
            L.pre_delay_slices.count[(size_t)f] = (int)L.pre_delay_slices.synapse_ids.size() - L.pre_delay_slices.start[(size_t)f];
        }
        N.row_ptr[(size_t)p + 1] = (int64_t)N.post.size();
    }
    // model specs (the dense models' parameter names)
    npgr::DenseNeuronModelSpec lif;
    lif.factory_model_id = npgr::DenseNeuronModelFactory::kLifVoltageJumpModelId;
    lif.legacy_model_name = "TimeDrivenLIF_Voltage_jump";
    lif.model_id = 0;
    lif.range = {0, kLif};
    lif.per_neuron_params = {{"tau", {20.0f}}, {"V_rest", {-65.0f}}, {"V_reset", {-65.0f}}, {"V_th", {-50.0f}},
                             {"R", {1.0f}}, {"t_ref", {20.0f}}};
    npgr::DenseNeuronModelSpec izh;
    izh.factory_model_id = npgr::DenseNeuronModelFactory::kIzhikevichExponentialDecayModelId;
    izh.legacy_model_name = "TimeDrivenIzhikevic_Exponential_Decay";
    izh.model_id = 1;
    izh.range = {kLif, kIzh};
    for (int k = 0; k < kIzh; ++k) {
        const float r = U(rng);
        const bool exc = k < kIzh * 8 / 10;
        N.a.push_back(exc ? 0.02f : 0.02f + 0.08f * r);
        N.b.push_back(exc ? 0.2f : 0.25f - 0.05f * r);
        N.c.push_back(exc ? -65.0f + 15.0f * r * r : -65.0f);
        N.d.push_back(exc ? 8.0f - 6.0f * r * r : 2.0f);
    }
    izh.per_neuron_params = {{"a", N.a}, {"b", N.b}, {"c", N.c}, {"d", N.d}, {"V_rest", {-65.0f}}, {"V_th", {30.0f}},
                             {"R", {1.0f}}, {"E_ampa", {0.0f}}, {"E_gaba", {-80.0f}}, {"ampa_tau", {5.0f}},
                             {"gaba_tau", {10.0f}}, {"nmda_tau", {100.0f}}};
    N.specs = {izh, lif};  // any order: the runtime sorts the ranges
    return N;
}

orc_model lif_model() {
    orc_model m{};
    m.model = ORC_LIF_VJ;
    m.integrator = ORC_EULER;
    m.n = kLif;
    m.t_ref = 20;
    m.dt = 0.1f;
    m.v_rest = -65.0f;
    m.tau = 20.0f;
    m.v_th = -50.0f;
    m.r = 1.0f;
    m.v_reset = -65.0f;
    return m;
}

orc_model izh_model(const Net& N) {
    orc_model m{};
    m.model = ORC_IZH;
    m.integrator = ORC_EULER;
    m.n = kIzh;
    m.dt = 0.1f;
    m.v_rest = -65.0f;
    m.izh_v_th = 30.0f;
    m.izh_r = 1.0f;
    m.e_ampa = 0.0f;
    m.e_gaba = -80.0f;
    m.ampa_tau = 5.0f;
    m.gaba_tau = 10.0f;
    m.nmda_tau = 100.0f;
    m.a = N.a.data();
    m.b = N.b.data();
    m.c = N.c.data();
    m.d = N.d.data();
    return m;
}

int kind_of(uint8_t ch, int target) {  // reference channel -> oracle input kind, -1: not an input
    const bool izh = target >= kLif;
    switch ((npgr::PendingChannel)ch) {
        case npgr::PendingChannel::ExcitatoryConductance: return 0;
        case npgr::PendingChannel::InhibitoryConductance: return 1;
        case npgr::PendingChannel::NmdaConductance: return izh ? 2 : -1;
        case npgr::PendingChannel::Current: return ORC_INPUT_CURRENT;
        default: return -1;
    }
}

struct Run {
    std::vector<std::vector<int>> fired;  // per step, ascending
    std::vector<std::vector<int>> outputs;
};

}  // namespace

int main() {
    if (nbx_device_count() <= 0) return fail("no HIP device");
    std::mt19937 rng(0x5EED0009u);
    Net N = build(rng);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);

    // interface: 80 slots onto random neurons, each with a fixed channel; some unmapped; Izhikevich
    // current slots replace I_ext
    std::vector<int> slot_target(kNSlots);
    std::vector<uint8_t> slot_chan(kNSlots);
    for (int s = 0; s < kNSlots; ++s) {
        slot_target[s] = (s % 11 == 10) ? -1 : (int)(rng() % kN);
        const int r = (int)(rng() % 10);
        slot_chan[s] = (uint8_t)(r < 6 ? npgr::PendingChannel::ExcitatoryConductance
                                 : r < 8 ? npgr::PendingChannel::InhibitoryConductance
                                         : npgr::PendingChannel::NmdaConductance);
    }
    for (int s = 0; s < 6; ++s) {  // current slots on Izhikevich neurons
        slot_target[s] = kLif + (int)(rng() % kIzh);
        slot_chan[s] = (uint8_t)npgr::PendingChannel::Current;
    }
    std::vector<npgr::DeviceToDeviceInterfaceCopySpec> copies;
    for (int k = 0; k < 12; ++k) {
        npgr::DeviceToDeviceInterfaceCopySpec c;
        c.interface_slot = -1;
        c.target_neuron = (int)(rng() % kN);
        c.pending_channel = (uint8_t)(k % 3 == 2 ? npgr::PendingChannel::InhibitoryConductance
                                                 : npgr::PendingChannel::ExcitatoryConductance);
        c.source_index = (int)(rng() % kNSrc);
        c.scale = 0.5f;
        c.overwrite = false;
        copies.push_back(c);
    }
    // per-step inputs, drawn once and replayed for the repeated run
    std::vector<std::vector<npgr::InterfaceChannelValue>> recs(kT);
    std::vector<std::vector<float>> srcs(kT);
    for (int t = 0; t < kT; ++t) {
        for (int s = 0; s < kNSlots; ++s) {
            npgr::InterfaceChannelValue v;
            v.interface_slot = s;
            v.target_neuron = -1;
            v.pending_channel = slot_chan[s];
            const bool cur = slot_chan[s] == (uint8_t)npgr::PendingChannel::Current;
            v.value = cur ? 8.0f * U(rng) : (U(rng) < 0.5f ? 2.5f * U(rng) : 0.0f);
            v.scale = cur ? 1.0f : 1.5f;
            v.overwrite = cur ? 1 : 0;
            recs[(size_t)t].push_back(v);
        }
        for (int k = 0; k < kNSrc; ++k) srcs[(size_t)t].push_back(U(rng) < 0.5f ? 2.0f * U(rng) : 0.0f);
    }
    float* d_src = nullptr;
    if (hipMalloc((void**)&d_src, sizeof(float) * kNSrc) != hipSuccess) return fail("hipMalloc");

    npgr::GpuPropagationRuntime prop;
    npgr::DenseUnifiedNeuronRuntime neur;
    npgr::RuntimeConfig cfg;
    cfg.steps_to_keep = 40;
    cfg.dt_ms = 0.1f;
    std::string why;
    {  // the reference rejects a history shorter than delay_slot_count + 1
        npgr::GpuPropagationRuntime p2;
        npgr::RuntimeConfig short_cfg = cfg;
        short_cfg.steps_to_keep = kSlots;
        if (p2.Initialize(N.layout, short_cfg, &why)) return fail("steps_to_keep below delay_slot_count + 1 accepted");
    }
    // DenseSubnetworkRuntimeGpu::Initialize order
    if (!prop.Initialize(N.layout, cfg, &why)) return fail("propagation Initialize: " + why);
    if (!prop.IsCarlsimLikeGpuPathActive()) return fail("GPU path not active");
    if (!prop.SetInterfaceTargetNeurons(slot_target, &why)) return fail("SetInterfaceTargetNeurons: " + why);
    std::vector<unsigned char> mask(kN, 0);
    for (int i = kLif; i < kN; ++i) mask[(size_t)i] = 1;
    if (!neur.Initialize(N.specs, kN, cfg.dt_ms, &why) || !neur.ConfigureOutputNeuronMask(mask, &why) ||
        !neur.AllocateDeviceBuffers(&why) || !neur.UploadHostToDevice(&why))
        return fail("neuron runtime: " + why);
    nbx_net* h = prop.nbx_handle();
    if (!h) return fail("no shared network");
    int32_t shift = 0;
    nbx_net_get_fixed_shift(h, &shift);

    auto run_steps = [&](Run* out, int t0, int nsteps) -> std::string {
        for (int t = t0; t < t0 + nsteps; ++t) {
            if (!prop.BeginStep(t, &why)) return "BeginStep: " + why;
            if (!prop.UploadInterfaceChannelValues(recs[(size_t)t], &why)) return "UploadInterfaceChannelValues: " + why;
            if (!prop.ApplyInterfaceStateVectors(&why)) return "ApplyInterfaceStateVectors: " + why;
            if (hipMemcpy(d_src, srcs[(size_t)t].data(), sizeof(float) * kNSrc, hipMemcpyHostToDevice) != hipSuccess)
                return "hipMemcpy";
            if (!prop.UploadInterfaceChannelValuesDevice(d_src, copies, &why)) return "UploadInterfaceChannelValuesDevice: " + why;
            if (!prop.RunPropagationOnly(&why)) return "RunPropagationOnly: " + why;
            const npgr::DeviceCommonBuffersView view = prop.GetDeviceCommonBuffersView();
            if (!view.d_pending_channels || view.neuron_count != kN) return "device common buffers view";
            std::vector<int> outs, full;
            if (!neur.Step(view, t, &outs, &full, true, &why)) return "Step: " + why;
            if (!prop.ApplyPostLearningFromDeviceFirings(neur.device_current_did_fire(), &why))
                return "ApplyPostLearningFromDeviceFirings: " + why;
            if (!prop.CommitCurrentDeviceFiringsFromExternal(t, neur.device_full_firing_ids(),
                                                             neur.device_full_firing_count(), &why))
                return "CommitCurrentDeviceFiringsFromExternal: " + why;
            out->fired.push_back(full);
            out->outputs.push_back(outs);
        }
        return "";
    };
    Run gpu;
    std::string err = run_steps(&gpu, 0, kT);
    if (!err.empty()) return fail(err);

    // ---- the oracle on the same network and the inputs the reference's interface kernels apply
    orc_model pops[2] = {lif_model(), izh_model(N)};
    const int32_t pop_begin[3] = {0, kLif, kN};
    std::vector<orc_input> inputs;
    std::vector<int32_t> input_ptr(1, 0);
    for (int t = 0; t < kT; ++t) {
        for (const npgr::InterfaceChannelValue& v : recs[(size_t)t]) {  // ApplyBoundaryStateVectorKernel
            if (v.value == 0.0f) continue;
            const int tgt = v.target_neuron >= 0 ? v.target_neuron : slot_target[(size_t)v.interface_slot];
            if (tgt < 0) continue;
            const int kind = kind_of(v.pending_channel, tgt);
            if (kind < 0) continue;
            inputs.push_back(orc_input{tgt, kind, v.value * v.scale});
        }
        for (const npgr::DeviceToDeviceInterfaceCopySpec& c : copies) {  // ApplyDeviceInterfaceChannelValuesKernel
            const float val = srcs[(size_t)t][(size_t)c.source_index] * c.scale;
            if (val == 0.0f && !c.overwrite) continue;
            const int kind = kind_of(c.pending_channel, c.target_neuron);
            if (kind < 0) continue;
            inputs.push_back(orc_input{c.target_neuron, kind, val});
        }
        input_ptr.push_back((int32_t)inputs.size());
    }
    orc_net on{};
    on.m = pops[0];
    on.d_slots = kSlots;
    on.fixed_shift = shift;
    on.row_ptr = N.row_ptr.data();
    on.post = N.post.data();
    on.delay = N.delay.data();
    on.type = N.type.data();
    on.w = N.w.data();
    on.npops = 2;
    on.pops = pops;
    on.pop_begin = pop_begin;
    on.inputs = inputs.data();
    on.input_ptr = input_ptr.data();
    on.input_t0 = 0;
    on.input_nsteps = kT;
    const int ns = 6;
    std::vector<float> x((size_t)kN * ns, 0.0f);
    std::vector<int32_t> last(kN, 10000), hist((size_t)kSlots * kN), hcnt(kSlots, 0), counts(kT);
    for (int i = 0; i < kLif; ++i) x[(size_t)i * ns] = -65.0f;  // InitNeuronState, sigma 0
    for (int k = 0; k < kIzh; ++k) {
        x[(size_t)(kLif + k) * ns] = -65.0f;
        x[(size_t)(kLif + k) * ns + 1] = N.b[(size_t)k] * -65.0f;
    }
    const int words = (kN + 31) / 32;
    std::vector<uint32_t> raster((size_t)kT * words, 0u);
    int64_t events = 0;
    if (orc_run(&on, ORC_ACC_FIXED, x.data(), last.data(), hist.data(), hcnt.data(), nullptr, 0, kT, counts.data(),
                raster.data(), &events) != 0)
        return fail("oracle run");

    // every step's firings
    long total = 0, izh_total = 0;
    for (int t = 0; t < kT; ++t) {
        std::vector<int> ref;
        for (int i = 0; i < kN; ++i)
            if (raster[(size_t)t * words + i / 32] >> (i % 32) & 1u) ref.push_back(i);
        if (ref != gpu.fired[(size_t)t]) return fail("firings differ at step " + std::to_string(t));
        std::vector<int> ref_out;
        for (int i : ref)
            if (i >= kLif) ref_out.push_back(i);
        if (ref_out != gpu.outputs[(size_t)t]) return fail("output firings differ at step " + std::to_string(t));
        total += (long)ref.size();
        izh_total += (long)ref_out.size();
    }
    if (total < 200 || izh_total < 20 || total - izh_total < 20) return fail("network too quiet to test");
    // final state bit for bit
    std::vector<float> xg((size_t)kN * nbx_net_nstate(h));
    std::vector<int32_t> lg(kN);
    if (nbx_net_nstate(h) != ns || nbx_net_get_state(h, xg.data(), lg.data()) != NBX_OK) return fail("get_state");
    if (std::memcmp(xg.data(), x.data(), sizeof(float) * x.size()) != 0) return fail("state differs from the oracle");
    if (lg != last) return fail("LastSpike differs from the oracle");
    int64_t ev = 0;
    nbx_net_events(h, &ev);
    if (ev != events) return fail("synaptic events differ");
    // firing history over steps_to_keep steps
    std::vector<npgr::FiringStepView> hv;
    if (!prop.ExportFiringHistoryRing(&hv, &why) || (int)hv.size() != cfg.steps_to_keep) return fail("history export");
    for (int t = kT - cfg.steps_to_keep; t < kT; ++t) {
        const npgr::FiringStepView& v = hv[(size_t)(t % cfg.steps_to_keep)];
        std::vector<int> ids = v.firing_ids;
        std::sort(ids.begin(), ids.end());
        if (v.time_step != t || ids != gpu.fired[(size_t)t]) return fail("history of step " + std::to_string(t));
    }
    // pending channels: the reference's five planes
    std::vector<float> pend;
    int pc = 0, pstride = 0;
    if (!prop.DownloadPendingChannels(&pend, &pc, &pstride, &why) || pc != 5 || pstride != kN ||
        pend.size() != (size_t)5 * kN)
        return fail("DownloadPendingChannels: " + why);
    // weights by reference synapse id
    std::vector<float> wd;
    if (!prop.DownloadSynapticWeights(&wd, &why) || wd != N.layout.synapses.weight) return fail("weights: " + why);
    float got = 0.0f;
    if (!prop.SetSynapseWeight(17, 0.25f, &why) || !prop.GetSynapseWeight(17, &got, &why) || got != 0.25f)
        return fail("Set/GetSynapseWeight: " + why);
    if (!prop.SetSynapseWeight(17, N.layout.synapses.weight[17], &why)) return fail("SetSynapseWeight: " + why);
    // ResetState, then the same inputs give the same run
    if (!prop.ResetState(&why) || !neur.ResetState(&why)) return fail("ResetState: " + why);
    Run again;
    err = run_steps(&again, 0, 60);
    if (!err.empty()) return fail("after reset: " + err);
    for (int t = 0; t < 60; ++t)
        if (again.fired[(size_t)t] != gpu.fired[(size_t)t]) return fail("run after ResetState differs at step " + std::to_string(t));
    hipFree(d_src);
    std::printf("steps=%d spikes=%ld izhikevich_outputs=%ld events=%lld result=PASS\n", kT, total, izh_total,
                (long long)events);
    return 0;
}
