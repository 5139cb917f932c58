// Current synapses (synapse type 3) of the legacy engine.
//
// Reference: a model's CheckType registers each type-3 Interconnection as the next input-current
// slot of its target (TimeDrivenLIF_Voltage_jump.cpp:127-147, CurrentSynapse's per-target slot
// counters, Current/src/CurrentSynapes.cpp); an InputCurrent of a source neuron
// (Event/src/Current/InputCurrent.cpp:13-24, from ArrayInputCurrentDriver) becomes, after each
// synapse's delay, a PropogatedCurrent (Event/src/Current/PropogatedCurrent.cpp:15-25) whose
// ProcessCurrent sets that synapse's slot to the current (the synapse weight is not applied) and
// writes the sum of the target's slots, in slot order, into its I_ext
// (TimeDrivenLIF_Voltage_jump.cpp:117-125).
//
// Here the slots live on the device; injected currents wait on the host keyed by arrival step, and
// the arrivals of a step are applied before its neuron update: the run loop stops at such steps
// (the engines' launches compose like separate runs). One step's arrivals are applied in injection
// order (a later current for the same synapse wins), then every target they touch recomputes
// I_ext as the float sum of its slots in ascending slot order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nbx_net.h"

using namespace nbx;

namespace {

__global__ void nbx_cs_set_slots(const int32_t* pairs, int32_t count, float* slots) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    slots[pairs[2 * k]] = __int_as_float(pairs[2 * k + 1]);
}

__global__ void nbx_cs_sum_slots(const int32_t* targets, int32_t count, const int32_t* in_ptr, const float* slots,
                                 float* iext) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int32_t i = targets[k];
    float c = 0.0f;  // CurrentSynapse::GetTotalInputCurrentPerNeuron: float adds in slot order
    for (int32_t j = in_ptr[i]; j < in_ptr[i + 1]; ++j) c += slots[j];
    iext[i] = c;
}

int32_t cs_alloc(CurrentSynapses& C, void** p, size_t bytes) {
    NBX_HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 4)));
    C.allocs.push_back(*p);
    return NBX_OK;
}

int32_t check_net(nbx_net* net) {
    if (!net) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    if (!net->finalized) {
        set_error("network is not finalized");
        return NBX_ERR_STATE;
    }
    NBX_HIP_TRY(hipSetDevice(net->device));
    return NBX_OK;
}

}  // namespace

namespace nbx {

int32_t currents_due_next(const nbx_net* net, int32_t t) {
    const CurrentSynapses& C = net->cs;
    for (int32_t k = std::max(0, t - C.base); k < (int32_t)C.pending.size(); ++k)
        if (!C.pending[(size_t)k].empty()) return C.base + k;
    return -1;
}

int32_t apply_currents(nbx_net* net, int32_t t) {
    CurrentSynapses& C = net->cs;
    const int32_t k = t - C.base;
    if (C.count == 0 || k < 0 || k >= (int32_t)C.pending.size() || C.pending[(size_t)k].empty()) return NBX_OK;
    // injection order, last value per synapse wins; the targets they touch, ascending
    std::vector<std::pair<int32_t, float>>& ev = C.pending[(size_t)k];
    std::vector<int32_t> last_of(ev.size());
    std::vector<std::pair<int32_t, int32_t>> by_syn;  // (synapse, event index)
    by_syn.reserve(ev.size());
    for (size_t e = 0; e < ev.size(); ++e) by_syn.emplace_back(ev[e].first, (int32_t)e);
    std::stable_sort(by_syn.begin(), by_syn.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int32_t> host;  // pairs (slot, value bits) then targets
    std::vector<int32_t> targets;
    for (size_t q = 0; q < by_syn.size(); ++q) {
        if (q + 1 < by_syn.size() && by_syn[q + 1].first == by_syn[q].first) continue;  // a later event wins
        const int32_t s = by_syn[q].first;
        int32_t bits;
        std::memcpy(&bits, &ev[(size_t)by_syn[q].second].second, 4);
        host.push_back(C.slot[(size_t)s]);
        host.push_back(bits);
        targets.push_back(C.target[(size_t)s]);
    }
    std::sort(targets.begin(), targets.end());
    targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
    const int32_t npairs = (int32_t)host.size() / 2, ntg = (int32_t)targets.size();
    host.insert(host.end(), targets.begin(), targets.end());
    if ((int64_t)host.size() > C.stage_cap) {
        if (C.d_stage) {
            auto it = std::find(C.allocs.begin(), C.allocs.end(), (void*)C.d_stage);
            if (it != C.allocs.end()) C.allocs.erase(it);
            NBX_HIP_TRY(hipStreamSynchronize(net->stream));
            NBX_HIP_TRY(hipFree(C.d_stage));
            C.d_stage = nullptr;
        }
        const int64_t cap = std::max<int64_t>(1024, (int64_t)host.size() * 2);
        int32_t rc = cs_alloc(C, (void**)&C.d_stage, (size_t)cap * 4);
        if (rc) return rc;
        C.stage_cap = cap;
    }
    NBX_HIP_TRY(hipMemcpyAsync(C.d_stage, host.data(), host.size() * 4, hipMemcpyHostToDevice, net->stream));
    nbx_cs_set_slots<<<(npairs + 255) / 256, 256, 0, net->stream>>>(C.d_stage, npairs, C.d_slots);
    NBX_HIP_TRY(hipGetLastError());
    nbx_cs_sum_slots<<<(ntg + 255) / 256, 256, 0, net->stream>>>(C.d_stage + 2 * npairs, ntg, C.d_in_ptr, C.d_slots,
                                                                 net->d_iext);
    NBX_HIP_TRY(hipGetLastError());
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));  // the host staging stays alive until the copy ran
    ev.clear();
    // drop the leading empty steps
    size_t z = 0;
    while (z < C.pending.size() && C.pending[z].empty()) ++z;
    C.pending.erase(C.pending.begin(), C.pending.begin() + (ptrdiff_t)z);
    C.base += (int32_t)z;
    return NBX_OK;
}

void clear_currents(nbx_net* net) {
    CurrentSynapses& C = net->cs;
    C.pending.clear();
    C.base = 0;
    if (C.d_slots && C.count > 0) {
        (void)launch_fill_i32(reinterpret_cast<int32_t*>(C.d_slots), 0, C.count, net->stream);
        (void)hipStreamSynchronize(net->stream);
    }
}

}  // namespace nbx

extern "C" {

int32_t nbx_net_set_current_synapses(nbx_net* net, const int32_t* src, const int32_t* dst, const uint8_t* delay,
                                     int32_t count) {
    int32_t rc = check_net(net);
    if (rc) return rc;
    if (count < 0 || (count > 0 && (!src || !dst || !delay))) {
        set_error("nbx_net_set_current_synapses: bad arguments");
        return NBX_ERR_INVALID;
    }
    if (net->phase != 0) {
        set_error("a per-phase step is open (finish it with nbx_net_commit_firings)");
        return NBX_ERR_STATE;
    }
    const int32_t n = net->n;
    for (int32_t k = 0; k < count; ++k) {
        if (src[k] < 0 || src[k] >= n || dst[k] < 0 || dst[k] >= n) {
            set_error("current synapse neuron id out of range");
            return NBX_ERR_INVALID;
        }
        if (delay[k] < 1 || delay[k] > 63) {
            set_error("current synapse delay must be 1..63 steps");
            return NBX_ERR_INVALID;
        }
    }
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    CurrentSynapses C;
    C.count = count;
    C.delay.assign(delay, delay + count);
    C.target.assign(dst, dst + count);
    C.out_ptr.assign((size_t)n + 1, 0);
    for (int32_t k = 0; k < count; ++k) C.out_ptr[(size_t)src[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) C.out_ptr[(size_t)i + 1] += C.out_ptr[(size_t)i];
    C.out_syn.assign((size_t)count, 0);
    {
        std::vector<int64_t> fill(C.out_ptr.begin(), C.out_ptr.end() - 1);
        for (int32_t k = 0; k < count; ++k) C.out_syn[(size_t)fill[(size_t)src[k]]++] = k;
    }
    // slots: target-major, each target's in declaration order (its CheckType order)
    std::vector<int32_t> in_ptr((size_t)n + 1, 0);
    for (int32_t k = 0; k < count; ++k) in_ptr[(size_t)dst[k] + 1]++;
    for (int32_t i = 0; i < n; ++i) in_ptr[(size_t)i + 1] += in_ptr[(size_t)i];
    C.slot.assign((size_t)count, 0);
    {
        std::vector<int32_t> fill(in_ptr.begin(), in_ptr.end() - 1);
        for (int32_t k = 0; k < count; ++k) C.slot[(size_t)k] = fill[(size_t)dst[k]]++;
    }
    rc = cs_alloc(C, (void**)&C.d_in_ptr, ((size_t)n + 1) * 4);
    if (!rc) rc = cs_alloc(C, (void**)&C.d_slots, (size_t)std::max(count, 1) * 4);
    if (rc) {
        for (void* p : C.allocs) (void)hipFree(p);
        return rc;
    }
    NBX_HIP_TRY(hipMemcpy(C.d_in_ptr, in_ptr.data(), in_ptr.size() * 4, hipMemcpyHostToDevice));
    NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(C.d_slots), 0, std::max(count, 1), net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    for (void* p : net->cs.allocs) (void)hipFree(p);
    net->cs = std::move(C);
    return NBX_OK;
}

int32_t nbx_net_inject_currents(nbx_net* net, const int32_t* src, const float* current, int32_t count, int32_t t) {
    int32_t rc = check_net(net);
    if (rc) return rc;
    CurrentSynapses& C = net->cs;
    if (count < 0 || (count > 0 && (!src || !current))) {
        set_error("nbx_net_inject_currents: bad arguments");
        return NBX_ERR_INVALID;
    }
    if (t < net->t) {
        set_error("currents are injected at the current step or later");
        return NBX_ERR_INVALID;
    }
    for (int32_t k = 0; k < count; ++k) {
        if (src[k] < 0 || src[k] >= net->n) {
            set_error("current source id out of range");
            return NBX_ERR_INVALID;
        }
        if (!std::isfinite(current[k])) {
            set_error("currents must be finite");
            return NBX_ERR_INVALID;
        }
    }
    if (C.count == 0) return NBX_OK;  // no current synapses: nothing is reached
    if (C.pending.empty()) C.base = net->t;
    for (int32_t k = 0; k < count; ++k)
        for (int64_t e = C.out_ptr[(size_t)src[k]]; e < C.out_ptr[(size_t)src[k] + 1]; ++e) {
            const int32_t s = C.out_syn[(size_t)e];
            int32_t at = t + C.delay[(size_t)s] - C.base;
            if (at < 0) {  // earlier than the first pending step
                C.pending.insert(C.pending.begin(), (size_t)(-at), {});
                C.base += at;
                at = 0;
            }
            if (at >= (int32_t)C.pending.size()) C.pending.resize((size_t)at + 1);
            C.pending[(size_t)at].emplace_back(s, current[k]);
        }
    return NBX_OK;
}

}  // extern "C"
