// Monitor export in the reference's DebugMonitor directory schema
// (src/native/gpu/src/debug_monitor/DebugMonitorWriter.cpp:193-249: meta.json and eight CSV
// tables with fixed headers; MainNetworkDebugSource.cpp:95-145: main-network rows, state fields
// "v", "state_1", ...), so the reference's readers (python/src/neuronbridge/results.py) work on
// it unchanged.
//
// Two ways to fill it:
//  - nbx_net_monitor_start / _flush / _stop: the reference's SimulationDebugMonitor
//    (SimulationDebugMonitor.cpp:57-96): every step whose number is a multiple of the sample
//    interval is captured once its update is done, as a frame of device-to-device copies (state,
//    the step's firing list, the arrivals pending for the next step, weights) into frame buffers
//    on the device; the held frames are written as CSV rows every flush_interval_steps frames
//    (sooner when the buffers, at most kFrameBudget bytes, are full), at a flush and at stop.
//  - nbx_net_export_monitor: one call after a recorded run; spikes come from the recorded raster
//    (every recorded step that is a multiple of the sample interval), state, pending arrivals and
//    weights are one snapshot at export time.
// Weights of a network loaded from .nbnet are listed in file record order (the reference's
// connection order); otherwise in CSR order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "nbx_net.h"

using namespace nbx;

extern "C" int32_t nbx_net_get_state(nbx_net* net, float* x, int32_t* last);
extern "C" int32_t nbx_net_get_weights(nbx_net* net, float* w, int64_t count);

namespace {

bool ok(std::ofstream& f) { return f.is_open() && f.good(); }

constexpr size_t kFrameBudget = size_t(512) << 20;  // device bytes of the time-series frame buffers

struct Table {
    const char* file;
    const char* header;
};
// DebugMonitorWriter.cpp:227-249
const Table kTables[8] = {
    {"neuron_state.csv",
     "time_step,component_kind,component_index,component_name,global_neuron_id,local_neuron_id,field_name,value"},
    {"spikes.csv", "time_step,component_kind,component_index,component_name,global_neuron_id,local_neuron_id"},
    {"pending_channels.csv", "time_step,subnetwork_index,subnetwork_name,channel,local_neuron_id,global_neuron_id,value"},
    {"inputconv_outputs.csv", "time_step,inputconv_index,inputconv_name,output_index,value"},
    {"inputconv_inputs.csv", "time_step,inputconv_index,inputconv_name,input_index,value"},
    {"inputconv_state.csv", "time_step,inputconv_index,inputconv_name,field_name,state_index,value"},
    {"outer_dynamic_state.csv", "time_step,component_index,component_name,field_name,index,value"},
    {"weights.csv", "time_step,component_kind,component_index,component_name,synapse_id,value"},
};
enum { kState = 0, kSpikes = 1, kPending = 2, kWeights = 7 };

// Checks the descriptor, builds the neuron selection, creates the directory and writes meta.json
// and the eight tables' headers (truncating earlier contents).
int32_t open_monitor_dir(const nbx_net* net, const nbx_monitor_desc* d, std::vector<uint8_t>* sel) {
    if (d->sample_interval_steps < 1) {
        set_error("sample_interval_steps must be >= 1");
        return NBX_ERR_INVALID;
    }
    const int32_t n = net->n;
    sel->assign((size_t)n, d->neuron_ids ? 0 : 1);
    if (d->neuron_ids) {
        for (int32_t k = 0; k < d->neuron_count; ++k) {
            if (d->neuron_ids[k] < 0 || d->neuron_ids[k] >= n) {
                set_error("monitored neuron id out of range");
                return NBX_ERR_INVALID;
            }
            (*sel)[(size_t)d->neuron_ids[k]] = 1;
        }
    }
    namespace fs = std::filesystem;
    std::error_code ec;
    fs::create_directories(fs::path(d->output_dir), ec);
    if (ec) {
        set_error("failed to create debug monitor output directory: " + ec.message());
        return NBX_ERR_INVALID;
    }
    const fs::path root(d->output_dir);
    auto tf = [](int32_t v) { return v ? "true" : "false"; };
    std::ofstream meta((root / "meta.json").string(), std::ios::out | std::ios::trunc);
    if (!meta.is_open()) {
        set_error("failed to open debug monitor meta.json");
        return NBX_ERR_INVALID;
    }
    meta << "{\n";
    meta << "  \"sample_interval_steps\": " << d->sample_interval_steps << ",\n";
    meta << "  \"flush_interval_steps\": " << d->flush_interval_steps << ",\n";
    meta << "  \"record_state\": " << tf(d->record_state) << ",\n";
    meta << "  \"record_spikes\": " << tf(d->record_spikes) << ",\n";
    meta << "  \"record_weights\": " << tf(d->record_weights) << ",\n";
    meta << "  \"record_pending_channels\": " << tf(d->record_pending_channels) << ",\n";
    meta << "  \"record_inputconv_outputs\": false,\n";
    meta << "  \"record_outer_dynamic_state\": false\n";
    meta << "}\n";
    meta.flush();
    if (!ok(meta)) {
        set_error("failed to write debug monitor meta.json");
        return NBX_ERR_INVALID;
    }
    for (const Table& t : kTables) {
        std::ofstream f((root / t.file).string(), std::ios::out | std::ios::trunc);
        if (!f.is_open()) {
            set_error("failed to open one or more debug monitor output files");
            return NBX_ERR_INVALID;
        }
        f << t.header << '\n';
        f.flush();
        if (!ok(f)) {
            set_error("failed to write debug monitor output");
            return NBX_ERR_INVALID;
        }
    }
    return NBX_OK;
}

// Rows of one neuron's printable state at step `step`: its state variables, the step of its last
// update, then two unused entries reading -1 (Neuron_State_Vector.cpp:127-151).
void state_rows(std::ostream& o, int32_t step, const std::string& name, int32_t i, int32_t ns, const float* x,
                int64_t stride) {
    const int32_t np = ns + 3;
    for (int32_t f = 0; f < np; ++f) {
        const float v = f < ns ? x[(int64_t)f * stride] : f == ns ? (float)step : -1.0f;
        o << step << ",MainNetwork,0," << name << ',' << i << ',' << i << ','
          << (f == 0 ? std::string("v") : "state_" + std::to_string(f)) << ',' << v << '\n';
    }
}

// Writes the held frames of the time-series monitor and empties the frame buffers.
int32_t write_frames(nbx_net* net) {
    NetMonitor& M = net->mon;
    if (M.held == 0) return NBX_OK;
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const int32_t n = net->n, ns = net->ns, H = M.held;
    std::vector<float> st;
    std::vector<int32_t> spk, pend;
    std::vector<float> w;
    if (M.state) {
        st.resize((size_t)H * ns * n);
        NBX_HIP_TRY(hipMemcpy(st.data(), M.d_state, sizeof(float) * st.size(), hipMemcpyDeviceToHost));
    }
    if (M.spikes) {
        spk.resize((size_t)H * (n + 1));
        NBX_HIP_TRY(hipMemcpy(spk.data(), M.d_spk, sizeof(int32_t) * spk.size(), hipMemcpyDeviceToHost));
    }
    if (M.pending) {
        pend.resize((size_t)H * net->nch * n);
        NBX_HIP_TRY(hipMemcpy(pend.data(), M.d_pend, sizeof(int32_t) * pend.size(), hipMemcpyDeviceToHost));
    }
    if (M.weights) {
        w.resize((size_t)H * net->S);
        NBX_HIP_TRY(hipMemcpy(w.data(), M.d_w, sizeof(float) * w.size(), hipMemcpyDeviceToHost));
    }
    namespace fs = std::filesystem;
    const fs::path root(M.dir);
    std::ofstream out[8];
    for (int k : {kState, kSpikes, kPending, kWeights}) {
        out[k].open((root / kTables[k].file).string(), std::ios::out | std::ios::app);
        if (!out[k].is_open()) {
            set_error("failed to open one or more debug monitor output files");
            return NBX_ERR_INVALID;
        }
    }
    const float inv = std::ldexp(1.0f, -net->shift);
    const bool file_order = net->h_src_index.size() == (size_t)net->S;
    std::vector<float> by_file;
    std::vector<int32_t> ids;
    for (int32_t f = 0; f < H; ++f) {
        const int32_t s = M.steps[(size_t)f];
        if (M.spikes) {
            const int32_t* row = spk.data() + (size_t)f * (n + 1);
            const int32_t cnt = std::min(std::max(row[0], 0), n);
            ids.assign(row + 1, row + 1 + cnt);
            std::sort(ids.begin(), ids.end());
            for (int32_t id : ids)
                if (id >= 0 && id < n && M.sel[(size_t)id])
                    out[kSpikes] << s << ",MainNetwork,0," << M.name << ',' << id << ',' << id << '\n';
        }
        if (M.state) {
            const float* x = st.data() + (size_t)f * ns * n;
            for (int32_t i = 0; i < n; ++i)
                if (M.sel[(size_t)i]) state_rows(out[kState], s, M.name, i, ns, x + i, n);
        }
        if (M.pending) {
            const int32_t* r = pend.data() + (size_t)f * net->nch * n;
            for (int32_t ch = 0; ch < net->nch; ++ch)
                for (int32_t i = 0; i < n; ++i)
                    if (M.sel[(size_t)i])
                        out[kPending] << s << ",0," << M.name << ',' << ch << ',' << i << ',' << i << ','
                                      << (float)r[(size_t)ch * n + i] * inv << '\n';
        }
        if (M.weights) {
            const float* wf = w.data() + (size_t)f * net->S;
            if (file_order) {
                by_file.resize((size_t)net->S);
                for (size_t j = 0; j < (size_t)net->S; ++j) by_file[net->h_src_index[j]] = wf[j];
                wf = by_file.data();
            }
            for (int64_t j = 0; j < net->S; ++j)
                out[kWeights] << s << ",MainNetwork,0," << M.name << ',' << j << ',' << wf[j] << '\n';
        }
    }
    for (int k : {kState, kSpikes, kPending, kWeights}) {
        out[k].flush();
        if (!ok(out[k])) {
            set_error("failed to write debug monitor output");
            return NBX_ERR_INVALID;
        }
    }
    M.frames_written += H;
    M.held = 0;
    M.steps.clear();
    return NBX_OK;
}

void free_monitor(nbx_net* net) {
    NetMonitor& M = net->mon;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (void* p : M.allocs) (void)hipFree(p);
    M = NetMonitor{};
}

}  // namespace

namespace nbx {

int32_t monitor_capture(nbx_net* net, int32_t step) {
    NetMonitor& M = net->mon;
    if (!M.on) return NBX_OK;
    const int32_t n = net->n, ns = net->ns, f = M.held;
    hipStream_t st = net->stream;
    if (M.state) {
        float* dst = M.d_state + (size_t)f * ns * n;
        // columns a population's model does not have read 0
        NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(dst), 0, (int64_t)ns * n, st));
        for (const NetPop& pp : net->pops) {
            const std::vector<float*> cols = state_columns(net, pp.md.model);
            for (size_t j = 0; j < cols.size() && (int32_t)j < ns; ++j)
                if (cols[j])
                    NBX_HIP_TRY(hipMemcpyAsync(dst + (size_t)j * n + pp.begin, cols[j] + pp.begin,
                                               sizeof(float) * (size_t)pp.md.n, hipMemcpyDeviceToDevice, st));
        }
    }
    if (M.spikes) {
        const int32_t slot = step % net->D;
        int32_t* dst = M.d_spk + (size_t)f * (n + 1);
        NBX_HIP_TRY(hipMemcpyAsync(dst, net->d_hist_cnt + slot, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        NBX_HIP_TRY(hipMemcpyAsync(dst + 1, net->d_hist + (size_t)slot * n, sizeof(int32_t) * (size_t)n,
                                   hipMemcpyDeviceToDevice, st));
    }
    if (M.pending) {
        const int32_t slot = (step + 1) % net->D;
        NBX_HIP_TRY(hipMemcpyAsync(M.d_pend + (size_t)f * net->nch * n, net->d_ring + (size_t)slot * net->nch * n,
                                   sizeof(int32_t) * (size_t)net->nch * n, hipMemcpyDeviceToDevice, st));
    }
    if (M.weights)
        NBX_HIP_TRY(hipMemcpyAsync(M.d_w + (size_t)f * net->S, net->d_w, sizeof(float) * (size_t)net->S,
                                   hipMemcpyDeviceToDevice, st));
    M.steps.push_back(step);
    M.held += 1;
    if (M.held == M.cap) return write_frames(net);
    return NBX_OK;
}

}  // namespace nbx

extern "C" int32_t nbx_net_export_monitor(nbx_net* net, const nbx_monitor_desc* d) {
    if (!net || !d || !d->output_dir) {
        set_error("nbx_net_export_monitor: null argument");
        return NBX_ERR_INVALID;
    }
    if (!net->finalized) {
        set_error("network is not finalized");
        return NBX_ERR_STATE;
    }
    if (net->phase != 0) {
        set_error("a per-phase step is open (finish it with nbx_net_commit_firings)");
        return NBX_ERR_STATE;
    }
    const int32_t n = net->n;
    std::vector<uint8_t> sel;
    {
        const int32_t rc = open_monitor_dir(net, d, &sel);
        if (rc) return rc;
    }
    namespace fs = std::filesystem;
    const fs::path root(d->output_dir);
    const std::string name = d->component_name ? d->component_name : "main";
    std::ofstream out[8];
    for (int k : {kState, kSpikes, kPending, kWeights}) {
        out[k].open((root / kTables[k].file).string(), std::ios::out | std::ios::app);
        if (!out[k].is_open()) {
            set_error("failed to open one or more debug monitor output files");
            return NBX_ERR_INVALID;
        }
    }
    std::ofstream &state_out = out[0], &spikes_out = out[1], &pending_out = out[2], &weights_out = out[7];
    NBX_HIP_TRY(hipSetDevice(net->device));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const int32_t snap = net->t - 1;  // last completed step

    if (d->record_spikes && net->d_raster && net->rec_n > 0) {
        std::vector<uint32_t> r((size_t)net->rec_n * net->words);
        NBX_HIP_TRY(hipMemcpy(r.data(), net->d_raster, sizeof(uint32_t) * r.size(), hipMemcpyDeviceToHost));
        const int32_t t_end = std::min(net->rec_t0 + net->rec_n, net->t);
        for (int32_t t = net->rec_t0; t < t_end; ++t) {
            if (t % d->sample_interval_steps != 0) continue;
            const uint32_t* row = r.data() + (size_t)(t - net->rec_t0) * net->words;
            for (int32_t w = 0; w < net->words; ++w)
                for (uint32_t m = row[w]; m; m &= m - 1u) {
                    const int32_t id = w * 32 + __builtin_ctz(m);
                    if (id < n && sel[(size_t)id])
                        spikes_out << t << ",MainNetwork,0," << name << ',' << id << ',' << id << '\n';
                }
        }
    }
    if (d->record_state && snap >= 0) {
        std::vector<float> x((size_t)n * net->ns);
        std::vector<int32_t> last((size_t)n);
        const int32_t rc = nbx_net_get_state(net, x.data(), last.data());
        if (rc) return rc;
        for (int32_t i = 0; i < n; ++i)
            if (sel[(size_t)i]) state_rows(state_out, snap, name, i, net->ns, x.data() + (size_t)i * net->ns, 1);
    }
    if (d->record_pending_channels) {
        // arrivals already in the delay ring for the next step, per channel
        const int32_t slot = net->t % net->D;
        std::vector<int32_t> ring((size_t)net->nch * n);
        NBX_HIP_TRY(hipMemcpy(ring.data(), net->d_ring + (size_t)slot * net->nch * n, sizeof(int32_t) * ring.size(),
                              hipMemcpyDeviceToHost));
        const float inv = std::ldexp(1.0f, -net->shift);
        for (int32_t ch = 0; ch < net->nch; ++ch)
            for (int32_t i = 0; i < n; ++i)
                if (sel[(size_t)i])
                    pending_out << net->t << ",0," << name << ',' << ch << ',' << i << ',' << i << ','
                                << (float)ring[(size_t)ch * n + i] * inv << '\n';
    }
    if (d->record_weights && net->has_sparse && snap >= 0) {
        std::vector<float> w((size_t)net->S);
        const int32_t rc = nbx_net_get_weights(net, w.data(), net->S);
        if (rc) return rc;
        if (net->h_src_index.size() == w.size()) {  // file record order
            std::vector<float> by_file(w.size());
            for (size_t j = 0; j < w.size(); ++j) by_file[net->h_src_index[j]] = w[j];
            w.swap(by_file);
        }
        for (size_t j = 0; j < w.size(); ++j)
            weights_out << snap << ",MainNetwork,0," << name << ',' << j << ',' << w[j] << '\n';
    }
    for (int k : {kState, kSpikes, kPending, kWeights}) {
        out[k].flush();
        if (!ok(out[k])) {
            set_error("failed to write debug monitor output");
            return NBX_ERR_INVALID;
        }
    }
    return NBX_OK;
}

extern "C" int32_t nbx_net_monitor_start(nbx_net* net, const nbx_monitor_desc* d) {
    if (!net || !d || !d->output_dir) {
        set_error("nbx_net_monitor_start: null argument");
        return NBX_ERR_INVALID;
    }
    if (!net->finalized) {
        set_error("network is not finalized");
        return NBX_ERR_STATE;
    }
    if (net->mon.on) {
        set_error("the monitor is already running (nbx_net_monitor_stop first)");
        return NBX_ERR_STATE;
    }
    if (d->record_weights && !net->has_sparse) {
        set_error("record_weights needs sparse synapses");
        return NBX_ERR_INVALID;
    }
    NBX_HIP_TRY(hipSetDevice(net->device));
    NetMonitor M;
    {
        const int32_t rc = open_monitor_dir(net, d, &M.sel);
        if (rc) return rc;
    }
    M.dir = d->output_dir;
    M.name = d->component_name ? d->component_name : "main";
    M.interval = d->sample_interval_steps;
    M.flush_every = d->flush_interval_steps > 0 ? d->flush_interval_steps : 100;  // the reference's default
    M.spikes = d->record_spikes != 0;
    M.state = d->record_state != 0;
    M.pending = d->record_pending_channels != 0;
    M.weights = d->record_weights != 0;
    const size_t n = (size_t)net->n;
    const size_t frame = (M.state ? (size_t)net->ns * n * 4 : 0) + (M.spikes ? (n + 1) * 4 : 0) +
                         (M.pending ? (size_t)net->nch * n * 4 : 0) + (M.weights ? (size_t)net->S * 4 : 0);
    M.cap = (int32_t)std::max<size_t>(1, std::min<size_t>((size_t)M.flush_every, frame ? kFrameBudget / frame : 1));
    auto get = [&](auto** p, size_t count) -> int32_t {
        void* q = nullptr;
        NBX_HIP_TRY(hipMalloc(&q, std::max<size_t>(count, 1)));
        M.allocs.push_back(q);
        *p = static_cast<std::remove_reference_t<decltype(*p)>>(q);
        return NBX_OK;
    };
    int32_t rc = NBX_OK;
    if (M.state) rc = get(&M.d_state, (size_t)M.cap * net->ns * n * 4);
    if (!rc && M.spikes) rc = get(&M.d_spk, (size_t)M.cap * (n + 1) * 4);
    if (!rc && M.pending) rc = get(&M.d_pend, (size_t)M.cap * net->nch * n * 4);
    if (!rc && M.weights) rc = get(&M.d_w, (size_t)M.cap * (size_t)net->S * 4);
    if (rc) {
        for (void* p : M.allocs) (void)hipFree(p);
        return rc;
    }
    M.on = true;
    net->mon = std::move(M);
    return NBX_OK;
}

extern "C" int32_t nbx_net_monitor_flush(nbx_net* net) {
    if (!net) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    if (!net->mon.on) {
        set_error("the monitor is not running");
        return NBX_ERR_STATE;
    }
    NBX_HIP_TRY(hipSetDevice(net->device));
    return write_frames(net);
}

extern "C" int32_t nbx_net_monitor_stop(nbx_net* net) {
    if (!net) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    if (!net->mon.on) return NBX_OK;
    NBX_HIP_TRY(hipSetDevice(net->device));
    const int32_t rc = write_frames(net);
    free_monitor(net);
    return rc;
}

extern "C" int32_t nbx_net_monitor_frames(const nbx_net* net, int64_t* written, int32_t* held) {
    if (!net || !written || !held) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    *written = net->mon.frames_written;
    *held = net->mon.held;
    return NBX_OK;
}
