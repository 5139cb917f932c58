// Monitor export in the reference's DebugMonitor directory schema
// (src/native/gpu/src/debug_monitor/DebugMonitorWriter.cpp:193-235: meta.json and eight CSV
// tables with fixed headers; MainNetworkDebugSource.cpp:95-145: main-network rows, state fields
// "v", "state_1", ...), so the reference's readers (python/src/neuronbridge/results.py) work on
// it unchanged. Spikes come from the raster recorded on the device (every recorded step that is a
// multiple of the sample interval); state, pending arrivals and weights are one snapshot at
// export time. Weights of a network loaded from .nbnet are listed in file record order (the
// reference's connection order); otherwise in CSR order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "nbx_net.h"

using namespace nbx;

extern "C" int32_t nbx_net_get_state(nbx_net* net, float* x, int32_t* last);
extern "C" int32_t nbx_net_get_weights(nbx_net* net, float* w, int64_t count);

namespace {

bool ok(std::ofstream& f) { return f.is_open() && f.good(); }

}  // namespace

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
    if (d->sample_interval_steps < 1) {
        set_error("sample_interval_steps must be >= 1");
        return NBX_ERR_INVALID;
    }
    const int32_t n = net->n;
    std::vector<uint8_t> sel((size_t)n, d->neuron_ids ? 0 : 1);
    if (d->neuron_ids) {
        for (int32_t k = 0; k < d->neuron_count; ++k) {
            if (d->neuron_ids[k] < 0 || d->neuron_ids[k] >= n) {
                set_error("monitored neuron id out of range");
                return NBX_ERR_INVALID;
            }
            sel[(size_t)d->neuron_ids[k]] = 1;
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
    const std::string name = d->component_name ? d->component_name : "main";
    auto tf = [](int32_t v) { return v ? "true" : "false"; };
    {
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
    }
    struct Table {
        const char* file;
        const char* header;
    };
    static const Table tables[] = {
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
    std::ofstream out[8];
    for (int k = 0; k < 8; ++k) {
        out[k].open((root / tables[k].file).string(), std::ios::out | std::ios::trunc);
        if (!out[k].is_open()) {
            set_error("failed to open one or more debug monitor output files");
            return NBX_ERR_INVALID;
        }
        out[k] << tables[k].header << '\n';
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
        // printable values of a time-driven neuron: its state variables, the step of its last
        // update, then two unused entries reading -1 (Neuron_State_Vector.cpp:127-151)
        const int32_t np = net->ns + 3;
        for (int32_t i = 0; i < n; ++i) {
            if (!sel[(size_t)i]) continue;
            for (int32_t f = 0; f < np; ++f) {
                const float v = f < net->ns ? x[(size_t)i * net->ns + f] : f == net->ns ? (float)snap : -1.0f;
                state_out << snap << ",MainNetwork,0," << name << ',' << i << ',' << i << ','
                          << (f == 0 ? std::string("v") : "state_" + std::to_string(f)) << ',' << v << '\n';
            }
        }
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
    for (auto& f : out) {
        f.flush();
        if (!ok(f)) {
            set_error("failed to write debug monitor output");
            return NBX_ERR_INVALID;
        }
    }
    return NBX_OK;
}
