// Host side of the C ABI declared in include/nbx_core.h: validation, device
// allocation and upload, engine orchestration (per-phase launches, hipGraph
// capture of many steps, the persistent engine) and downloads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "nbx_net.h"

namespace nbx {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

}  // namespace nbx

using namespace nbx;

#define NBX_CHECK(cond, msg)                     \
    do {                                         \
        if (!(cond)) {                           \
            ::nbx::set_error(msg);               \
            return NBX_ERR_INVALID;              \
        }                                        \
    } while (0)

namespace {

// exp(-dt * (1/tau)) with float arithmetic and libm expf, the reference's lookup value
// (TimeDrivenLIF_Exponential_Decay.cpp:37-39).
float decay_factor(float dt, float tau) {
    const float inv = (tau != 0.0f) ? (1.0f / tau) : 0.0f;
    return std::exp(-dt * inv);
}

// TimeDrivenLIF_Exponential_triple.cpp:56-60 divides instead of multiplying by 1/tau
float decay_factor_div(float dt, float tau) { return std::exp(-dt / tau); }

int32_t model_nstate(int32_t model) {
    switch (model) {
        case NBX_MODEL_LIF_VJ: return 2;
        case NBX_MODEL_LIF_EXP: return 3;
        case NBX_MODEL_LIF_EXP2: return 4;
        case NBX_MODEL_LIF_EXP3: return 5;
        case NBX_MODEL_POISSON: return 2;
        case NBX_MODEL_TRIGGER_RELAY: return 1;
        default: return 6;
    }
}

// Device array behind each column of the reference's AoS state of `model` (null: a column this
// build does not keep, read as 0)
std::vector<float*> model_columns(nbx_net* net, int32_t model) {
    switch (model) {
        case NBX_MODEL_LIF_VJ: return {net->d_v, net->d_iext};
        case NBX_MODEL_LIF_EXP: return {net->d_v, net->d_s1, net->d_iext};
        case NBX_MODEL_LIF_EXP2: return {net->d_v, net->d_ga, net->d_gg, net->d_iext};
        case NBX_MODEL_LIF_EXP3: return {net->d_v, net->d_ga, net->d_gg, net->d_s1, net->d_iext};
        case NBX_MODEL_POISSON: return {net->d_v, net->d_iext};
        case NBX_MODEL_TRIGGER_RELAY: return {net->d_v};
        case NBX_MODEL_IZH: return {net->d_v, net->d_s1, net->d_ga, net->d_gg, net->d_g3, net->d_iext};
        default: return {net->d_v, net->d_s1, net->d_ga, net->d_gg, nullptr, net->d_iext};
    }
}

// Parameters of one population as the kernels take them.
ModelParams make_params(const nbx_model_desc* md, int32_t begin) {
    ModelParams P{};
    P.model = md->model;
    P.integrator = md->integrator;
    P.n = md->n;
    P.begin = begin;
    P.t_ref = md->t_ref;
    P.dt = md->dt;
    P.hdt = md->dt * 0.5f;
    P.dt6 = md->dt / 6.0f;
    P.v_rest = md->v_rest;
    P.tau = md->tau;
    P.v_th = md->v_th;
    P.r = md->r;
    P.v_reset = md->v_reset;
    P.e_rev = md->e_rev;
    P.izh_v_th = md->izh_v_th;
    P.izh_r = md->izh_r;
    P.e_ampa = md->e_ampa;
    P.e_gaba = md->e_gaba;
    P.g_decay = decay_factor(md->dt, md->g_tau);
    P.ampa_decay = decay_factor(md->dt, md->ampa_tau);
    P.gaba_decay = decay_factor(md->dt, md->gaba_tau);
    // Izhikevich NMDA: the single-entry lookup exp(-dt * (1/nmda_tau))
    // (TimeDrivenIzhikevic_Exponential_Decay.cpp:32-39)
    P.nmda_decay = md->model == NBX_MODEL_IZH ? decay_factor(md->dt, md->nmda_tau) : 1.0f;
    if (md->model == NBX_MODEL_LIF_EXP3) {
        P.ampa_decay = decay_factor_div(md->dt, md->ampa_tau);
        P.gaba_decay = decay_factor_div(md->dt, md->gaba_tau);
        P.nmda_decay = decay_factor_div(md->dt, md->nmda_tau);
    }
    P.rate_bias = md->rate_bias_hz;
    P.rate_gain = md->rate_gain_hz;
    P.cond = 0;
    return P;
}

const char* model_desc_error(const nbx_model_desc* md) {
    if (!(md->model >= NBX_MODEL_LIF_VJ && md->model <= NBX_MODEL_TRIGGER_RELAY)) return "unknown model id";
    if (!(md->integrator == NBX_EULER || md->integrator == NBX_RK4)) return "unknown integrator";
    if (!(md->n > 0 && md->n < kMaxNeurons)) return "neuron count must be in [1, 2^24)";
    if (!(md->dt > 0.0f)) return "dt must be positive";
    const bool izh = md->model == NBX_MODEL_IZH || md->model == NBX_MODEL_IZH_CUR;
    if (izh && !(md->izh_a && md->izh_b && md->izh_c && md->izh_d)) return "Izhikevich models need per-neuron a,b,c,d";
    const bool lif = md->model == NBX_MODEL_LIF_VJ || md->model == NBX_MODEL_LIF_EXP ||
                     md->model == NBX_MODEL_LIF_EXP2 || md->model == NBX_MODEL_LIF_EXP3;
    if (lif && md->tau == 0.0f) return "LIF tau must be non-zero";
    if (md->model == NBX_MODEL_LIF_EXP3 && !(md->ampa_tau != 0.0f && md->gaba_tau != 0.0f && md->nmda_tau != 0.0f))
        return "LIF_EXP3 conductance time constants must be non-zero";
    return nullptr;
}

// Trace decay tables by integer step delta: exp(-(delta * dt) * (1/tau)) in float, the arithmetic
// of STDP_State::SetUpdate (STDP_State.cpp:17-26); the last entry is 0 or covers every delta a
// run can reach.
void trace_tables(float dt, float tau_a, float tau_b, std::vector<float>* ta, std::vector<float>* tb) {
    const float inv_a = (float)(1.0 / (double)tau_a);
    const float inv_b = (float)(1.0 / (double)tau_b);
    ta->clear();
    tb->clear();
    for (int32_t k = 0;; ++k) {
        const float dT = (float)k * dt;
        ta->push_back(std::exp(-dT * inv_a));
        tb->push_back(std::exp(-dT * inv_b));
        if ((ta->back() == 0.0f && tb->back() == 0.0f) || k >= (1 << 22)) break;
    }
}

// Cerebellar kernel (CerebullarLearningRule.cpp:56-88): the window is grown from maxpos in
// kernel steps until the kernel falls below 1e-2, then shifted by initpos; all in float.
void cerebellar_kernel(const nbx_rule_desc& q, std::vector<float>* tab, float* maxtime, float* inv_step) {
    float maxpos = q.maxpos;
    if (maxpos <= q.initpos) maxpos = q.initpos + 1e-6f;
    const float width = maxpos - q.initpos;
    float mt = maxpos;
    for (;;) {
        const float value = (1.0f / width) * mt * std::exp(-(mt / width) + 1.0f);
        if (value < 1e-2f) break;
        mt += q.kernel_step;
    }
    mt += q.initpos;
    *maxtime = mt;
    *inv_step = 1.0f / q.kernel_step;
    const int32_t len = (int32_t)(mt * *inv_step) + 1;
    tab->resize((size_t)len);
    for (int32_t i = 0; i < len; ++i) {
        const float time = (float)i * q.kernel_step;
        if (time < q.initpos) {
            (*tab)[(size_t)i] = 0.0f;
        } else {
            const float shifted = time - q.initpos;
            (*tab)[(size_t)i] = (1.0f / width) * shifted * std::exp(-(shifted / width) + 1.0f);
        }
    }
}

// Device tables of the trigger-driven rules: rule records with their trace tables, per-synapse
// rule / trigger ids and weight bounds, rule-plastic inputs per (post, rule) and trigger synapses
// per post, both in ascending synapse order.
int32_t finalize_rules(nbx_net* net) {
    const int32_t n = net->n, nr = (int32_t)net->rules.size();
    const size_t S = (size_t)net->S;
    std::vector<RuleDev> rd((size_t)nr);
    std::vector<float> tabs;
    for (int32_t r = 0; r < nr; ++r) {
        const nbx_rule_desc& q = net->rules[(size_t)r];
        RuleDev& d = rd[(size_t)r];
        d.kind = q.kind;
        d.clear = q.clear_eligibility ? 1 : 0;
        d.max_ltp = q.max_ltp;
        d.max_ltd = q.max_ltd;
        d.reward = q.reward;
        d.punishment = q.punishment;
        d.a1pre = q.a1pre;
        d.a2prepre = q.a2prepre;
        std::vector<float> ta, tb;  // the additive rule's state decays both traces with ltp_tau
        trace_tables(net->md.dt, q.ltp_tau, q.kind == NBX_RULE_ADDITIVE ? q.ltp_tau : q.ltd_tau, &ta, &tb);
        d.tab_off = (int32_t)tabs.size();
        d.tab_len = (int32_t)ta.size();
        tabs.insert(tabs.end(), ta.begin(), ta.end());
        tabs.insert(tabs.end(), tb.begin(), tb.end());
        if (q.kind == NBX_RULE_RSTDP) net->rstdp = true;
        d.min_weight = q.min_weight;
        d.seed = q.seed;
        d.kern_off = d.kern_len = 0;
        d.maxtime = d.inv_step = 0.0f;
        if (q.kind == NBX_RULE_CEREBELLAR) {
            std::vector<float> kt;
            cerebellar_kernel(q, &kt, &d.maxtime, &d.inv_step);
            d.kern_off = (int32_t)tabs.size();
            d.kern_len = (int32_t)kt.size();
            tabs.insert(tabs.end(), kt.begin(), kt.end());
            net->ccap = std::max(net->ccap, q.history > 0 ? q.history : 16);
        }
    }
    NBX_PASS(net->upload(&net->d_rules, rd.data(), rd.size()));
    NBX_PASS(net->upload(&net->d_rtab, tabs.data(), tabs.size()));
    NBX_PASS(net->upload(&net->d_syn_rule, net->h_rule.data(), S));
    NBX_PASS(net->upload(&net->d_syn_trig, net->h_trig.data(), S));
    std::vector<int32_t> stamp(std::max(S, (size_t)n), -1);
    NBX_PASS(net->upload(&net->d_tstamp, stamp.data(), S));
    NBX_PASS(net->upload(&net->d_tpost_stamp, stamp.data(), (size_t)n));
    NBX_PASS(net->alloc(&net->d_tposts, (size_t)2 * n));
    const int32_t zero2[2] = {0, 0};
    NBX_PASS(net->upload(&net->d_tcount, zero2, 2));
    std::vector<int64_t> rin_ptr((size_t)n * nr + 1, 0), tin_ptr((size_t)n + 1, 0);
    for (size_t j = 0; j < S; ++j) {
        const int32_t post = syn_post(net->h_syn[j]);
        if (net->h_rule[j] >= 0) rin_ptr[(size_t)post * nr + net->h_rule[j] + 1]++;
        if (net->h_trig[j] >= 0) tin_ptr[(size_t)post + 1]++;
    }
    for (size_t k = 1; k < rin_ptr.size(); ++k) rin_ptr[k] += rin_ptr[k - 1];
    for (size_t k = 1; k < tin_ptr.size(); ++k) tin_ptr[k] += tin_ptr[k - 1];
    std::vector<int64_t> rfill(rin_ptr.begin(), rin_ptr.end() - 1), tfill(tin_ptr.begin(), tin_ptr.end() - 1);
    std::vector<int64_t> rin_syn((size_t)rin_ptr.back() + 1), tin_syn((size_t)tin_ptr.back() + 1);
    for (size_t j = 0; j < S; ++j) {  // ascending j within every list
        const int32_t post = syn_post(net->h_syn[j]);
        if (net->h_rule[j] >= 0) rin_syn[(size_t)rfill[(size_t)post * nr + net->h_rule[j]]++] = (int64_t)j;
        if (net->h_trig[j] >= 0) tin_syn[(size_t)tfill[(size_t)post]++] = (int64_t)j;
    }
    if (net->ccap > 0) {  // cerebellar arrival histories, one row per cerebellar synapse
        std::vector<int32_t> cidx(S, -1), conn(S);
        int32_t rows = 0;
        for (size_t j = 0; j < S; ++j) {
            conn[j] = net->h_src_index.size() == S ? (int32_t)net->h_src_index[j] : (int32_t)j;
            if (net->h_rule[j] >= 0 && rd[(size_t)net->h_rule[j]].kind == NBX_RULE_CEREBELLAR) cidx[j] = rows++;
        }
        net->crows = rows;
        NBX_PASS(net->upload(&net->d_cidx, cidx.data(), S));
        NBX_PASS(net->upload(&net->d_conn_index, conn.data(), S));
        std::vector<int32_t> zero((size_t)std::max(rows, 1), 0);
        NBX_PASS(net->upload(&net->d_chead, zero.data(), zero.size()));
        NBX_PASS(net->alloc(&net->d_chist, (size_t)std::max(rows, 1) * net->ccap));
    }
    NBX_PASS(net->upload(&net->d_rin_ptr, rin_ptr.data(), rin_ptr.size()));
    NBX_PASS(net->upload(&net->d_rin_syn, rin_syn.data(), rin_syn.size()));
    NBX_PASS(net->upload(&net->d_tin_ptr, tin_ptr.data(), tin_ptr.size()));
    NBX_PASS(net->upload(&net->d_tin_syn, tin_syn.data(), tin_syn.size()));
    return NBX_OK;
}

// Per-synapse learning records in their initial state (traces 0, last update 0, the weight bounds):
// one 32-byte record per synapse for networks with learning rules (STDP_State traces shared with
// STDP), one 16-byte head for pair STDP alone. alloc: allocate (finalize), else overwrite (reset).
int32_t upload_learning_records(nbx_net* net, bool alloc) {
    const size_t S = (size_t)net->S;
    if (!net->rules.empty()) {
        std::vector<SynTrace> tr(S);
        for (size_t j = 0; j < S; ++j)
            tr[j] = SynTrace{0.0f, 0.0f, 0, net->h_smaxw.empty() ? net->sd.w_max : net->h_smaxw[j], 0.0f,
                             net->h_maxw.empty() ? 0.0f : net->h_maxw[j], 0, 0};
        if (alloc) {
            NBX_PASS(net->upload(&net->d_trace, tr.data(), S));
            net->d_heads = reinterpret_cast<TraceHead*>(net->d_trace);
            net->head_stride = (int32_t)(sizeof(SynTrace) / sizeof(TraceHead));
        } else {
            NBX_HIP_TRY(hipMemcpy(net->d_trace, tr.data(), sizeof(SynTrace) * S, hipMemcpyHostToDevice));
        }
    } else if (net->stdp) {
        std::vector<TraceHead> hd(S);
        for (size_t j = 0; j < S; ++j) hd[j] = TraceHead{0.0f, 0.0f, 0, net->h_smaxw.empty() ? net->sd.w_max : net->h_smaxw[j]};
        if (alloc) {
            NBX_PASS(net->upload(&net->d_heads, hd.data(), S));
            net->head_stride = 1;
        } else {
            NBX_HIP_TRY(hipMemcpy(net->d_heads, hd.data(), sizeof(TraceHead) * S, hipMemcpyHostToDevice));
        }
    }
    return NBX_OK;
}

int32_t poisson_table(double lambda, std::vector<uint32_t>* tab) {
    tab->clear();
    double p = std::exp(-lambda), cdf = 0.0;
    for (int k = 0; k < 64; ++k) {
        cdf += p;
        const double v = std::floor(cdf * 4294967296.0);
        const uint32_t th = v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
        tab->push_back(th);
        if (th == 0xFFFFFFFFu) return k;
        p = p * lambda / (double)(k + 1);
    }
    return 63;
}

int32_t device_scope(nbx_net* net) {
    NBX_HIP_TRY(hipSetDevice(net->device));
    return NBX_OK;
}

int32_t require_finalized(const nbx_net* net) {
    if (!net) {
        set_error("null network handle");
        return NBX_ERR_INVALID;
    }
    if (!net->finalized) {
        set_error("network is not finalized (call nbx_net_finalize)");
        return NBX_ERR_STATE;
    }
    return NBX_OK;
}

int32_t check_device_error(nbx_net* net) {
    int32_t e = 0;
    NBX_HIP_TRY(hipMemcpy(&e, net->d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (e & kErrHistory) {
        set_error("cerebellar arrival history overflow: an arrival still inside the kernel window was "
                  "overwritten (raise nbx_rule_desc.history)");
        return NBX_ERR_STATE;
    }
    if (e & kErrInput) {
        NBX_PASS(clear_error_bits(net, kErrInput));
        set_error("device interface values could overflow the int32 arrival sums; those values were dropped");
        return NBX_ERR_STATE;
    }
    if (e & kErrResidency) {
        set_error("persistent engine: its workgroups could not all be resident at once (other work held compute "
                  "units); the run took no step on the device; nbx_net_reset_state clears the error");
        return NBX_ERR_STATE;
    }
    if (e & kErrRaster) {  // an argument error of one replay call: reported once
        NBX_PASS(clear_raster_error(net));
        set_error("replay STDP: invalid raster (a spike id out of range, a neuron listed twice in one step, "
                  "tick_ptr not ascending or more spikes than nspikes); no synapse was changed");
        return NBX_ERR_INVALID;
    }
    if (e != 0) {
        set_error("device-side error flag " + std::to_string(e) + " (a wait of the persistent engine timed out)");
        return NBX_ERR_STATE;
    }
    return NBX_OK;
}

// Plain STDP (no rules): the post-side updates of step k run inside the pull of step k + 1,
// so a sequence of steps ends with one post-side launch of its own.
bool merged_stdp(const nbx_net* net) {
    return net->stdp && net->d_in_ptr && net->d_fmask && net->rules.empty() && net->pull_mode && net->has_sparse;
}

// Launch the kernels of one step; tbase == nullptr means k is the absolute step. first / last:
// the step opens / closes a sequence (the post-side updates of the step before the sequence are
// done, those of its last step are launched at its end).
hipError_t enqueue_step(nbx_net* net, const StepArgs& A, const int32_t* tbase, int32_t k, bool first, bool last) {
    hipError_t e = hipSuccess;
    const bool merged = merged_stdp(net);
    if (net->pull_mode && net->has_sparse) {
        e = (merged && !first) ? launch_pull_ltp(A, tbase, k, net->stream) : launch_pull(A, tbase, k, net->stream);
        if (e != hipSuccess) return e;
    }
    if (!net->rules.empty()) {
        e = launch_trigger(A, tbase, k, net->stream);
        if (e != hipSuccess) return e;
    }
    e = launch_update(A, tbase, k, net->stream);
    if (e != hipSuccess) return e;
    if (((net->stdp && net->d_in_ptr) || net->rstdp) && (!merged || last)) {
        e = launch_ltp(A, tbase, k, net->stream);
        if (e != hipSuccess) return e;
    }
    if (!net->pull_mode && net->has_sparse) {
        e = launch_deliver_push(A, tbase, k, net->stream);
        if (e != hipSuccess) return e;
    }
    if (net->has_dense) e = launch_deliver_dense(A, tbase, k, net->stream);
    return e;
}

constexpr int32_t kGraphChunk = 64;

int32_t run_graph(nbx_net* net, int32_t nsteps) {
    const StepArgs A = net->args();
    int32_t done = 0;
    if (nsteps >= kGraphChunk) {
        if (!net->gexec) {
            hipGraph_t g = nullptr;
            NBX_HIP_TRY(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (int32_t k = 0; k < kGraphChunk && e == hipSuccess; ++k)
                e = enqueue_step(net, A, net->d_tbase, k, k == 0, k == kGraphChunk - 1);
            if (e == hipSuccess) e = launch_advance(net->d_tbase, kGraphChunk, net->stream);
            hipError_t ec = hipStreamEndCapture(net->stream, &g);
            if (e != hipSuccess || ec != hipSuccess) {
                if (g) hipGraphDestroy(g);
                set_error(std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : ec));
                return NBX_ERR_HIP;
            }
            hipError_t ei = hipGraphInstantiate(&net->gexec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            NBX_HIP_TRY(ei);
            net->gchunk = kGraphChunk;
        }
        NBX_HIP_TRY(hipMemcpyAsync(net->d_tbase, &net->t, sizeof(int32_t), hipMemcpyHostToDevice, net->stream));
        // the host value must stay alive until the copy has run
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        while (nsteps - done >= net->gchunk) {
            NBX_HIP_TRY(hipGraphLaunch(net->gexec, net->stream));
            done += net->gchunk;
        }
    }
    for (const int32_t rest = done; done < nsteps; ++done)
        NBX_HIP_TRY(enqueue_step(net, A, nullptr, net->t + done, done == rest, done == nsteps - 1));
    net->t += nsteps;
    return NBX_OK;
}

int32_t run_kernels(nbx_net* net, int32_t nsteps) {
    const StepArgs A = net->args();
    for (int32_t k = 0; k < nsteps; ++k) NBX_HIP_TRY(enqueue_step(net, A, nullptr, net->t + k, k == 0, k == nsteps - 1));
    net->t += nsteps;
    return NBX_OK;
}

// Largest sum of |contributions| that can land in one (step, channel, neuron) for any spike
// pattern: every synapse into the neuron (plastic ones at their weight bound), the dense column and
// the largest Poisson count. `w` holds the sparse weights in CSR order.
double arrival_bound(const nbx_net* net, const float* w_csr) {
    std::vector<double> bound((size_t)net->n, 0.0);
    if (net->has_sparse) {
        for (int64_t j = 0; j < net->S; ++j) {
            double w = std::fabs((double)w_csr[(size_t)j]);
            if (net->stdp && !net->h_plastic.empty() && net->h_plastic[(size_t)j])
                w = std::max(w, net->h_smaxw.empty() ? (double)net->sd.w_max : std::fabs((double)net->h_smaxw[(size_t)j]));
            if (!net->h_rule.empty() && net->h_rule[(size_t)j] >= 0)
                w = std::max(w, std::fabs((double)net->h_maxw[(size_t)j]));
            bound[(size_t)syn_post(net->h_syn[(size_t)j])] += w;
        }
    }
    if (net->has_dense) {
        for (int32_t p = 0; p < net->n; ++p) {
            const float* col = net->h_wt.data() + (size_t)p * net->n;
            for (int32_t i = 0; i < net->n; ++i) bound[(size_t)i] += std::fabs((double)col[i]);
        }
    }
    double B = 0.0;
    for (double b : bound) B = std::max(B, b);
    if (net->pois) B += (double)net->kmax * std::fabs((double)net->w_bg);
    return B;
}

// Largest shift (at most 24) whose int32 sums cannot overflow for the bound B; the margin leaves
// room for the rounding of each quantized term (0.5 ulp of 2^-s each).
int32_t shift_for_bound(double B) {
    int32_t s = 24;
    while (s > 0 && B * std::ldexp(1.0, s) + 1e6 >= 2147483647.0) --s;
    return s;
}

// Whether shift s keeps every int32 sum of the bound B in range.
bool shift_fits(double B, int32_t s) { return B * std::ldexp(1.0, s) + 1e6 < 2147483647.0; }

int32_t auto_shift(const nbx_net* net) {
    return shift_for_bound(arrival_bound(net, net->has_sparse ? net->h_w.data() : nullptr));
}

}  // namespace

extern "C" {

const char* nbx_last_error(void) { return g_error.c_str(); }

int32_t nbx_version(void) { return 10000; }

int32_t nbx_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int32_t nbx_net_create(const nbx_model_desc* md, int32_t device, nbx_net** out) {
    return nbx_net_create_populations(md, 1, device, out);
}

int32_t nbx_net_create_populations(const nbx_model_desc* pops, int32_t count, int32_t device, nbx_net** out) {
    NBX_CHECK(pops && out, "null argument");
    *out = nullptr;
    NBX_CHECK(count >= 1 && count <= kMaxPops, "population count must be in [1, 64]");
    int64_t total = 0;
    for (int32_t k = 0; k < count; ++k) {
        const char* why = model_desc_error(&pops[k]);
        if (why) {
            set_error(count > 1 ? "population " + std::to_string(k) + ": " + why : std::string(why));
            return NBX_ERR_INVALID;
        }
        NBX_CHECK(pops[k].dt == pops[0].dt, "every population must step with the same dt");
        total += pops[k].n;
    }
    NBX_CHECK(total < kMaxNeurons, "neuron count must be in [1, 2^24)");
    int ndev = 0;
    NBX_HIP_TRY(hipGetDeviceCount(&ndev));
    NBX_CHECK(device >= 0 && device < ndev, "device index out of range");
    std::unique_ptr<nbx_net> net(new nbx_net());
    net->device = device;
    net->n = (int32_t)total;
    bool any_izh = false;
    for (int32_t k = 0; k < count; ++k) any_izh |= pops[k].model == NBX_MODEL_IZH || pops[k].model == NBX_MODEL_IZH_CUR;
    if (any_izh) {
        net->h_a.assign((size_t)total, 0.0f);
        net->h_b.assign((size_t)total, 0.0f);
        net->h_c.assign((size_t)total, 0.0f);
        net->h_d.assign((size_t)total, 0.0f);
    }
    int32_t begin = 0;
    for (int32_t k = 0; k < count; ++k) {
        const nbx_model_desc* md = &pops[k];
        NetPop pp;
        pp.begin = begin;
        pp.md = *md;
        pp.md.izh_a = pp.md.izh_b = pp.md.izh_c = pp.md.izh_d = nullptr;
        pp.P = make_params(md, begin);
        if (md->model == NBX_MODEL_IZH || md->model == NBX_MODEL_IZH_CUR) {
            std::copy(md->izh_a, md->izh_a + md->n, net->h_a.begin() + begin);
            std::copy(md->izh_b, md->izh_b + md->n, net->h_b.begin() + begin);
            std::copy(md->izh_c, md->izh_c + md->n, net->h_c.begin() + begin);
            std::copy(md->izh_d, md->izh_d + md->n, net->h_d.begin() + begin);
        }
        net->pops.push_back(pp);
        begin += md->n;
    }
    net->md = net->pops[0].md;
    net->P = net->pops[0].P;
    int32_t ns = 0;
    for (const NetPop& pp : net->pops) ns = std::max(ns, model_nstate(pp.md.model));
    net->ns = ns;
    if (net->multi()) {
        net->md.n = net->n;
        net->P.model = kModelTyped;
        net->P.n = net->n;
        net->nch = model_channels(kModelTyped);
        net->h_pop_of.resize((size_t)total);
        for (int32_t k = 0; k < count; ++k)
            std::fill(net->h_pop_of.begin() + net->pops[(size_t)k].begin,
                      net->h_pop_of.begin() + net->pops[(size_t)k].begin + net->pops[(size_t)k].md.n, (uint8_t)k);
    } else {
        net->nch = model_channels(net->md.model);
    }
    NBX_HIP_TRY(hipSetDevice(device));
    NBX_HIP_TRY(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    *out = net.release();
    return NBX_OK;
}

int32_t nbx_net_destroy(nbx_net* net) {
    if (!net) return NBX_OK;
    hipSetDevice(net->device);
    if (net->mon.on) (void)nbx_net_monitor_stop(net);  // writes the frames still held
    if (net->stream) hipStreamSynchronize(net->stream);
    net->release();
    delete net;
    return NBX_OK;
}

int32_t nbx_net_set_sparse(nbx_net* net, const nbx_sparse_desc* syn) {
    NBX_CHECK(net && syn, "null argument");
    NBX_CHECK(!net->finalized, "synapses must be set before nbx_net_finalize");
    NBX_CHECK(syn->count >= 0, "negative synapse count");
    NBX_CHECK(syn->max_delay >= 1 && syn->max_delay <= kMaxDelay, "max_delay must be in [1, 63]");
    NBX_CHECK(syn->row_ptr && (syn->count == 0 || (syn->post && syn->delay && syn->type && syn->weight)),
              "missing synapse arrays");
    const int32_t n = net->n;
    NBX_CHECK(syn->row_ptr[0] == 0 && syn->row_ptr[n] == syn->count, "row_ptr must start at 0 and end at count");
    net->h_row_ptr.assign(syn->row_ptr, syn->row_ptr + n + 1);
    net->h_src_index.clear();
    net->h_file_rule.clear();
    net->h_file_trig.clear();
    net->h_file_maxw.clear();
    net->h_smaxw.clear();
    net->h_syn.resize((size_t)syn->count);
    net->h_w.assign(syn->weight, syn->weight + syn->count);
    if (syn->plastic) net->h_plastic.assign(syn->plastic, syn->plastic + syn->count);
    else net->h_plastic.clear();
    for (int32_t p = 0; p < n; ++p) {
        const int64_t b = syn->row_ptr[p], e = syn->row_ptr[p + 1];
        NBX_CHECK(b <= e, "row_ptr must be non-decreasing");
        int32_t prev_d = 0;
        for (int64_t j = b; j < e; ++j) {
            const int32_t d = syn->delay[j], post = syn->post[j], ty = syn->type[j];
            NBX_CHECK(d >= 1 && d <= syn->max_delay, "synapse delay outside [1, max_delay]");
            NBX_CHECK(post >= 0 && post < n, "synapse post index out of range");
            const int32_t tm = net->model_of(post);
            NBX_CHECK(ty >= 0 && ty <= model_max_type(tm),
                      model_max_type(tm) == 2 ? "synapse type must be 0 (AMPA), 1 (GABA) or 2 (NMDA)"
                                              : "synapse type must be 0 (excitatory) or 1 (inhibitory)");
            NBX_CHECK(d >= prev_d, "each row must be sorted by delay");
            prev_d = d;
            net->h_syn[(size_t)j] = pack_syn(post, d, ty);
        }
    }
    net->S = syn->count;
    net->max_delay = std::max(net->max_delay, syn->max_delay);
    net->has_sparse = true;
    return NBX_OK;
}

int32_t nbx_net_set_dense(nbx_net* net, const float* wt, int32_t delay) {
    NBX_CHECK(net && wt, "null argument");
    NBX_CHECK(!net->finalized, "dense weights must be set before nbx_net_finalize");
    NBX_CHECK(delay >= 1 && delay <= kMaxDelay, "dense delay must be in [1, 63]");
    NBX_CHECK(!net->multi(), "dense all-to-all weights need a one-population network (use sparse synapses)");
    net->h_wt.assign(wt, wt + (size_t)net->n * net->n);
    net->dense_delay = delay;
    net->max_delay = std::max(net->max_delay, delay);
    net->has_dense = true;
    return NBX_OK;
}

int32_t nbx_net_set_poisson(nbx_net* net, uint32_t seed_lo, uint32_t seed_hi, double lambda, float w_bg) {
    NBX_CHECK(net, "null argument");
    NBX_CHECK(!net->finalized, "Poisson input must be set before nbx_net_finalize");
    NBX_CHECK(lambda >= 0.0 && lambda <= 20.0, "Poisson lambda must be in [0, 20]");
    net->pois = lambda > 0.0;
    net->seed_lo = seed_lo;
    net->seed_hi = seed_hi;
    net->lambda = lambda;
    net->w_bg = w_bg;
    net->kmax = poisson_table(lambda, &net->pois_tab);
    return NBX_OK;
}

int32_t nbx_net_set_stdp(nbx_net* net, const nbx_stdp_desc* sd) {
    NBX_CHECK(net && sd, "null argument");
    NBX_CHECK(!net->finalized, "STDP must be set before nbx_net_finalize");
    NBX_CHECK(sd->ltp_tau > 0.0f && sd->ltd_tau > 0.0f, "STDP time constants must be positive");
    net->stdp = true;
    net->sd = *sd;
    net->h_smaxw.clear();
    return NBX_OK;
}

}  // extern "C"

namespace nbx {
const char* rule_desc_error(const nbx_rule_desc& q, bool allow_stdp) {
    switch (q.kind) {
        case NBX_RULE_STDP:
            if (!allow_stdp) return "pair STDP is set with nbx_net_set_stdp";
            if (!(q.ltp_tau > 0.0f && q.ltd_tau > 0.0f)) return "learning rule time constants must be positive";
            return nullptr;
        case NBX_RULE_RSTDP:
            if (!(q.ltp_tau > 0.0f && q.ltd_tau > 0.0f)) return "learning rule time constants must be positive";
            return nullptr;
        case NBX_RULE_ADDITIVE:
            if (!(q.ltp_tau > 0.0f)) return "learning rule time constants must be positive";
            return nullptr;
        case NBX_RULE_CEREBELLAR:
            if (!(q.kernel_step > 0.0f) || !std::isfinite(q.initpos) || !std::isfinite(q.maxpos))
                return "cerebellar kernel needs a positive step and finite positions";
            if (q.history < 0 || q.history > 4096) return "cerebellar history must be in [0, 4096]";
            if ((q.maxpos - q.initpos) / q.kernel_step > 1.0e6f)
                return "cerebellar kernel window too long for its step";
            return nullptr;
        default: return "unknown learning rule kind";
    }
}
std::vector<float*> state_columns(nbx_net* net, int32_t model) { return model_columns(net, model); }
}  // namespace nbx

extern "C" {

int32_t nbx_net_set_rules(nbx_net* net, const nbx_rule_desc* rules, int32_t count, const int32_t* rule,
                          const int32_t* trigger, const float* max_weight) {
    NBX_CHECK(net && (rules || count == 0), "null argument");
    NBX_CHECK(!net->finalized, "learning rules must be set before nbx_net_finalize");
    NBX_CHECK(net->has_sparse, "learning rules need sparse synapses (set them first)");
    NBX_CHECK(count >= 0 && count <= kMaxRules, "rule count must be in [0, 64]");
    NBX_CHECK(count == 0 || (rule && trigger && max_weight), "missing per-synapse rule arrays");
    for (int32_t r = 0; r < count; ++r) {
        const char* why = rule_desc_error(rules[r], false);
        NBX_CHECK(why == nullptr, why ? why : "");
    }
    const size_t S = (size_t)net->S;
    std::vector<int8_t> hr(S, -1), ht(S, -1);
    std::vector<float> hm(S, 0.0f);
    for (size_t j = 0; j < S && count > 0; ++j) {
        NBX_CHECK(rule[j] >= -1 && rule[j] < count && trigger[j] >= -1 && trigger[j] < count,
                  "synapse refers to an invalid learning rule");
        NBX_CHECK(!(rule[j] >= 0 && trigger[j] >= 0), "a synapse is either rule-plastic or a trigger");
        NBX_CHECK(std::isfinite(max_weight[j]), "max_weight must be finite");
        hr[j] = (int8_t)rule[j];
        ht[j] = (int8_t)trigger[j];
        hm[j] = max_weight[j];
    }
    net->rules.assign(rules, rules + count);
    if (count == 0) {
        net->h_rule.clear();
        net->h_trig.clear();
        net->h_maxw.clear();
    } else {
        net->h_rule.swap(hr);
        net->h_trig.swap(ht);
        net->h_maxw.swap(hm);
    }
    return NBX_OK;
}

int32_t nbx_net_set_current(nbx_net* net, const float* cur, int32_t t_begin, int32_t nsteps) {
    NBX_CHECK(net && (cur || nsteps == 0), "null argument");
    NBX_CHECK(nsteps >= 0, "negative step count");
    net->h_cur.assign(cur, cur + (size_t)nsteps * net->n);
    net->cur_t0 = t_begin;
    net->cur_n = nsteps;
    if (net->finalized) {
        int32_t rc = device_scope(net);
        if (rc) return rc;
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        if (net->d_cur) {
            auto it = std::find(net->allocs.begin(), net->allocs.end(), (void*)net->d_cur);
            if (it != net->allocs.end()) net->allocs.erase(it);
            NBX_HIP_TRY(hipFree(net->d_cur));
            net->d_cur = nullptr;
        }
        if (nsteps) {
            rc = net->upload(&net->d_cur, net->h_cur.data(), net->h_cur.size());
            if (rc) return rc;
        }
        if (net->gexec) {
            hipGraphExecDestroy(net->gexec);
            net->gexec = nullptr;
        }
    }
    return NBX_OK;
}

int32_t nbx_net_set_fixed_shift(nbx_net* net, int32_t shift) {
    NBX_CHECK(net, "null argument");
    NBX_CHECK(!net->finalized, "fixed shift must be set before nbx_net_finalize");
    NBX_CHECK(shift >= 0 && shift <= 30, "fixed shift must be in [0, 30]");
    net->shift_req = shift;
    return NBX_OK;
}

int32_t nbx_net_get_fixed_shift(const nbx_net* net, int32_t* shift) {
    NBX_CHECK(net && shift, "null argument");
    *shift = net->finalized ? net->shift : (net->shift_req ? net->shift_req : -1);
    return NBX_OK;
}

int32_t nbx_net_nstate(const nbx_net* net) { return net ? net->ns : 0; }

int32_t nbx_net_finalize(nbx_net* net) {
    NBX_CHECK(net, "null argument");
    NBX_CHECK(!net->finalized, "network already finalized");
    int32_t rc = device_scope(net);
    if (rc) return rc;
    const int32_t n = net->n;
    // max_delay + 1 slots (the reference's delay_slot_count, DenseSubnetworkBuildFinalizer.cpp:62),
    // at least 3 so the slot being filled, the one being read and the one being freed differ
    net->D = std::max(net->max_delay + 1, 3);
    net->pull_mode = net->stdp || !net->rules.empty();
    if (net->multi())  // each synapse's type must be one its target's model takes (.nbnet tables too)
        for (int64_t j = 0; j < net->S; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            NBX_CHECK(syn_type(e) <= model_max_type(net->model_of(syn_post(e))),
                      "a synapse type is not an input of its target population's model");
        }
    NBX_CHECK(!(net->stdp && !net->has_sparse), "STDP needs sparse synapses");
    NBX_CHECK(net->rules.empty() || (net->has_sparse && net->h_rule.size() == (size_t)net->S),
              "learning rules must be set after the synapses they refer to");
    if (!net->rules.empty() && net->stdp && !net->h_plastic.empty())
        for (size_t j = 0; j < (size_t)net->S; ++j)
            NBX_CHECK(!(net->h_plastic[j] && (net->h_rule[j] >= 0 || net->h_trig[j] >= 0)),
                      "a synapse is either STDP-plastic, rule-plastic or a trigger");
    if (net->shift_req) {
        const double B = arrival_bound(net, net->has_sparse ? net->h_w.data() : nullptr);
        if (!shift_fits(B, net->shift_req)) {
            set_error("fixed shift " + std::to_string(net->shift_req) +
                      " can overflow the int32 arrival sums of this network (largest safe shift " +
                      std::to_string(shift_for_bound(B)) + ")");
            return NBX_ERR_INVALID;
        }
    }
    net->shift = net->shift_req ? net->shift_req : auto_shift(net);
    net->words = (n + 31) / 32;
    // conductance inputs present in each population (the reference models' CheckType flags): the
    // types of the synapses into it, dense weight signs and the (excitatory) background
    {
        std::vector<int32_t> cond(net->pops.size(), net->pois ? kCondExc : 0);
        for (int64_t j = 0; j < net->S; ++j) {
            const uint32_t e = net->h_syn[(size_t)j];
            cond[net->multi() ? net->h_pop_of[(size_t)syn_post(e)] : 0] |= 1 << syn_type(e);
        }
        if (net->has_dense)
            for (float w : net->h_wt) cond[0] |= (w < 0.0f) ? kCondInh : kCondExc;
        for (size_t k = 0; k < net->pops.size(); ++k)
            net->pops[k].P.cond = cond[k] & (kCondExc | kCondInh | kCondNmda);
        if (!net->multi()) net->P.cond = net->pops[0].P.cond;
    }
    // Izhikevich NMDA synapses (type 2): a third ring channel and the g_nmda state array
    bool izh_nmda = false;
    for (const NetPop& pp : net->pops) izh_nmda |= pp.md.model == NBX_MODEL_IZH && (pp.P.cond & kCondNmda);
    if (!net->multi() && izh_nmda) net->nch = 3;
    // state (SoA) initialised like Neuron_State_Vector::InitNeuronState with sigma 0
    std::vector<float> v((size_t)n), s1((size_t)n, 0.0f), zero((size_t)n, 0.0f);
    std::vector<int32_t> last((size_t)n, 10000);
    for (const NetPop& pp : net->pops) {
        for (int32_t i = pp.begin; i < pp.begin + pp.md.n; ++i) {
            switch (pp.md.model) {
                case NBX_MODEL_LIF_VJ:  // TimeDrivenLIF_Voltage_jump.cpp:37
                case NBX_MODEL_LIF_EXP3: v[(size_t)i] = pp.md.v_reset; break;  // ..._triple.cpp:71
                case NBX_MODEL_POISSON:  // PoissonRate.cpp:45
                case NBX_MODEL_TRIGGER_RELAY: v[(size_t)i] = 0.0f; break;
                default: v[(size_t)i] = pp.md.v_rest;
            }
            if (pp.md.model == NBX_MODEL_IZH || pp.md.model == NBX_MODEL_IZH_CUR)
                s1[(size_t)i] = net->h_b[(size_t)i] * pp.md.v_rest;
        }
    }
    net->h_init_v = v;
    net->h_init_s1 = s1;
#define NBX_RC(x)            \
    do {                     \
        rc = (x);            \
        if (rc) return rc;   \
    } while (0)
    NBX_RC(net->upload(&net->d_v, v.data(), (size_t)n));
    NBX_RC(net->upload(&net->d_s1, s1.data(), (size_t)n));
    NBX_RC(net->upload(&net->d_ga, zero.data(), (size_t)n));
    NBX_RC(net->upload(&net->d_gg, zero.data(), (size_t)n));
    NBX_RC(net->upload(&net->d_iext, zero.data(), (size_t)n));
    NBX_RC(net->upload(&net->d_last, last.data(), (size_t)n));
    if (izh_nmda) NBX_RC(net->upload(&net->d_g3, zero.data(), (size_t)n));
    if (net->multi()) {
        std::vector<ModelParams> pp;
        for (const NetPop& q : net->pops) pp.push_back(q.P);
        NBX_RC(net->upload(&net->d_pops, pp.data(), pp.size()));
        NBX_RC(net->upload(&net->d_pop_of, net->h_pop_of.data(), net->h_pop_of.size()));
    }
    if (!net->h_a.empty()) {
        NBX_RC(net->upload(&net->d_a, net->h_a.data(), (size_t)n));
        NBX_RC(net->upload(&net->d_b, net->h_b.data(), (size_t)n));
        NBX_RC(net->upload(&net->d_c, net->h_c.data(), (size_t)n));
        NBX_RC(net->upload(&net->d_d, net->h_d.data(), (size_t)n));
    }
    NBX_RC(net->alloc(&net->d_ring, (size_t)net->D * net->nch * n));
    NBX_HIP_TRY(hipMemsetAsync(net->d_ring, 0, sizeof(int32_t) * (size_t)net->D * net->nch * n, net->stream));
    NBX_RC(net->alloc(&net->d_hist, (size_t)net->D * n));
    NBX_RC(net->alloc(&net->d_hist_cnt, (size_t)net->D));
    NBX_HIP_TRY(hipMemsetAsync(net->d_hist_cnt, 0, sizeof(int32_t) * net->D, net->stream));
    if (net->stdp) {  // firing history of the merged pre/post STDP kernel
        NBX_RC(net->alloc(&net->d_fmask, (size_t)n));
        NBX_HIP_TRY(hipMemsetAsync(net->d_fmask, 0, sizeof(uint64_t) * (size_t)n, net->stream));
    }
    NBX_RC(net->alloc(&net->d_events, 1));
    NBX_HIP_TRY(hipMemsetAsync(net->d_events, 0, sizeof(unsigned long long), net->stream));
    NBX_RC(net->alloc(&net->d_tbase, 1));
    NBX_RC(net->alloc(&net->d_stage, (size_t)n));
    NBX_RC(net->alloc(&net->d_err, 1));
    NBX_HIP_TRY(hipMemsetAsync(net->d_err, 0, sizeof(int32_t), net->stream));
    // the stream is non-blocking: finish the clears before anything else touches these buffers
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    if (net->has_sparse) {
        NBX_RC(net->upload(&net->d_row_ptr, net->h_row_ptr.data(), net->h_row_ptr.size()));
        NBX_RC(net->upload(&net->d_syn, net->h_syn.data(), net->h_syn.size()));
        NBX_RC(net->upload(&net->d_w, net->h_w.data(), net->h_w.size()));
        // start of each delay inside each row: dptr[p * (D + 1) + d]
        const int32_t D = net->D;
        std::vector<int64_t> dptr((size_t)n * (D + 1));
        for (int32_t p = 0; p < n; ++p) {
            int64_t j = net->h_row_ptr[(size_t)p];
            const int64_t e = net->h_row_ptr[(size_t)p + 1];
            for (int32_t d = 0; d <= D; ++d) {
                while (j < e && syn_delay(net->h_syn[(size_t)j]) < d) ++j;
                dptr[(size_t)p * (D + 1) + d] = j;
            }
            dptr[(size_t)p * (D + 1) + D] = e;
        }
        NBX_RC(net->upload(&net->d_dptr, dptr.data(), dptr.size()));
        if (!net->h_plastic.empty()) NBX_RC(net->upload(&net->d_plastic, net->h_plastic.data(), net->h_plastic.size()));
    }
    if (net->has_dense) NBX_RC(net->upload(&net->d_wt, net->h_wt.data(), net->h_wt.size()));
    if (net->pois) NBX_RC(net->upload(&net->d_pois_tab, net->pois_tab.data(), net->pois_tab.size()));
    if (net->cur_n) NBX_RC(net->upload(&net->d_cur, net->h_cur.data(), net->h_cur.size()));
    NBX_RC(upload_learning_records(net, true));
    if (!net->rules.empty()) NBX_RC(finalize_rules(net));
    if (net->stdp) {
        NBX_CHECK(!net->h_plastic.empty(), "STDP needs per-synapse plastic flags");
        const size_t S = (size_t)net->S;
        std::vector<float> tp, td;
        trace_tables(net->md.dt, net->sd.ltp_tau, net->sd.ltd_tau, &tp, &td);
        net->tab_len = (int32_t)tp.size();
        NBX_RC(net->upload(&net->d_ltp_tab, tp.data(), tp.size()));
        NBX_RC(net->upload(&net->d_ltd_tab, td.data(), td.size()));
        // plastic inputs of each post neuron (post-side STDP index)
        std::vector<int64_t> in_ptr((size_t)n + 1, 0);
        for (size_t j = 0; j < S; ++j)
            if (net->h_plastic[j]) in_ptr[(size_t)syn_post(net->h_syn[j]) + 1]++;
        for (int32_t i = 0; i < n; ++i) in_ptr[(size_t)i + 1] += in_ptr[(size_t)i];
        std::vector<int64_t> fill(in_ptr.begin(), in_ptr.end() - 1), in_syn((size_t)in_ptr[(size_t)n]);
        std::vector<int32_t> in_pd(in_syn.size());
        for (int32_t p = 0; p < n; ++p)
            for (int64_t j = net->h_row_ptr[(size_t)p]; j < net->h_row_ptr[(size_t)p + 1]; ++j) {
                if (!net->h_plastic[(size_t)j]) continue;
                const uint32_t e = net->h_syn[(size_t)j];
                const int64_t k = fill[(size_t)syn_post(e)]++;
                in_syn[(size_t)k] = j;
                in_pd[(size_t)k] = p | (syn_delay(e) << 24);
            }
        NBX_RC(net->upload(&net->d_in_ptr, in_ptr.data(), in_ptr.size()));
        NBX_RC(net->upload(&net->d_in_syn, in_syn.data(), in_syn.size()));
        NBX_RC(net->upload(&net->d_in_pd, in_pd.data(), in_pd.size()));
    }
    net->finalized = true;
    return NBX_OK;
}

int32_t nbx_net_set_state(nbx_net* net, const float* x, const int32_t* last) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(x && last, "null argument");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const int32_t ns = net->ns;
    std::vector<float> col;
    for (const NetPop& pp : net->pops) {  // each population's columns (several: x[i * ns + j], ns the largest)
        const std::vector<float*> dst = model_columns(net, pp.md.model);
        col.resize((size_t)pp.md.n);
        for (size_t j = 0; j < dst.size(); ++j) {
            if (!dst[j]) continue;
            for (int32_t i = 0; i < pp.md.n; ++i) col[(size_t)i] = x[(size_t)(pp.begin + i) * ns + j];
            NBX_HIP_TRY(hipMemcpy(dst[j] + pp.begin, col.data(), sizeof(float) * pp.md.n, hipMemcpyHostToDevice));
        }
    }
    NBX_HIP_TRY(hipMemcpy(net->d_last, last, sizeof(int32_t) * net->n, hipMemcpyHostToDevice));
    return NBX_OK;
}

int32_t nbx_net_get_state(nbx_net* net, float* x, int32_t* last) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(x && last, "null argument");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const int32_t ns = net->ns;
    for (size_t k = 0; k < (size_t)net->n * ns; ++k) x[k] = 0.0f;
    std::vector<float> col;
    for (const NetPop& pp : net->pops) {
        const std::vector<float*> src = model_columns(net, pp.md.model);
        col.resize((size_t)pp.md.n);
        for (size_t j = 0; j < src.size(); ++j) {
            if (!src[j]) continue;
            NBX_HIP_TRY(hipMemcpy(col.data(), src[j] + pp.begin, sizeof(float) * pp.md.n, hipMemcpyDeviceToHost));
            for (int32_t i = 0; i < pp.md.n; ++i) x[(size_t)(pp.begin + i) * ns + j] = col[(size_t)i];
        }
    }
    NBX_HIP_TRY(hipMemcpy(last, net->d_last, sizeof(int32_t) * net->n, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_net_record(nbx_net* net, int32_t t_begin, int32_t nsteps, int32_t flags) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(nsteps >= 0 && t_begin >= 0, "negative record window");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    for (void** p : {(void**)&net->d_counts, (void**)&net->d_raster}) {
        if (*p) {
            auto it = std::find(net->allocs.begin(), net->allocs.end(), *p);
            if (it != net->allocs.end()) net->allocs.erase(it);
            NBX_HIP_TRY(hipFree(*p));
            *p = nullptr;
        }
    }
    net->rec_t0 = t_begin;
    net->rec_n = nsteps;
    if (nsteps && (flags & NBX_RECORD_COUNTS)) {
        NBX_RC(net->alloc(&net->d_counts, (size_t)nsteps));
        NBX_HIP_TRY(hipMemsetAsync(net->d_counts, 0, sizeof(int32_t) * (size_t)nsteps, net->stream));
    }
    if (nsteps && (flags & NBX_RECORD_RASTER)) {
        const size_t nw = (size_t)nsteps * net->words;
        NBX_RC(net->alloc(&net->d_raster, nw));
        NBX_HIP_TRY(hipMemsetAsync(net->d_raster, 0, sizeof(uint32_t) * nw, net->stream));
    }
    // ordered with the network's (non-blocking) stream, and done before the call returns
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    if (net->gexec) {  // captured kernels hold the old recording pointers
        hipGraphExecDestroy(net->gexec);
        net->gexec = nullptr;
    }
    return NBX_OK;
}

int32_t nbx_net_get_counts(nbx_net* net, int32_t* counts, int32_t nsteps) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->d_counts, "spike counts are not being recorded");
    NBX_CHECK(counts && nsteps >= 0 && nsteps <= net->rec_n, "bad counts buffer");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    NBX_HIP_TRY(hipMemcpy(counts, net->d_counts, sizeof(int32_t) * (size_t)nsteps, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_net_get_raster(nbx_net* net, uint32_t* raster, int64_t nwords) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->d_raster, "raster is not being recorded");
    NBX_CHECK(raster && nwords >= 0 && nwords <= (int64_t)net->rec_n * net->words, "bad raster buffer");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    NBX_HIP_TRY(hipMemcpy(raster, net->d_raster, sizeof(uint32_t) * (size_t)nwords, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_net_pick_engine(const nbx_net* net, int32_t* engine) {
    NBX_CHECK(net && engine, "null argument");
    *engine = persistent_supported(net) ? NBX_ENGINE_PERSISTENT : NBX_ENGINE_GRAPH;
    return NBX_OK;
}

static int32_t run_engine(nbx_net* net, int32_t nsteps, int32_t engine) {
    switch (engine) {
        case NBX_ENGINE_KERNELS: return run_kernels(net, nsteps);
        case NBX_ENGINE_GRAPH: return run_graph(net, nsteps);
        case NBX_ENGINE_PERSISTENT: return run_persistent(net, nsteps);
        default: set_error("unknown engine"); return NBX_ERR_INVALID;
    }
}

int32_t nbx_net_run_async(nbx_net* net, int32_t nsteps, int32_t engine) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(nsteps >= 0, "negative step count");
    NBX_CHECK(net->phase == 0, "a per-phase step is open (finish it with nbx_net_commit_firings)");
    NBX_RC(device_scope(net));
    if (engine == NBX_ENGINE_AUTO) NBX_RC(nbx_net_pick_engine(net, &engine));
    NBX_CHECK(engine >= NBX_ENGINE_KERNELS && engine <= NBX_ENGINE_PERSISTENT, "unknown engine");
    if (engine == NBX_ENGINE_PERSISTENT && !persistent_supported(net)) {
        set_error("persistent engine unsupported for this network: " + persistent_reason(net));
        return NBX_ERR_UNSUPPORTED;
    }
    if (!net->mon.on && currents_due_next(net, net->t) < 0) return run_engine(net, nsteps, engine);
    // The run is cut before every step with current-synapse arrivals (applied before it runs) and,
    // with the time-series monitor on, after every sampled step (its frame is captured there); the
    // same engine runs each piece, and pieces compose like separate runs.
    const int32_t iv = net->mon.interval;
    for (int32_t done = 0; done < nsteps;) {
        const int32_t t = net->t;
        if (currents_due_next(net, t) == t) NBX_RC(apply_currents(net, t));
        int32_t len = nsteps - done;
        int32_t s = -1;
        if (net->mon.on) {
            s = ((t + iv - 1) / iv) * iv;  // first sampled step >= t
            len = std::min(len, s - t + 1);
        }
        const int32_t due = currents_due_next(net, t + 1);
        if (due >= 0) len = std::min(len, due - t);
        NBX_RC(run_engine(net, len, engine));
        done += len;
        if (net->mon.on && net->t - 1 == s) NBX_RC(monitor_capture(net, s));
    }
    return NBX_OK;
}

int32_t nbx_net_sync(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    return check_device_error(net);
}

int32_t nbx_net_run(nbx_net* net, int32_t nsteps, int32_t engine) {
    int32_t rc = nbx_net_run_async(net, nsteps, engine);
    if (rc) return rc;
    return nbx_net_sync(net);
}

int32_t nbx_net_time(const nbx_net* net, int32_t* t) {
    NBX_CHECK(net && t, "null argument");
    *t = net->t;
    return NBX_OK;
}

int32_t nbx_net_events(nbx_net* net, int64_t* events) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(events, "null argument");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    unsigned long long e = 0;
    NBX_HIP_TRY(hipMemcpy(&e, net->d_events, sizeof(e), hipMemcpyDeviceToHost));
    *events = (int64_t)e;
    return NBX_OK;
}

// ---------------------------------------------------------------- per-phase calls

int32_t nbx_net_begin_step(nbx_net* net, int32_t t) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 0) {
        set_error("previous step not committed");
        return NBX_ERR_STATE;
    }
    if (t != net->t) {
        set_error("BeginStep time " + std::to_string(t) + " != network time " + std::to_string(net->t));
        return NBX_ERR_STATE;
    }
    if (currents_due_next(net, t) == t) {  // current-synapse arrivals of this step, before its update
        NBX_RC(device_scope(net));
        NBX_RC(apply_currents(net, t));
    }
    net->phase = 1;
    return NBX_OK;
}

int32_t nbx_net_run_propagation(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 1) {
        set_error("RunPropagation must follow BeginStep");
        return NBX_ERR_STATE;
    }
    NBX_RC(device_scope(net));
    if (net->pull_mode && net->has_sparse) NBX_HIP_TRY(launch_pull(net->args(), nullptr, net->t, net->stream));
    if (!net->rules.empty()) NBX_HIP_TRY(launch_trigger(net->args(), nullptr, net->t, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    if (net->ccap > 0) NBX_RC(check_device_error(net));
    net->phase = 2;
    return NBX_OK;
}

int32_t nbx_net_neuron_step(nbx_net* net, int32_t* ids_out, int32_t ids_cap, int32_t* count_out) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 2) {
        set_error("neuron step must follow RunPropagation");
        return NBX_ERR_STATE;
    }
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(launch_update(net->args(), nullptr, net->t, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    net->phase = 3;
    if (ids_out || count_out) {
        int32_t cnt = 0;
        NBX_RC(nbx_net_get_history(net, net->t, ids_out, ids_cap, &cnt));
        if (count_out) *count_out = cnt;
    }
    return NBX_OK;
}

int32_t nbx_net_inject_spikes(nbx_net* net, const int32_t* ids, int32_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 3 && net->phase != 4) {
        set_error("inject spikes between the neuron step and the commit");
        return NBX_ERR_STATE;
    }
    NBX_CHECK(count >= 0 && (ids || count == 0), "bad spike list");
    if (count == 0) return NBX_OK;
    for (int32_t k = 0; k < count; ++k) NBX_CHECK(ids[k] >= 0 && ids[k] < net->n, "injected id out of range");
    NBX_CHECK(count <= net->n, "too many spikes in one step");
    NBX_RC(device_scope(net));
    // A neuron fires at most once per step: an id repeated in the call, or already in the step's
    // list (it fired in the update or was injected before), would be delivered twice.
    std::vector<int32_t> listed((size_t)net->n);
    int32_t cur = 0;
    NBX_RC(nbx_net_get_history(net, net->t, listed.data(), net->n, &cur));
    NBX_CHECK(cur + count <= net->n, "too many spikes in one step");
    std::vector<uint8_t> seen((size_t)net->n, 0);
    for (int32_t k = 0; k < cur; ++k) seen[(size_t)listed[(size_t)k]] = 1;
    for (int32_t k = 0; k < count; ++k) {
        NBX_CHECK(!seen[(size_t)ids[k]], "injected neuron already fires in this step (listed twice or fired in the "
                                         "update): a neuron fires at most once per step");
        seen[(size_t)ids[k]] = 1;
    }
    NBX_HIP_TRY(hipMemcpy(net->d_stage, ids, sizeof(int32_t) * count, hipMemcpyHostToDevice));
    NBX_HIP_TRY(launch_inject(net->args(), net->d_stage, count, net->t, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    return NBX_OK;
}

int32_t nbx_net_firing_device_view(nbx_net* net, int32_t** d_ids, int32_t** d_count, uint8_t** d_did_fire) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 3 && net->phase != 4) {
        set_error("the step's firings exist after its neuron update");
        return NBX_ERR_STATE;
    }
    NBX_RC(device_scope(net));
    const int32_t slot = net->t % net->D;
    if (d_ids) *d_ids = net->d_hist + (size_t)slot * net->n;
    if (d_count) *d_count = net->d_hist_cnt + slot;
    if (d_did_fire) {
        if (!net->d_fire_flags) NBX_RC(net->alloc(&net->d_fire_flags, (size_t)net->n));
        NBX_HIP_TRY(launch_fire_flags(net->args(), net->t, net->d_fire_flags, net->stream));
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        *d_did_fire = net->d_fire_flags;
    }
    return NBX_OK;
}

int32_t nbx_net_commit_external(nbx_net* net, int32_t t, const int32_t* d_ids, const int32_t* d_count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 3 && net->phase != 4) {
        set_error("commit must follow the neuron step");
        return NBX_ERR_STATE;
    }
    if (t != net->t) {
        set_error("commit for step " + std::to_string(t) + " while the network is at step " + std::to_string(net->t));
        return NBX_ERR_STATE;
    }
    NBX_CHECK(d_ids && d_count, "null firing list");
    NBX_RC(device_scope(net));
    const int32_t slot = net->t % net->D;
    if (d_ids != net->d_hist + (size_t)slot * net->n || d_count != net->d_hist_cnt + slot) {
        int32_t cnt = 0;
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        NBX_HIP_TRY(hipMemcpy(&cnt, d_count, sizeof(int32_t), hipMemcpyDeviceToHost));
        NBX_CHECK(cnt >= 0 && cnt <= net->n, "external firing count out of range");
        std::vector<int32_t> ext((size_t)std::max(cnt, 1)), own((size_t)net->n);
        if (cnt) NBX_HIP_TRY(hipMemcpy(ext.data(), d_ids, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
        int32_t own_cnt = 0;
        NBX_RC(nbx_net_get_history(net, net->t, own.data(), net->n, &own_cnt));
        std::vector<uint8_t> in_ext((size_t)net->n, 0);
        for (int32_t k = 0; k < cnt; ++k) {
            NBX_CHECK(ext[(size_t)k] >= 0 && ext[(size_t)k] < net->n, "external firing id out of range");
            NBX_CHECK(!in_ext[(size_t)ext[(size_t)k]], "external firing list repeats an id");
            in_ext[(size_t)ext[(size_t)k]] = 1;
        }
        std::vector<uint8_t> in_own((size_t)net->n, 0);
        for (int32_t k = 0; k < own_cnt; ++k) {
            NBX_CHECK(in_ext[(size_t)own[(size_t)k]], "external firing list omits a neuron that fired in the update");
            in_own[(size_t)own[(size_t)k]] = 1;
        }
        std::vector<int32_t> extra;
        for (int32_t k = 0; k < cnt; ++k)
            if (!in_own[(size_t)ext[(size_t)k]]) extra.push_back(ext[(size_t)k]);
        if (!extra.empty()) NBX_RC(nbx_net_inject_spikes(net, extra.data(), (int32_t)extra.size()));
    }
    return nbx_net_commit_firings(net);
}

int32_t nbx_net_apply_post_learning(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 3) {
        set_error("post learning must follow the neuron step");
        return NBX_ERR_STATE;
    }
    NBX_RC(device_scope(net));
    if ((net->stdp && net->d_in_ptr) || net->rstdp) NBX_HIP_TRY(launch_ltp(net->args(), nullptr, net->t, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    net->phase = 4;
    return NBX_OK;
}

int32_t nbx_net_commit_firings(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    if (net->phase != 3 && net->phase != 4) {
        set_error("commit must follow the neuron step");
        return NBX_ERR_STATE;
    }
    NBX_RC(device_scope(net));
    const StepArgs A = net->args();
    if (net->phase == 3 && ((net->stdp && net->d_in_ptr) || net->rstdp))
        NBX_HIP_TRY(launch_ltp(A, nullptr, net->t, net->stream));
    if (!net->pull_mode && net->has_sparse) NBX_HIP_TRY(launch_deliver_push(A, nullptr, net->t, net->stream));
    if (net->has_dense) NBX_HIP_TRY(launch_deliver_dense(A, nullptr, net->t, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    net->t += 1;
    net->phase = 0;
    if (net->mon.on && (net->t - 1) % net->mon.interval == 0) NBX_RC(monitor_capture(net, net->t - 1));
    return NBX_OK;
}

int32_t nbx_net_set_output_mask(nbx_net* net, const uint8_t* mask) {
    NBX_CHECK(net, "null argument");
    if (!mask) net->out_mask.clear();
    else net->out_mask.assign(mask, mask + net->n);
    return NBX_OK;
}

int32_t nbx_net_neuron_step_outputs(nbx_net* net, int32_t* out_ids, int32_t out_cap, int32_t* out_count,
                                    int32_t* full_ids, int32_t full_cap, int32_t* full_count) {
    std::vector<int32_t> ids((size_t)std::max(net ? net->n : 1, 1));
    int32_t cnt = 0;
    int32_t rc = nbx_net_neuron_step(net, ids.data(), (int32_t)ids.size(), &cnt);
    if (rc) return rc;
    if (full_count) *full_count = cnt;
    if (full_ids) {
        NBX_CHECK(full_cap >= cnt, "full firing id buffer too small");
        std::copy(ids.begin(), ids.begin() + cnt, full_ids);
    }
    int32_t k = 0;
    for (int32_t j = 0; j < cnt; ++j) {
        if (!net->out_mask.empty() && !net->out_mask[(size_t)ids[(size_t)j]]) continue;
        if (out_ids) {
            NBX_CHECK(k < out_cap, "output firing id buffer too small");
            out_ids[k] = ids[(size_t)j];
        }
        ++k;
    }
    if (out_count) *out_count = k;
    return NBX_OK;
}

}  // extern "C"

namespace {

// Room left in every int32 arrival sum after the largest synaptic input (the fixed-point shift
// keeps B * 2^shift + 1e6 below 2^31): the bound on the fixed-point image of a step's external
// inputs into one neuron and channel.
double input_headroom(const nbx_net* net) {
    std::vector<float> w;
    const float* wp = nullptr;
    if (net->has_sparse) {
        if (net->pull_mode) {
            w.resize((size_t)net->S);
            if (hipMemcpy(w.data(), net->d_w, sizeof(float) * w.size(), hipMemcpyDeviceToHost) != hipSuccess) return 0.0;
            wp = w.data();
        } else {
            wp = net->h_w.data();
        }
    }
    const double B = arrival_bound(net, wp);
    return std::max(0.0, 2147483647.0 - 1e6 - B * std::ldexp(1.0, net->shift));
}

// Target of an interface record: the direct target, else the neuron of its slot; -1 for none.
int32_t resolve_target(const nbx_net* net, int32_t slot, int32_t target) {
    if (target >= 0) return target < net->n ? target : -1;
    if (slot >= 0 && slot < (int32_t)net->iface_targets.size()) return net->iface_targets[(size_t)slot];
    return -1;
}

// Input kind of a pending channel for a target neuron: synapse type 0..2, kInputCurrent, or -1
// when the target's model takes no such input.
int32_t channel_kind(const nbx_net* net, int32_t channel, int32_t target) {
    int32_t type;
    switch (channel) {
        case NBX_PENDING_EXC: type = 0; break;
        case NBX_PENDING_INH: type = 1; break;
        case NBX_PENDING_NMDA: type = 2; break;
        case NBX_PENDING_CURRENT: return net->model_of(target) == NBX_MODEL_TRIGGER_RELAY ? -1 : kInputCurrent;
        default: return -1;
    }
    if (type > model_max_type(net->model_of(target))) return -1;
    if (type >= net->nch && arrival_channel(net->P.model, type) >= net->nch) return -1;
    return type;
}

int32_t ensure_input_stage(nbx_net* net, int64_t count) {
    if (count <= net->inputs_cap) return NBX_OK;
    if (net->d_inputs) {
        auto it = std::find(net->allocs.begin(), net->allocs.end(), (void*)net->d_inputs);
        if (it != net->allocs.end()) net->allocs.erase(it);
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        NBX_HIP_TRY(hipFree(net->d_inputs));
        net->d_inputs = nullptr;
        net->inputs_cap = 0;
    }
    NBX_PASS(net->alloc(&net->d_inputs, (size_t)count));
    net->inputs_cap = count;
    return NBX_OK;
}

// The step external inputs go to now: the step being run before its update, else the next one.
int32_t input_step(const nbx_net* net, int32_t* t) {
    if (net->phase >= 3) {
        set_error("external inputs of a step come before its neuron update");
        return NBX_ERR_STATE;
    }
    *t = net->t;
    return NBX_OK;
}

// Checks host records' fixed-point images against the headroom (per target and channel) and
// launches them for step t.
int32_t run_inputs(nbx_net* net, const std::vector<InputRec>& recs, int32_t t) {
    if (recs.empty()) return NBX_OK;
    const double H = input_headroom(net);
    const double scale = std::ldexp(1.0, net->shift);
    std::vector<std::pair<int64_t, double>> sums;
    sums.reserve(recs.size());
    for (const InputRec& r : recs)
        if (r.kind != kInputCurrent && r.src < 0)
            sums.push_back({(int64_t)r.target * 4 + r.kind, std::fabs((double)r.value * scale) + 0.5});
    std::sort(sums.begin(), sums.end());
    for (size_t a = 0; a < sums.size();) {
        double tot = 0.0;
        size_t b = a;
        for (; b < sums.size() && sums[b].first == sums[a].first; ++b) tot += sums[b].second;
        if (tot > H) {
            set_error("external inputs into neuron " + std::to_string(sums[a].first / 4) +
                      " can overflow the int32 arrival sums at fixed shift " + std::to_string(net->shift));
            return NBX_ERR_INVALID;
        }
        a = b;
    }
    NBX_PASS(ensure_input_stage(net, (int64_t)recs.size()));
    NBX_HIP_TRY(hipMemcpy(net->d_inputs, recs.data(), sizeof(InputRec) * recs.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(launch_inputs(net->args(), net->d_inputs, (int32_t)recs.size(), nullptr, t, (float)H, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    return NBX_OK;
}

}  // namespace

extern "C" {

int32_t nbx_net_set_interface_targets(nbx_net* net, const int32_t* neuron_ids, int32_t count) {
    NBX_CHECK(net && (neuron_ids || count == 0) && count >= 0, "bad interface target list");
    for (int32_t k = 0; k < count; ++k)
        NBX_CHECK(neuron_ids[k] >= -1 && neuron_ids[k] < net->n, "interface target neuron out of range");
    net->iface_targets.assign(neuron_ids, neuron_ids + count);
    return NBX_OK;
}

int32_t nbx_net_upload_interface_values(nbx_net* net, const nbx_interface_value* values, int32_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(count >= 0 && (values || count == 0), "bad interface record list");
    NBX_CHECK(count == 0 || net->iface_targets.empty() || count == (int32_t)net->iface_targets.size(),
              "host interface record count must match interface target mapping count");
    std::vector<InputRec> st;
    for (int32_t k = 0; k < count; ++k) {
        const nbx_interface_value& v = values[k];
        NBX_CHECK(std::isfinite(v.value) && std::isfinite(v.scale), "interface values must be finite");
        NBX_CHECK(v.pending_channel != NBX_PENDING_CURRENT || v.overwrite,
                  "current inputs replace I_ext: set overwrite on current records");
        if (v.value == 0.0f) continue;  // ApplyBoundaryStateVectorKernel skips zero values
        const int32_t tgt = resolve_target(net, v.interface_slot, v.target_neuron);
        if (tgt < 0) continue;
        const int32_t kind = channel_kind(net, v.pending_channel, tgt);
        if (kind < 0) continue;
        st.push_back(InputRec{tgt, kind, -1, v.overwrite ? 1 : 0, v.value * v.scale});
    }
    net->iface_staged.swap(st);
    return NBX_OK;
}

int32_t nbx_net_apply_interface(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    int32_t t = 0;
    NBX_RC(input_step(net, &t));
    NBX_RC(device_scope(net));
    return run_inputs(net, net->iface_staged, t);
}

int32_t nbx_net_upload_interface_values_device(nbx_net* net, const float* d_values, const nbx_interface_copy* specs,
                                               int32_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(d_values, "device_values must not be null");
    NBX_CHECK(count >= 0 && (specs || count == 0), "bad interface copy list");
    int32_t t = 0;
    NBX_RC(input_step(net, &t));
    std::vector<InputRec> recs;
    for (int32_t k = 0; k < count; ++k) {
        const nbx_interface_copy& c = specs[k];
        NBX_CHECK(std::isfinite(c.scale), "interface scales must be finite");
        NBX_CHECK(c.pending_channel != NBX_PENDING_CURRENT || c.overwrite,
                  "current inputs replace I_ext: set overwrite on current records");
        const int32_t tgt = resolve_target(net, c.interface_slot, c.target_neuron);
        if (tgt < 0 || c.source_index < 0) continue;
        const int32_t kind = channel_kind(net, c.pending_channel, tgt);
        if (kind < 0) continue;
        recs.push_back(InputRec{tgt, kind, c.source_index, c.overwrite ? 1 : 0, c.scale});
    }
    if (recs.empty()) return NBX_OK;
    NBX_RC(device_scope(net));
    NBX_RC(ensure_input_stage(net, (int64_t)recs.size()));
    const double H = input_headroom(net);
    NBX_HIP_TRY(hipMemcpy(net->d_inputs, recs.data(), sizeof(InputRec) * recs.size(), hipMemcpyHostToDevice));
    NBX_HIP_TRY(launch_inputs(net->args(), net->d_inputs, (int32_t)recs.size(), d_values, t, (float)H, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    int32_t e = 0;
    NBX_HIP_TRY(hipMemcpy(&e, net->d_err, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (e & kErrInput) {
        NBX_PASS(clear_error_bits(net, kErrInput));
        set_error("device interface values could overflow the int32 arrival sums; those values were dropped");
        return NBX_ERR_STATE;
    }
    return NBX_OK;
}

int32_t nbx_net_add_arrivals(nbx_net* net, const int32_t* targets, const uint8_t* types, const float* weights,
                             int32_t count, int32_t t) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(count >= 0 && (count == 0 || (targets && types && weights)), "bad arrival list");
    const int32_t lo = net->phase >= 3 ? net->t + 1 : net->t;
    NBX_CHECK(t >= lo && t <= net->t + net->D - 1,
              "arrival step outside the delay ring's reach (current step .. current step + slots - 1)");
    std::vector<InputRec> recs((size_t)count);
    for (int32_t k = 0; k < count; ++k) {
        NBX_CHECK(targets[k] >= 0 && targets[k] < net->n, "arrival target out of range");
        NBX_CHECK(types[k] <= model_max_type(net->model_of(targets[k])), "arrival type is not an input of its target's model");
        NBX_CHECK(arrival_channel(net->P.model, types[k]) < net->nch,
                  "the network keeps no channel for this arrival type (no synapse of the type was given)");
        NBX_CHECK(std::isfinite(weights[k]), "arrival weights must be finite");
        recs[(size_t)k] = InputRec{targets[k], types[k], -1, 0, weights[k]};
    }
    NBX_RC(device_scope(net));
    return run_inputs(net, recs, t);
}

int32_t nbx_net_reset_state(nbx_net* net) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const size_t n = (size_t)net->n;
    std::vector<float> zero(n, 0.0f);
    std::vector<int32_t> last(n, 10000);
    NBX_HIP_TRY(hipMemcpy(net->d_v, net->h_init_v.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(net->d_s1, net->h_init_s1.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    for (float* p : {net->d_ga, net->d_gg, net->d_iext, net->d_g3})
        if (p) NBX_HIP_TRY(hipMemcpy(p, zero.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    NBX_HIP_TRY(hipMemcpy(net->d_last, last.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    hipStream_t st = net->stream;
    NBX_HIP_TRY(launch_fill_i32(net->d_ring, 0, (int64_t)net->D * net->nch * net->n, st));
    NBX_HIP_TRY(launch_fill_i32(net->d_hist_cnt, 0, net->D, st));
    if (net->d_fmask) NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(net->d_fmask), 0, 2 * (int64_t)n, st));
    NBX_HIP_TRY(launch_fill_i32(reinterpret_cast<int32_t*>(net->d_events), 0, 2, st));
    NBX_HIP_TRY(launch_fill_i32(net->d_err, 0, 1, st));
    net->pl.t_next = -1;  // the persistent engine's counter banks are cleared before its next launch
    if (!net->rules.empty()) {
        NBX_HIP_TRY(launch_fill_i32(net->d_tstamp, -1, net->S, st));
        NBX_HIP_TRY(launch_fill_i32(net->d_tpost_stamp, -1, net->n, st));
        NBX_HIP_TRY(launch_fill_i32(net->d_tcount, 0, 2, st));
        if (net->d_chead) NBX_HIP_TRY(launch_fill_i32(net->d_chead, 0, std::max(net->crows, 1), st));
    }
    NBX_HIP_TRY(hipStreamSynchronize(st));
    NBX_RC(upload_learning_records(net, false));
    clear_currents(net);  // current-synapse slots and pending currents
    net->t = 0;
    net->phase = 0;
    return NBX_OK;
}

int32_t nbx_net_download_pending(nbx_net* net, float* pending, int64_t count, int32_t* channel_count,
                                 int32_t* channel_stride) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(pending && count == (int64_t)net->nch * net->n, "pending buffer size must be channels * n");
    NBX_RC(device_scope(net));
    float* d = nullptr;
    NBX_RC(nbx_net_pending_device_float(net, &d));
    NBX_HIP_TRY(hipMemcpy(pending, d, sizeof(float) * (size_t)count, hipMemcpyDeviceToHost));
    if (channel_count) *channel_count = net->nch;
    if (channel_stride) *channel_stride = net->n;
    return NBX_OK;
}

int32_t nbx_net_pending_device_float(nbx_net* net, float** d_pending) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(d_pending, "null argument");
    NBX_RC(device_scope(net));
    if (!net->d_pending_f) NBX_RC(net->alloc(&net->d_pending_f, (size_t)net->nch * net->n));
    NBX_HIP_TRY(launch_pending_float(net->args(), net->t, net->d_pending_f, net->stream));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    *d_pending = net->d_pending_f;
    return NBX_OK;
}

int32_t nbx_net_ring_shape(const nbx_net* net, int32_t* slots, int32_t* channels) {
    NBX_CHECK(net && slots && channels, "null argument");
    if (!net->finalized) {
        set_error("network is not finalized");
        return NBX_ERR_STATE;
    }
    *slots = net->D;
    *channels = net->nch;
    return NBX_OK;
}

int32_t nbx_net_get_ring(nbx_net* net, int32_t* ring, int64_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(ring && count == (int64_t)net->D * net->nch * net->n, "ring buffer size must be slots*channels*n");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    NBX_HIP_TRY(hipMemcpy(ring, net->d_ring, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_net_get_history(nbx_net* net, int32_t t, int32_t* ids, int32_t cap, int32_t* count_out) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(count_out, "null argument");
    // the update of step u writes slot u % D and frees slot (u + 1) % D, so D - 1 steps are held
    const int32_t newest = net->phase >= 3 ? net->t : net->t - 1;
    if (t < 0 || t > newest || t < newest - net->D + 2) {
        set_error("step " + std::to_string(t) + " is not held in the firing-history ring");
        return NBX_ERR_INVALID;
    }
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const int32_t slot = t % net->D;
    int32_t cnt = 0;
    NBX_HIP_TRY(hipMemcpy(&cnt, net->d_hist_cnt + slot, sizeof(int32_t), hipMemcpyDeviceToHost));
    *count_out = cnt;
    if (ids) {
        NBX_CHECK(cap >= cnt, "id buffer too small");
        NBX_HIP_TRY(hipMemcpy(ids, net->d_hist + (size_t)slot * net->n, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost));
        std::sort(ids, ids + cnt);
    }
    return NBX_OK;
}

int32_t nbx_net_pending_view(nbx_net* net, nbx_pending_view* view) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(view, "null argument");
    view->d_pending = net->d_ring + (size_t)(net->t % net->D) * net->nch * net->n;
    view->channel_count = net->nch;
    view->channel_stride = net->n;
    view->neuron_count = net->n;
    view->inv_scale = 1.0f / std::ldexp(1.0f, net->shift);
    return NBX_OK;
}

int32_t nbx_net_get_weights(nbx_net* net, float* w, int64_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "network has no sparse synapses");
    NBX_CHECK(w && count == net->S, "weight buffer size must equal the synapse count");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    NBX_HIP_TRY(hipMemcpy(w, net->d_w, sizeof(float) * (size_t)count, hipMemcpyDeviceToHost));
    return NBX_OK;
}

}  // extern "C"

namespace {

// Replaces the weights of synapses [j0, j0 + count) (CSR order). Push-mode networks hold the
// arrivals of spikes already fired in the delay ring, quantized with the weights of their spike
// step; the reference delivers each arrival with the weight current at its arrival
// (PropogatedSpikeGroup.cpp:33-71 reads Interconnections::weight when the group is processed),
// so the pending arrivals of changed synapses are moved to the new weights: every spike of the
// last D - 1 steps (the firing-history ring) whose arrival step is not yet past adds
// q(new) - q(old) to its ring entry. Pull-mode networks read the weights at arrival already.
int32_t set_weight_range(nbx_net* net, int64_t j0, int64_t count, const float* w) {
    for (int64_t k = 0; k < count; ++k) NBX_CHECK(std::isfinite(w[k]), "weights must be finite");
    if (net->phase != 0) {
        set_error("weights change between steps (the per-phase step is open)");
        return NBX_ERR_STATE;
    }
    std::vector<float> nw;
    if (!net->pull_mode) {
        nw = net->h_w;  // host copy is current: no learning changes push-mode weights
    } else {
        nw.resize((size_t)net->S);
        NBX_HIP_TRY(hipStreamSynchronize(net->stream));
        NBX_HIP_TRY(hipMemcpy(nw.data(), net->d_w, sizeof(float) * nw.size(), hipMemcpyDeviceToHost));
    }
    std::copy(w, w + count, nw.begin() + j0);
    {  // the fixed-point shift was chosen at finalize: the new weights must not overflow its sums
        const double B = arrival_bound(net, nw.data());
        if (!shift_fits(B, net->shift)) {
            set_error("new weights can overflow the int32 arrival sums at fixed shift " + std::to_string(net->shift) +
                      " (largest safe shift " + std::to_string(shift_for_bound(B)) + "); weights unchanged");
            return NBX_ERR_INVALID;
        }
    }
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    if (!net->pull_mode && net->t > 0) {
        const int32_t n = net->n, D = net->D, t = net->t;
        std::vector<int32_t> hist((size_t)D * n), cnt((size_t)D);
        std::vector<int32_t> ring((size_t)D * net->nch * n);
        NBX_HIP_TRY(hipMemcpy(cnt.data(), net->d_hist_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(hist.data(), net->d_hist, sizeof(int32_t) * hist.size(), hipMemcpyDeviceToHost));
        NBX_HIP_TRY(hipMemcpy(ring.data(), net->d_ring, sizeof(int32_t) * ring.size(), hipMemcpyDeviceToHost));
        const float scale = std::ldexp(1.0f, net->shift);
        bool touched = false;
        for (int32_t st = std::max(0, t - D + 1); st < t; ++st) {
            const int32_t slot = st % D;
            for (int32_t k = 0; k < cnt[(size_t)slot]; ++k) {
                const int32_t pre = hist[(size_t)slot * n + k];
                const int64_t b = std::max(net->h_row_ptr[(size_t)pre], j0);
                const int64_t e = std::min(net->h_row_ptr[(size_t)pre + 1], j0 + count);
                for (int64_t j = b; j < e; ++j) {
                    const uint32_t en = net->h_syn[(size_t)j];
                    const int32_t arr = st + syn_delay(en);
                    if (arr < t) continue;  // already delivered
                    int32_t dq = (int32_t)std::nearbyint(nw[(size_t)j] * scale) -
                                 (int32_t)std::nearbyint(net->h_w[(size_t)j] * scale);
                    if (dq == 0) continue;
                    const int32_t ty = syn_type(en);
                    if (arrival_negated(net->P.model, ty)) dq = -dq;
                    ring[((size_t)(arr % D) * net->nch + arrival_channel(net->P.model, ty)) * n + syn_post(en)] += dq;
                    touched = true;
                }
            }
        }
        if (touched)
            NBX_HIP_TRY(hipMemcpy(net->d_ring, ring.data(), sizeof(int32_t) * ring.size(), hipMemcpyHostToDevice));
    }
    NBX_HIP_TRY(hipMemcpy(net->d_w + j0, w, sizeof(float) * (size_t)count, hipMemcpyHostToDevice));
    std::copy(w, w + count, net->h_w.begin() + j0);
    return refresh_persistent_weights(net);
}

}  // namespace

extern "C" {

int32_t nbx_net_set_weights(nbx_net* net, const float* w, int64_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "network has no sparse synapses");
    NBX_CHECK(w && count == net->S, "weight buffer size must equal the synapse count");
    NBX_RC(device_scope(net));
    return set_weight_range(net, 0, count, w);
}

int32_t nbx_net_get_synapse_weight(nbx_net* net, int64_t synapse, float* w) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "network has no sparse synapses");
    NBX_CHECK(w, "null argument");
    NBX_CHECK(synapse >= 0 && synapse < net->S, "synapse index out of range");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    NBX_HIP_TRY(hipMemcpy(w, net->d_w + synapse, sizeof(float), hipMemcpyDeviceToHost));
    return NBX_OK;
}

int32_t nbx_net_set_synapse_weight(nbx_net* net, int64_t synapse, float w) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "network has no sparse synapses");
    NBX_CHECK(synapse >= 0 && synapse < net->S, "synapse index out of range");
    NBX_RC(device_scope(net));
    return set_weight_range(net, synapse, 1, &w);
}

int32_t nbx_net_get_stdp_state(nbx_net* net, float* apre, float* apost, int32_t* last_update, int64_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->stdp, "STDP is not enabled");
    NBX_CHECK(count == net->S, "buffer size must equal the synapse count");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const size_t st = (size_t)net->head_stride;
    std::vector<TraceHead> hd((size_t)count * st);
    NBX_HIP_TRY(hipMemcpy(hd.data(), net->d_heads, sizeof(TraceHead) * hd.size(), hipMemcpyDeviceToHost));
    for (size_t j = 0; j < (size_t)count; ++j) {
        if (apre) apre[j] = hd[j * st].apre;
        if (apost) apost[j] = hd[j * st].apost;
        if (last_update) last_update[j] = hd[j * st].slast;
    }
    return NBX_OK;
}

int32_t nbx_net_get_rule_state(nbx_net* net, float* apre, float* apost, float* eligibility, int32_t* last_update,
                               int64_t count) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(!net->rules.empty(), "no learning rules are set");
    NBX_CHECK(count == net->S, "buffer size must equal the synapse count");
    NBX_RC(device_scope(net));
    NBX_HIP_TRY(hipStreamSynchronize(net->stream));
    const size_t S = (size_t)count;
    std::vector<SynTrace> tr(S);
    NBX_HIP_TRY(hipMemcpy(tr.data(), net->d_trace, sizeof(SynTrace) * S, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < S; ++j) {
        const bool on = net->h_rule[j] >= 0;
        if (apre) apre[j] = on ? tr[j].apre : 0.0f;
        if (apost) apost[j] = on ? tr[j].apost : 0.0f;
        if (eligibility) eligibility[j] = on ? tr[j].elig : 0.0f;
        if (last_update) last_update[j] = on ? tr[j].slast : 0;
    }
    return NBX_OK;
}

void* nbx_net_stream(nbx_net* net) { return net ? (void*)net->stream : nullptr; }

int32_t nbx_replay_deliver(nbx_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t nsteps,
                           int32_t* out) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "replay delivery needs sparse synapses");
    NBX_CHECK(tick_ptr && out && nsteps > 0 && t0 >= 0, "bad replay arguments");
    NBX_CHECK(tick_ptr[0] == 0, "tick_ptr must start at 0");
    const int32_t total = tick_ptr[nsteps];
    for (int32_t k = 0; k < nsteps; ++k) NBX_CHECK(tick_ptr[k] <= tick_ptr[k + 1], "tick_ptr must be non-decreasing");
    for (int32_t j = 0; j < total; ++j) NBX_CHECK(ids[j] >= 0 && ids[j] < net->n, "spike id out of range");
    NBX_RC(device_scope(net));
    int32_t *d_ids = nullptr, *d_tp = nullptr, *d_out = nullptr;
    const size_t out_count = (size_t)nsteps * net->nch * net->n;
    hipError_t e = hipMalloc(&d_ids, sizeof(int32_t) * std::max(total, 1));
    if (e == hipSuccess) e = hipMalloc(&d_tp, sizeof(int32_t) * (nsteps + 1));
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(int32_t) * out_count);
    if (e == hipSuccess && total) e = hipMemcpy(d_ids, ids, sizeof(int32_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tp, tick_ptr, sizeof(int32_t) * (nsteps + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = nbx_replay_deliver_device(net, d_ids, d_tp, nsteps, d_out);
        if (rc == NBX_OK) e = hipStreamSynchronize(net->stream);
    }
    if (e == hipSuccess && rc == NBX_OK) e = hipMemcpy(out, d_out, sizeof(int32_t) * out_count, hipMemcpyDeviceToHost);
    hipFree(d_ids);
    hipFree(d_tp);
    hipFree(d_out);
    if (rc) return rc;
    NBX_HIP_TRY(e);
    return check_device_error(net);  // a bounded wait of the streamed replay kernel that gave up
}

int32_t nbx_replay_stdp(nbx_net* net, const int32_t* ids, const int32_t* tick_ptr, int32_t t0, int32_t lead,
                        int32_t nsteps, int64_t* updates) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->stdp && net->has_sparse, "replay STDP needs sparse synapses with STDP (nbx_net_set_stdp)");
    NBX_CHECK(net->rules.empty(), "replay STDP does not run trigger-driven learning rules");
    NBX_CHECK(tick_ptr && nsteps > 0 && lead >= 0 && lead <= t0, "bad replay arguments");
    NBX_CHECK((int64_t)lead + nsteps <= 65535, "replay STDP covers at most 65535 steps per call (chain calls)");
    const int32_t K = lead + nsteps;
    NBX_CHECK(tick_ptr[0] == 0, "tick_ptr must start at 0");
    for (int32_t k = 0; k < K; ++k) NBX_CHECK(tick_ptr[k] <= tick_ptr[k + 1], "tick_ptr must be non-decreasing");
    const int32_t total = tick_ptr[K];
    NBX_CHECK(total == 0 || ids, "null spike ids");
    {  // ids in range and distinct within a step (a neuron fires at most once per step)
        std::vector<int32_t> seen((size_t)net->n, -1);
        for (int32_t k = 0; k < K; ++k)
            for (int32_t f = tick_ptr[k]; f < tick_ptr[k + 1]; ++f) {
                NBX_CHECK(ids[f] >= 0 && ids[f] < net->n, "spike id out of range");
                NBX_CHECK(seen[(size_t)ids[f]] != k, "a neuron is listed twice in one step");
                seen[(size_t)ids[f]] = k;
            }
    }
    NBX_RC(device_scope(net));
    if (!net->rsw.upd) NBX_RC(net->alloc(&net->rsw.upd, 1));
    int32_t *d_ids = nullptr, *d_tp = nullptr;
    hipError_t e = hipMalloc(&d_ids, sizeof(int32_t) * std::max(total, 1));
    if (e == hipSuccess) e = hipMalloc(&d_tp, sizeof(int32_t) * (K + 1));
    if (e == hipSuccess && total) e = hipMemcpy(d_ids, ids, sizeof(int32_t) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tp, tick_ptr, sizeof(int32_t) * (K + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemsetAsync(net->rsw.upd, 0, sizeof(unsigned long long), net->stream);
    if (e == hipSuccess) {
        rc = launch_replay_stdp(net, d_ids, d_tp, t0, lead, nsteps, total, net->rsw.upd);
        if (rc == NBX_OK) e = hipStreamSynchronize(net->stream);
    }
    unsigned long long upd = 0;
    if (e == hipSuccess && rc == NBX_OK)
        e = hipMemcpy(&upd, net->rsw.upd, sizeof(upd), hipMemcpyDeviceToHost);
    hipFree(d_ids);
    hipFree(d_tp);
    if (rc) return rc;
    NBX_HIP_TRY(e);
    NBX_RC(check_device_error(net));
    if (updates) *updates = (int64_t)upd;
    return refresh_persistent_weights(net);
}

int32_t nbx_replay_stdp_device(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0, int32_t lead,
                               int32_t nsteps, int64_t nspikes, uint64_t* d_updates) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->stdp && net->has_sparse, "replay STDP needs sparse synapses with STDP (nbx_net_set_stdp)");
    NBX_CHECK(net->rules.empty(), "replay STDP does not run trigger-driven learning rules");
    NBX_CHECK(d_tick_ptr && nsteps > 0 && lead >= 0 && lead <= t0 && nspikes >= 0, "bad replay arguments");
    NBX_CHECK((int64_t)lead + nsteps <= 65535, "replay STDP covers at most 65535 steps per call (chain calls)");
    NBX_CHECK(nspikes == 0 || d_ids, "null spike ids");
    NBX_RC(device_scope(net));
    NBX_RC(refresh_persistent_weights(net));
    static_assert(sizeof(uint64_t) == sizeof(unsigned long long), "64-bit counter");
    return launch_replay_stdp(net, d_ids, d_tick_ptr, t0, lead, nsteps, nspikes,
                              reinterpret_cast<unsigned long long*>(d_updates));
}

int32_t nbx_replay_stdp_device_noalloc(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0,
                                       int32_t lead, int32_t nsteps, int64_t nspikes, uint64_t* d_updates) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->stdp && net->has_sparse, "replay STDP needs sparse synapses with STDP (nbx_net_set_stdp)");
    NBX_CHECK(net->rules.empty(), "replay STDP does not run trigger-driven learning rules");
    NBX_CHECK(d_tick_ptr && nsteps > 0 && lead >= 0 && lead <= t0 && nspikes >= 0, "bad replay arguments");
    NBX_CHECK((int64_t)lead + nsteps <= 65535, "replay STDP covers at most 65535 steps per call (chain calls)");
    NBX_CHECK(nspikes == 0 || d_ids, "null spike ids");
    // the blocked layouts carry pre-quantized weights; invalidating them would free memory
    NBX_CHECK(!net->pl.built && !net->rs.built, "blocked delivery layouts are built for this network");
    NBX_RC(device_scope(net));
    return launch_replay_stdp_noalloc(net, d_ids, d_tick_ptr, t0, lead, nsteps, nspikes,
                                      reinterpret_cast<unsigned long long*>(d_updates));
}

int32_t nbx_replay_path(nbx_net* net, int32_t* path) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(path, "null argument");
    NBX_CHECK(net->has_sparse, "replay delivery needs sparse synapses");
    NBX_RC(device_scope(net));
    *path = replay_path(net);
    return NBX_OK;
}

int32_t nbx_replay_deliver_device(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t nsteps,
                                  int32_t* d_out) {
    int32_t rc = require_finalized(net);
    if (rc) return rc;
    NBX_CHECK(net->has_sparse, "replay delivery needs sparse synapses");
    NBX_CHECK(d_tick_ptr && d_out && nsteps > 0, "bad replay arguments");
    NBX_RC(device_scope(net));
    return launch_replay(net, d_ids, d_tick_ptr, nsteps, d_out);
}

}  // extern "C"
