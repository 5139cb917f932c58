// The network object behind the opaque nbx_net handle of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nbx_internal.h"

using nbx::ModelParams;
using nbx::StepArgs;
using nbx::RuleDev;
using nbx::SynTrace;
using nbx::TraceHead;

// Post-blocked layout of the persistent engine and of replay delivery: block b owns
// neurons [b * P, min((b + 1) * P, n)). Blocks are grouped into regions of rb blocks
// (rb = B / 8 rounded up to an even count, so pairs of blocks never straddle a region).
// The synapses of presynaptic neuron i that target block b are seg[o0 .. o1) with
//   o0 = off[i * nr * gs + (b / rb) * gs + b % rb],   o1 = the next entry of that region,
// i.e. each region keeps rb + 1 running offsets in its own 128-byte aligned group of gs
// ints, and each region's synapses start on a 128-byte boundary in seg (padding entries
// between regions are never inside any [o0, o1)).
struct PersistentLayout {
    bool built = false;
    bool fits = false;
    std::string reason;
    int32_t B = 0, P = 0, threads = 0, npt = 0, lds_bytes = 0;
    int32_t rb = 0, nr = 0, gs = 0;  // blocks per region, regions, ints per region in an offset row
    int64_t seg_count = 0;           // entries in seg, region padding included
    int64_t max_row = 0;             // longest presynaptic row
    int32_t* d_off = nullptr;        // [n][nr * gs]
    uint2* d_seg = nullptr;          // {post_local | delay << 16 (6 bits) | channel << 22 (2 bits) | (block & 127) << 25, q}
    unsigned long long* d_glist = nullptr;  // [D][kGroups][gcap] the steps' spike lists {tag << 32 | id}
    unsigned long long* d_gctr = nullptr;   // [2 banks][D][kGroups][16] their counters {spikes << 32 | arrivals}
    int32_t gcap = 0;                       // list capacity: the neurons of one group
    int32_t t_next = -1;                    // step at which the last persistent launch ended (-1: none yet)
    int32_t bank = 0;                       // step-counter bank of the next launch (the other is cleared by it)
    uint32_t epoch = 0;                     // the next launch tags its steps epoch + 1, epoch + 2, ...
    bool launch_checked = false;     // LDS attribute set and co-residency of the B workgroups checked
};

// Fixed-slot layout of replay delivery: the synapses of presynaptic neuron i onto the m*P
// neurons of post block q sit in one 128-byte slot of 16 entries, slots[(q * n + i) * 16 ..]
// (block-major: one workgroup's slots are contiguous), addressable without an offset lookup. Unused entries hold kSlotEmpty; a segment longer than
// 16 keeps 15 entries in the slot and a link in the last one ({kSlotLink | count, start}) to the
// rest, stored contiguously in ovf. Entry: {channel*m*P + post - q*m*P | (delay - 1) << 16, q}:
// the entry's LDS word within one ring step, and its output step offset.
struct ReplaySlots {
    bool built = false;
    bool packed = false;       // 6-byte entries (20 per slot), meta = (delay - 1) << wbits | LDS word
    int32_t m = 0, Bq = 0, wbits = 0;
    uint2* d_slots = nullptr;  // [Bq][n][16] 8-byte entries, or [Bq][n] 128-byte packed slots
    uint2* d_ovf = nullptr;
    int64_t ovf_count = 0;
};

// Replay-STDP workspace: per-neuron spike lists of the replayed raster (grown on demand).
struct ReplayStdpWork {
    int32_t* cnt = nullptr;  // [n + 1] spike counts
    int32_t* ptr = nullptr;  // [n + 1] list starts
    int32_t* k = nullptr;    // [cap] local steps, ascending per neuron
    int64_t cap = 0;
    unsigned long long* mask = nullptr;  // [blocks][n] spike bits per 64-step block
    int32_t* blk = nullptr;              // [blocks][n] list position of each block's first spike
    int64_t mcap = 0;                    // n * blocks allocated
    float2* tab2 = nullptr;              // [tab_len] {ltp_tab, ltd_tab} interleaved
    uint16_t* fix = nullptr;             // [n][20] first spike steps of each neuron, 0xFFFF after the last
    uint8_t* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    unsigned long long* upd = nullptr;  // update counter of the host entry point
};

// Time-series monitor (nbx_net_monitor_start): after every sampled step a frame of device copies
// (state, the step's firing list, pending arrivals, weights) goes into the next free
// slot of the frame buffers; when they are full, or at a flush, the held frames are written to the
// DebugMonitor CSV tables in dir.
struct NetMonitor {
    bool on = false;
    std::string dir, name;
    int32_t interval = 1, flush_every = 100;
    bool spikes = false, state = false, weights = false, pending = false;
    std::vector<uint8_t> sel;    // [n] selected neurons
    int32_t cap = 0, held = 0;   // frames the buffers hold; frames captured and not yet written
    std::vector<int32_t> steps;  // step of each held frame
    float* d_state = nullptr;    // [cap][ns][n]
    int32_t* d_spk = nullptr;    // [cap][n + 1]: the step's firing count, then its ids
    int32_t* d_pend = nullptr;   // [cap][nch][n] arrivals pending for the next step
    float* d_w = nullptr;        // [cap][S]
    std::vector<void*> allocs;   // freed at stop
    int64_t frames_written = 0;
};

// Current synapses (synapse type 3): per source the synapses it drives, per target its slots in
// declaration order; currents injected at a source wait on the host, keyed by arrival step, until
// that step is run.
struct CurrentSynapses {
    int32_t count = 0;
    std::vector<int64_t> out_ptr;   // [n + 1] synapses of each source
    std::vector<int32_t> out_syn;   // synapse ids by source
    std::vector<int32_t> delay;     // [count]
    std::vector<int32_t> target;    // [count]
    std::vector<int32_t> slot;      // [count] global slot (target-major, declaration order)
    int32_t* d_in_ptr = nullptr;    // [n + 1] slots of each target (int32: count < 2^31)
    float* d_slots = nullptr;       // [count] the slot currents
    std::vector<std::vector<std::pair<int32_t, float>>> pending;  // arrival step - base: (synapse, value)
    int32_t base = 0;               // step of pending[0]
    int32_t* d_stage = nullptr;     // staging of one step's (slot, value-bits) pairs and targets
    int64_t stage_cap = 0;
    std::vector<void*> allocs;
};

// One neuron population: neurons [begin, begin + md.n) with one model and its parameters.
struct NetPop {
    int32_t begin = 0;
    nbx_model_desc md{};
    ModelParams P{};
};

struct nbx_net {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    nbx_model_desc md{};  // the first population's (dt, integrator and the model of a one-population net)
    ModelParams P{};      // its parameters; P.model = kModelTyped (typed delivery) for several populations
    std::vector<NetPop> pops;
    std::vector<uint8_t> h_pop_of;  // [n] population of each neuron (several populations)
    bool multi() const { return pops.size() > 1; }
    int32_t model_of(int32_t i) const {
        return multi() ? pops[h_pop_of[(size_t)i]].md.model : md.model;
    }
    // largest synapse type any population accepts (each synapse is checked against its target's
    // population at finalize)
    friend int32_t net_max_type(const nbx_net* net) {
        int32_t m = 0;
        for (const NetPop& p : net->pops) m = std::max(m, nbx::model_max_type(p.md.model));
        return m;
    }
    int32_t n = 0, ns = 0, nch = 1;
    std::vector<float> h_a, h_b, h_c, h_d;

    // sparse synapses (host copies until finalize)
    bool has_sparse = false;
    int64_t S = 0;
    int32_t max_delay = 1;
    std::vector<int64_t> h_row_ptr;
    std::vector<uint32_t> h_syn;
    std::vector<float> h_w;
    std::vector<uint8_t> h_plastic;
    std::vector<uint32_t> h_src_index;  // file record index of each CSR synapse (networks loaded from .nbnet)
    // dense
    bool has_dense = false;
    std::vector<float> h_wt;
    int32_t dense_delay = 1;
    // Poisson background
    bool pois = false;
    uint32_t seed_lo = 0, seed_hi = 0;
    double lambda = 0.0;
    float w_bg = 0.0f;
    std::vector<uint32_t> pois_tab;
    int32_t kmax = 0;
    // STDP
    bool stdp = false;
    nbx_stdp_desc sd{};
    // trigger-driven rules
    std::vector<nbx_rule_desc> rules;
    std::vector<int8_t> h_rule, h_trig;
    std::vector<float> h_maxw;
    std::vector<int32_t> h_file_rule, h_file_trig;  // a loaded .nbnet table's rule ids (CSR order)
    std::vector<float> h_file_maxw;                 // and its records' max_weight
    std::vector<float> h_smaxw;                     // per-synapse STDP weight bound (else sd.w_max)
    // current
    std::vector<float> h_cur;
    int32_t cur_t0 = 0, cur_n = 0;

    int32_t shift_req = 0, shift = 0;
    bool finalized = false;
    bool pull_mode = false;
    int32_t D = 2;

    // device buffers
    std::vector<void*> allocs;
    float *d_v = nullptr, *d_s1 = nullptr, *d_ga = nullptr, *d_gg = nullptr, *d_iext = nullptr;
    int32_t* d_last = nullptr;
    float *d_a = nullptr, *d_b = nullptr, *d_c = nullptr, *d_d = nullptr;
    float* d_g3 = nullptr;          // Izhikevich NMDA conductance (networks with such synapses)
    uint8_t* d_pop_of = nullptr;    // several populations: [n] population of each neuron
    ModelParams* d_pops = nullptr;  // and their parameters
    std::vector<float> h_init_v, h_init_s1;  // state at finalize (ResetState)
    // external inputs: interface slot map, staged host records, device staging, output mask
    std::vector<int32_t> iface_targets;
    std::vector<nbx::InputRec> iface_staged;
    nbx::InputRec* d_inputs = nullptr;  // device staging of input records (grown on demand)
    int64_t inputs_cap = 0;
    float* d_pending_f = nullptr;       // float snapshot of a step's pending arrivals
    uint8_t* d_fire_flags = nullptr;    // did-fire flags of a step (device view)
    std::vector<uint8_t> out_mask;      // output neurons (empty: all)
    int32_t crows = 0;                  // cerebellar arrival-history rows
    int32_t *d_ring = nullptr, *d_hist = nullptr, *d_hist_cnt = nullptr;
    uint64_t* d_fmask = nullptr;
    int64_t* d_row_ptr = nullptr;
    uint32_t* d_syn = nullptr;
    float* d_w = nullptr;
    uint8_t* d_plastic = nullptr;
    int64_t* d_dptr = nullptr;
    float* d_wt = nullptr;
    uint32_t* d_pois_tab = nullptr;
    float* d_cur = nullptr;
    SynTrace* d_trace = nullptr;   // networks with learning rules
    TraceHead* d_heads = nullptr;  // pair-STDP records: d_trace's heads (stride 2) or an array of their own
    int32_t head_stride = 1;
    float *d_ltp_tab = nullptr, *d_ltd_tab = nullptr;
    int32_t tab_len = 1;
    int64_t *d_in_ptr = nullptr, *d_in_syn = nullptr;
    int32_t* d_in_pd = nullptr;
    RuleDev* d_rules = nullptr;
    int8_t *d_syn_rule = nullptr, *d_syn_trig = nullptr;
    float* d_rtab = nullptr;
    int64_t *d_rin_ptr = nullptr, *d_rin_syn = nullptr, *d_tin_ptr = nullptr, *d_tin_syn = nullptr;
    int32_t *d_tstamp = nullptr, *d_tpost_stamp = nullptr, *d_tposts = nullptr, *d_tcount = nullptr;
    int32_t *d_conn_index = nullptr, *d_cidx = nullptr, *d_chist = nullptr, *d_chead = nullptr;
    int32_t ccap = 0;
    bool rstdp = false;
    unsigned long long* d_events = nullptr;
    int32_t* d_tbase = nullptr;
    int32_t* d_stage = nullptr;  // [n] staging of host-given neuron ids (injected spikes)
    int32_t* d_err = nullptr;
    int32_t* d_counts = nullptr;
    uint32_t* d_raster = nullptr;
    int32_t rec_t0 = 0, rec_n = 0, words = 0;

    int32_t t = 0;
    // graph engine
    hipGraphExec_t gexec = nullptr;
    int32_t gchunk = 0;
    // per-phase sequence: 0 idle, 1 begun, 2 propagated, 3 updated, 4 learned
    int32_t phase = 0;
    // persistent engine
    PersistentLayout pl;
    // replay delivery
    ReplaySlots rs;
    // replay STDP
    ReplayStdpWork rsw;
    // time-series monitor
    NetMonitor mon;
    // current synapses (type 3)
    CurrentSynapses cs;

    StepArgs args() const {
        StepArgs A{};
        A.P = P;
        A.n = n;
        A.nch = nch;
        A.D = D;
        A.words = words;
        A.v = d_v;
        A.s1 = d_s1;
        A.ga = d_ga;
        A.gg = d_gg;
        A.iext = d_iext;
        A.g3 = d_g3;
        A.last = d_last;
        A.pop_of = d_pop_of;
        A.pops = d_pops;
        A.a = d_a;
        A.b = d_b;
        A.c = d_c;
        A.d = d_d;
        A.ring = d_ring;
        A.hist = d_hist;
        A.hist_cnt = d_hist_cnt;
        A.fmask = d_fmask;
        A.row_ptr = d_row_ptr;
        A.syn = d_syn;
        A.w = d_w;
        A.dptr = d_dptr;
        A.plastic = d_plastic;
        A.wt = d_wt;
        A.dense_delay = dense_delay;
        A.scale = std::ldexp(1.0f, shift);
        A.inv_scale = 1.0f / A.scale;
        A.events = d_events;
        A.pois = pois ? 1 : 0;
        A.seed_lo = seed_lo;
        A.seed_hi = seed_hi;
        A.kmax = kmax;
        A.pois_tab = d_pois_tab;
        A.q_bg = pois ? (int32_t)std::lrint(w_bg * A.scale) : 0;
        A.cur = d_cur;
        A.cur_t0 = cur_t0;
        A.cur_n = cur_n;
        A.stdp = stdp ? 1 : 0;
        A.max_ltp = sd.max_ltp;
        A.max_ltd = sd.max_ltd;
        A.w_max = sd.w_max;
        A.trace = d_trace;
        A.heads = d_heads;
        A.head_stride = head_stride;
        A.ltp_tab = d_ltp_tab;
        A.ltd_tab = d_ltd_tab;
        A.tab_len = tab_len;
        A.in_ptr = d_in_ptr;
        A.in_syn = d_in_syn;
        A.in_pd = d_in_pd;
        A.nrules = (int32_t)rules.size();
        A.rstdp = rstdp ? 1 : 0;
        A.rules = d_rules;
        A.syn_rule = d_syn_rule;
        A.syn_trig = d_syn_trig;
        A.rtab = d_rtab;
        A.rin_ptr = d_rin_ptr;
        A.rin_syn = d_rin_syn;
        A.tin_ptr = d_tin_ptr;
        A.tin_syn = d_tin_syn;
        A.tstamp = d_tstamp;
        A.tpost_stamp = d_tpost_stamp;
        A.tposts = d_tposts;
        A.tcount = d_tcount;
        A.conn_index = d_conn_index;
        A.cidx = d_cidx;
        A.chist = d_chist;
        A.chead = d_chead;
        A.ccap = ccap;
        A.counts = d_counts;
        A.raster = d_raster;
        A.rec_t0 = rec_t0;
        A.rec_n = rec_n;
        A.err = d_err;
        return A;
    }

    template <typename T>
    int32_t alloc(T** p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        void* q = nullptr;
        NBX_HIP_TRY(hipMalloc(&q, count * sizeof(T)));
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return NBX_OK;
    }
    template <typename T>
    int32_t upload(T** p, const T* src, size_t count) {
        int32_t rc = alloc(p, count);
        if (rc) return rc;
        if (count) NBX_HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
        return NBX_OK;
    }
    void release() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        for (void* p : mon.allocs) (void)hipFree(p);
        mon.allocs.clear();
        mon.on = false;
        for (void* p : cs.allocs) (void)hipFree(p);
        cs.allocs.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};


namespace nbx {
bool persistent_supported(const nbx_net* net);
std::string persistent_reason(const nbx_net* net);
int32_t build_persistent_layout(nbx_net* net);
int32_t refresh_persistent_weights(nbx_net* net);
int32_t run_persistent(nbx_net* net, int32_t nsteps);
std::vector<float*> state_columns(nbx_net* net, int32_t model);  // device array of each state column
int32_t monitor_capture(nbx_net* net, int32_t step);  // frame of the completed step (flushes when full)
int32_t currents_due_next(const nbx_net* net, int32_t t);  // first step >= t with current arrivals, or -1
int32_t apply_currents(nbx_net* net, int32_t t);            // arrivals of step t (before its neuron update)
void clear_currents(nbx_net* net);                          // slots and pending currents (ResetState)
std::string small_dense_reason(const nbx_net* net);  // empty when the single-workgroup engine fits
int32_t run_small_dense(nbx_net* net, int32_t nsteps);
int32_t launch_replay(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t nsteps, int32_t* d_out);
int32_t replay_path(const nbx_net* net);  // NBX_REPLAY_GLOBAL / _OFFSETS / _SLOTS
int32_t clear_raster_error(nbx_net* net);
int32_t clear_error_bits(nbx_net* net, int32_t bits);  // clears device error flags (a kernel, then a sync)
int32_t launch_replay_stdp(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0, int32_t lead,
                           int32_t nsteps, int64_t nspikes, unsigned long long* d_updates);
// Same, with the workspace in the library's static device array (no allocation); rejects calls
// whose workspace exceeds it.
constexpr size_t kReplayStdpScratchBytes = NBX_REPLAY_STDP_SCRATCH_BYTES;
int32_t launch_replay_stdp_noalloc(nbx_net* net, const int32_t* d_ids, const int32_t* d_tick_ptr, int32_t t0,
                                   int32_t lead, int32_t nsteps, int64_t nspikes, unsigned long long* d_updates);
// Workspace bytes of one replay-STDP call (-1 when the scan storage query fails).
int64_t replay_stdp_scratch_bytes(const nbx_net* net, int32_t K, int64_t nspikes, size_t* scan_bytes);
const char* rule_desc_error(const nbx_rule_desc& q, bool allow_stdp);  // null when valid
}  // namespace nbx
