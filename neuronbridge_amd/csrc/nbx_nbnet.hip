// .nbnet connection tables -> the network's sparse synapses, sorted on the device.
//
// File format (python/src/neuronbridge/nbnet.py:34-38): a 68-byte little-endian header
// {magic "NBNET01\0", u16 major, u16 minor, u32 flags, u32 record size, u64 metadata bytes,
// u64 connection count, 32-byte SHA-256}, the metadata JSON, then fixed 32-byte records
// {u32 source, u32 target, i32 synapse type, f32 weight, f32 max weight, u32 delay,
// i32 synapse rule, i32 trigger rule}. The SHA-256 (flag bit 0) covers metadata + records.
// Validation follows the reference reader (src/native/shared/src/streaming_build/
// NbnetReader.cpp:533-580 header and size checks, :630-640 per-record checks).
//
// Records stream through pinned host buffers into a device extract kernel (validation,
// sort key, per-presynaptic counts), one radix sort by (pre, delay, post) — stable, so equal
// keys keep file order — and a gather into the packed CSR the engines use. Loading 1e8
// records costs one pass over the file and a device sort instead of a host object graph.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nbx_net.h"

using namespace nbx;

namespace {

constexpr size_t kHeaderSize = 68;
constexpr size_t kRecordSize = 32;
constexpr uint64_t kMaxMetadata = 64ull << 20;
constexpr uint32_t kFlagSha256 = 1;
constexpr size_t kChunkRecords = 1u << 22;  // 128 MiB of records per transfer

// ---------------------------------------------------------------- SHA-256 (FIPS 180-4)
class Sha256 {
public:
    Sha256() { reset(); }
    void reset() {
        static const uint32_t h0[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                       0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
        std::memcpy(h_, h0, sizeof(h_));
        len_ = 0;
        fill_ = 0;
    }
    void update(const uint8_t* p, size_t n) {
        len_ += n;
        if (fill_) {
            const size_t take = std::min(n, (size_t)64 - fill_);
            std::memcpy(buf_ + fill_, p, take);
            fill_ += take;
            p += take;
            n -= take;
            if (fill_ == 64) {
                block(buf_);
                fill_ = 0;
            }
        }
        for (; n >= 64; p += 64, n -= 64) block(p);
        if (n) {
            std::memcpy(buf_, p, n);
            fill_ = n;
        }
    }
    void final(uint8_t out[32]) {
        const uint64_t bits = len_ * 8;
        const uint8_t one = 0x80;
        update(&one, 1);
        const uint8_t zero = 0;
        while (fill_ != 56) update(&zero, 1);
        uint8_t lb[8];
        for (int i = 0; i < 8; ++i) lb[i] = (uint8_t)(bits >> (56 - 8 * i));
        update(lb, 8);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h_[i] >> (24 - 8 * j));
    }

private:
    static uint32_t rotr(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }
    void block(const uint8_t* p) {
        // round constants: first 32 bits of the fractional parts of the cube roots of the first 64 primes
        static const uint32_t K[64] = {
            0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
            0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
            0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
            0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
            0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
            0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
            0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
            0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = h + S1 + ch + K[i] + w[i];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
            const uint32_t t2 = S0 + mj;
            h = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h_[0] += a;
        h_[1] += b;
        h_[2] += c;
        h_[3] += d;
        h_[4] += e;
        h_[5] += f;
        h_[6] += g;
        h_[7] += h;
    }
    uint32_t h_[8];
    uint64_t len_;
    uint8_t buf_[64];
    size_t fill_;
};

template <typename T>
T rd(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Header {
    uint16_t major = 0, minor = 0;
    uint32_t flags = 0, record_size = 0;
    uint64_t metadata = 0, count = 0, file_bytes = 0;
    uint8_t sha[32] = {};
};

// Opens and validates the header and the file size; on success the file is positioned at
// the metadata.
int32_t open_nbnet(const char* path, File& file, Header& h) {
    if (!path) {
        set_error("null path");
        return NBX_ERR_INVALID;
    }
    file.f = std::fopen(path, "rb");
    if (!file.f) {
        set_error(std::string("cannot open nbnet file: ") + path);
        return NBX_ERR_INVALID;
    }
    if (fseeko(file.f, 0, SEEK_END) != 0) {
        set_error(std::string("cannot determine nbnet file size: ") + path);
        return NBX_ERR_INVALID;
    }
    h.file_bytes = (uint64_t)ftello(file.f);
    fseeko(file.f, 0, SEEK_SET);
    uint8_t hb[kHeaderSize];
    if (h.file_bytes < kHeaderSize || std::fread(hb, 1, kHeaderSize, file.f) != kHeaderSize) {
        set_error("truncated nbnet header");
        return NBX_ERR_INVALID;
    }
    if (std::memcmp(hb, "NBNET01\0", 8) != 0) {
        set_error("invalid nbnet magic");
        return NBX_ERR_INVALID;
    }
    h.major = rd<uint16_t>(hb + 8);
    h.minor = rd<uint16_t>(hb + 10);
    h.flags = rd<uint32_t>(hb + 12);
    h.record_size = rd<uint32_t>(hb + 16);
    h.metadata = rd<uint64_t>(hb + 20);
    h.count = rd<uint64_t>(hb + 28);
    std::memcpy(h.sha, hb + 36, 32);
    if (h.major != 1) {
        set_error("unsupported nbnet major version");
        return NBX_ERR_INVALID;
    }
    if (h.record_size != kRecordSize) {
        set_error("unsupported nbnet connection record size");
        return NBX_ERR_INVALID;
    }
    if (h.metadata > kMaxMetadata) {
        set_error("nbnet metadata exceeds 64 MiB");
        return NBX_ERR_INVALID;
    }
    if (h.count > (UINT64_MAX - kHeaderSize - h.metadata) / kRecordSize) {
        set_error("nbnet connection table size overflows uint64");
        return NBX_ERR_INVALID;
    }
    if (kHeaderSize + h.metadata + h.count * kRecordSize != h.file_bytes) {
        set_error("nbnet file size mismatch");
        return NBX_ERR_INVALID;
    }
    return NBX_OK;
}

// SHA-256 over metadata + records; leaves the file positioned at the records.
int32_t verify_checksum(File& file, const Header& h) {
    if ((h.flags & kFlagSha256) == 0) {
        set_error("nbnet SHA-256 checksum is missing");
        return NBX_ERR_INVALID;
    }
    Sha256 sha;
    std::vector<uint8_t> buf(8u << 20);
    fseeko(file.f, (off_t)kHeaderSize, SEEK_SET);
    uint64_t left = h.file_bytes - kHeaderSize;
    while (left) {
        const size_t take = (size_t)std::min<uint64_t>(left, buf.size());
        if (std::fread(buf.data(), 1, take, file.f) != take) {
            set_error("truncated nbnet file");
            return NBX_ERR_INVALID;
        }
        sha.update(buf.data(), take);
        left -= take;
    }
    uint8_t dig[32];
    sha.final(dig);
    if (std::memcmp(dig, h.sha, 32) != 0) {
        set_error("nbnet checksum mismatch");
        return NBX_ERR_INVALID;
    }
    return NBX_OK;
}

enum : uint32_t {
    kBadId = 1,
    kBadWeight = 2,
    kBadDelay = 4,
    kBadType = 8,
    kBadRule = 16,
    kHasRule = 32,  // not an error: some record names a learning or trigger rule
};

// Record j of a chunk -> sort key (pre << 30 | delay << 24 | post), file index, type, weight,
// plastic flag, rule / trigger ids and weight bound; counts per presynaptic neuron; validation bits.
__global__ void nbx_nbnet_extract(const uint32_t* __restrict__ rec, int64_t count, int64_t base, int32_t n,
                                  uint64_t* __restrict__ key, uint32_t* __restrict__ idx, uint8_t* __restrict__ type,
                                  float* __restrict__ w, uint8_t* __restrict__ plastic, int3* __restrict__ rules,
                                  int32_t* __restrict__ cnt, uint32_t* __restrict__ bad,
                                  int32_t* __restrict__ max_delay, int32_t max_type) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t* r = rec + j * 8;
        const uint32_t src = r[0], dst = r[1], dly = r[5];
        const int32_t ty = (int32_t)r[2], rule = (int32_t)r[6], trig = (int32_t)r[7];
        const float wt = __uint_as_float(r[3]), mw = __uint_as_float(r[4]);
        uint32_t b = 0;
        if (src >= (uint32_t)n || dst >= (uint32_t)n) b |= kBadId;
        if (!isfinite(wt) || !isfinite(mw)) b |= kBadWeight;
        if (dly < 1u || dly > (uint32_t)kMaxDelay) b |= kBadDelay;
        if (ty < 0 || ty > max_type) b |= kBadType;
        if (rule < -1 || trig < -1) b |= kBadRule;  // Network.cu:392-399
        if (b) {
            atomicOr(bad, b);
            continue;
        }
        const int64_t i = base + j;
        key[i] = (uint64_t)src << 30 | (uint64_t)dly << 24 | (uint64_t)dst;
        idx[i] = (uint32_t)i;
        type[i] = (uint8_t)ty;
        w[i] = wt;
        plastic[i] = rule >= 0 ? 1 : 0;
        rules[i] = make_int3(rule, trig, (int32_t)r[4]);
        if (rule >= 0 || trig >= 0) atomicOr(bad, (uint32_t)kHasRule);
        atomicAdd(&cnt[src], 1);
        atomicMax(max_delay, (int32_t)dly);
    }
}

__global__ void nbx_nbnet_gather(const uint64_t* __restrict__ key, const uint32_t* __restrict__ idx, int64_t count,
                                 const uint8_t* __restrict__ type, const float* __restrict__ w_in,
                                 const uint8_t* __restrict__ pl_in, const int3* __restrict__ rules_in,
                                 uint32_t* __restrict__ syn, float* __restrict__ w, uint8_t* __restrict__ plastic,
                                 int3* __restrict__ rules) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        const uint32_t r = idx[i];
        syn[i] = pack_syn((int32_t)(k & 0xFFFFFFu), (int32_t)((k >> 24) & 0x3Fu), type[r]);
        w[i] = w_in[r];
        plastic[i] = pl_in[r];
        rules[i] = rules_in[r];
    }
}

__global__ void nbx_nbnet_widen(const int32_t* __restrict__ cnt, int64_t* __restrict__ out, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) out[i] = i < n ? cnt[i] : 0;
}

// Device buffers of one load, freed on every exit path.
struct Scratch {
    std::vector<void*> dev;
    void* pinned[2] = {nullptr, nullptr};
    ~Scratch() {
        for (void* p : dev) (void)hipFree(p);
        for (void* p : pinned)
            if (p) (void)hipHostFree(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) dev.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
};

int blocks_for(int64_t count) { return (int)std::min<int64_t>((count + 255) / 256, 8192); }

}  // namespace

extern "C" {

int32_t nbx_nbnet_info(const char* path, int32_t verify_checksum_flag, nbx_nbnet_header* info) {
    if (!info) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    File file;
    Header h;
    int32_t rc = open_nbnet(path, file, h);
    if (rc) return rc;
    if (verify_checksum_flag && (rc = verify_checksum(file, h))) return rc;
    info->file_bytes = (int64_t)h.file_bytes;
    info->metadata_bytes = (int64_t)h.metadata;
    info->connection_count = (int64_t)h.count;
    info->version_major = h.major;
    info->version_minor = h.minor;
    info->flags = h.flags;
    info->checksum_verified = verify_checksum_flag ? 1 : 0;
    return NBX_OK;
}

int32_t nbx_nbnet_metadata(const char* path, char* buf, int64_t cap, int64_t* len) {
    if (!len) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    File file;
    Header h;
    int32_t rc = open_nbnet(path, file, h);
    if (rc) return rc;
    *len = (int64_t)h.metadata;
    if (!buf) return NBX_OK;  // size query
    if (cap < (int64_t)h.metadata) {
        set_error("metadata buffer too small");
        return NBX_ERR_INVALID;
    }
    if (std::fread(buf, 1, (size_t)h.metadata, file.f) != (size_t)h.metadata) {
        set_error("truncated nbnet metadata");
        return NBX_ERR_INVALID;
    }
    return NBX_OK;
}

int32_t nbx_net_load_nbnet(nbx_net* net, const char* path, int32_t verify_checksum_flag) {
    if (!net) {
        set_error("null network");
        return NBX_ERR_INVALID;
    }
    if (net->finalized) {
        set_error("synapses must be loaded before nbx_net_finalize");
        return NBX_ERR_STATE;
    }
    File file;
    Header h;
    int32_t rc = open_nbnet(path, file, h);
    if (rc) return rc;
    if (verify_checksum_flag && (rc = verify_checksum(file, h))) return rc;
    if (h.count >= ((uint64_t)1 << 31)) {
        set_error("more than 2^31 - 1 connections");
        return NBX_ERR_UNSUPPORTED;
    }
    const int64_t S = (int64_t)h.count;
    const int32_t n = net->n;
    NBX_HIP_TRY(hipSetDevice(net->device));
    Scratch sc;
    uint64_t *key = nullptr, *key2 = nullptr;
    uint32_t *idx = nullptr, *idx2 = nullptr, *syn = nullptr, *bad = nullptr, *rec = nullptr;
    uint8_t *type = nullptr, *pl = nullptr, *pl_out = nullptr;
    float *w = nullptr, *w_out = nullptr;
    int32_t *cnt = nullptr, *maxd = nullptr;
    int64_t* row = nullptr;
    NBX_HIP_TRY(sc.alloc(&key, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&key2, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&idx, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&idx2, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&type, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&pl, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&w, (size_t)S));
    int3 *rules = nullptr, *rules_out = nullptr;
    NBX_HIP_TRY(sc.alloc(&rules, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&cnt, (size_t)n));
    NBX_HIP_TRY(sc.alloc(&row, (size_t)n + 1));
    NBX_HIP_TRY(sc.alloc(&bad, 1));
    NBX_HIP_TRY(sc.alloc(&maxd, 1));
    NBX_HIP_TRY(sc.alloc(&rec, kChunkRecords * 8));
    for (int i = 0; i < 2; ++i) NBX_HIP_TRY(hipHostMalloc(&sc.pinned[i], kChunkRecords * kRecordSize));
    hipStream_t st = net->stream;
    NBX_HIP_TRY(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (size_t)n, st));
    NBX_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(uint32_t), st));
    NBX_HIP_TRY(hipMemsetAsync(maxd, 0, sizeof(int32_t), st));

    // stream the records: the file read of chunk k+1 overlaps the copy and extraction of chunk
    // k (two pinned buffers; a pinned buffer is reused once its copy has completed)
    fseeko(file.f, (off_t)(kHeaderSize + h.metadata), SEEK_SET);
    hipEvent_t copied[2] = {nullptr, nullptr};
    struct Events {
        hipEvent_t* e;
        ~Events() {
            for (int i = 0; i < 2; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } ev_guard{copied};
    for (int i = 0; i < 2; ++i) NBX_HIP_TRY(hipEventCreateWithFlags(&copied[i], hipEventDisableTiming));
    int cur = 0;
    for (int64_t base = 0; base < S; base += (int64_t)kChunkRecords) {
        const int64_t c = std::min<int64_t>((int64_t)kChunkRecords, S - base);
        NBX_HIP_TRY(hipEventSynchronize(copied[cur]));  // the copy that last read this buffer is done
        if (std::fread(sc.pinned[cur], kRecordSize, (size_t)c, file.f) != (size_t)c) {
            set_error("truncated nbnet connection table");
            return NBX_ERR_INVALID;
        }
        // rec is reused in stream order: the previous chunk's extraction precedes this copy
        NBX_HIP_TRY(hipMemcpyAsync(rec, sc.pinned[cur], (size_t)c * kRecordSize, hipMemcpyHostToDevice, st));
        NBX_HIP_TRY(hipEventRecord(copied[cur], st));
        nbx_nbnet_extract<<<blocks_for(c), 256, 0, st>>>(rec, c, base, n, key, idx, type, w, pl, rules, cnt, bad, maxd,
                                                          net_max_type(net));
        NBX_HIP_TRY(hipGetLastError());
        cur ^= 1;
    }
    NBX_HIP_TRY(hipStreamSynchronize(st));
    uint32_t bad_h = 0;
    int32_t maxd_h = 0;
    NBX_HIP_TRY(hipMemcpy(&bad_h, bad, sizeof(bad_h), hipMemcpyDeviceToHost));
    NBX_HIP_TRY(hipMemcpy(&maxd_h, maxd, sizeof(maxd_h), hipMemcpyDeviceToHost));
    if (bad_h & kBadId) {
        set_error("nbnet connection neuron id is out of range");
        return NBX_ERR_INVALID;
    }
    if (bad_h & kBadWeight) {
        set_error("nbnet connection weight is not finite");
        return NBX_ERR_INVALID;
    }
    if (bad_h & kBadDelay) {
        set_error("nbnet connection delay outside [1, 63] steps");
        return NBX_ERR_UNSUPPORTED;
    }
    if (bad_h & kBadType) {
        set_error(net_max_type(net) == 2
                      ? "nbnet synapse type must be 0 (AMPA), 1 (GABA) or 2 (NMDA)"
                      : "nbnet synapse type must be 0 (excitatory) or 1 (inhibitory)");
        return NBX_ERR_UNSUPPORTED;
    }
    if (bad_h & kBadRule) {
        set_error("nbnet connection references an invalid learning rule");
        return NBX_ERR_INVALID;
    }

    // sort by (pre, delay, post); radix sort is stable, so equal keys keep file order
    size_t tmp_bytes = 0;
    hipcub::DoubleBuffer<uint64_t> kb(key, key2);
    hipcub::DoubleBuffer<uint32_t> vb(idx, idx2);
    if (S) {
        NBX_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, kb, vb, (int)S, 0, 54, st));
        uint8_t* tmp = nullptr;
        NBX_HIP_TRY(sc.alloc(&tmp, tmp_bytes));
        NBX_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kb, vb, (int)S, 0, 54, st));
    }
    NBX_HIP_TRY(sc.alloc(&syn, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&w_out, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&pl_out, (size_t)S));
    NBX_HIP_TRY(sc.alloc(&rules_out, (size_t)S));
    if (S) {
        nbx_nbnet_gather<<<blocks_for(S), 256, 0, st>>>(kb.Current(), vb.Current(), S, type, w, pl, rules, syn, w_out,
                                                        pl_out, rules_out);
        NBX_HIP_TRY(hipGetLastError());
    }
    // row starts: exclusive scan of the per-presynaptic counts
    int64_t* cnt64 = nullptr;
    NBX_HIP_TRY(sc.alloc(&cnt64, (size_t)n + 1));
    nbx_nbnet_widen<<<blocks_for(n + 1), 256, 0, st>>>(cnt, cnt64, n);
    NBX_HIP_TRY(hipGetLastError());
    size_t scan_bytes = 0;
    NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cnt64, row, n + 1, st));
    uint8_t* scan_tmp = nullptr;
    NBX_HIP_TRY(sc.alloc(&scan_tmp, scan_bytes));
    NBX_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, cnt64, row, n + 1, st));

    std::vector<int64_t> h_row((size_t)n + 1);
    std::vector<uint32_t> h_syn((size_t)S), h_src((size_t)S);
    std::vector<float> h_w((size_t)S);
    std::vector<uint8_t> h_pl((size_t)S);
    NBX_HIP_TRY(hipMemcpyAsync(h_row.data(), row, sizeof(int64_t) * h_row.size(), hipMemcpyDeviceToHost, st));
    if (S) {
        NBX_HIP_TRY(hipMemcpyAsync(h_syn.data(), syn, sizeof(uint32_t) * (size_t)S, hipMemcpyDeviceToHost, st));
        NBX_HIP_TRY(hipMemcpyAsync(h_w.data(), w_out, sizeof(float) * (size_t)S, hipMemcpyDeviceToHost, st));
        NBX_HIP_TRY(hipMemcpyAsync(h_pl.data(), pl_out, (size_t)S, hipMemcpyDeviceToHost, st));
        NBX_HIP_TRY(hipMemcpyAsync(h_src.data(), vb.Current(), sizeof(uint32_t) * (size_t)S, hipMemcpyDeviceToHost, st));
    }
    const bool has_rules = (bad_h & kHasRule) != 0;
    std::vector<int3> h_rules(has_rules ? (size_t)S : 0);
    if (has_rules && S)
        NBX_HIP_TRY(hipMemcpyAsync(h_rules.data(), rules_out, sizeof(int3) * (size_t)S, hipMemcpyDeviceToHost, st));
    NBX_HIP_TRY(hipStreamSynchronize(st));
    net->h_file_rule.clear();
    net->h_file_trig.clear();
    net->h_file_maxw.clear();
    if (has_rules) {  // kept in CSR order for nbx_net_assign_nbnet_rules
        net->h_file_rule.resize((size_t)S);
        net->h_file_trig.resize((size_t)S);
        net->h_file_maxw.resize((size_t)S);
        for (size_t j = 0; j < (size_t)S; ++j) {
            net->h_file_rule[j] = h_rules[j].x;
            net->h_file_trig[j] = h_rules[j].y;
            std::memcpy(&net->h_file_maxw[j], &h_rules[j].z, sizeof(float));
        }
    }
    bool any_plastic = false;
    for (uint8_t v : h_pl) any_plastic |= v != 0;
    net->h_row_ptr = std::move(h_row);
    net->h_syn = std::move(h_syn);
    net->h_w = std::move(h_w);
    net->h_src_index = std::move(h_src);
    if (any_plastic) net->h_plastic = std::move(h_pl);
    else net->h_plastic.clear();
    net->S = S;
    net->max_delay = std::max(net->max_delay, std::max(1, maxd_h));
    net->has_sparse = true;
    return NBX_OK;
}

int32_t nbx_net_assign_nbnet_rules(nbx_net* net, const nbx_rule_desc* rules, int32_t count) {
    if (!net || (!rules && count > 0)) {
        set_error("null argument");
        return NBX_ERR_INVALID;
    }
    if (net->finalized) {
        set_error("learning rules must be assigned before nbx_net_finalize");
        return NBX_ERR_STATE;
    }
    if (!net->has_sparse || net->h_src_index.size() != (size_t)net->S) {
        set_error("assign .nbnet rules after nbx_net_load_nbnet");
        return NBX_ERR_STATE;
    }
    if (count < 0 || count > kMaxRules) {
        set_error("rule count must be in [0, 64]");
        return NBX_ERR_INVALID;
    }
    int32_t stdp_rule = -1;
    for (int32_t r = 0; r < count; ++r) {
        const nbx_rule_desc& q = rules[r];
        if (const char* why = rule_desc_error(q, true)) {
            set_error(why);
            return NBX_ERR_INVALID;
        }
        if (q.kind != NBX_RULE_STDP) continue;
        if (stdp_rule >= 0) {
            const nbx_rule_desc& a = rules[stdp_rule];
            if (a.max_ltp != q.max_ltp || a.ltp_tau != q.ltp_tau || a.max_ltd != q.max_ltd || a.ltd_tau != q.ltd_tau) {
                set_error("STDP rules with different parameters in one network");
                return NBX_ERR_UNSUPPORTED;
            }
        }
        stdp_rule = r;
    }
    const size_t S = (size_t)net->S;
    const bool has = !net->h_file_rule.empty();
    std::vector<uint8_t> plastic(S, 0);
    std::vector<int32_t> rule(S, -1), trig(S, -1);
    bool any_stdp = false, any_rule = false;
    for (size_t j = 0; j < S && has; ++j) {
        const int32_t r = net->h_file_rule[j], tr = net->h_file_trig[j];
        if (r >= count || tr >= count) {
            set_error("nbnet connection references an invalid learning rule");
            return NBX_ERR_INVALID;
        }
        if (r >= 0 && rules[r].kind == NBX_RULE_STDP) {
            plastic[j] = 1;
            any_stdp = true;
        } else if (r >= 0) {
            rule[j] = r;
            any_rule = true;
        }
        if (tr >= 0 && rules[tr].kind != NBX_RULE_STDP) {
            trig[j] = tr;
            any_rule = true;
        }
        if (trig[j] >= 0 && r >= 0) {
            set_error("a synapse that is both plastic and a trigger is not supported");
            return NBX_ERR_UNSUPPORTED;
        }
    }
    net->stdp = any_stdp;
    net->h_smaxw.clear();
    if (any_stdp) {
        const nbx_rule_desc& q = rules[stdp_rule];
        net->sd.max_ltp = q.max_ltp;
        net->sd.ltp_tau = q.ltp_tau;
        net->sd.max_ltd = q.max_ltd;
        net->sd.ltd_tau = q.ltd_tau;
        net->sd.w_max = 0.0f;  // bounds are per synapse
        net->h_plastic.swap(plastic);
        net->h_smaxw = net->h_file_maxw;
    } else {
        net->h_plastic.clear();
    }
    net->rules.clear();
    net->h_rule.clear();
    net->h_trig.clear();
    net->h_maxw.clear();
    if (any_rule) {
        net->rules.assign(rules, rules + count);
        net->h_rule.resize(S);
        net->h_trig.resize(S);
        for (size_t j = 0; j < S; ++j) {
            net->h_rule[j] = (int8_t)rule[j];
            net->h_trig[j] = (int8_t)trig[j];
        }
        net->h_maxw = net->h_file_maxw;
    }
    return NBX_OK;
}

}  // extern "C"
