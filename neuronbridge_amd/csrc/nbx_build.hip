// Host-side network builder: fixed in-degree random topology straight into the
// CSR-by-presynaptic-neuron layout the delivery kernels read, rows sorted by
// (delay, post). Plays the role of the reference's Network::CaculateOutputConnection
// (Network.cu:483-560: synapses grouped by source, then delay) and of the dense build
// finalizer's per-delay slices (DenseSubnetworkBuildFinalizer.cpp:310-330), without an
// object per synapse: two passes over counter-based random draws and one counting
// sort over (pre, delay) buckets, split over host threads.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "nbx_internal.h"

namespace {

void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]) {
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

constexpr uint32_t kTopoStream = 0x544F504Fu; /* "TOPO" */

// Presynaptic sources and delays of post neuron j: Floyd's sampling of k distinct
// ids from [0, n), one Philox draw per sample (counter = (j, m, stream, 0)).
struct Drawer {
    int32_t n, k, dmin, span;
    uint32_t s0, s1;
    std::vector<uint64_t> bits;
    Drawer(int32_t n_, int32_t k_, int32_t dmin_, int32_t dmax_, uint32_t s0_, uint32_t s1_)
        : n(n_), k(k_), dmin(dmin_), span(dmax_ - dmin_ + 1), s0(s0_), s1(s1_), bits(((size_t)n_ + 63) / 64, 0) {}
    void draw(int32_t j, int32_t* pre, uint8_t* del) {
        for (int32_t m = 0; m < k; ++m) {
            uint32_t o[4];
            philox((uint32_t)j, (uint32_t)m, kTopoStream, 0u, s0, s1, o);
            const int32_t i = n - k + m;
            int32_t r = (int32_t)(((uint64_t)o[0] * (uint64_t)(i + 1)) >> 32);
            if (bits[(size_t)r >> 6] >> (r & 63) & 1u) r = i;
            bits[(size_t)r >> 6] |= 1ull << (r & 63);
            pre[m] = r;
            del[m] = (uint8_t)(dmin + (int32_t)(((uint64_t)o[1] * (uint64_t)span) >> 32));
        }
        for (int32_t m = 0; m < k; ++m) bits[(size_t)pre[m] >> 6] = 0;
    }
};

}  // namespace

extern "C" int32_t nbx_build_fixed_indegree(int32_t n, int32_t k_in, int32_t d_min, int32_t d_max, uint32_t seed_lo,
                                            uint32_t seed_hi, int32_t n_exc, int32_t nthreads, int64_t* row_ptr,
                                            int32_t* post, uint8_t* delay, uint8_t* type) {
    if (n <= 0 || k_in < 0 || k_in > n || d_min < 1 || d_max < d_min || d_max > nbx::kMaxDelay || !row_ptr ||
        (k_in > 0 && (!post || !delay || !type))) {
        nbx::set_error("nbx_build_fixed_indegree: bad arguments");
        return NBX_ERR_INVALID;
    }
    const int32_t D = d_max + 1;
    int T = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    T = std::min(T, 64);
    T = std::max(1, std::min(T, n));
    // per-thread bucket counts over (pre, delay); thread t owns posts [lo_t, hi_t)
    std::vector<std::vector<int64_t>> cnt((size_t)T);
    auto range = [&](int t, int32_t* lo, int32_t* hi) {
        *lo = (int32_t)((int64_t)n * t / T);
        *hi = (int32_t)((int64_t)n * (t + 1) / T);
    };
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                cnt[(size_t)t].assign((size_t)n * D, 0);
                int32_t lo, hi;
                range(t, &lo, &hi);
                Drawer dr(n, k_in, d_min, d_max, seed_lo, seed_hi);
                std::vector<int32_t> pre((size_t)k_in);
                std::vector<uint8_t> del((size_t)k_in);
                for (int32_t j = lo; j < hi; ++j) {
                    dr.draw(j, pre.data(), del.data());
                    for (int32_t m = 0; m < k_in; ++m) cnt[(size_t)t][(size_t)pre[(size_t)m] * D + del[(size_t)m]]++;
                }
            });
        for (auto& x : th) x.join();
    }
    // exclusive prefix in (pre, delay, thread) order: bucket (p, d) holds the posts in
    // ascending order because threads own ascending post ranges
    int64_t pos = 0;
    for (int32_t p = 0; p < n; ++p) {
        row_ptr[p] = pos;
        for (int32_t d = 0; d < D; ++d)
            for (int t = 0; t < T; ++t) {
                const int64_t c = cnt[(size_t)t][(size_t)p * D + d];
                cnt[(size_t)t][(size_t)p * D + d] = pos;
                pos += c;
            }
    }
    row_ptr[n] = pos;
    {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                int32_t lo, hi;
                range(t, &lo, &hi);
                Drawer dr(n, k_in, d_min, d_max, seed_lo, seed_hi);
                std::vector<int32_t> pre((size_t)k_in);
                std::vector<uint8_t> del((size_t)k_in);
                for (int32_t j = lo; j < hi; ++j) {
                    dr.draw(j, pre.data(), del.data());
                    for (int32_t m = 0; m < k_in; ++m) {
                        const int32_t p = pre[(size_t)m];
                        const int64_t at = cnt[(size_t)t][(size_t)p * D + del[(size_t)m]]++;
                        post[at] = j;
                        delay[at] = del[(size_t)m];
                        type[at] = (uint8_t)(p >= n_exc ? 1 : 0);
                    }
                }
            });
        for (auto& x : th) x.join();
    }
    return NBX_OK;
}

extern "C" int32_t nbx_build_dense_normal_bf16(int32_t n, float sigma, uint32_t seed_lo, uint32_t seed_hi,
                                               int32_t nthreads, uint16_t* out) {
    if (n <= 0 || !out) {
        nbx::set_error("nbx_build_dense_normal_bf16: bad arguments");
        return NBX_ERR_INVALID;
    }
    int T = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    T = std::max(1, std::min(T, std::min(64, n)));
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([=] {
            const int32_t lo = (int32_t)((int64_t)n * t / T), hi = (int32_t)((int64_t)n * (t + 1) / T);
            for (int32_t p = lo; p < hi; ++p)
                for (int32_t q = 0; q < n; ++q) {
                    uint32_t o[4];
                    philox((uint32_t)p, (uint32_t)q, 0x57474854u /* "WGHT" */, 0u, seed_lo, seed_hi, o);
                    const double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
                    const double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
                    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    const float f = (float)(z * (double)sigma);
                    uint32_t bits;
                    std::memcpy(&bits, &f, 4);
                    bits += 0x7FFFu + ((bits >> 16) & 1u);  // round to nearest even
                    out[(size_t)p * n + q] = (uint16_t)(bits >> 16);
                }
        });
    for (auto& x : th) x.join();
    return NBX_OK;
}
