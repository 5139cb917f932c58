// Differentiable recurrent LIF: forward step and surrogate-gradient adjoint step.
//
// Per step (all neurons i, trials b; FP32 state):
//   Vm = g*V + (1 - g)*(V_rest + R*I)      RK4 on dV/dt = (V_rest - V + R*I)/tau with I held over
//                                          the step is exactly this affine map, g = 1 - h + h^2/2
//                                          - h^3/6 + h^4/24, h = dt/tau (g computed on the host)
//   s  = Vm > V_th                         (strict, as the reference's LIF spike test)
//   V  = Vm - (V_th - V_reset)*s           subtractive reset
// Adjoint of one step, given the carried dL/dV of this step's output V (dV), the readout
// term R_t = (1-beta) W_out^T dy_t and the recurrent term G_t = W_rec^T dI_{t+1}:
//   ds  = R_t + G_t - (V_th - V_reset)*dV
//   dVm = dV + sigma'(Vm - V_th)*ds,      sigma'(x) = 1/(1 + slope*|x|)^2 (fast sigmoid)
//   dI  = (1 - g)*R*dVm,   dV_prev = g*dVm
// mode 0 uses sigma' = 0 (the exact gradient away from threshold crossings); mode 2 is a
// smooth relaxation whose forward spike is the fast sigmoid itself,
//   s = 1/2 + (1/2) u/(1 + |u|),  u = slope*(Vm - V_th),  with its exact derivative
//   sigma'(x) = (slope/2)/(1 + |u|)^2,
// so finite differences can check the whole recurrence.
// The step kernels take the recurrent term from the caller (a library GEMM between steps); the fused
// segment entry points compute it on the matrix cores inside the step kernel instead.
#include <hip/hip_runtime.h>

#include "nbx_internal.h"

using namespace nbx;

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void nbx_lif_fwd_kernel(nbx_lif_adj_params p, const float* __restrict__ I,
                                                               float* __restrict__ V, float* __restrict__ Vm_out,
                                                               float* __restrict__ s_out, int64_t count) {
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= count) return;
    const float vm = p.g * V[j] + p.one_minus_g * (p.v_rest + p.r * I[j]);
    float s;
    if (p.mode == 2) {
        const float u = p.slope * (vm - p.v_th);
        s = 0.5f + 0.5f * u / (1.0f + fabsf(u));
    } else {
        s = vm > p.v_th ? 1.0f : 0.0f;
    }
    V[j] = vm - p.delta_reset * s;
    if (Vm_out) Vm_out[j] = vm;
    s_out[j] = s;
}

__global__ __launch_bounds__(kThreads) void nbx_lif_bwd_kernel(nbx_lif_adj_params p, const float* __restrict__ G,
                                                               const float* __restrict__ R,
                                                               const float* __restrict__ Vm, float* __restrict__ dV,
                                                               float* __restrict__ dI_out, int64_t count) {
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= count) return;
    const float dv = dV[j];
    float ds = -p.delta_reset * dv;
    if (R) ds += R[j];
    if (G) ds += G[j];
    const float den = 1.0f + p.slope * fabsf(Vm[j] - p.v_th);
    float sg = 0.0f;
    if (p.mode == 1) sg = 1.0f / (den * den);
    else if (p.mode == 2) sg = 0.5f * p.slope / (den * den);
    const float dvm = dv + sg * ds;
    dI_out[j] = p.one_minus_g * p.r * dvm;
    dV[j] = p.g * dvm;
}

// y_t = beta*y_{t-1} + (1-beta)*Z_t over t for m independent channels (Z, y: [T][m]).
__global__ __launch_bounds__(kThreads) void nbx_readout_fwd_kernel(int32_t T, int64_t m, float beta,
                                                                   const float* __restrict__ Z,
                                                                   float* __restrict__ y, float* __restrict__ y_state) {
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    float acc = y_state ? y_state[j] : 0.0f;
    const float a = 1.0f - beta;
    for (int32_t t = 0; t < T; ++t) {
        acc = beta * acc + a * Z[(int64_t)t * m + j];
        y[(int64_t)t * m + j] = acc;
    }
    if (y_state) y_state[j] = acc;
}

// dZ_t = (1-beta)*dy~_t with dy~_t = g_t + beta*dy~_{t+1} (reverse scan; g = dL/dy).
__global__ __launch_bounds__(kThreads) void nbx_readout_bwd_kernel(int32_t T, int64_t m, float beta,
                                                                   const float* __restrict__ gy,
                                                                   float* __restrict__ dZ, float* __restrict__ carry) {
    const int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    float acc = carry ? carry[j] : 0.0f;
    const float a = 1.0f - beta;
    for (int32_t t = T - 1; t >= 0; --t) {
        acc = gy[(int64_t)t * m + j] + beta * acc;
        dZ[(int64_t)t * m + j] = a * acc;
    }
    if (carry) carry[j] = acc;
}

// ---------------------------------------------------------------- fused recurrent steps
//
// One step of the recurrence with its product on the matrix cores: out[i][b] = sum_j M[i][j] X[j][b]
// (M = W_rec forward, W_rec^T backward; X = the previous step's spikes / the next step's dI, both
// [n][batch]) in FP32 MFMA 16x16x4 tiles, then the step's elementwise update on the tile.
// Workgroup = 16 rows x 32 trials (two MFMA tiles); the reduction runs in chunks of kMmK: each
// chunk's M block (16 x kMmK) and X block (kMmK x 32) are loaded with coalesced 16-byte loads into
// registers while the previous chunk is on the matrix cores, then stored to LDS, from which the
// eight waves each reduce their own kMmK / 8 slice of the chunk. Every M row is read once per
// trial tile and every X row once per row block, with one load round trip per chunk in flight.
// The eight partial tiles are summed in a fixed order through LDS (so recomputation is
// bit-identical). MFMA step s of wave w reduces k = k0 + w * kMmKw + 16 kb + s over the four lane
// groups kb: lane (r, kb) reads M[i0 + r][that k] (16 consecutive floats per chunk: four LDS
// 16-byte reads) and lane (c, kb) reads X[that k][b0 + 16 ct + c]. Row blocks are placed on one XCD
// with their two trial tiles so each M row is read into one L2.
constexpr int kRecWaves = 8;
constexpr int kRecThreads = 64 * kRecWaves;
constexpr int kMmRows = 16;                  // output rows per workgroup
constexpr int kMmCols = 32;                  // output trials per workgroup
constexpr int kMmK = 512;                    // reduction chunk
constexpr int kMmKw = kMmK / kRecWaves;      // per wave and chunk: 16 MFMA steps x 4
constexpr int kMmAS = kMmK + 4;              // LDS row stride of the M block (16-byte aligned rows)
constexpr int kMmXS = kMmCols + 1;           // LDS row stride of the X block (odd: lane groups kb 16 rows
                                             // apart fall on different banks)
constexpr int kMmA4 = kMmRows * kMmK / 4 / kRecThreads;  // 16-byte M loads per thread and chunk
constexpr int kMmX4 = kMmK * kMmCols / 4 / kRecThreads;  // 16-byte X loads per thread and chunk
constexpr int kMmLds = (kMmRows * kMmAS + kMmK * kMmXS) * (int)sizeof(float);
static_assert(kMmRows * kMmCols == kRecThreads, "one output per thread in the epilogue");
static_assert(kRecWaves * kMmRows * kMmCols * (int)sizeof(float) <= kMmLds, "the partial tiles reuse the chunk LDS");

struct RecArgs {
    nbx_lif_adj_params p;
    const float* M;   // [n][n]
    const float* X;   // [n][batch] or null (no recurrent term)
    const float* add; // forward: IN_t [n][batch]; backward: R_t or null
    const float* Vm;  // backward: Vm_t
    float* V;         // forward: V (in/out); backward: dV (in/out)
    float* out;       // forward: s_t; backward: dI_t
    float* Vm_out;    // forward: Vm_t or null
    int32_t tiles_c;  // trial tiles: ceil(batch / kMmCols)
};

typedef float f4 __attribute__((ext_vector_type(4)));

// Operand loads are raw buffer loads: an offset at or past the resource's byte count (kMmOff for
// an element outside the matrix) reads zeros without a memory request, so the loads need no
// branches and all of a chunk's loads stay in flight together.
typedef uint32_t mm_u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kMmOff = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t mm_rsrc(const float* base, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0,
                                             (int32_t)(bytes < 0x7FFFFFFF ? bytes : 0x7FFFFFFF), 0x00020000);
}
// Four consecutive floats at element e of the resource: one 16-byte load when the rows are 16-byte
// aligned (AL), else four 4-byte loads; an element at or past `lim` (the end of its row) reads 0.
template <bool AL>
__device__ inline float4 mm_load4(__amdgpu_buffer_rsrc_t r, uint32_t e, int32_t first, int32_t lim) {
    if (AL) {
        const mm_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, first + 3 < lim ? 4u * e : kMmOff, 0, 0);
        return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
    }
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
        v[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, first + q < lim ? 4u * (e + q) : kMmOff, 0, 0));
    return make_float4(v[0], v[1], v[2], v[3]);
}

template <bool BWD, bool AL>
__global__ __launch_bounds__(kRecThreads) void nbx_lif_rec_kernel(RecArgs a) {
    extern __shared__ float s_mm[];
    float* sA = s_mm;                      // [kMmRows][kMmAS]
    float* sX = s_mm + kMmRows * kMmAS;    // [kMmK][kMmXS]
    const nbx_lif_adj_params& p = a.p;
    const int32_t n = p.n, B = p.batch;
    const int32_t G = gridDim.x;
    const int32_t lin = (G % 8 == 0) ? (int32_t)(blockIdx.x % 8) * (G / 8) + (int32_t)blockIdx.x / 8
                                     : (int32_t)blockIdx.x;
    const int32_t i0 = (lin / a.tiles_c) * kMmRows, b0 = (lin % a.tiles_c) * kMmCols;
    const int32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int32_t r = lane & 15, kb = lane >> 4;
    f4 acc[2] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
    if (a.X) {
        // this workgroup's rows of M, and (per chunk) the chunk's rows of X: rows past n read zeros
        const __amdgpu_buffer_rsrc_t ma = mm_rsrc(a.M + (int64_t)i0 * n, (int64_t)min(kMmRows, n - i0) * n * 4);
        float4 ra[kMmA4], rx[kMmX4];
        auto fetch = [&](int32_t k0) {
#pragma unroll
            for (int h = 0; h < kMmA4; ++h) {
                const int32_t idx = tid + h * kRecThreads;
                const int32_t row = idx / (kMmK / 4), k = k0 + 4 * (idx % (kMmK / 4));
                ra[h] = mm_load4<AL>(ma, (uint32_t)(row * n + k), k, n);
            }
            const __amdgpu_buffer_rsrc_t xa = mm_rsrc(a.X + (int64_t)k0 * B, (int64_t)min(kMmK, n - k0) * B * 4);
#pragma unroll
            for (int h = 0; h < kMmX4; ++h) {
                const int32_t idx = tid + h * kRecThreads;
                const int32_t kr = idx / (kMmCols / 4), c = b0 + 4 * (idx % (kMmCols / 4));
                rx[h] = mm_load4<AL>(xa, (uint32_t)(kr * B + c), c, B);
            }
        };
        auto stage = [&]() {
#pragma unroll
            for (int h = 0; h < kMmA4; ++h) {
                const int32_t idx = tid + h * kRecThreads;
                *reinterpret_cast<float4*>(sA + (idx / (kMmK / 4)) * kMmAS + 4 * (idx % (kMmK / 4))) = ra[h];
            }
#pragma unroll
            for (int h = 0; h < kMmX4; ++h) {
                const int32_t idx = tid + h * kRecThreads;
                float* q = sX + (idx / (kMmCols / 4)) * kMmXS + 4 * (idx % (kMmCols / 4));
                q[0] = rx[h].x;
                q[1] = rx[h].y;
                q[2] = rx[h].z;
                q[3] = rx[h].w;
            }
        };
        fetch(0);
        for (int32_t k0 = 0; k0 < n; k0 += kMmK) {
            if (k0 > 0) __syncthreads();  // the previous chunk's LDS reads are done
            stage();
            __syncthreads();
            if (k0 + kMmK < n) fetch(k0 + kMmK);  // in flight while this chunk is reduced
            const int32_t kw = wave * kMmKw + 16 * kb;
            float am[16];
            {
                const float4* q = reinterpret_cast<const float4*>(sA + r * kMmAS + kw);
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const float4 v = q[h];
                    am[4 * h] = v.x;
                    am[4 * h + 1] = v.y;
                    am[4 * h + 2] = v.z;
                    am[4 * h + 3] = v.w;
                }
            }
            float xv[16][2];  // every LDS read of the slice is issued before the first MFMA
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const float* xr = sX + (kw + s) * kMmXS + r;
                xv[s][0] = xr[0];
                xv[s][1] = xr[16];
            }
            __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the MFMAs
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[s], xv[s][0], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[s], xv[s][1], acc[1], 0, 0, 0);
            }
        }
        __syncthreads();  // the chunk LDS is reused for the partial tiles
    }
    // lane holds rows 4 kb .. 4 kb + 3 of columns r and 16 + r of this wave's partial tile
    float* red = s_mm;  // [kRecWaves][kMmRows][kMmCols]
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int e = 0; e < 4; ++e) red[(wave * kMmRows + 4 * kb + e) * kMmCols + 16 * ct + r] = acc[ct][e];
    __syncthreads();
    const int32_t row = tid / kMmCols, col = tid % kMmCols;
    const int32_t i = i0 + row, b = b0 + col;
    if (i >= n || b >= B) return;
    float prod = red[tid];
#pragma unroll
    for (int w = 1; w < kRecWaves; ++w) prod += red[w * kMmRows * kMmCols + tid];  // fixed order
    const int64_t j = (int64_t)i * B + b;
    if (!BWD) {  // the forward step of nbx_lif_fwd_kernel on I = IN_t + W_rec s_{t-1}
        const float I = a.add[j] + prod;
        const float vm = p.g * a.V[j] + p.one_minus_g * (p.v_rest + p.r * I);
        float sp;
        if (p.mode == 2) {
            const float u = p.slope * (vm - p.v_th);
            sp = 0.5f + 0.5f * u / (1.0f + fabsf(u));
        } else {
            sp = vm > p.v_th ? 1.0f : 0.0f;
        }
        a.V[j] = vm - p.delta_reset * sp;
        if (a.Vm_out) a.Vm_out[j] = vm;
        a.out[j] = sp;
    } else {  // the adjoint step of nbx_lif_bwd_kernel with G = W_rec^T dI_{t+1}
        const float dv = a.V[j];
        float ds = -p.delta_reset * dv;
        if (a.add) ds += a.add[j];
        if (a.X) ds += prod;
        const float den = 1.0f + p.slope * fabsf(a.Vm[j] - p.v_th);
        float sg = 0.0f;
        if (p.mode == 1) sg = 1.0f / (den * den);
        else if (p.mode == 2) sg = 0.5f * p.slope / (den * den);
        const float dvm = dv + sg * ds;
        a.out[j] = p.one_minus_g * p.r * dvm;
        a.V[j] = p.g * dvm;
    }
}

hipError_t launch_rec(bool bwd, RecArgs a, hipStream_t st) {
    static bool lds_set = false;
    if (!lds_set) {  // more than the default 64 KB of dynamic LDS
        for (const void* f : {(const void*)nbx_lif_rec_kernel<true, true>, (const void*)nbx_lif_rec_kernel<false, true>,
                              (const void*)nbx_lif_rec_kernel<true, false>,
                              (const void*)nbx_lif_rec_kernel<false, false>}) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMmLds);
            if (e != hipSuccess) return e;
        }
        lds_set = true;
    }
    a.tiles_c = (a.p.batch + kMmCols - 1) / kMmCols;
    const int32_t G = (a.p.n + kMmRows - 1) / kMmRows * a.tiles_c;
    // 16-byte loads need 16-byte aligned rows of both operands
    const bool al = (a.p.n % 4) == 0 && (a.p.batch % 4) == 0 && (reinterpret_cast<uintptr_t>(a.M) & 15u) == 0 &&
                    (!a.X || (reinterpret_cast<uintptr_t>(a.X) & 15u) == 0);
    if (bwd && al) nbx_lif_rec_kernel<true, true><<<G, kRecThreads, kMmLds, st>>>(a);
    else if (bwd) nbx_lif_rec_kernel<true, false><<<G, kRecThreads, kMmLds, st>>>(a);
    else if (al) nbx_lif_rec_kernel<false, true><<<G, kRecThreads, kMmLds, st>>>(a);
    else nbx_lif_rec_kernel<false, false><<<G, kRecThreads, kMmLds, st>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------- sparse forward step (binary spikes)
//
// With Heaviside spikes (modes 0 and 1) W_rec s_{t-1} is a sum of the columns of the neurons that
// fired at t-1: a trial has ~0.5-2 % of its neurons firing per step, so the product reads those
// rows of W_rec^T (coalesced, 4 B per neuron each) instead of the whole matrix. Workgroup = 64
// neurons x 8 trials: wave w takes trial b, lists the trial's fired neurons j in ascending order
// from its spike words (one word per lane, popcount prefix over the wave), and lane l sums
// WT[j][i0 + l] over the list in that order (deterministic, so a recomputed segment is
// bit-identical). The step itself runs on a transposed mapping (8 consecutive threads = one
// neuron's 8 trials, contiguous in the [n][batch] arrays), and the step's spikes go out as words
// again, [batch][ceil(n/32)], for the next step.
constexpr int kSpTrials = 8;                     // trials per workgroup (one wave each)
constexpr int kSpNeurons = 64;                   // neurons per workgroup (one per lane)
constexpr int kSpThreads = 64 * kSpTrials;
constexpr int kSpWords = 64;                     // spike words listed per pass (one per lane)
constexpr int kSpListCap = kSpWords * 32;        // list entries per wave and pass
constexpr int kSpRows = 32;                      // rows of W_rec^T loaded together per lane

struct SpArgs {
    nbx_lif_adj_params p;
    const float* WT;         // [n][n]: WT[j * n + i] = W_rec[i][j]
    const uint32_t* bits;    // spikes of the previous step, [batch][nw]: bit i % 32 of word i / 32
    const float* IN;         // [n][batch]
    float* V;                // [n][batch], in/out
    float* S;                // [n][batch], out
    float* Vm;               // [n][batch], out or null
    uint32_t* bits_out;      // spikes of this step, [batch][nw]
    int32_t nw;
};

__global__ __launch_bounds__(kSpThreads) void nbx_lif_spk_kernel(SpArgs a) {
    __shared__ int32_t s_list[kSpTrials][kSpListCap];
    __shared__ float s_acc[kSpTrials][kSpNeurons];
    __shared__ uint8_t s_fire[kSpTrials][kSpNeurons];
    const nbx_lif_adj_params& p = a.p;
    const int32_t n = p.n, B = p.batch;
    const int32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int32_t ntg = (B + kSpTrials - 1) / kSpTrials;
    const int32_t nb = (int32_t)blockIdx.x / ntg, tg = (int32_t)blockIdx.x % ntg;
    const int32_t i0 = nb * kSpNeurons;
    const int32_t b = tg * kSpTrials + wave;
    const int32_t i = i0 + lane;
    int32_t* list = s_list[wave];
    // this thread's output of the step (transposed mapping below): its input current and potential
    // are loaded first, so their latency overlaps the list and the row loads
    const int32_t il = tid / kSpTrials, bl = tid % kSpTrials;
    const int32_t ii = i0 + il, bb = tg * kSpTrials + bl;
    const bool mine = ii < n && bb < B;
    const int64_t jo = (int64_t)ii * B + bb;
    const float in_j = mine ? a.IN[jo] : 0.0f, v_j = mine ? a.V[jo] : 0.0f;
    float acc = 0.0f;
    for (int32_t w0 = 0; w0 < a.nw; w0 += kSpWords) {
        const int32_t wi = w0 + lane;
        const uint32_t word = (b < B && wi < a.nw) ? a.bits[(int64_t)b * a.nw + wi] : 0u;
        const int32_t c = __popc(word);
        int32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t z = __shfl_up(incl, o, 64);
            if (lane >= o) incl += z;
        }
        const int32_t total = __shfl(incl, 63, 64);
        int32_t pos = incl - c;
        for (uint32_t m = word; m; m &= m - 1u) list[pos++] = wi * 32 + __builtin_ctz(m);
        __syncthreads();
        if (i < n && total > 0) {
            // kSpRows rows per pass, all loads in flight together (a typical step's ~13 rows per
            // trial take one round trip): entries past the list re-read its last row, and their
            // values are not added, so the sum keeps ascending order exactly
            const float* col = a.WT + i;
            for (int32_t k = 0; k < total; k += kSpRows) {
                float w[kSpRows];
#pragma unroll
                for (int u = 0; u < kSpRows; ++u) w[u] = col[(int64_t)list[min(k + u, total - 1)] * n];
#pragma unroll
                for (int u = 0; u < kSpRows; ++u) acc = (k + u < total) ? acc + w[u] : acc;
            }
        }
        __syncthreads();  // the list is read before the next pass overwrites it
    }
    s_acc[wave][lane] = acc;
    __syncthreads();
    bool fired = false;
    if (mine) {  // the forward step of nbx_lif_fwd_kernel on I = IN_t + W_rec s_{t-1}
        const float I = in_j + s_acc[bl][il];
        const float vm = p.g * v_j + p.one_minus_g * (p.v_rest + p.r * I);
        fired = vm > p.v_th;
        const float sp = fired ? 1.0f : 0.0f;
        a.V[jo] = vm - p.delta_reset * sp;
        if (a.Vm) a.Vm[jo] = vm;
        a.S[jo] = sp;
    }
    s_fire[bl][il] = fired ? 1 : 0;
    __syncthreads();
    if (tid < 2 * kSpTrials) {  // this block's two spike words of each of its trials
        const int32_t bl2 = tid >> 1, h = tid & 1;
        const int32_t bb2 = tg * kSpTrials + bl2, wi = (i0 >> 5) + h;
        if (bb2 < B && wi < a.nw) {
            uint32_t word = 0;
#pragma unroll
            for (int q = 0; q < 32; ++q) word |= (uint32_t)s_fire[bl2][h * 32 + q] << q;
            a.bits_out[(int64_t)bb2 * a.nw + wi] = word;
        }
    }
}

// Spike words [batch][nw] of a float spike matrix s [n][batch] (s > 0: fired); null s: none fired.
__global__ __launch_bounds__(kThreads) void nbx_spike_words_kernel(const float* __restrict__ s, int32_t n, int32_t B,
                                                                   int32_t nw, uint32_t* __restrict__ bits) {
    const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= (int64_t)B * nw) return;
    const int32_t wi = (int32_t)(k / B), b = (int32_t)(k % B);  // consecutive threads: consecutive trials
    uint32_t word = 0;
    if (s)
        for (int q = 0; q < 32; ++q) {
            const int32_t i = wi * 32 + q;
            if (i < n && s[(int64_t)i * B + b] > 0.0f) word |= 1u << q;
        }
    bits[(int64_t)b * nw + wi] = word;
}

int blocks_for(int64_t count) { return (int)((count + kThreads - 1) / kThreads); }

int32_t check(const nbx_lif_adj_params* p) {
    if (!p || p->n <= 0 || p->batch <= 0) {
        set_error("nbx_lif: bad parameters");
        return NBX_ERR_INVALID;
    }
    return NBX_OK;
}

}  // namespace

extern "C" {

int32_t nbx_lif_fwd_step(const nbx_lif_adj_params* p, const float* d_I, float* d_V, float* d_Vm, float* d_s,
                         void* stream) {
    int32_t rc = check(p);
    if (rc) return rc;
    if (!d_I || !d_V || !d_s) {
        set_error("nbx_lif_fwd_step: null buffer");
        return NBX_ERR_INVALID;
    }
    const int64_t count = (int64_t)p->n * p->batch;
    nbx_lif_fwd_kernel<<<blocks_for(count), kThreads, 0, (hipStream_t)stream>>>(*p, d_I, d_V, d_Vm, d_s, count);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

int32_t nbx_lif_bwd_step(const nbx_lif_adj_params* p, const float* d_G, const float* d_R, const float* d_Vm,
                         float* d_dV, float* d_dI, void* stream) {
    int32_t rc = check(p);
    if (rc) return rc;
    if (!d_Vm || !d_dV || !d_dI) {
        set_error("nbx_lif_bwd_step: null buffer");
        return NBX_ERR_INVALID;
    }
    const int64_t count = (int64_t)p->n * p->batch;
    nbx_lif_bwd_kernel<<<blocks_for(count), kThreads, 0, (hipStream_t)stream>>>(*p, d_G, d_R, d_Vm, d_dV, d_dI, count);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

int32_t nbx_lif_fwd_segment(const nbx_lif_adj_params* p, const float* d_W, const float* d_IN, const float* d_s_init,
                            float* d_V, float* d_S, float* d_Vm, int32_t nsteps, void* stream) {
    int32_t rc = check(p);
    if (rc) return rc;
    if (!d_W || !d_IN || !d_V || !d_S || nsteps < 0) {
        set_error("nbx_lif_fwd_segment: bad arguments");
        return NBX_ERR_INVALID;
    }
    const int64_t nb = (int64_t)p->n * p->batch;
    RecArgs a{};
    a.p = *p;
    a.M = d_W;
    for (int32_t k = 0; k < nsteps; ++k) {
        a.X = k == 0 ? d_s_init : d_S + (k - 1) * nb;
        a.add = d_IN + k * nb;
        a.V = d_V;
        a.out = d_S + k * nb;
        a.Vm_out = d_Vm ? d_Vm + k * nb : nullptr;
        NBX_HIP_TRY(launch_rec(false, a, (hipStream_t)stream));
    }
    return NBX_OK;
}

int32_t nbx_lif_fwd_segment_sparse(const nbx_lif_adj_params* p, const float* d_WT, const float* d_IN,
                                   const float* d_s_init, float* d_V, float* d_S, float* d_Vm, uint32_t* d_bits,
                                   int32_t nsteps, void* stream) {
    int32_t rc = check(p);
    if (rc) return rc;
    if (!d_WT || !d_IN || !d_V || !d_S || !d_bits || nsteps < 0) {
        set_error("nbx_lif_fwd_segment_sparse: bad arguments");
        return NBX_ERR_INVALID;
    }
    if (p->mode != 0 && p->mode != 1) {
        set_error("nbx_lif_fwd_segment_sparse: binary spikes only (modes 0 and 1); mode 2 uses nbx_lif_fwd_segment");
        return NBX_ERR_INVALID;
    }
    hipStream_t st = (hipStream_t)stream;
    const int32_t nw = (p->n + 31) / 32;
    const int64_t nb = (int64_t)p->n * p->batch, words = (int64_t)p->batch * nw;
    nbx_spike_words_kernel<<<blocks_for(words), kThreads, 0, st>>>(d_s_init, p->n, p->batch, nw, d_bits);
    NBX_HIP_TRY(hipGetLastError());
    SpArgs a{};
    a.p = *p;
    a.WT = d_WT;
    a.nw = nw;
    const int32_t grid = ((p->n + kSpNeurons - 1) / kSpNeurons) * ((p->batch + kSpTrials - 1) / kSpTrials);
    for (int32_t k = 0; k < nsteps; ++k) {
        a.bits = d_bits + (k & 1) * words;
        a.bits_out = d_bits + ((k + 1) & 1) * words;
        a.IN = d_IN + k * nb;
        a.V = d_V;
        a.S = d_S + k * nb;
        a.Vm = d_Vm ? d_Vm + k * nb : nullptr;
        nbx_lif_spk_kernel<<<grid, kSpThreads, 0, st>>>(a);
        NBX_HIP_TRY(hipGetLastError());
    }
    return NBX_OK;
}

int32_t nbx_lif_bwd_segment(const nbx_lif_adj_params* p, const float* d_WT, const float* d_dI_after, const float* d_R,
                            const float* d_Vm, float* d_dV, float* d_dI, int32_t nsteps, void* stream) {
    int32_t rc = check(p);
    if (rc) return rc;
    if (!d_WT || !d_Vm || !d_dV || !d_dI || nsteps < 0) {
        set_error("nbx_lif_bwd_segment: bad arguments");
        return NBX_ERR_INVALID;
    }
    const int64_t nb = (int64_t)p->n * p->batch;
    RecArgs a{};
    a.p = *p;
    a.M = d_WT;
    for (int32_t k = nsteps - 1; k >= 0; --k) {
        a.X = k == nsteps - 1 ? d_dI_after : d_dI + (k + 1) * nb;
        a.add = d_R ? d_R + k * nb : nullptr;
        a.Vm = d_Vm + k * nb;
        a.V = d_dV;
        a.out = d_dI + k * nb;
        NBX_HIP_TRY(launch_rec(true, a, (hipStream_t)stream));
    }
    return NBX_OK;
}

int32_t nbx_readout_fwd(int32_t T, int64_t m, float beta, const float* d_Z, float* d_y, float* d_state, void* stream) {
    if (T <= 0 || m <= 0 || !d_Z || !d_y) {
        set_error("nbx_readout_fwd: bad arguments");
        return NBX_ERR_INVALID;
    }
    nbx_readout_fwd_kernel<<<blocks_for(m), kThreads, 0, (hipStream_t)stream>>>(T, m, beta, d_Z, d_y, d_state);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

int32_t nbx_readout_bwd(int32_t T, int64_t m, float beta, const float* d_gy, float* d_dZ, float* d_carry,
                        void* stream) {
    if (T <= 0 || m <= 0 || !d_gy || !d_dZ) {
        set_error("nbx_readout_bwd: bad arguments");
        return NBX_ERR_INVALID;
    }
    nbx_readout_bwd_kernel<<<blocks_for(m), kThreads, 0, (hipStream_t)stream>>>(T, m, beta, d_gy, d_dZ, d_carry);
    NBX_HIP_TRY(hipGetLastError());
    return NBX_OK;
}

}  // extern "C"
