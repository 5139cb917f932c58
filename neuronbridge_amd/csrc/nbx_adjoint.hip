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
// The products with W_rec, W_in, W_out are GEMMs run by the caller between these steps.
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
