// cell_ops.hip -- per-cell RiemannSolver operations over batches (FP64,
// contraction off, the reference's association): get_time_derivative,
// get_W, get_Q, and TimeLine::calculate_timestep.  Solver batches are
// component-major SoA: W, Q [ndim+2][N], grad [ndim][ndim+2][N]; the time
// step reads the mesh arrays, cell-major.
#include "flux_pass_impl.h"
#include "sfx_launch.h"
#include "shadowfax_amd.h"

namespace sfx {
namespace {

template <int ND>
__global__ __launch_bounds__(256) void time_derivative_kernel(double gamma, int64_t N,
                                                              const double* __restrict__ W,
                                                              const double* __restrict__ grad,
                                                              double* __restrict__ out) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= N) return;
    double w[NC], g[ND][NC], d[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) w[k] = W[k * N + i];
#pragma unroll
    for(int l = 0; l < ND; l++)
#pragma unroll
        for(int k = 0; k < NC; k++) g[l][k] = grad[(l * NC + k) * N + i];
    flux::time_derivative<ND, double>(gamma, w, g, d);
#pragma unroll
    for(int k = 0; k < NC; k++) out[k * N + i] = d[k];
}

template <int ND>
__global__ __launch_bounds__(256) void get_W_kernel(double gamma, int64_t N,
                                                    const double* __restrict__ V,
                                                    double* __restrict__ Q,
                                                    double* __restrict__ W) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= N) return;
    double q[NC], w[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) q[k] = Q[k * N + i];
    flux::get_W_state<ND>(gamma, V[i], q, w);
#pragma unroll
    for(int k = 0; k < NC; k++) {
        Q[k * N + i] = q[k];
        W[k * N + i] = w[k];
    }
}

template <int ND, int KIND>
__global__ __launch_bounds__(256) void get_Q_kernel(double gamma, int64_t N,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ W,
                                                    double* __restrict__ Q) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= N) return;
    double q[NC], w[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) w[k] = W[k * N + i];
    flux::get_Q_state<ND, KIND>(gamma, V[i], w, q);
#pragma unroll
    for(int k = 0; k < NC; k++) Q[k * N + i] = q[k];
}

inline dim3 blocks(int64_t N) { return dim3((unsigned)((N + 255) / 256)); }

// TimeLine::get_realtime / get_realtime_interval (src/TimeLine.hpp:145-166),
// no cosmology
__device__ __forceinline__ double tl_realtime(const TimelineArgs& tl, uint64_t t) {
    double x = (double)t / (double)tl.integer_maxtime;
    return tl.tottime * x + tl.mintime;
}
__device__ __forceinline__ double tl_interval(const TimelineArgs& tl, uint64_t t) {
    double x = (double)t / (double)tl.integer_maxtime;
    return tl.tottime * x;
}

__global__ void timestep_init_kernel(uint64_t* result, uint64_t smax) {
    if(threadIdx.x == 0) {
        result[0] = smax;
        result[1] = 0;
    }
}

// One lane per cell: TimeLine::calculate_timestep's gas loop
// (src/TimeLine.cpp:248-320).  The smallest step and the active count go to
// result through one atomic per block.
template <int ND, int KIND>
__global__ __launch_bounds__(256) void timestep_kernel(
        TimelineArgs tl, double gamma, int64_t C, const double* __restrict__ W,
        const double* __restrict__ vel, const double* __restrict__ volume,
        uint64_t* __restrict__ starttime, uint64_t* __restrict__ endtime, double* __restrict__ dt,
        uint8_t* __restrict__ active, uint64_t* __restrict__ result, uint8_t* __restrict__ status) {
    constexpr int NC = ND + 2, P = ND + 1;
    __shared__ uint64_t s_min[256];
    __shared__ unsigned s_cnt;
    if(threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t Smax = tl.max_timestep;
    uint64_t mine = Smax;
    if(i < C) {
        uint8_t st = 0;
        const uint64_t end0 = endtime[i];
        const bool act = end0 == tl.current_time + tl.timestep;
        active[i] = act ? 1 : 0;
        if(act) {
            atomicAdd(&s_cnt, 1u);
            const double* w = W + i * NC;
            // v = W.v - particle velocity; Vec::norm (src/Vec.hpp:140,378)
            const double vx = w[1] - vel[i * ND], vy = w[2] - vel[i * ND + 1];
            double vi;
            if(ND == 3) {
                const double vz = w[3] - vel[i * ND + 2];
                vi = sqrt(vx * vx + vy * vy + vz * vz);
            } else {
                vi = sqrt(vx * vx + vy * vy);
            }
            // get_soundspeed: the exact solver guards rho = 0, HLLC does not
            const double ci = (KIND == 0 && !(w[0] != 0.)) ? 0. : sqrt(gamma * w[P] / w[0]);
            // VorCell::set_h (src/VorCell.cpp:930-938)
            const double V = volume[i];
            const double pi = 3.14159265358979323846;
            const double h = (ND == 3) ? cbrt(V * 3. / 4. / pi) : sqrt(V / pi);
            const double S = tl.cfl * (h / (vi + ci));
            uint64_t S2;
            if(tl_interval(tl, Smax) < S) {
                S2 = Smax;
            } else {
                S2 = tl.integer_maxtime;
                while(tl_interval(tl, S2) > S) S2 >>= 1;  // at most 64 halvings
                if(!S2) {
                    st = SFX_ST_TIMESTEP;
                } else {
                    // S2 = 1 divides everything: at most 64 halvings
                    while((tl.integer_maxtime - end0) % S2) S2 >>= 1;
                    S2 = (S2 < tl.min_timestep) ? tl.min_timestep : S2;
                    S2 = (S2 > tl.max_timestep) ? tl.max_timestep : S2;
                }
            }
            if(!st) {
                mine = S2;
                if(!tl.global_timestep) {
                    starttime[i] = end0;
                    endtime[i] = end0 + S2;
                }
            }
        }
        if(status) status[i] = st;
        if(!tl.global_timestep) dt[i] = tl_realtime(tl, endtime[i] - starttime[i]);
    }
    s_min[threadIdx.x] = mine;
    __syncthreads();
    for(int k = 128; k > 0; k >>= 1) {
        if((int)threadIdx.x < k && s_min[threadIdx.x + k] < s_min[threadIdx.x])
            s_min[threadIdx.x] = s_min[threadIdx.x + k];
        __syncthreads();
    }
    if(threadIdx.x == 0) {
        if(s_min[0] < Smax) atomicMin((unsigned long long*)result, (unsigned long long)s_min[0]);
        if(s_cnt) atomicAdd((unsigned long long*)(result + 1), (unsigned long long)s_cnt);
    }
}

// global time step (src/TimeLine.cpp:418-431): every cell takes the system step
__global__ __launch_bounds__(256) void timestep_global_kernel(
        TimelineArgs tl, int64_t C, const uint64_t* __restrict__ result,
        uint64_t* __restrict__ starttime, uint64_t* __restrict__ endtime, double* __restrict__ dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= C) return;
    const uint64_t S = result[0];
    const uint64_t e = endtime[i];
    starttime[i] = e;
    endtime[i] = e + S;
    dt[i] = tl_realtime(tl, S);
}

}  // namespace

hipError_t launch_hydro_timestep(int kind, int ndim, double gamma, int64_t C,
                                 const TimelineArgs& tl, const double* W, const double* vel,
                                 const double* volume, uint64_t* starttime, uint64_t* endtime,
                                 double* dt, uint8_t* active, uint64_t* result,
                                 uint8_t* status, hipStream_t s) {
    hipLaunchKernelGGL(timestep_init_kernel, dim3(1), dim3(64), 0, s, result, tl.max_timestep);
    if(C > 0) {
#define SFX_TS(ND, K)                                                                       \
    hipLaunchKernelGGL((timestep_kernel<ND, K>), blocks(C), dim3(256), 0, s, tl, gamma, C, W, \
                       vel, volume, starttime, endtime, dt, active, result, status)
        if(ndim == 3) {
            if(kind == 1)
                SFX_TS(3, 1);
            else
                SFX_TS(3, 0);
        } else {
            if(kind == 1)
                SFX_TS(2, 1);
            else
                SFX_TS(2, 0);
        }
#undef SFX_TS
        if(tl.global_timestep)
            hipLaunchKernelGGL(timestep_global_kernel, blocks(C), dim3(256), 0, s, tl, C, result,
                               starttime, endtime, dt);
    }
    return hipGetLastError();
}

hipError_t launch_cell_op(int op, int kind, int ndim, double gamma, int64_t N,
                          const double* a, const double* b, double* c, double* d,
                          hipStream_t s) {
    if(N <= 0) return hipSuccess;
    switch(op) {
        case 0:  // time derivative: a = W, b = grad, c = out
            if(ndim == 3)
                hipLaunchKernelGGL(time_derivative_kernel<3>, blocks(N), dim3(256), 0, s, gamma,
                                   N, a, b, c);
            else
                hipLaunchKernelGGL(time_derivative_kernel<2>, blocks(N), dim3(256), 0, s, gamma,
                                   N, a, b, c);
            break;
        case 1:  // get_W: a = V, d = Q (in/out), c = W
            if(ndim == 3)
                hipLaunchKernelGGL(get_W_kernel<3>, blocks(N), dim3(256), 0, s, gamma, N, a, d, c);
            else
                hipLaunchKernelGGL(get_W_kernel<2>, blocks(N), dim3(256), 0, s, gamma, N, a, d, c);
            break;
        default:  // get_Q: a = V, b = W, c = Q
            if(ndim == 3) {
                if(kind == 1)
                    hipLaunchKernelGGL((get_Q_kernel<3, 1>), blocks(N), dim3(256), 0, s, gamma,
                                       N, a, b, c);
                else
                    hipLaunchKernelGGL((get_Q_kernel<3, 0>), blocks(N), dim3(256), 0, s, gamma,
                                       N, a, b, c);
            } else {
                if(kind == 1)
                    hipLaunchKernelGGL((get_Q_kernel<2, 1>), blocks(N), dim3(256), 0, s, gamma,
                                       N, a, b, c);
                else
                    hipLaunchKernelGGL((get_Q_kernel<2, 0>), blocks(N), dim3(256), 0, s, gamma,
                                       N, a, b, c);
            }
    }
    return hipGetLastError();
}

}  // namespace sfx
