// cell_ops.hip -- per-cell RiemannSolver operations over batches (FP64,
// contraction off, the reference's association): get_time_derivative,
// get_W, get_Q.  Batches are component-major SoA like the solver batches:
// W, Q [ndim+2][N], grad [ndim][ndim+2][N].
#include "flux_pass_impl.h"
#include "sfx_launch.h"

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

}  // namespace

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
