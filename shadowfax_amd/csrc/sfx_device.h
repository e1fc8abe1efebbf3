// sfx_device.h -- device-side building blocks shared by the HIP kernels.
//
// Layout conventions (all kernels):
//   states are SoA, component-major: X[k*N + i], k in [0, ndim+2)
//   W = [rho, v_1..v_ndim, p], Q = [m, p_1..p_ndim, e]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfx {

// Solver constants derived from gamma exactly as the reference derives them
// (src/riemann/ExactRiemannSolver.cpp:52-61), computed on the host.
struct Consts {
    double gamma;
    double gp1d2g, gm1d2g, gm1dgp1, tdgp1, tdgm1, gm1d2, tgdgm1, ginv;
    double tol, cutoff;
    int itercap;   // bound on Newton and on Brent iterations (reference: none)
    int fast53;    // gamma == 5/3: exponents 0.2, 0.6, 0.8, 3, 5 (mixed path)
};

// status bits, one byte per problem / face / cell
enum : uint8_t {
    ST_SOLVER = 1,   // NaN or negative solution (RiemannSolverException)
    ST_NANFLUX = 2,  // NaN mass flux (FluxCalculationException)
    ST_ITERCAP = 4,  // iteration cap reached
    ST_UPDATEQ = 8,  // negative mass or energy after update (UpdateQException)
};

// counters[] slots accumulated by the batch kernels
enum { CNT_NEVAL = 0, CNT_NBRENT, CNT_NVACUUM, CNT_NITER, CNT_NFAIL, CNT_N };

// std::max / std::min as the reference's StateVector uses them
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sminf(float a, float b) { return (b < a) ? b : a; }

// Sum of a per-lane counter over the wavefront, added once per wave.
__device__ __forceinline__ void wave_add(unsigned long long* ctr, unsigned v) {
    if(!ctr) return;
#pragma unroll
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, (unsigned long long)v);
}

}  // namespace sfx
