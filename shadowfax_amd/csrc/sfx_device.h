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

// counters accumulated by the batch kernels; the device array holds
// kCtrSlots copies (kCtrSlots x CNT_N) that blocks add into round-robin, so
// atomics from thousands of blocks do not queue on one address; readers sum
// the copies.
enum { CNT_NEVAL = 0, CNT_NBRENT, CNT_NVACUUM, CNT_NITER, CNT_NFAIL, CNT_N };
constexpr int kCtrSlots = 256;

// std::max / std::min as the reference's StateVector uses them
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sminf(float a, float b) { return (b < a) ? b : a; }

// Adds per-lane counters v[0..CNT_N) over the whole block: wave sums by
// shuffles, block sums in LDS, then one global atomic per counter into slot
// blockIdx.x % kCtrSlots.  Every thread of the block must call it (it
// synchronizes the block); ctr == nullptr is a uniform no-op.
__device__ __forceinline__ void block_add_counters(unsigned long long* ctr,
                                                   const unsigned (&v)[CNT_N]) {
    if(!ctr) return;
    __shared__ unsigned long long s_ctr[CNT_N];
    if(threadIdx.x < CNT_N) s_ctr[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for(int k = 0; k < CNT_N; k++) {
        unsigned x = v[k];
#pragma unroll
        for(int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if((threadIdx.x & 63) == 0 && x) atomicAdd(&s_ctr[k], (unsigned long long)x);
    }
    __syncthreads();
    if(threadIdx.x < CNT_N && s_ctr[threadIdx.x])
        atomicAdd(ctr + (blockIdx.x % kCtrSlots) * CNT_N + threadIdx.x, s_ctr[threadIdx.x]);
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin
// (dispatch index b runs on XCD b % 8) and every XCD has its own L2.  This
// maps the dispatch index to a logical block index such that each XCD works
// through one contiguous range of blocks: kernels whose neighbouring items
// share data (cells and the faces between them) then find it in their own
// L2.  A bijection on [0, gridDim.x) for any grid size.
__device__ __forceinline__ int64_t xcd_block() {
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t q = nb >> 3, r = nb & 7, x = b & 7, t = b >> 3;
    return x * q + (x < r ? x : r) + t;
}

}  // namespace sfx
