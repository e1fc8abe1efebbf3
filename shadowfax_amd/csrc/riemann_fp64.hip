// riemann_fp64.hip -- plain FP64 Riemann solvers for gfx950.
//
// Compiled with -ffp-contract=off: every expression keeps the association of
// the reference (Shadowfax src/riemann/ExactRiemannSolver.cpp,
// src/riemann/HLLCRiemannSolver.cpp), so results differ from the CPU
// reference only where the device pow() rounds differently from glibc.
//
// One lane per problem, SoA loads/stores (fully coalesced, 144 B per 3D
// problem with v = 0).  The exact solver's Newton/Brent loops run per lane
// with an iteration cap; divergence handling is left to the hardware here
// (this is the parity path; the mixed path is the fast one).
#include "riemann_fp64_impl.h"
#include "sfx_device.h"
#include "sfx_launch.h"

namespace sfx {
namespace fp64 {

// --------------------------------------------------------------- kernels --

template <int ND, int KIND>
__global__ __launch_bounds__(256) void solve_for_flux_kernel(
        Consts c, int64_t N, const double* __restrict__ WL, const double* __restrict__ WR,
        const double* __restrict__ n, const double* __restrict__ v,
        const uint8_t* __restrict__ refl, double* __restrict__ F,
        uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned iters = 0, brent = 0, vac = 0, fail = 0, valid = 0;
    if(i < N) {
        valid = 1;
        double wl[NC], wr[NC], nn[ND], vv[ND], f[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            wl[k] = WL[k * N + i];
            wr[k] = WR[k * N + i];
        }
#pragma unroll
        for(int d = 0; d < ND; d++) {
            nn[d] = n[d * N + i];
            vv[d] = v ? v[d * N + i] : 0.;
        }
        const bool r = refl ? refl[i] != 0 : false;
        uint8_t st;
        if(KIND == 0)
            st = ex_solve_for_flux<ND>(c, wl, wr, nn, vv, r, f, iters, brent, vac);
        else
            st = hllc_solve_for_flux<ND>(c, wl, wr, nn, vv, r, f, vac);
#pragma unroll
        for(int k = 0; k < NC; k++) F[k * N + i] = f[k];
        if(status) status[i] = st;
        fail = st != 0;
    }
    const unsigned cv[CNT_N] = {valid, brent, vac, iters, fail};
    block_add_counters(ctr, cv);
}

template <int ND>
__global__ __launch_bounds__(256) void exact_solve_kernel(
        Consts c, int64_t N, const double* __restrict__ WL, const double* __restrict__ WR,
        const double* __restrict__ n, const uint8_t* __restrict__ refl,
        double* __restrict__ Whalf, double* __restrict__ WRout, double* __restrict__ mach_io,
        uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned iters = 0, brent = 0, vac = 0, fail = 0, valid = 0;
    if(i < N) {
        valid = 1;
        double wl[NC], wr[NC], nn[ND], w[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            wl[k] = WL[k * N + i];
            wr[k] = WR[k * N + i];
        }
#pragma unroll
        for(int d = 0; d < ND; d++) nn[d] = n[d * N + i];
        double mach = mach_io ? mach_io[i] : 0.;
        uint8_t st = ex_solve<ND>(c, wl, wr, nn, refl ? refl[i] != 0 : false, w, &mach,
                                  iters, brent, vac);
#pragma unroll
        for(int k = 0; k < NC; k++) {
            Whalf[k * N + i] = w[k];
            if(WRout) WRout[k * N + i] = wr[k];
        }
        if(mach_io) mach_io[i] = mach;
        if(status) status[i] = st;
        fail = st != 0;
    }
    const unsigned cv[CNT_N] = {valid, brent, vac, iters, fail};
    block_add_counters(ctr, cv);
}

}  // namespace fp64

// ------------------------------------------------------------ launchers --

static inline unsigned nblocks(int64_t N, int bs) { return (unsigned)((N + bs - 1) / bs); }

hipError_t launch_solve_for_flux_fp64(int kind, int ndim, const Consts& c, int64_t N,
                                      const double* WL, const double* WR, const double* n,
                                      const double* v, const uint8_t* refl, double* F,
                                      uint8_t* status, unsigned long long* ctr,
                                      hipStream_t stream) {
    if(N <= 0) return hipSuccess;
    const int bs = 256;
    dim3 grid(nblocks(N, bs)), block(bs);
    if(ndim == 3 && kind == 0)
        hipLaunchKernelGGL((fp64::solve_for_flux_kernel<3, 0>), grid, block, 0, stream, c, N,
                           WL, WR, n, v, refl, F, status, ctr);
    else if(ndim == 3)
        hipLaunchKernelGGL((fp64::solve_for_flux_kernel<3, 1>), grid, block, 0, stream, c, N,
                           WL, WR, n, v, refl, F, status, ctr);
    else if(kind == 0)
        hipLaunchKernelGGL((fp64::solve_for_flux_kernel<2, 0>), grid, block, 0, stream, c, N,
                           WL, WR, n, v, refl, F, status, ctr);
    else
        hipLaunchKernelGGL((fp64::solve_for_flux_kernel<2, 1>), grid, block, 0, stream, c, N,
                           WL, WR, n, v, refl, F, status, ctr);
    return hipGetLastError();
}

hipError_t launch_exact_solve_fp64(int ndim, const Consts& c, int64_t N, const double* WL,
                                   const double* WR, const double* n, const uint8_t* refl,
                                   double* Whalf, double* WRout, double* mach, uint8_t* status,
                                   unsigned long long* ctr, hipStream_t stream) {
    if(N <= 0) return hipSuccess;
    const int bs = 256;
    dim3 grid(nblocks(N, bs)), block(bs);
    if(ndim == 3)
        hipLaunchKernelGGL((fp64::exact_solve_kernel<3>), grid, block, 0, stream, c, N, WL,
                           WR, n, refl, Whalf, WRout, mach, status, ctr);
    else
        hipLaunchKernelGGL((fp64::exact_solve_kernel<2>), grid, block, 0, stream, c, N, WL,
                           WR, n, refl, Whalf, WRout, mach, status, ctr);
    return hipGetLastError();
}

}  // namespace sfx
