// riemann_mixed.hip -- batched mixed-precision solve_for_flux kernels; the
// algorithm and its precision split are described in riemann_mixed_impl.h.
#include "riemann_mixed_impl.h"
#include "sfx_launch.h"

namespace sfx {
namespace mixed {

template <int ND, int KIND, typename T>
__global__ __launch_bounds__(256) void solve_for_flux_kernel(
        CF c, Consts cd, int64_t N, const T* __restrict__ WL, const T* __restrict__ WR,
        const T* __restrict__ n, const T* __restrict__ v, const uint8_t* __restrict__ refl,
        T* __restrict__ F, uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned iters = 0, vac = 0, fail = 0, valid = 0;
    if(i < N) {
        valid = 1;
        double wl64[NC], wr64[NC], n64[ND], v64[ND], f[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            wl64[k] = (double)WL[k * N + i];
            wr64[k] = (double)WR[k * N + i];
        }
#pragma unroll
        for(int d = 0; d < ND; d++) {
            n64[d] = (double)n[d * N + i];
            v64[d] = v ? (double)v[d * N + i] : 0.0;
        }
        const bool r = refl ? refl[i] != 0 : false;
        uint8_t st;
        if(KIND == 0) {
            float wl[NC], wr[NC], nn[ND];
#pragma unroll
            for(int k = 0; k < NC; k++) {
                wl[k] = (float)wl64[k];
                wr[k] = (float)wr64[k];
            }
#pragma unroll
            for(int d = 0; d < ND; d++) nn[d] = (float)n64[d];
            st = exact_mixed<ND>(c, cd, wl, wr, nn, r, wl64, wr64, n64, v64, f, iters, vac);
        } else {
            st = fp64::hllc_solve_for_flux<ND>(cd, wl64, wr64, n64, v64, r, f, vac);
        }
#pragma unroll
        for(int k = 0; k < NC; k++) F[k * N + i] = (T)f[k];
        if(status) status[i] = st;
        fail = st != 0;
    }
    const unsigned cv[CNT_N] = {valid, 0u, vac, iters, fail};
    block_add_counters(ctr, cv);
}

}  // namespace mixed

template <typename T>
hipError_t launch_solve_for_flux_mixed(int kind, int ndim, const Consts& c, int64_t N,
                                       const T* WL, const T* WR, const T* n, const T* v,
                                       const uint8_t* refl, T* F, uint8_t* status,
                                       unsigned long long* ctr, hipStream_t stream) {
    if(N <= 0) return hipSuccess;
    const mixed::CF cf = mixed::make_cf(c);
    const int bs = 256;
    dim3 grid((unsigned)((N + bs - 1) / bs)), block(bs);
#define SFX_LAUNCH(ND, K)                                                                 \
    hipLaunchKernelGGL((mixed::solve_for_flux_kernel<ND, K, T>), grid, block, 0, stream, cf, \
                       c, N, WL, WR, n, v, refl, F, status, ctr)
    if(ndim == 3 && kind == 0) SFX_LAUNCH(3, 0);
    else if(ndim == 3) SFX_LAUNCH(3, 1);
    else if(kind == 0) SFX_LAUNCH(2, 0);
    else SFX_LAUNCH(2, 1);
#undef SFX_LAUNCH
    return hipGetLastError();
}

template hipError_t launch_solve_for_flux_mixed<float>(int, int, const Consts&, int64_t,
                                                       const float*, const float*, const float*,
                                                       const float*, const uint8_t*, float*,
                                                       uint8_t*, unsigned long long*,
                                                       hipStream_t);
template hipError_t launch_solve_for_flux_mixed<double>(int, int, const Consts&, int64_t,
                                                        const double*, const double*,
                                                        const double*, const double*,
                                                        const uint8_t*, double*, uint8_t*,
                                                        unsigned long long*, hipStream_t);

}  // namespace sfx
