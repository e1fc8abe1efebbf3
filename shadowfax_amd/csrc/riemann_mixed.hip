// riemann_mixed.hip -- batched mixed-precision solve_for_flux kernels; the
// algorithm and its precision split are described in riemann_mixed_impl.h.
#include "riemann_mixed_impl.h"
#include "sfx_launch.h"

#include <algorithm>

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
            st = hllc_mixed<ND>(c, cd, wl64, wr64, n64, v64, r, f, vac);
        }
#pragma unroll
        for(int k = 0; k < NC; k++) F[k * N + i] = (T)f[k];
        if(status) status[i] = st;
        fail = st != 0;
    }
    const unsigned cv[CNT_N] = {valid, 0u, vac, iters, fail};
    block_add_counters(ctr, cv);
}

// ---- exact solver in two phases (wave-level compaction) -----------------
//
// Phase 1 runs every problem's setup and at most `stop` FP32 Newton
// iterations; problems still iterating after that (strong shocks, near
// vacuum) are written, with their iteration state, to their block's segment
// of a queue (slots from wave ballots and one LDS counter per block -- no
// global atomics), and the wave finishes the converged ones together instead
// of idling while a few lanes iterate on.  Phase 2 gives each block the
// stragglers of kSegs consecutive phase-1 segments, packed densely through a
// prefix sum of their counts, resumes each problem where it stopped and
// finishes it.  The iteration sequence per problem is unchanged, so results
// are bit-identical to the one-kernel path.

struct Straggler {
    int32_t i;
    float p, lo, hi;
    int32_t it;
};

constexpr int kSegs = 16;  // phase-1 segments per phase-2 block

template <int ND, typename T>
__device__ __forceinline__ void load_problem(int64_t N, int64_t i, const T* __restrict__ WL,
                                             const T* __restrict__ WR, const T* __restrict__ n,
                                             const T* __restrict__ v, double* wl64, double* wr64,
                                             double* n64, double* v64, float* wl, float* wr,
                                             float* nn) {
    const int NC = ND + 2;
#pragma unroll
    for(int k = 0; k < NC; k++) {
        wl64[k] = (double)WL[k * N + i];
        wr64[k] = (double)WR[k * N + i];
        wl[k] = (float)wl64[k];
        wr[k] = (float)wr64[k];
    }
#pragma unroll
    for(int d = 0; d < ND; d++) {
        n64[d] = (double)n[d * N + i];
        v64[d] = v ? (double)v[d * N + i] : 0.0;
        nn[d] = (float)n64[d];
    }
}

template <int ND, typename T>
__global__ __launch_bounds__(256) void exact_phase1_kernel(
        CF c, Consts cd, int64_t N, const T* __restrict__ WL, const T* __restrict__ WR,
        const T* __restrict__ n, const T* __restrict__ v, const uint8_t* __restrict__ refl,
        T* __restrict__ F, uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr,
        Straggler* __restrict__ queue, unsigned* __restrict__ counts, int stop) {
    const int NC = ND + 2;
    __shared__ unsigned s_cnt;
    if(threadIdx.x == 0) s_cnt = 0u;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < N;
    unsigned iters = 0, vac = 0, fail = 0, done = 0;
    double wl64[NC], wr64[NC], n64[ND], v64[ND];
    float wl[NC], wr[NC], nn[ND];
    ExactCase s;
    s.phase = EX_DONE;
    if(valid) {
        load_problem<ND, T>(N, i, WL, WR, n, v, wl64, wr64, n64, v64, wl, wr, nn);
        exact_setup<ND>(c, cd, wl, wr, nn, refl ? refl[i] != 0 : false, wl64, wr64, n64, s);
        if(s.phase == EX_ITER) exact_iterate(c, s, stop);
    }
    const bool strag = valid && s.phase == EX_ITER;
    // slots in the block's segment: a ballot per wave, one LDS atomic per wave
    const unsigned long long mask = __ballot(strag);
    const int lane = __lane_id();
    unsigned base = 0;
    if(lane == 0 && mask) base = atomicAdd(&s_cnt, (unsigned)__popcll(mask));
    base = __shfl(base, 0);
    if(strag) {
        const unsigned slot = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
        Straggler g;
        g.i = (int32_t)i;
        g.p = s.p;
        g.lo = s.lo;
        g.hi = s.hi;
        g.it = s.it;
        queue[(int64_t)blockIdx.x * blockDim.x + slot] = g;
    } else if(valid) {
        double f[NC];
        const uint8_t st = exact_finish<ND>(c, cd, s, wl64, wr64, n64, v64, f, iters, vac);
#pragma unroll
        for(int k = 0; k < NC; k++) F[k * N + i] = (T)f[k];
        if(status) status[i] = st;
        fail = st != 0;
        done = 1;
    }
    __syncthreads();
    if(threadIdx.x == 0) counts[blockIdx.x] = s_cnt;
    const unsigned cv[CNT_N] = {done, 0u, vac, iters, fail};
    block_add_counters(ctr, cv);
}

template <int ND, typename T>
__global__ __launch_bounds__(256) void exact_phase2_kernel(
        CF c, Consts cd, int64_t N, const T* __restrict__ WL, const T* __restrict__ WR,
        const T* __restrict__ n, const T* __restrict__ v, const uint8_t* __restrict__ refl,
        T* __restrict__ F, uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr,
        const Straggler* __restrict__ queue, const unsigned* __restrict__ counts,
        int64_t nsegs) {
    const int NC = ND + 2;
    __shared__ unsigned s_pre[kSegs + 1];
    const int64_t seg0 = (int64_t)blockIdx.x * kSegs;
    if(threadIdx.x == 0) {
        unsigned acc = 0;
        for(int q = 0; q < kSegs; q++) {
            s_pre[q] = acc;
            acc += (seg0 + q < nsegs) ? counts[seg0 + q] : 0u;
        }
        s_pre[kSegs] = acc;
    }
    __syncthreads();
    const unsigned total = s_pre[kSegs];
    unsigned iters = 0, vac = 0, fail = 0, done = 0;
    for(unsigned k = threadIdx.x; k < total; k += blockDim.x) {
        int q = 0;
        while(q + 1 < kSegs && s_pre[q + 1] <= k) q++;
        const Straggler g = queue[(seg0 + q) * blockDim.x + (k - s_pre[q])];
        const int64_t i = g.i;
        double wl64[NC], wr64[NC], n64[ND], v64[ND], f[NC];
        float wl[NC], wr[NC], nn[ND];
        load_problem<ND, T>(N, i, WL, WR, n, v, wl64, wr64, n64, v64, wl, wr, nn);
        ExactCase s;
        exact_setup<ND>(c, cd, wl, wr, nn, refl ? refl[i] != 0 : false, wl64, wr64, n64, s);
        s.p = g.p;
        s.lo = g.lo;
        s.hi = g.hi;
        s.it = g.it;
        s.phase = EX_ITER;
        exact_iterate(c, s, c.itmax);
        const uint8_t st = exact_finish<ND>(c, cd, s, wl64, wr64, n64, v64, f, iters, vac);
#pragma unroll
        for(int r = 0; r < NC; r++) F[r * N + i] = (T)f[r];
        if(status) status[i] = st;
        fail += st != 0;
        done++;
    }
    const unsigned cv[CNT_N] = {done, 0u, vac, iters, fail};
    block_add_counters(ctr, cv);
}

}  // namespace mixed

size_t exact_split_scratch_bytes(int64_t N) {
    const int64_t nb = (std::max<int64_t>(N, 1) + 255) / 256;
    return (size_t)(((4 * nb + 255) / 256) * 256) + sizeof(mixed::Straggler) * (size_t)(nb * 256);
}

template <typename T>
hipError_t launch_exact_mixed_split(int ndim, const Consts& c, int64_t N, const T* WL,
                                    const T* WR, const T* n, const T* v, const uint8_t* refl,
                                    T* F, uint8_t* status, unsigned long long* ctr,
                                    void* scratch, int stop, int persistent_blocks,
                                    hipStream_t stream) {
    (void)persistent_blocks;
    if(N <= 0) return hipSuccess;
    const mixed::CF cf = mixed::make_cf(c);
    const int64_t nb = (N + 255) / 256;
    unsigned* counts = (unsigned*)scratch;
    mixed::Straggler* queue =
            (mixed::Straggler*)((char*)scratch + ((4 * nb + 255) / 256) * 256);
    const dim3 grid((unsigned)nb), block(256),
            grid2((unsigned)((nb + mixed::kSegs - 1) / mixed::kSegs));
#define SFX_SPLIT(ND)                                                                          \
    do {                                                                                       \
        hipLaunchKernelGGL((mixed::exact_phase1_kernel<ND, T>), grid, block, 0, stream, cf, c, N, \
                           WL, WR, n, v, refl, F, status, ctr, queue, counts, stop);            \
        hipLaunchKernelGGL((mixed::exact_phase2_kernel<ND, T>), grid2, block, 0, stream, cf, c,  \
                           N, WL, WR, n, v, refl, F, status, ctr, queue, counts, nb);           \
    } while(0)
    if(ndim == 3)
        SFX_SPLIT(3);
    else
        SFX_SPLIT(2);
#undef SFX_SPLIT
    return hipGetLastError();
}

template hipError_t launch_exact_mixed_split<float>(int, const Consts&, int64_t, const float*,
                                                    const float*, const float*, const float*,
                                                    const uint8_t*, float*, uint8_t*,
                                                    unsigned long long*, void*, int, int,
                                                    hipStream_t);
template hipError_t launch_exact_mixed_split<double>(int, const Consts&, int64_t, const double*,
                                                     const double*, const double*, const double*,
                                                     const uint8_t*, double*, uint8_t*,
                                                     unsigned long long*, void*, int, int,
                                                     hipStream_t);

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
