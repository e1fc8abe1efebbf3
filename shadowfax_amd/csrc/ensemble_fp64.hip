// ensemble_fp64.hip -- FP64 instantiation of the ensemble kernels (the
// parity flux pass per item, compiled like flux_pass.hip with contraction
// off), the plan kernel and the assembly helpers.
#include "ensemble_launch.h"
#include "sfx_launch.h"

namespace sfx {
namespace ens {

// batch: [n, id_0 .. id_{n-1}] (n <= kMaxBatch).  One block of kMaxBatch
// threads.
__global__ __launch_bounds__(kMaxBatch) void plan_kernel(const int32_t* __restrict__ batch,
                                                         const int64_t* __restrict__ mcell0,
                                                         const int64_t* __restrict__ mncell,
                                                         const int64_t* __restrict__ mface0,
                                                         const int64_t* __restrict__ mnface,
                                                         int32_t tile, Plan* __restrict__ plan) {
    __shared__ int64_t s[4][kMaxBatch];
    const int t = threadIdx.x;
    const int n = batch[0];
    int64_t v[4] = {0, 0, 0, 0};
    int64_t c0 = 0, f0 = 0;
    if(t < n) {
        const int32_t id = batch[1 + t];
        c0 = mcell0[id];
        f0 = mface0[id];
        v[0] = mncell[id];
        v[1] = mnface[id];
        v[2] = (v[0] + tile - 1) / tile;
        v[3] = (v[1] + tile - 1) / tile;
    }
#pragma unroll
    for(int q = 0; q < 4; q++) s[q][t] = v[q];
    __syncthreads();
    // inclusive Hillis-Steele scans of the four columns
    for(int d = 1; d < kMaxBatch; d <<= 1) {
        int64_t add[4] = {0, 0, 0, 0};
        if(t >= d)
#pragma unroll
            for(int q = 0; q < 4; q++) add[q] = s[q][t - d];
        __syncthreads();
#pragma unroll
        for(int q = 0; q < 4; q++) s[q][t] += add[q];
        __syncthreads();
    }
    if(t < n) {
        Slot sl;
        sl.cell0 = c0;
        sl.ncell = v[0];
        sl.face0 = f0;
        sl.nface = v[1];
        sl.gshift = (s[0][t] - v[0]) - c0;
        sl.qshift = (s[1][t] - v[1]) - f0;
        sl.ctile0 = s[2][t] - v[2];
        sl.ftile0 = s[3][t] - v[3];
        plan->slot[t] = sl;
    }
    if(t == 0) {
        plan->n = n;
        plan->tile = tile;
        plan->ctiles = n > 0 ? s[2][n - 1] : 0;
        plan->ftiles = n > 0 ? s[3][n - 1] : 0;
    }
    if(t < kPhases * kXcd) plan->ctr[t / kXcd][t % kXcd] = 0u;
}

// ---- assembling an ensemble: copy a mesh with its indices offset ---------

// dst[i] = src[i] + (src[i] >= 0 ? off : 0)   (int32 indices; -1 kept)
__global__ __launch_bounds__(256) void offset_copy_i32(const int32_t* __restrict__ src,
                                                       int32_t* __restrict__ dst, int64_t n,
                                                       int64_t off) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n) {
        const int32_t v = src[k];
        dst[k] = v >= 0 ? (int32_t)(v + off) : v;
    }
}

__global__ __launch_bounds__(256) void offset_copy_i64(const int64_t* __restrict__ src,
                                                       int64_t* __restrict__ dst, int64_t n,
                                                       int64_t off) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(k < n) dst[k] = src[k] + off;
}

}  // namespace ens

template <>
hipError_t ens_launch_step<double>(int ndim, int kind, const EnsArgs<double>& a, hipStream_t s) {
    return ens_step_impl<double, 0>(ndim, kind, a, s);
}

template <>
hipError_t ens_launch_range<double>(int ndim, int kind, const EnsArgs<double>& a, int64_t cell0,
                                    int64_t ncell, int64_t face0, int64_t nface, int64_t gshift,
                                    int64_t qshift, hipStream_t s) {
    return ens_range_impl<double, 0>(ndim, kind, a, cell0, ncell, face0, nface, gshift, qshift, s);
}

template <>
hipError_t ens_launch_queue<double>(int ndim, int kind, const EnsArgs<double>& a, const QueueArgs& qa,
                                int ncu, hipStream_t s, int* blocks) {
    return ens_queue_impl<double, 0>(ndim, kind, a, qa, ncu, s, blocks);
}

template <>
hipError_t ens_launch_entries<double>(int ndim, const flux::MeshView<double>& m, int64_t E,
                                      double* egeo, hipStream_t s) {
    return launch_entry_geometry<double>(ndim, m, E, egeo, s);
}

hipError_t ens_launch_plan(const int32_t* batch, const int64_t* mcell0, const int64_t* mncell,
                           const int64_t* mface0, const int64_t* mnface, int32_t tile,
                           ens::Plan* plan, hipStream_t s) {
    hipLaunchKernelGGL(ens::plan_kernel, dim3(1), dim3(ens::kMaxBatch), 0, s, batch, mcell0,
                       mncell, mface0, mnface, tile, plan);
    return hipGetLastError();
}

hipError_t ens_launch_offset_i32(const int32_t* src, int32_t* dst, int64_t n, int64_t off,
                                 hipStream_t s) {
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ens::offset_copy_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       src, dst, n, off);
    return hipGetLastError();
}

hipError_t ens_launch_offset_i64(const int64_t* src, int64_t* dst, int64_t n, int64_t off,
                                 hipStream_t s) {
    if(n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ens::offset_copy_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       src, dst, n, off);
    return hipGetLastError();
}

}  // namespace sfx
