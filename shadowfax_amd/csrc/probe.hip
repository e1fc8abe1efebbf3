// probe.hip -- measured ceilings of the GPU the roofline figures are priced
// against: a streaming copy (HBM bandwidth) and FP64 / FP32 FMA loops (VALU
// rate).  Built as its own library (libsfxprobe.so): these kernels are
// measurement tools, not part of the face-flux path's interface.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// out[i] = in[i] over n float4 (16-byte aligned), grid-stride with four
// loads in flight a lane before their stores; non-temporal, as a stream
typedef float f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_kernel(const f4* __restrict__ in,
                                                   f4* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for(; i + 3 * stride < n; i += 4 * stride) {
        f4 v[4];
#pragma unroll
        for(int u = 0; u < 4; u++) v[u] = __builtin_nontemporal_load(in + i + u * stride);
#pragma unroll
        for(int u = 0; u < 4; u++) __builtin_nontemporal_store(v[u], out + i + u * stride);
    }
    for(; i < n; i += stride) out[i] = in[i];
}

// 8 independent chains of iters fused multiply-adds a lane (2 flops each);
// the result is stored so that nothing is dropped
template <typename T>
__global__ __launch_bounds__(256) void fma_kernel(T* __restrict__ out, int iters, T b, T c) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    T x[8];
#pragma unroll
    for(int k = 0; k < 8; k++) x[k] = (T)(threadIdx.x + k) * (T)1e-3;
#pragma unroll 4
    for(int i = 0; i < iters; i++) {
#pragma unroll
        for(int k = 0; k < 8; k++) x[k] = fma(x[k], b, c);
    }
    T s = 0;
#pragma unroll
    for(int k = 0; k < 8; k++) s += x[k];
    out[g] = s;
}

}  // namespace

extern "C" {

// Copies n4 float4 from in to out on stream; 0 on success, else the HIP error.
int sfxp_copy(const void* in, void* out, int64_t n4, int blocks, void* stream) {
    if(n4 <= 0 || blocks <= 0 || !in || !out) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const f4*)in, (f4*)out, n4);
    return (int)hipGetLastError();
}

// blocks x 256 lanes, each 8 x iters FMAs in FP64 (fp64 != 0) or FP32;
// out holds blocks x 256 values of that type.  Flops: 16 x iters a lane.
int sfxp_fma(int fp64, void* out, int blocks, int iters, void* stream) {
    if(blocks <= 0 || iters <= 0 || !out) return (int)hipErrorInvalidValue;
    if(fp64)
        hipLaunchKernelGGL(fma_kernel<double>, dim3((unsigned)blocks), dim3(256), 0,
                           (hipStream_t)stream, (double*)out, iters, 0.999999, 1e-7);
    else
        hipLaunchKernelGGL(fma_kernel<float>, dim3((unsigned)blocks), dim3(256), 0,
                           (hipStream_t)stream, (float*)out, iters, 0.999999f, 1e-7f);
    return (int)hipGetLastError();
}

}  // extern "C"
