// sfx_launch.h -- host-side launchers of the HIP kernels (internal to the
// library; the public surface is include/shadowfax_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sfx_device.h"

namespace sfx {

// riemann_fp64.hip
hipError_t launch_solve_for_flux_fp64(int kind, int ndim, const Consts& c, int64_t N,
                                      const double* WL, const double* WR, const double* n,
                                      const double* v, const uint8_t* refl, double* F,
                                      uint8_t* status, unsigned long long* ctr,
                                      hipStream_t stream);
hipError_t launch_exact_solve_fp64(int ndim, const Consts& c, int64_t N, const double* WL,
                                   const double* WR, const double* n, const uint8_t* refl,
                                   double* Whalf, double* WRout, double* mach, uint8_t* status,
                                   unsigned long long* ctr, hipStream_t stream);

// riemann_mixed.hip: T = float (FP32 storage) or double (FP64 storage)
template <typename T>
hipError_t launch_solve_for_flux_mixed(int kind, int ndim, const Consts& c, int64_t N,
                                       const T* WL, const T* WR, const T* n, const T* v,
                                       const uint8_t* refl, T* F, uint8_t* status,
                                       unsigned long long* ctr, hipStream_t stream);

}  // namespace sfx
