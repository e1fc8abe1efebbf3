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

// cell_ops.hip: op 0 time derivative (a = W, b = grad, c = dW/dt), 1 get_W
// (a = volume, d = Q in/out, c = W), 2 get_Q (a = volume, b = W, c = Q)
hipError_t launch_cell_op(int op, int kind, int ndim, double gamma, int64_t N,
                          const double* a, const double* b, double* c, double* d,
                          hipStream_t s);

// cell_ops.hip: TimeLine::calculate_timestep over C cells (gas, no gravity)
struct TimelineArgs {
    uint64_t integer_maxtime, current_time, timestep, max_timestep, min_timestep;
    double tottime, mintime, cfl;
    int global_timestep;
};
hipError_t launch_hydro_timestep(int kind, int ndim, double gamma, int64_t C,
                                 const TimelineArgs& tl, const double* W, const double* vel,
                                 const double* volume, uint64_t* starttime, uint64_t* endtime,
                                 double* dt, uint8_t* active, uint64_t* result,
                                 uint8_t* status, hipStream_t s);

// riemann_mixed.hip: the exact solver in two phases (stop = FP32 iterations
// before a problem is queued for the persistent second phase)
size_t exact_split_scratch_bytes(int64_t N);
template <typename T>
hipError_t launch_exact_mixed_split(int ndim, const Consts& c, int64_t N, const T* WL,
                                    const T* WR, const T* n, const T* v, const uint8_t* refl,
                                    T* F, uint8_t* status, unsigned long long* ctr,
                                    void* scratch, int stop, int persistent_blocks,
                                    hipStream_t stream);

// flux pass (flux_pass.hip: T = double; flux_pass_mixed.hip: T = float)
namespace flux {
template <typename T>
struct MeshView;
}
template <typename T>
hipError_t launch_entry_geometry(int ndim, const flux::MeshView<T>& m, int64_t E, T* egeo,
                                 hipStream_t s);
template <typename T>
hipError_t launch_gradients(int ndim, const flux::MeshView<T>& m, T* grad, hipStream_t s);
template <typename T>
hipError_t launch_face_fluxes(int ndim, int kind, const flux::MeshView<T>& m, const T* grad,
                              const Consts& c, T* dQL, uint8_t* fstatus,
                              unsigned long long* ctr, hipStream_t s);
// endtime null: every cell's W is rebuilt; else those with endtime == tnow;
// mass (may be null) takes the new Q.m
hipError_t launch_update(int ndim, const Consts& c, int64_t C, double* Q, double* dQ,
                         const double* volume, double* W, const uint64_t* endtime,
                         uint64_t tnow, uint8_t* cstatus, double* mass,
                         unsigned long long* ctr, hipStream_t s);
hipError_t launch_gather_fp64(int ndim, int64_t C, int64_t F, const int64_t* goff,
                              const int32_t* gent, const double* dQL, double* dQ,
                              hipStream_t s);
// flux_pass_mixed.hip: the mixed face kernel's per-face records (frec: 3 or
// 4 planes of m.F float4, see flux::face_record_kernel)
hipError_t launch_face_records(int ndim, const flux::MeshView<float>& m, float* frec,
                               hipStream_t s);
hipError_t launch_gather_mixed(int ndim, int64_t C, int64_t F, const int64_t* goff,
                               const int32_t* gent, const uint8_t* ghost, const int32_t* twin,
                               const float* dQL, double* dQ, hipStream_t s);

}  // namespace sfx
