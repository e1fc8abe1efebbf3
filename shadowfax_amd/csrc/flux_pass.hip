// flux_pass.hip -- FP64 parity instantiation of the flux-pass kernels
// (flux_pass_impl.h), compiled with -ffp-contract=off so every expression
// keeps the reference's association.
#include "flux_pass_impl.h"
#include "sfx_launch.h"

namespace sfx {

static inline dim3 grid_for(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

template <>
hipError_t launch_entry_geometry<double>(int ndim, const flux::MeshView<double>& m, int64_t E,
                                     double* egeo, hipStream_t s) {
    if(m.egrp) {
        // grouped layout: a block per group of 64 cells (the last one padded)
        if(m.C <= 0) return hipSuccess;
        const dim3 g((unsigned)((m.C + 63) / 64));  // one block per group
        const unsigned egeo_lds = flux::occupancy_lds_bytes(SFX_EGEO_WAVES, m.C, SFX_GATHER_BIG, 1024);
        if(ndim == 3)
            hipLaunchKernelGGL((flux::entry_geometry_grouped_kernel<3, double>), g, dim3(256), egeo_lds,
                               s, m, egeo);
        else
            hipLaunchKernelGGL((flux::entry_geometry_grouped_kernel<2, double>), g, dim3(256), egeo_lds,
                               s, m, egeo);
        return hipGetLastError();
    }
    if(E <= 0) return hipSuccess;
    if(ndim == 3)
        hipLaunchKernelGGL((flux::entry_geometry_kernel<3, double>), grid_for(E, 256), dim3(256), 0,
                           s, m, E, egeo);
    else
        hipLaunchKernelGGL((flux::entry_geometry_kernel<2, double>), grid_for(E, 256), dim3(256), 0,
                           s, m, E, egeo);
    return hipGetLastError();
}

template <>
hipError_t launch_gradients<double>(int ndim, const flux::MeshView<double>& m, double* grad,
                                    hipStream_t s) {
    if(m.C <= 0) return hipSuccess;
    if(ndim == 3)
        hipLaunchKernelGGL((flux::gradient_kernel<3, double, double, false>), grid_for(m.C, 256),
                           dim3(256), 0, s, m, grad);
    else
        hipLaunchKernelGGL((flux::gradient_kernel<2, double, double, false>), grid_for(m.C, 256),
                           dim3(256), 0, s, m, grad);
    return hipGetLastError();
}

template <>
hipError_t launch_face_fluxes<double>(int ndim, int kind, const flux::MeshView<double>& m,
                                      const double* grad, const Consts& c, double* dQL,
                                      uint8_t* fstatus, unsigned long long* ctr,
                                      hipStream_t s) {
    if(m.F <= 0) return hipSuccess;
    const mixed::CF cf = mixed::make_cf(c);
#define SFX_FACE(ND, K)                                                                    \
    hipLaunchKernelGGL((flux::face_flux_kernel<ND, K, 0, double, double>), grid_for(m.F, 256), \
                       dim3(256), 0, s, m, grad, c, cf, dQL, fstatus, ctr)
    if(ndim == 3 && kind == 0) SFX_FACE(3, 0);
    else if(ndim == 3) SFX_FACE(3, 1);
    else if(kind == 0) SFX_FACE(2, 0);
    else SFX_FACE(2, 1);
#undef SFX_FACE
    return hipGetLastError();
}

hipError_t launch_gather_fp64(int ndim, int64_t C, int64_t F, const int64_t* goff,
                              const int32_t* gent, const double* dQL, double* dQ,
                              hipStream_t s) {
    if(C <= 0) return hipSuccess;
    const unsigned lds = flux::gather_lds_bytes<double>(C);
    // large meshes (lds > 0): the variant with more rows in flight
    if(ndim == 3)
        if(lds)
            hipLaunchKernelGGL((flux::gather_kernel<3, 0, double, SFX_GATHER_KB_BIG_F64>), grid_for(C, 256), dim3(256), lds,
                               s, C, F, goff, gent, nullptr, nullptr, dQL, dQ);
        else
            hipLaunchKernelGGL((flux::gather_kernel<3, 0, double>), grid_for(C, 256), dim3(256), lds, s,
                           C, F, goff, gent, nullptr, nullptr, dQL, dQ);
    else
        if(lds)
            hipLaunchKernelGGL((flux::gather_kernel<2, 0, double, SFX_GATHER_KB_BIG_F64>), grid_for(C, 256), dim3(256), lds,
                               s, C, F, goff, gent, nullptr, nullptr, dQL, dQ);
        else
            hipLaunchKernelGGL((flux::gather_kernel<2, 0, double>), grid_for(C, 256), dim3(256), lds, s,
                           C, F, goff, gent, nullptr, nullptr, dQL, dQ);
    return hipGetLastError();
}

hipError_t launch_update(int ndim, const Consts& c, int64_t C, double* Q, double* dQ,
                         const double* volume, double* W, const uint64_t* endtime,
                         uint64_t tnow, uint8_t* cstatus, double* mass,
                         unsigned long long* ctr, hipStream_t s) {
    if(C <= 0) return hipSuccess;
    if(ndim == 3)
        hipLaunchKernelGGL((flux::update_kernel<3>), grid_for(C, 256), dim3(256), 0, s, c, C, Q,
                           dQ, volume, W, endtime, tnow, cstatus, mass, ctr);
    else
        hipLaunchKernelGGL((flux::update_kernel<2>), grid_for(C, 256), dim3(256), 0, s, c, C, Q,
                           dQ, volume, W, endtime, tnow, cstatus, mass, ctr);
    return hipGetLastError();
}

}  // namespace sfx
