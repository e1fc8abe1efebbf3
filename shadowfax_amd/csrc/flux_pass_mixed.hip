// flux_pass_mixed.hip -- mixed instantiation of the flux-pass kernels
// (flux_pass_impl.h): FP32 storage of generator velocities, gradients and
// face geometry (half the bytes of those gathers); the primitive states stay
// FP64 (a face flux driven by the small difference of two large pressures,
// as between the hot cells of a Sedov blast, cannot be recovered from
// FP32-rounded states: rounding the states alone gave 6e-6 face errors);
// gradient estimation (in the closure-free difference form), limiting and
// the MUSCL-Hancock predictor evaluated in FP64 registers; the mixed Riemann
// solvers (riemann_mixed_impl.h); face fluxes dt A F formed in FP64 and
// stored in FP32 (a 6e-8 relative rounding per face; each stored value
// enters its two cells with opposite signs); FP64 accumulation.  Each
// periodic boundary face is evaluated once and applied with opposite signs
// to its two cells, so mass, momentum and energy are conserved to FP64
// round-off (the reference evaluates both copies, src/VorFace.cpp:913-948).
#include "flux_pass_impl.h"
#include "sfx_launch.h"

namespace sfx {

static inline dim3 grid_for(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

template <>
hipError_t launch_entry_geometry<float>(int ndim, const flux::MeshView<float>& m, int64_t E,
                                     float* egeo, hipStream_t s) {
    if(m.egrp) {
        // grouped layout: a block per group of 64 cells (the last one padded)
        if(m.C <= 0) return hipSuccess;
        const dim3 g((unsigned)((m.C + 63) / 64));  // one block per group
        const unsigned egeo_lds = flux::occupancy_lds_bytes(SFX_EGEO_WAVES, m.C, SFX_GATHER_BIG, 1024);
        if(ndim == 3)
            hipLaunchKernelGGL((flux::entry_geometry_grouped_kernel<3, float>), g, dim3(256), egeo_lds,
                               s, m, egeo);
        else
            hipLaunchKernelGGL((flux::entry_geometry_grouped_kernel<2, float>), g, dim3(256), egeo_lds,
                               s, m, egeo);
        return hipGetLastError();
    }
    if(E <= 0) return hipSuccess;
    if(ndim == 3)
        hipLaunchKernelGGL((flux::entry_geometry_kernel<3, float>), grid_for(E, 256), dim3(256), 0,
                           s, m, E, egeo);
    else
        hipLaunchKernelGGL((flux::entry_geometry_kernel<2, float>), grid_for(E, 256), dim3(256), 0,
                           s, m, E, egeo);
    return hipGetLastError();
}

template <>
hipError_t launch_gradients<float>(int ndim, const flux::MeshView<float>& m, float* grad,
                                   hipStream_t s) {
    if(m.C <= 0) return hipSuccess;
    // the conditions under which gradient_kernel reads grouped records
    // through the buffer descriptor (rec_buffer)
    const bool buf = SFX_REC_BUFFER && m.egrp && m.egeo_bytes > 0 && m.egeo_bytes <= 0x7FFFFFF0ull;
    if(ndim == 3 && buf && SFX_GRAD_LDS_K > 0)
        hipLaunchKernelGGL((flux::gradient_lds_kernel<3>), grid_for(m.C, 256), dim3(256), 0, s, m,
                           grad);
    else if(ndim == 2 && buf && SFX_GRAD_LDS_K_2D > 0)
        hipLaunchKernelGGL((flux::gradient_lds_kernel<2>), grid_for(m.C, 256), dim3(256), 0, s, m,
                           grad);
    else if(ndim == 3)
        hipLaunchKernelGGL((flux::gradient_kernel<3, float, double, true>), grid_for(m.C, 256),
                           dim3(256), 0, s, m, grad);
    else
        hipLaunchKernelGGL((flux::gradient_kernel<2, float, double, true>), grid_for(m.C, 256),
                           dim3(256), 0, s, m, grad);
    return hipGetLastError();
}

template <>
hipError_t launch_face_fluxes<float>(int ndim, int kind, const flux::MeshView<float>& m,
                                     const float* grad, const Consts& c, float* dQL,
                                     uint8_t* fstatus, unsigned long long* ctr, hipStream_t s) {
    if(m.F <= 0) return hipSuccess;
    const mixed::CF cf = mixed::make_cf(c);
    // with the per-face records, the variant that reads them
#define SFX_FACE(ND, K)                                                                          \
    do {                                                                                         \
        if(m.frec)                                                                               \
            hipLaunchKernelGGL((flux::face_flux_kernel<ND, K, 1, float, double, true>),          \
                               grid_for(m.F, 256), dim3(256), 0, s, m, grad, c, cf, dQL, fstatus, \
                               ctr);                                                             \
        else                                                                                     \
            hipLaunchKernelGGL((flux::face_flux_kernel<ND, K, 1, float, double>),                \
                               grid_for(m.F, 256), dim3(256), 0, s, m, grad, c, cf, dQL, fstatus, \
                               ctr);                                                             \
    } while(0)
    if(ndim == 3 && kind == 0) SFX_FACE(3, 0);
    else if(ndim == 3) SFX_FACE(3, 1);
    else if(kind == 0) SFX_FACE(2, 0);
    else SFX_FACE(2, 1);
#undef SFX_FACE
    return hipGetLastError();
}

hipError_t launch_face_records(int ndim, const flux::MeshView<float>& m, float* frec,
                               hipStream_t s) {
    if(m.F <= 0) return hipSuccess;
    float4* out = reinterpret_cast<float4*>(frec);
    if(ndim == 3)
        hipLaunchKernelGGL((flux::face_record_kernel<3>), grid_for(m.F, 256), dim3(256), 0, s, m,
                           out);
    else
        hipLaunchKernelGGL((flux::face_record_kernel<2>), grid_for(m.F, 256), dim3(256), 0, s, m,
                           out);
    return hipGetLastError();
}

hipError_t launch_gather_mixed(int ndim, int64_t C, int64_t F, const int64_t* goff,
                               const int32_t* gent, const uint8_t* ghost, const int32_t* twin,
                               const float* dQL, double* dQ, hipStream_t s) {
    if(C <= 0) return hipSuccess;
    const unsigned lds = flux::gather_lds_bytes<float>(C);
    // large meshes (lds > 0): the variant with more rows in flight
    if(ndim == 3)
        if(lds)
            hipLaunchKernelGGL((flux::gather_kernel<3, 1, float, SFX_GATHER_KB_BIG>), grid_for(C, 256), dim3(256), lds,
                               s, C, F, goff, gent, ghost, twin, dQL, dQ);
        else
            hipLaunchKernelGGL((flux::gather_kernel<3, 1, float>), grid_for(C, 256), dim3(256), lds, s,
                           C, F, goff, gent, ghost, twin, dQL, dQ);
    else
        if(lds)
            hipLaunchKernelGGL((flux::gather_kernel<2, 1, float, SFX_GATHER_KB_BIG>), grid_for(C, 256), dim3(256), lds,
                               s, C, F, goff, gent, ghost, twin, dQL, dQ);
        else
            hipLaunchKernelGGL((flux::gather_kernel<2, 1, float>), grid_for(C, 256), dim3(256), lds, s,
                           C, F, goff, gent, ghost, twin, dQL, dQ);
    return hipGetLastError();
}


}  // namespace sfx
