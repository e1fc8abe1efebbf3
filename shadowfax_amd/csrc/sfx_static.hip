// sfx_static.hip -- the face-flux pass for callers that cannot allocate:
// the per-face fluxes dQL, the pass's only scratch, live in a file-scope
// device array instead of a caller workspace.
//
// The array's name is a host-side stand-in, not a device address: host code
// uses only the address hipGetSymbolAddress returns for it, fetched once per
// device (with that device current) on the entry point's first call there.
// Every call on a device shares its copy; the face kernel writes every row
// of the mesh's fluxes before the gather reads them (one stream), so nothing
// must start at zero, but calls that overlap on one device (other streams or
// threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sfx_capi_internal.h"
#include "shadowfax_amd.h"

using namespace sfx;

namespace {

static __device__ __align__(256) unsigned char scratch[SFX_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// the device address of this device's scratch (the device must be current)
int scratch_ptr(int dev, void** out) {
    if(dev < 0 || dev >= kMaxDevices)
        return set_error(SFX_ERR_ARG, "sfx_face_flux_pass_static: device %d", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if(!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p)
            return hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                             "sfx_face_flux_pass_static: scratch address");
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return SFX_OK;
}

}  // namespace

extern "C" {

int sfx_face_flux_pass_static(sfx_solver* h, const sfx_mesh* m, void* grad, double* dQ,
                              uint8_t* face_status, void* stream) {
    if(!h || !m) return set_error(SFX_ERR_ARG, "sfx_face_flux_pass_static: NULL handle or mesh");
    if(m->F < 0 || m->ndim != h->ndim)
        return set_error(SFX_ERR_ARG, "sfx_face_flux_pass_static: bad mesh");
    const size_t tsz = h->precision == SFX_PREC_FP64 ? 8 : 4;
    const size_t need = (size_t)m->F * (size_t)(m->ndim + 2) * tsz;
    if(need > (size_t)SFX_STATIC_SCRATCH_BYTES)
        return set_error(SFX_ERR_ARG,
                         "sfx_face_flux_pass_static: %lld faces need %zu bytes of face-flux "
                         "scratch, the static scratch holds %zu (sfx_face_flux_pass takes a "
                         "workspace of any size)",
                         (long long)m->F, need, (size_t)SFX_STATIC_SCRATCH_BYTES);
    int rc = bind_device(h);
    if(rc) return rc;
    void* dQL = nullptr;
    if((rc = scratch_ptr(h->device, &dQL))) return rc;
    return sfx_face_flux_pass(h, m, grad, dQL, dQ, face_status, stream);
}

}  // extern "C"
