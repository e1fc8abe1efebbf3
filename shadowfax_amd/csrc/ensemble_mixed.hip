// ensemble_mixed.hip -- mixed instantiation of the ensemble kernels (the
// mixed flux pass per item, compiled like flux_pass_mixed.hip).
#include "ensemble_launch.h"
#include "sfx_launch.h"

namespace sfx {

template <>
hipError_t ens_launch_step<float>(int ndim, int kind, const EnsArgs<float>& a, hipStream_t s) {
    return ens_step_impl<float, 1>(ndim, kind, a, s);
}

template <>
hipError_t ens_launch_range<float>(int ndim, int kind, const EnsArgs<float>& a, int64_t cell0,
                                   int64_t ncell, int64_t face0, int64_t nface, int64_t gshift,
                                   int64_t qshift, hipStream_t s) {
    return ens_range_impl<float, 1>(ndim, kind, a, cell0, ncell, face0, nface, gshift, qshift, s);
}

template <>
hipError_t ens_launch_queue<float>(int ndim, int kind, const EnsArgs<float>& a, const QueueArgs& qa,
                                int ncu, hipStream_t s, int* blocks) {
    return ens_queue_impl<float, 1>(ndim, kind, a, qa, ncu, s, blocks);
}

template <>
hipError_t ens_launch_entries<float>(int ndim, const flux::MeshView<float>& m, int64_t E,
                                     float* egeo, hipStream_t s) {
    return launch_entry_geometry<float>(ndim, m, E, egeo, s);
}

}  // namespace sfx
