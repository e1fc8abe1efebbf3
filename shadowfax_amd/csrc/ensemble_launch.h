// ensemble_launch.h -- launch sequences of the ensemble kernels, shared by
// the FP64 (ensemble_fp64.hip) and mixed (ensemble_mixed.hip) translation
// units, which instantiate them with the same compile flags as the
// standalone flux pass of that precision.
#pragma once

#include "ensemble_impl.h"
#include "ensemble_queue.h"

namespace sfx {

template <typename T, int MIXED>
hipError_t ens_step_impl(int ndim, int kind, const EnsArgs<T>& a, hipStream_t s) {
    const mixed::CF cf = mixed::make_cf(a.c);
    const int grid = a.nblocks == 0;
    // grid mode: a multiple of 8 blocks covering every XCD chunk of the
    // largest tile count
    auto cover = [](int64_t tiles) {
        return (unsigned)(ens::kXcd * ((tiles + ens::kXcd - 1) / ens::kXcd + 1));
    };
    const dim3 gc(grid ? cover(a.max_ctiles) : (unsigned)a.nblocks),
            gf(grid ? cover(a.max_ftiles) : (unsigned)a.nblocks), b((unsigned)a.block);
#define SFX_ENS(ND)                                                                            \
    do {                                                                                       \
        hipLaunchKernelGGL((ens::drain_gradients<ND, T, double, MIXED != 0>), gc, b, 0, s, a.m, \
                           a.grad, a.plan, grid);                                               \
        if(kind == 0)                                                                          \
            hipLaunchKernelGGL((ens::drain_faces<ND, 0, MIXED, T, double>), gf, b, 0, s, a.m,   \
                               a.grad, a.c, cf, a.dQL, a.fstatus, a.ctr, a.plan, grid);         \
        else                                                                                   \
            hipLaunchKernelGGL((ens::drain_faces<ND, 1, MIXED, T, double>), gf, b, 0, s, a.m,   \
                               a.grad, a.c, cf, a.dQL, a.fstatus, a.ctr, a.plan, grid);         \
        hipLaunchKernelGGL((ens::drain_gather<ND, MIXED, T>), gc, b, 0, s, a.m.gather_off,         \
                           a.m.gather_entry, a.m.ghost, a.m.twin, a.dQL, a.dQ, a.plan, grid);   \
        hipLaunchKernelGGL((ens::drain_update<ND>), gc, b, 0, s, a.c, a.Q, a.dQ, a.m.volume,    \
                           a.W, a.cstatus, a.mass, a.ctr, a.plan, grid);                        \
    } while(0)
    if(ndim == 3)
        SFX_ENS(3);
    else
        SFX_ENS(2);
#undef SFX_ENS
    return hipGetLastError();
}

template <typename T, int MIXED>
hipError_t ens_range_impl(int ndim, int kind, const EnsArgs<T>& a, int64_t cell0, int64_t ncell,
                          int64_t face0, int64_t nface, int64_t gshift, int64_t qshift,
                          hipStream_t s) {
    const mixed::CF cf = mixed::make_cf(a.c);
    const dim3 b(256);
    const dim3 gc((unsigned)((ncell + 255) / 256)), gf((unsigned)((nface + 255) / 256));
    if(ncell <= 0) return hipSuccess;
#define SFX_RNG(ND)                                                                            \
    do {                                                                                       \
        hipLaunchKernelGGL((ens::range_gradients<ND, T, double, MIXED != 0>), gc, b, 0, s, a.m, \
                           a.grad, cell0, ncell, gshift);                                       \
        if(nface > 0) {                                                                        \
            if(kind == 0)                                                                      \
                hipLaunchKernelGGL((ens::range_faces<ND, 0, MIXED, T, double>), gf, b, 0, s,    \
                                   a.m, a.grad, a.c, cf, a.dQL, a.fstatus, a.ctr, face0,        \
                                   nface, gshift, qshift);                                      \
            else                                                                               \
                hipLaunchKernelGGL((ens::range_faces<ND, 1, MIXED, T, double>), gf, b, 0, s,    \
                                   a.m, a.grad, a.c, cf, a.dQL, a.fstatus, a.ctr, face0,        \
                                   nface, gshift, qshift);                                      \
        }                                                                                      \
        hipLaunchKernelGGL((ens::range_gather<ND, MIXED, T>), gc, b, 0, s, a.m.gather_off,         \
                           a.m.gather_entry, a.m.ghost, a.m.twin, a.dQL, a.dQ, cell0, ncell,    \
                           qshift);                                                             \
        hipLaunchKernelGGL((ens::range_update<ND>), gc, b, 0, s, a.c, a.Q, a.dQ, a.m.volume,    \
                           a.W, a.cstatus, a.mass, a.ctr, cell0, ncell);                        \
    } while(0)
    if(ndim == 3)
        SFX_RNG(3);
    else
        SFX_RNG(2);
#undef SFX_RNG
    return hipGetLastError();
}

template <typename T, int MIXED>
hipError_t ens_queue_impl(int ndim, int kind, const EnsArgs<T>& a, const QueueArgs& qa, int ncu,
                          hipStream_t s, int* blocks) {
    mixed::CF cf = mixed::make_cf(a.c);
    const void* fn = nullptr;
    if(ndim == 3)
        fn = kind == 0 ? (const void*)ens::queue_kernel<3, 0, MIXED, T, double>
                       : (const void*)ens::queue_kernel<3, 1, MIXED, T, double>;
    else
        fn = kind == 0 ? (const void*)ens::queue_kernel<2, 0, MIXED, T, double>
                       : (const void*)ens::queue_kernel<2, 1, MIXED, T, double>;
    const size_t shmem = 4 * (size_t)qa.nmesh_words;
    int per_cu = 0;
    hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, a.block, shmem);
    if(err != hipSuccess) return err;
    if(per_cu < 1) return hipErrorInvalidConfiguration;
    *blocks = per_cu * ncu;
    EnsArgs<T> aa = a;
    QueueArgs q = qa;
    void* args[] = {&aa,          &cf,        &q.q,           &q.box,         &q.box_mesh,
                    &q.box_done,  &q.mc0,     &q.mnc,         &q.mf0,         &q.mnf,
                    &q.arena_cells, &q.arena_faces, &q.nmesh_words};
    return hipLaunchCooperativeKernel(fn, dim3((unsigned)*blocks), dim3((unsigned)a.block), args,
                                      (unsigned)shmem, s);
}

}  // namespace sfx
