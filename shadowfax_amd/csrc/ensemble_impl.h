// ensemble_impl.h -- one hydro step for a batch of independent meshes held
// in HBM together (the serving path).
//
// All meshes of an ensemble live in one set of arrays with global cell, face
// and entry indices (mesh k owns the ranges [cell0, cell0 + ncell) etc.).
// Per-step scratch -- gradients [cell][ND][NC] and face fluxes dQL
// [face][NC] -- lives in an arena that each batch shares out: a mesh admitted
// to a batch gets a contiguous arena range, and its kernels address it with
// a constant shift (arena index = global index + shift).
//
// A step for a batch is: plan_kernel (ensemble_fp64.hip; one block: prefix sums over the
// batch give every mesh its arena offsets and its range of work tiles, and
// it zeroes the tile counters) followed by four persistent drains --
// gradients, face fluxes, gather, update -- each a fixed grid whose blocks
// take tiles of `tile` consecutive cells or faces (or, as the autotuner may
// choose, a grid of one block per possible tile, see drain).
// The tile range of a phase is cut into one contiguous chunk per XCD, and
// the blocks of the XCD a block runs on (dispatch index % 8) share its
// chunk, every block taking every nx-th tile, so neighbouring cells mostly
// meet in one L2.  (SFX_ENS_CLAIM: blocks claim tiles through one atomic
// counter per chunk instead, eight per claim, their own chunk first, then
// the others'.)
//
// The per-item work is exactly the standalone kernels' (grad_cell,
// face_item, gather_cell, update_cell in flux_pass_impl.h), so a mesh
// stepped inside an ensemble gets bit-identical results.
#pragma once

#include "flux_pass_impl.h"

namespace sfx {
namespace ens {

#ifndef SFX_ENS_CLAIM  // persistent drains: 1 = tiles claimed through atomic counters, 0 = static stride
#define SFX_ENS_CLAIM 0
#endif

constexpr int kMaxBatch = 1024;
constexpr int kXcd = 8;
constexpr int kPhases = 4;  // gradients, faces, gather, update

// One admitted mesh.
struct Slot {
    int64_t cell0, ncell, face0, nface;  // global ranges of the mesh
    int64_t gshift, qshift;              // arena index - global index
    int64_t ctile0, ftile0;              // first cell tile, first face tile
};

struct Plan {
    int32_t n;       // meshes in the batch
    int32_t tile;    // items per tile
    int64_t ctiles;  // cell tiles
    int64_t ftiles;  // face tiles
    unsigned ctr[kPhases][kXcd];
    Slot slot[kMaxBatch];
};

// the slot whose tile range holds tile t (tile0 ascending in the slot order)
template <bool FACES>
__device__ __forceinline__ int find_slot(const Plan& p, int64_t t) {
    int lo = 0, hi = p.n - 1;
    while(lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t t0 = FACES ? p.slot[mid].ftile0 : p.slot[mid].ctile0;
        if(t0 <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Drains the tiles of one phase: fn(slot, item) for every item of every
// admitted mesh.  Called by every thread of the block.
template <bool FACES, class Fn>
__device__ __forceinline__ void run_tile(const Plan* __restrict__ p, int64_t t, int tile, Fn& fn) {
    const int s = find_slot<FACES>(*p, t);
    const Slot& sl = p->slot[s];
    const int64_t base = FACES ? sl.face0 + (t - sl.ftile0) * tile
                               : sl.cell0 + (t - sl.ctile0) * tile;
    const int64_t end = FACES ? sl.face0 + sl.nface : sl.cell0 + sl.ncell;
    const int64_t item = base + threadIdx.x;
    if(threadIdx.x < tile && item < end) fn(sl, item);
}

// grid != 0: one tile per block instead -- the grid covers the largest tile
// count the arena allows, block b takes tile lo_x + b / 8 of the chunk of
// XCD x = b % 8 (the same chunks as the persistent drains), and blocks past
// their chunk's end exit at once; the hardware then keeps as many blocks
// in flight as registers allow, as for a standalone kernel.
template <bool FACES, class Fn>
__device__ __forceinline__ void drain(Plan* __restrict__ p, int phase, int grid, Fn fn) {
    __shared__ int64_t s_tile;
    // tiles claimed per atomic: the counters see tiles / kClaim requests
    constexpr int kClaim = 8;
    const int64_t T = FACES ? p->ftiles : p->ctiles;
    if(grid) {
        const int x = blockIdx.x % kXcd;
        const int64_t lo = T * x / kXcd, hi = T * (x + 1) / kXcd;
        const int64_t t = lo + blockIdx.x / kXcd;
        if(t < hi) run_tile<FACES>(p, t, p->tile, fn);
        return;
    }
    if constexpr(SFX_ENS_CLAIM) {
        // with fewer claims than blocks the first ones take them all
        if((int64_t)blockIdx.x * kClaim >= T) return;
        const int tile = p->tile;
        const int x0 = blockIdx.x % kXcd;
        for(int c = 0; c < kXcd; c++) {
            const int x = (x0 + c) % kXcd;
            const int64_t lo = T * x / kXcd, hi = T * (x + 1) / kXcd;
            for(;;) {
                if(threadIdx.x == 0) {
                    // counters only grow: a chunk seen exhausted by a plain
                    // load is exhausted, and costs no read-modify-write
                    const unsigned seen = __atomic_load_n(&p->ctr[phase][x], __ATOMIC_RELAXED);
                    s_tile = lo + (int64_t)seen >= hi
                                     ? hi
                                     : lo + (int64_t)atomicAdd(&p->ctr[phase][x], (unsigned)kClaim);
                }
                __syncthreads();
                const int64_t t0 = s_tile;
                __syncthreads();
                if(t0 >= hi) break;
                const int64_t t1 = (t0 + kClaim < hi) ? t0 + kClaim : hi;
                for(int64_t t = t0; t < t1; t++) run_tile<FACES>(p, t, tile, fn);
            }
        }
    } else {
        // static stride: the blocks of XCD x (dispatch index % 8) take every
        // nx-th tile of chunk x, nx = the number of such blocks; no atomics
        // (claims from every block through eight counters serialised where
        // those counters live, and a run's speed then hung on their
        // placement)
        (void)phase;
        (void)s_tile;
        const int tile = p->tile;
        if(gridDim.x < (unsigned)kXcd) {  // a chunk per XCD needs a block on each
            for(int64_t t = blockIdx.x; t < T; t += gridDim.x) run_tile<FACES>(p, t, tile, fn);
            return;
        }
        const int x = blockIdx.x % kXcd;
        const int64_t nx = ((int64_t)gridDim.x - x + kXcd - 1) / kXcd;
        const int64_t lo = T * x / kXcd, hi = T * (x + 1) / kXcd;
        for(int64_t t = lo + blockIdx.x / kXcd; t < hi; t += nx) run_tile<FACES>(p, t, tile, fn);
    }
}

template <int ND, typename T, typename R, bool DIFF>
__global__ __launch_bounds__(256) void drain_gradients(flux::MeshView<T> m, T* __restrict__ grad,
                                                       Plan* __restrict__ p, int grid) {
    drain<false>(p, 0, grid, [&](const Slot& sl, int64_t i) {
        flux::grad_cell<ND, T, R, DIFF>(m, i, grad, sl.gshift);
    });
}

template <int ND, int KIND, int MIXED, typename T, typename R>
__global__ __launch_bounds__(256) void drain_faces(flux::MeshView<T> m, const T* __restrict__ grad,
                                                   Consts cd, mixed::CF cf,
                                                   T* __restrict__ dQL,
                                                   uint8_t* __restrict__ fstatus,
                                                   unsigned long long* __restrict__ ctr,
                                                   Plan* __restrict__ p, int grid) {
    unsigned cnt[CNT_N] = {0u, 0u, 0u, 0u, 0u};
    drain<true>(p, 1, grid, [&](const Slot& sl, int64_t f) {
        flux::face_item<ND, KIND, MIXED, T, R>(m, f, grad, sl.gshift, cd, cf, dQL, sl.qshift,
                                               fstatus, cnt);
    });
    block_add_counters(ctr, cnt);
}

template <int ND, int MIXED, typename T>
__global__ __launch_bounds__(256) void drain_gather(const int64_t* __restrict__ goff,
                                                    const int32_t* __restrict__ gent,
                                                    const uint8_t* __restrict__ ghost,
                                                    const int32_t* __restrict__ twin,
                                                    const T* __restrict__ dQL,
                                                    double* __restrict__ dQ,
                                                    Plan* __restrict__ p, int grid) {
    drain<false>(p, 2, grid, [&](const Slot& sl, int64_t i) {
        flux::gather_cell<ND, MIXED, T>(i, goff, gent, ghost, twin, dQL, sl.qshift, dQ);
    });
}

template <int ND>
__global__ __launch_bounds__(256) void drain_update(Consts cd, double* __restrict__ Q,
                                                    double* __restrict__ dQ,
                                                    const double* __restrict__ volume,
                                                    double* __restrict__ W,
                                                    uint8_t* __restrict__ cstatus,
                                                    double* __restrict__ mass,
                                                    unsigned long long* __restrict__ ctr,
                                                    Plan* __restrict__ p, int grid) {
    unsigned fail = 0;
    drain<false>(p, 3, grid, [&](const Slot&, int64_t i) {
        fail += flux::update_cell<ND>(cd, i, Q, dQ, volume, W, cstatus, mass);
    });
    const unsigned cv[CNT_N] = {0u, 0u, 0u, 0u, fail};
    block_add_counters(ctr, cv);
}

// ---- one mesh, plain launches sized to the mesh (the unbatched path) ----

template <int ND, typename T, typename R, bool DIFF>
__global__ __launch_bounds__(256) void range_gradients(flux::MeshView<T> m, T* __restrict__ grad,
                                                       int64_t cell0, int64_t ncell,
                                                       int64_t gshift) {
    const int64_t k = xcd_block() * blockDim.x + threadIdx.x;
    if(k < ncell) flux::grad_cell<ND, T, R, DIFF>(m, cell0 + k, grad, gshift);
}

template <int ND, int KIND, int MIXED, typename T, typename R>
__global__ __launch_bounds__(256) void range_faces(flux::MeshView<T> m, const T* __restrict__ grad,
                                                   Consts cd, mixed::CF cf,
                                                   T* __restrict__ dQL,
                                                   uint8_t* __restrict__ fstatus,
                                                   unsigned long long* __restrict__ ctr,
                                                   int64_t face0, int64_t nface, int64_t gshift,
                                                   int64_t qshift) {
    const int64_t k = xcd_block() * blockDim.x + threadIdx.x;
    unsigned cnt[CNT_N] = {0u, 0u, 0u, 0u, 0u};
    if(k < nface)
        flux::face_item<ND, KIND, MIXED, T, R>(m, face0 + k, grad, gshift, cd, cf, dQL, qshift,
                                               fstatus, cnt);
    block_add_counters(ctr, cnt);
}

template <int ND, int MIXED, typename T>
__global__ __launch_bounds__(256) void range_gather(const int64_t* __restrict__ goff,
                                                    const int32_t* __restrict__ gent,
                                                    const uint8_t* __restrict__ ghost,
                                                    const int32_t* __restrict__ twin,
                                                    const T* __restrict__ dQL,
                                                    double* __restrict__ dQ, int64_t cell0,
                                                    int64_t ncell, int64_t qshift) {
    const int64_t k = xcd_block() * blockDim.x + threadIdx.x;
    if(k < ncell)
        flux::gather_cell<ND, MIXED, T>(cell0 + k, goff, gent, ghost, twin, dQL, qshift, dQ);
}

template <int ND>
__global__ __launch_bounds__(256) void range_update(Consts cd, double* __restrict__ Q,
                                                    double* __restrict__ dQ,
                                                    const double* __restrict__ volume,
                                                    double* __restrict__ W,
                                                    uint8_t* __restrict__ cstatus,
                                                    double* __restrict__ mass,
                                                    unsigned long long* __restrict__ ctr,
                                                    int64_t cell0, int64_t ncell) {
    const int64_t k = xcd_block() * blockDim.x + threadIdx.x;
    unsigned fail = 0;
    if(k < ncell) fail = flux::update_cell<ND>(cd, cell0 + k, Q, dQ, volume, W, cstatus, mass);
    const unsigned cv[CNT_N] = {0u, 0u, 0u, 0u, fail};
    block_add_counters(ctr, cv);
}

}  // namespace ens

// Everything one ensemble step touches.  m spans the whole ensemble (C, F
// are the capacities, the strides of the component-major face arrays).
template <typename T>
struct EnsArgs {
    flux::MeshView<T> m;
    T* grad;       // arena [cells][ND][NC]
    T* dQL;        // arena [faces][NC]
    double* dQ;    // [C][NC]
    double* Q;     // [C][NC]
    double* W;     // [C][NC] (the same array as m.W)
    double* mass;  // [C] (the same array as m.mass): the update writes Q.m
    uint8_t* fstatus;
    uint8_t* cstatus;
    Consts c;
    unsigned long long* ctr;
    ens::Plan* plan;
    int block;    // threads per block = items per tile
    int nblocks;  // persistent grid (0: one tile per block)
    int64_t max_ctiles, max_ftiles;  // tile counts the arena allows (grid mode)
};

// ensemble_fp64.hip (T = double) and ensemble_mixed.hip (T = float)
template <typename T>
hipError_t ens_launch_step(int ndim, int kind, const EnsArgs<T>& a, hipStream_t s);
template <typename T>
hipError_t ens_launch_range(int ndim, int kind, const EnsArgs<T>& a, int64_t cell0, int64_t ncell,
                            int64_t face0, int64_t nface, int64_t gshift, int64_t qshift,
                            hipStream_t s);
template <typename T>
hipError_t ens_launch_entries(int ndim, const flux::MeshView<T>& m, int64_t E, T* egeo,
                              hipStream_t s);
// The request queue of the persistent serving kernel (ensemble_queue.h).
namespace ens {
struct DevQueue;
struct HostBox;
}  // namespace ens
struct QueueArgs {
    ens::DevQueue* q;          // device
    ens::HostBox* box;         // the pinned mailbox, as the device sees it
    const int32_t* box_mesh;   // [nreq] mesh of each request (mailbox)
    int32_t* box_done;         // [nreq] batch number once done (mailbox)
    const int64_t *mc0, *mnc, *mf0, *mnf;  // mesh table (device)
    int64_t arena_cells, arena_faces;
    int nmesh_words;           // 32-bit words of the leader's mesh bitset
};
// launches the persistent kernel cooperatively, every block resident (as
// many per CU as occupancy allows); *blocks: the grid
template <typename T>
hipError_t ens_launch_queue(int ndim, int kind, const EnsArgs<T>& a, const QueueArgs& qa,
                            int ncu, hipStream_t s, int* blocks);

// ensemble_fp64.hip
hipError_t ens_launch_plan(const int32_t* batch, const int64_t* mcell0, const int64_t* mncell,
                           const int64_t* mface0, const int64_t* mnface, int32_t tile,
                           ens::Plan* plan, hipStream_t s);
hipError_t ens_launch_offset_i32(const int32_t* src, int32_t* dst, int64_t n, int64_t off,
                                 hipStream_t s);
hipError_t ens_launch_offset_i64(const int64_t* src, int64_t* dst, int64_t n, int64_t off,
                                 hipStream_t s);

}  // namespace sfx
