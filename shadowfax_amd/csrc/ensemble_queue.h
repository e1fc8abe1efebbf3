// ensemble_queue.h -- serving from a request queue on the device: one
// persistent kernel, launched cooperatively (every block resident), drains
// the requests the host posts, admits them by scratch footprint, plans each
// batch and runs the four phases of the batched step, with a grid barrier
// between phases.
//
// Host -> device: a mailbox in pinned, coherent host memory (the host writes
// request r's mesh id at mesh[r], then publishes tail = r + 1; `stop` once
// no request will follow).  The device's leader block (block 0) moves new
// requests into the queue proper, a ring in HBM (DevQueue), whose state only
// the device touches.  Device -> host: done[r] = the number (from 1) of the
// batch that completed request r's step (system-scope stores to the
// mailbox).
//
// One iteration of the loop:
//   leader: mark the previous batch's requests done; pull new requests
//     (waiting, bounded, while none is pending); admit -- oldest first, a
//     second request for a mesh already in the batch stays queued, the scan
//     stops at the first mesh whose cells or faces no longer fit the arena
//     (sfx_serve_admit's rule) -- and write the batch's plan (slots,
//     tiles, zeroed tile counters; plan_kernel's arithmetic);
//   grid barrier; then every block: gradients, barrier, faces, barrier,
//   gather, barrier, update, barrier -- the drains of ensemble_impl.h.
// Every wait (grid barrier, mailbox) is bounded in time: when one runs out
// the block sets `error` and leaves, so every block ends and the kernel
// completes; `stop` ends them normally.
#pragma once

#include "ensemble_impl.h"

#ifndef SFX_QUEUE_WAVES_2D  // waves per SIMD the 2D serving kernel is compiled for
#define SFX_QUEUE_WAVES_2D 2
#endif

namespace sfx {
namespace ens {

constexpr int kQueueCap = 1 << 16;   // requests pending on the device at once
constexpr int kStage = 256;          // candidates staged per admission round
// 100 MHz constant clock ticks a bounded wait may last without progress
constexpr uint64_t kWaitTicks = 100ull * 1000 * 1000 * 10;  // 10 s

struct HostBox {
    uint64_t tail;     // requests posted (host)
    int32_t stop;      // no request after tail (host)
    int32_t pad;
};

struct DevQueue {
    int32_t req[kQueueCap];
    int32_t mesh[kQueueCap];
    uint32_t taken[kQueueCap];
    uint64_t head, tail;      // pending entries: ring positions [head, tail)
    int32_t batch_req[kMaxBatch];
    int32_t batch_n;
    uint32_t bar_count, bar_gen;
    int32_t stop, error;
    uint64_t nbatch;
    uint64_t consumed;  // mailbox entries pulled so far (the leader's thread 0)
};

__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

// Grid barrier over `nblocks` resident blocks: thread 0 arrives on a
// counter; the last arrival resets it and bumps the generation the others
// wait on.  Returns false (error set) when the wait ran out.
__device__ __forceinline__ bool grid_barrier(DevQueue* q, unsigned nblocks, unsigned& gen) {
    __shared__ int s_ok;
    __syncthreads();
    if(threadIdx.x == 0) {
        int ok = 1;
        __threadfence();
        const unsigned arrived = __hip_atomic_fetch_add(&q->bar_count, 1u, __ATOMIC_ACQ_REL,
                                                        __HIP_MEMORY_SCOPE_AGENT);
        if(arrived == nblocks - 1) {
            __hip_atomic_store(&q->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&q->bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = now_ticks();
            while(__hip_atomic_load(&q->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
                  gen) {
                if(now_ticks() - t0 > kWaitTicks) {
                    __hip_atomic_store(&q->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __threadfence();
        gen++;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// ring entries are written by the leader's thread 0 and read by the other
// threads of its block: agent-scope accesses, past the CU's vector cache
template <typename V>
__device__ __forceinline__ V ld(const V* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V>
__device__ __forceinline__ void st(V* p, V v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Requests posted to the mailbox since the last pull, into the ring (as
// many as it has room for; the rest wait in the mailbox).  Thread 0 of the
// leader block.
__device__ __forceinline__ void pull(DevQueue* q, const HostBox* box, const int32_t* box_mesh,
                                     uint64_t& consumed) {
    const uint64_t t = __hip_atomic_load(&box->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    for(; consumed < t && q->tail - q->head < (uint64_t)kQueueCap; consumed++) {
        const uint64_t s = q->tail % kQueueCap;
        st(&q->req[s], (int32_t)consumed);
        st(&q->mesh[s], __hip_atomic_load(&box_mesh[consumed], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_SYSTEM));
        st(&q->taken[s], 0u);
        q->tail++;
    }
}

// The leader's part of one iteration (all threads of block 0 call it).
// bits: LDS bitset over mesh ids (the meshes already in the batch).  Its
// arguments are pointers and values, nothing the kernel's other phases
// read through a reference: where the compiler keeps it a call (it does
// for 2D), a reference to the kernel's argument block would put that block
// in scratch memory for the whole kernel.
static __device__ void leader(Plan* p, int tile, DevQueue* q, HostBox* box,
                              const int32_t* box_mesh, int32_t* box_done, const int64_t* mc0,
                              const int64_t* mnc, const int64_t* mf0, const int64_t* mnf,
                              int64_t arena_cells, int64_t arena_faces, uint32_t* bits) {
    __shared__ int32_t s_req[kStage], s_mesh[kStage];
    __shared__ int64_t s_nc[kStage], s_nf[kStage];
    __shared__ uint32_t s_taken[kStage];
    __shared__ int32_t s_n, s_stop, s_quit;
    __shared__ int64_t s_cells, s_faces;
    __shared__ uint64_t s_pos, s_tail;
    if(threadIdx.x == 0) {
        uint64_t consumed = q->consumed;
        // the previous batch is complete (the update's barrier is behind us)
        const int32_t nb = q->batch_n;
        for(int k = 0; k < nb; k++)
            __hip_atomic_store(&box_done[q->batch_req[k]], (int32_t)q->nbatch, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        q->batch_n = 0;
        // drop the taken entries at the front of the ring
        while(q->head < q->tail && ld(&q->taken[q->head % kQueueCap])) q->head++;
        // wait (bounded) until a request is pending or the host is done
        s_quit = 0;
        uint64_t t0 = now_ticks();
        for(;;) {
            pull(q, box, box_mesh, consumed);
            if(q->head < q->tail) break;
            const int32_t hstop = __hip_atomic_load(&box->stop, __ATOMIC_ACQUIRE,
                                                    __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t tl = __hip_atomic_load(&box->tail, __ATOMIC_ACQUIRE,
                                                  __HIP_MEMORY_SCOPE_SYSTEM);
            if(hstop && consumed >= tl) {
                s_quit = 1;
                break;
            }
            if(now_ticks() - t0 > kWaitTicks) {
                q->error = 1;
                s_quit = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        q->consumed = consumed;
        s_n = 0;
        s_cells = s_faces = 0;
        s_stop = 0;
        s_pos = q->head;
        s_tail = q->tail;
    }
    __syncthreads();
    if(s_quit) {
        if(threadIdx.x == 0) q->stop = 1;
        return;
    }
    // admission: candidates staged kStage at a time (their loads in
    // parallel), decided by thread 0 in queue order
    while(!s_stop) {
        const uint64_t pos = s_pos, tail = s_tail;
        const int cnt = (int)min((uint64_t)kStage, tail - pos);
        for(int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const uint64_t s = (pos + i) % kQueueCap;
            s_req[i] = ld(&q->req[s]);
            const int32_t m = ld(&q->mesh[s]);
            s_mesh[i] = m;
            s_taken[i] = ld(&q->taken[s]);
            s_nc[i] = mnc[m];
            s_nf[i] = mnf[m];
        }
        __syncthreads();
        if(threadIdx.x == 0) {
            int n = s_n;
            int64_t cells = s_cells, faces = s_faces;
            bool stop = cnt == 0;
            for(int k = 0; k < cnt && !stop; k++) {
                if(s_taken[k]) continue;
                const int32_t m = s_mesh[k];
                if(bits[m >> 5] & (1u << (m & 31))) continue;  // its mesh is in the batch
                if(cells + s_nc[k] > arena_cells || faces + s_nf[k] > arena_faces) {
                    stop = true;  // FIFO: no overtaking
                    break;
                }
                cells += s_nc[k];
                faces += s_nf[k];
                bits[m >> 5] |= 1u << (m & 31);
                st(&q->taken[(pos + k) % kQueueCap], 1u);
                q->batch_req[n] = s_req[k];
                p->slot[n].cell0 = m;  // the mesh id, replaced by the plan below
                n++;
                if(n == kMaxBatch) stop = true;
            }
            s_n = n;
            s_cells = cells;
            s_faces = faces;
            s_pos = pos + (uint64_t)cnt;
            s_stop = stop || s_pos >= tail;
        }
        __syncthreads();
    }
    // the plan (plan_kernel's arithmetic, sequential over the batch)
    if(threadIdx.x == 0) {
        const int n = s_n;
        int64_t c = 0, f = 0, ct = 0, ft = 0;
        for(int k = 0; k < n; k++) {
            const int32_t m = (int32_t)p->slot[k].cell0;
            bits[m >> 5] &= ~(1u << (m & 31));
            Slot sl;
            sl.cell0 = mc0[m];
            sl.ncell = mnc[m];
            sl.face0 = mf0[m];
            sl.nface = mnf[m];
            sl.gshift = c - sl.cell0;
            sl.qshift = f - sl.face0;
            sl.ctile0 = ct;
            sl.ftile0 = ft;
            p->slot[k] = sl;
            c += sl.ncell;
            f += sl.nface;
            ct += (sl.ncell + tile - 1) / tile;
            ft += (sl.nface + tile - 1) / tile;
        }
        p->n = n;
        p->tile = tile;
        p->ctiles = ct;
        p->ftiles = ft;
        for(int ph = 0; ph < kPhases; ph++)
            for(int x = 0; x < kXcd; x++) p->ctr[ph][x] = 0u;
        q->batch_n = n;
        q->nbatch++;
    }
    __syncthreads();
}

// The four phases of a batch: the drains of ensemble_impl.h in persistent
// mode (every block resident, tiles claimed eight per atomic).
template <int ND, typename T, typename R, bool DIFF>
__device__ __forceinline__ void q_gradients(const EnsArgs<T>& a) {
    drain<false>(a.plan, 0, 0, [&](const Slot& sl, int64_t i) {
        flux::grad_cell<ND, T, R, DIFF>(a.m, i, a.grad, sl.gshift);
    });
}
template <int ND, int KIND, int MIXED, typename T, typename R>
__device__ __forceinline__ void q_faces(const EnsArgs<T>& a, const mixed::CF& cf,
                                     unsigned (&cnt)[CNT_N]) {
    drain<true>(a.plan, 1, 0, [&](const Slot& sl, int64_t f) {
        flux::face_item<ND, KIND, MIXED, T, R>(a.m, f, a.grad, sl.gshift, a.c, cf, a.dQL,
                                               sl.qshift, a.fstatus, cnt);
    });
}
template <int ND, int MIXED, typename T>
__device__ __forceinline__ void q_gather(const EnsArgs<T>& a) {
    drain<false>(a.plan, 2, 0, [&](const Slot& sl, int64_t i) {
        flux::gather_cell<ND, MIXED, T>(i, a.m.gather_off, a.m.gather_entry, a.m.ghost, a.m.twin,
                                        a.dQL, sl.qshift, a.dQ);
    });
}
template <int ND, typename T>
__device__ __forceinline__ unsigned q_update(const EnsArgs<T>& a) {
    unsigned fail = 0;
    drain<false>(a.plan, 3, 0, [&](const Slot&, int64_t i) {
        fail += flux::update_cell<ND>(a.c, i, a.Q, a.dQ, a.m.volume, a.W, a.cstatus, a.mass);
    });
    return fail;
}

// The persistent serving kernel.  Launched cooperatively with every block
// resident; blockDim = a.block.  Dynamic LDS: the leader's bitset over mesh
// ids (4 bytes per 32 meshes).  One function holding the four phases needs
// more registers than any of them alone (3D mixed: 236 VGPRs against 191
// for the gradient drain, 2D: 208); it runs at two waves per SIMD (held to
// three or four, the 2D kernel spills and serves fewer requests).
template <int ND>
constexpr int queue_waves() {
    return ND == 2 ? SFX_QUEUE_WAVES_2D : 2;
}

template <int ND, int KIND, int MIXED, typename T, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(queue_waves<ND>()))) void
queue_kernel(EnsArgs<T> a, mixed::CF cf, DevQueue* q,
                                                    HostBox* box, const int32_t* box_mesh,
                                                    int32_t* box_done, const int64_t* mc0,
                                                    const int64_t* mnc, const int64_t* mf0,
                                                    const int64_t* mnf, int64_t arena_cells,
                                                    int64_t arena_faces, int nmesh_words) {
    extern __shared__ uint32_t bits[];
    for(int w = threadIdx.x; w < nmesh_words; w += blockDim.x) bits[w] = 0u;
    __syncthreads();
    const unsigned nb = gridDim.x;
    unsigned gen = 0;
    unsigned fail = 0;
    unsigned cnt[CNT_N] = {0u, 0u, 0u, 0u, 0u};
    for(;;) {
        if(blockIdx.x == 0)
            leader(a.plan, a.block, q, box, box_mesh, box_done, mc0, mnc, mf0, mnf,
                   arena_cells, arena_faces, bits);
        if(!grid_barrier(q, nb, gen)) break;
        if(__hip_atomic_load(&q->stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) break;
        q_gradients<ND, T, R, MIXED != 0>(a);
        if(!grid_barrier(q, nb, gen)) break;
        q_faces<ND, KIND, MIXED, T, R>(a, cf, cnt);
        if(!grid_barrier(q, nb, gen)) break;
        q_gather<ND, MIXED, T>(a);
        if(!grid_barrier(q, nb, gen)) break;
        fail += q_update<ND, T>(a);
        if(!grid_barrier(q, nb, gen)) break;
    }
    cnt[CNT_NFAIL] += fail;
    block_add_counters(a.ctr, cnt);
}

}  // namespace ens
}  // namespace sfx
