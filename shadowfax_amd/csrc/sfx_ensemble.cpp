// sfx_ensemble.cpp -- the serving runtime: an ensemble of independent meshes
// resident in HBM, stepped in batches through one captured HIP graph
// (ensemble_impl.h), with FIFO admission by scratch footprint and an
// open-loop request server that records per-request latency.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "ensemble_impl.h"
#include "ensemble_queue.h"
#include "sfx_capi_internal.h"
#include "shadowfax_amd.h"

using namespace sfx;

struct sfx_ensemble {
    sfx_solver* h = nullptr;
    int nd = 2, NC = 4, prec = SFX_PREC_FP64;
    size_t tsz = 8;
    int64_t cap_mesh = 0, cap_cells = 0, cap_faces = 0, cap_ent = 0, cap_gent = 0;
    int64_t nmesh = 0, ncells = 0, nfaces = 0, nent = 0, ngent = 0;
    // each mesh's first cell is a multiple of 64, so a mesh owns whole groups
    // of the grouped record layout: the global cell arrays hold cap_alloc =
    // cap_cells + 64 cap_mesh rows, cell_end is one past the last used row
    int64_t cap_alloc = 0, cell_end = 0;
    // global arrays
    double *W = nullptr, *Q = nullptr, *dQ = nullptr, *mass = nullptr, *total_area = nullptr,
           *dt = nullptr, *volume = nullptr;
    void* vel = nullptr;
    uint8_t *active = nullptr, *ghost = nullptr, *fstatus = nullptr, *cstatus = nullptr;
    int32_t *left = nullptr, *right = nullptr, *twin = nullptr, *cf_entry = nullptr,
            *gather_entry = nullptr;
    int64_t *cf_off = nullptr, *gather_off = nullptr, *egrp = nullptr;
    size_t egeo_bytes = 0;  // allocated
    uint64_t rec_bytes = 0;  // the grouped records built in it
    void *area = nullptr, *dL = nullptr, *dR = nullptr, *rRL = nullptr, *mc = nullptr,
         *rcent = nullptr, *egeo = nullptr;
    // mixed: the cells' closure residuals [nd][cap_alloc], allocated (zeroed)
    // when the first mesh that brings one is added
    void* closure = nullptr;
    // mesh table
    std::vector<int64_t> c0, nc, f0, nf;
    int64_t *d_c0 = nullptr, *d_nc = nullptr, *d_f0 = nullptr, *d_nf = nullptr;
    // scratch arena
    int64_t arena_cells = 0, arena_faces = 0;
    void* grad = nullptr;
    void* dQL = nullptr;
    // batching
    int32_t* h_batch = nullptr;  // pinned [1 + kMaxBatch]
    int32_t* d_batch = nullptr;
    ens::Plan* d_plan = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int block = 256, per_cu = 4, ncu = 256;
    bool entries_built = false;
    std::vector<void*> allocs;
};

namespace {

int ens_free(sfx_ensemble* e) {
    if(!e) return SFX_OK;
    if(e->h) bind_device(e->h);
    if(e->exec) hipGraphExecDestroy(e->exec);
    if(e->graph) hipGraphDestroy(e->graph);
    for(void* p : e->allocs) hipFree(p);
    if(e->h_batch) hipHostFree(e->h_batch);
    if(e->stream) hipStreamDestroy(e->stream);
    delete e;
    return SFX_OK;
}

template <typename P>
hipError_t dalloc(sfx_ensemble* e, P** p, size_t bytes) {
    void* q = nullptr;
    hipError_t err = hipMalloc(&q, std::max<size_t>(bytes, 256));
    if(err == hipSuccess) {
        e->allocs.push_back(q);
        *p = (P*)q;
    }
    return err;
}

template <typename T>
flux::MeshView<T> ens_view(const sfx_ensemble* e) {
    flux::MeshView<T> v;
    v.C = e->cap_alloc;
    v.F = e->cap_faces;
    v.W = e->W;
    v.vel = (const double*)e->vel;
    v.mass = e->mass;
    v.total_area = e->total_area;
    v.dt = e->dt;
    v.volume = e->volume;
    v.active = e->active;
    v.left = e->left;
    v.right = e->right;
    v.ghost = e->ghost;
    v.twin = e->prec == SFX_PREC_MIXED ? e->twin : nullptr;
    v.area = (const T*)e->area;
    v.dL = (const T*)e->dL;
    v.dR = (const T*)e->dR;
    v.rRL = (const T*)e->rRL;
    v.mc = (const T*)e->mc;
    v.rcent = (const T*)e->rcent;
    v.cf_off = e->cf_off;
    v.cf_entry = e->cf_entry;
    v.egeo = (const T*)e->egeo;
    v.egrp = e->egrp;
    v.egeo_bytes = e->rec_bytes;
    v.gather_off = e->gather_off;
    v.gather_entry = e->gather_entry;
    v.closure = (const T*)e->closure;
    return v;
}

template <typename T>
EnsArgs<T> ens_args(const sfx_ensemble* e) {
    EnsArgs<T> a;
    a.m = ens_view<T>(e);
    a.grad = (T*)e->grad;
    a.dQL = (T*)e->dQL;
    a.dQ = e->dQ;
    a.Q = e->Q;
    a.W = e->W;
    a.mass = e->mass;
    a.fstatus = e->fstatus;
    a.cstatus = e->cstatus;
    a.c = e->h->c;
    a.ctr = e->h->stats_on ? e->h->d_ctr : nullptr;
    a.plan = e->d_plan;
    a.block = e->block;
    a.nblocks = e->ncu * e->per_cu;
    // a batch's tiles: ceil(items / block) per mesh, at most kMaxBatch meshes
    a.max_ctiles = (e->arena_cells + e->block - 1) / e->block + ens::kMaxBatch;
    a.max_ftiles = (e->arena_faces + e->block - 1) / e->block + ens::kMaxBatch;
    return a;
}

void drop_graph(sfx_ensemble* e) {
    if(e->exec) hipGraphExecDestroy(e->exec);
    if(e->graph) hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// plan + drains for the batch already in d_batch
hipError_t launch_batch(sfx_ensemble* e, hipStream_t s) {
    hipError_t err = ens_launch_plan(e->d_batch, e->d_c0, e->d_nc, e->d_f0, e->d_nf, e->block,
                                     e->d_plan, s);
    if(err != hipSuccess) return err;
    return e->prec == SFX_PREC_FP64
                   ? ens_launch_step<double>(e->nd, e->h->kind, ens_args<double>(e), s)
                   : ens_launch_step<float>(e->nd, e->h->kind, ens_args<float>(e), s);
}

// The gradient records in the grouped layout (groups of 64 global cells,
// one row per entry of the group's fullest cell; flux_pass_impl.h) over
// cells 0 .. cell_end - 1, the alignment gaps included (no entries).
int build_entries(sfx_ensemble* e) {
    if(e->entries_built) return SFX_OK;
    const int64_t C = e->cell_end, ng = (C + 63) / 64;
    std::vector<int64_t> off(C + 1), grp(ng + 1);
    hipError_t err = hipStreamSynchronize(e->stream);
    if(err == hipSuccess)
        err = hipMemcpy(off.data(), e->cf_off, 8 * (C + 1), hipMemcpyDeviceToHost);
    if(err != hipSuccess) return hip_error(err, "ensemble entry groups");
    const int64_t rows = sfx_mesh_entry_groups(C, off.data(), grp.data());
    if(rows < 0) return SFX_ERR_ARG;
    const size_t bytes = (size_t)rows * 64 * flux::entry_rec(e->nd, e->tsz == 4) * e->tsz;
    if(bytes > e->egeo_bytes) {
        e->allocs.erase(std::find(e->allocs.begin(), e->allocs.end(), e->egeo));
        hipFree(e->egeo);
        e->egeo = nullptr;
        e->egeo_bytes = 0;
        err = dalloc(e, &e->egeo, bytes);
        if(err != hipSuccess)
            return err == hipErrorOutOfMemory
                           ? set_error(SFX_ERR_NOMEM, "ensemble entry records: out of device memory")
                           : hip_error(err, "ensemble entry records");
        e->egeo_bytes = bytes;
        drop_graph(e);
    }
    e->rec_bytes = bytes;
    err = hipMemcpy(e->egrp, grp.data(), 8 * (ng + 1), hipMemcpyHostToDevice);
    if(err == hipSuccess) {
        if(e->prec == SFX_PREC_FP64) {
            flux::MeshView<double> v = ens_view<double>(e);
            v.C = C;
            err = ens_launch_entries<double>(e->nd, v, e->nent, (double*)e->egeo, e->stream);
        } else {
            flux::MeshView<float> v = ens_view<float>(e);
            v.C = C;
            err = ens_launch_entries<float>(e->nd, v, e->nent, (float*)e->egeo, e->stream);
        }
    }
    if(err != hipSuccess) return hip_error(err, "ensemble entry records");
    e->entries_built = true;
    return SFX_OK;
}

// the captured step: batch list H2D, plan, four drains
int capture(sfx_ensemble* e) {
    if(e->exec) return SFX_OK;
    hipError_t err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if(err != hipSuccess) return hip_error(err, "graph capture begin");
    hipError_t e1 = hipMemcpyAsync(e->d_batch, e->h_batch,
                                   sizeof(int32_t) * (1 + ens::kMaxBatch),
                                   hipMemcpyHostToDevice, e->stream);
    hipError_t e2 = e1 == hipSuccess ? launch_batch(e, e->stream) : e1;
    hipGraph_t g = nullptr;
    err = hipStreamEndCapture(e->stream, &g);
    if(e2 != hipSuccess || err != hipSuccess) {
        if(g) hipGraphDestroy(g);
        return hip_error(e2 != hipSuccess ? e2 : err, "graph capture");
    }
    err = hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0);
    if(err != hipSuccess) {
        hipGraphDestroy(g);
        return hip_error(err, "graph instantiate");
    }
    e->graph = g;
    return SFX_OK;
}

int check_batch(const sfx_ensemble* e, const int32_t* ids, int32_t n) {
    if(n < 0 || n > ens::kMaxBatch)
        return set_error(SFX_ERR_ARG, "ensemble step: batch of %d meshes (max %d)", n,
                         ens::kMaxBatch);
    if(n > 0 && !ids) return set_error(SFX_ERR_ARG, "ensemble step: NULL ids");
    int64_t cells = 0, faces = 0;
    std::vector<int32_t> seen(ids, ids + n);
    std::sort(seen.begin(), seen.end());
    for(int32_t k = 0; k < n; k++) {
        if(ids[k] < 0 || ids[k] >= e->nmesh)
            return set_error(SFX_ERR_ARG, "ensemble step: mesh id %d out of range", ids[k]);
        if(k > 0 && seen[k] == seen[k - 1])
            return set_error(SFX_ERR_ARG, "ensemble step: mesh %d twice in one batch", seen[k]);
        cells += e->nc[ids[k]];
        faces += e->nf[ids[k]];
    }
    if(cells > e->arena_cells || faces > e->arena_faces)
        return set_error(SFX_ERR_ARG,
                         "ensemble step: batch needs %lld cells / %lld faces of scratch, the "
                         "arena holds %lld / %lld",
                         (long long)cells, (long long)faces, (long long)e->arena_cells,
                         (long long)e->arena_faces);
    return SFX_OK;
}

int run_batch(sfx_ensemble* e, const int32_t* ids, int32_t n) {
    e->h_batch[0] = n;
    std::memcpy(e->h_batch + 1, ids, sizeof(int32_t) * n);
    int rc = capture(e);
    if(rc) return rc;
    hipError_t err = hipGraphLaunch(e->exec, e->stream);
    if(err != hipSuccess) return hip_error(err, "graph launch");
    return SFX_OK;
}

int run_plain(sfx_ensemble* e, int32_t id) {
    const int64_t cell0 = e->c0[id], ncell = e->nc[id], face0 = e->f0[id], nface = e->nf[id];
    // scratch at the start of the arena
    const int64_t gshift = -cell0, qshift = -face0;
    hipError_t err = e->prec == SFX_PREC_FP64
                             ? ens_launch_range<double>(e->nd, e->h->kind, ens_args<double>(e),
                                                        cell0, ncell, face0, nface, gshift,
                                                        qshift, e->stream)
                             : ens_launch_range<float>(e->nd, e->h->kind, ens_args<float>(e),
                                                       cell0, ncell, face0, nface, gshift,
                                                       qshift, e->stream);
    if(err != hipSuccess) return hip_error(err, "plain step launch");
    return SFX_OK;
}

// SFX_SERVE_QUEUE: post the trace's requests to the pinned mailbox at their
// arrival times and watch their done flags while the persistent kernel
// serves them.
int serve_queue(sfx_ensemble* e, int64_t nreq, const double* arrival_s, const int32_t* mesh_id,
                double* latency_s, int32_t* batch_of, int64_t* nbatch_out, double* wall_out) {
    // mailbox: HostBox, mesh[nreq], done[nreq] (pinned, coherent)
    const size_t off_mesh = 64, off_done = off_mesh + 4 * (size_t)nreq;
    const size_t bytes = off_done + 4 * (size_t)nreq + 64;
    char* hb = nullptr;
    hipError_t err = hipHostMalloc((void**)&hb, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if(err != hipSuccess) return hip_error(err, "serve queue: mailbox");
    std::memset(hb, 0, bytes);
    auto* box = reinterpret_cast<ens::HostBox*>(hb);
    auto* bmesh = reinterpret_cast<int32_t*>(hb + off_mesh);
    auto* bdone = reinterpret_cast<int32_t*>(hb + off_done);
    char* dbox = nullptr;
    ens::DevQueue* dq = nullptr;
    err = hipHostGetDevicePointer((void**)&dbox, hb, 0);
    if(err == hipSuccess) err = hipMalloc(&dq, sizeof(ens::DevQueue));
    if(err == hipSuccess) err = hipMemsetAsync(dq, 0, sizeof(ens::DevQueue), e->stream);
    int blocks = 0;
    if(err == hipSuccess) {
        QueueArgs qa;
        qa.q = dq;
        qa.box = reinterpret_cast<ens::HostBox*>(dbox);
        qa.box_mesh = reinterpret_cast<const int32_t*>(dbox + off_mesh);
        qa.box_done = reinterpret_cast<int32_t*>(dbox + off_done);
        qa.mc0 = e->d_c0;
        qa.mnc = e->d_nc;
        qa.mf0 = e->d_f0;
        qa.mnf = e->d_nf;
        qa.arena_cells = e->arena_cells;
        qa.arena_faces = e->arena_faces;
        qa.nmesh_words = (int)((e->nmesh + 31) / 32);
        err = e->prec == SFX_PREC_FP64
                      ? ens_launch_queue<double>(e->nd, e->h->kind, ens_args<double>(e), qa, e->ncu,
                                                 e->stream, &blocks)
                      : ens_launch_queue<float>(e->nd, e->h->kind, ens_args<float>(e), qa, e->ncu,
                                                e->stream, &blocks);
    }
    if(err != hipSuccess) {
        hipStreamSynchronize(e->stream);
        if(dq) hipFree(dq);
        hipHostFree(hb);
        return hip_error(err, "serve queue: launch");
    }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    int64_t next = 0, first = 0, completed = 0, maxb = 0;
    double last_progress = 0.0;
    int rc = SFX_OK;
    while(completed < nreq) {
        const double tnow = now();
        bool progress = false;
        while(next < nreq && arrival_s[next] <= tnow) {
            bmesh[next] = mesh_id[next];
            __atomic_store_n(&box->tail, (uint64_t)(next + 1), __ATOMIC_RELEASE);
            next++;
            progress = true;
        }
        if(next == nreq) __atomic_store_n(&box->stop, 1, __ATOMIC_RELEASE);
        // done flags of the posted, unfinished requests (they finish in
        // batches, mostly in order)
        for(int64_t r = first; r < next; r++) {
            if(latency_s[r] >= 0.0) continue;
            const int32_t d = __atomic_load_n(&bdone[r], __ATOMIC_ACQUIRE);
            if(d == 0) continue;
            latency_s[r] = now() - arrival_s[r];
            if(batch_of) batch_of[r] = d - 1;
            maxb = std::max<int64_t>(maxb, d);
            completed++;
            progress = true;
        }
        while(first < next && latency_s[first] >= 0.0) first++;
        if(progress) {
            last_progress = now();
        } else if(now() - last_progress > 30.0) {
            rc = set_error(SFX_ERR_HIP, "serve queue: no request completed for 30 s");
            break;
        }
    }
    const double wall = now();
    __atomic_store_n(&box->stop, 1, __ATOMIC_RELEASE);
    err = hipStreamSynchronize(e->stream);
    int32_t qerr = 0;
    uint64_t nb = 0;
    if(err == hipSuccess) err = hipMemcpy(&qerr, &dq->error, sizeof qerr, hipMemcpyDeviceToHost);
    if(err == hipSuccess) err = hipMemcpy(&nb, &dq->nbatch, sizeof nb, hipMemcpyDeviceToHost);
    hipFree(dq);
    hipHostFree(hb);
    if(rc) return rc;
    if(err != hipSuccess) return hip_error(err, "serve queue");
    if(qerr) return set_error(SFX_ERR_HIP, "serve queue: a device wait ran out (%d)", qerr);
    *nbatch_out = (int64_t)nb;
    *wall_out = wall;
    (void)blocks;
    return SFX_OK;
}

double pct(std::vector<double> v, double q) {
    if(v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double r = q * (double)(v.size() - 1);
    size_t lo = (size_t)std::floor(r), hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (r - (double)lo) * (v[hi] - v[lo]);
}

}  // namespace

extern "C" {

int sfx_serve_admit(int64_t npending, const int32_t* pending, const int64_t* mesh_cells,
                    const int64_t* mesh_faces, int64_t arena_cells, int64_t arena_faces,
                    int32_t max_batch, int32_t* admitted, int32_t* nadmitted) {
    if(npending < 0 || max_batch < 0 || (npending > 0 && (!pending || !mesh_cells ||
                                                           !mesh_faces || !admitted)) ||
       !nadmitted)
        return set_error(SFX_ERR_ARG, "sfx_serve_admit: bad arguments");
    int32_t n = 0;
    int64_t cells = 0, faces = 0;
    std::vector<int32_t> taken;
    for(int64_t k = 0; k < npending && n < max_batch; k++) {
        const int32_t id = pending[k];
        if(std::find(taken.begin(), taken.end(), id) != taken.end())
            continue;  // a second step of a mesh waits for the first
        const int64_t c = mesh_cells[id], f = mesh_faces[id];
        if(cells + c > arena_cells || faces + f > arena_faces) break;  // FIFO: no overtaking
        cells += c;
        faces += f;
        taken.push_back(id);
        admitted[n++] = (int32_t)k;
    }
    *nadmitted = n;
    return SFX_OK;
}

int sfx_ensemble_create(sfx_solver* h, int64_t max_meshes, int64_t max_cells, int64_t max_faces,
                        int64_t max_entries, int64_t arena_cells, int64_t arena_faces,
                        sfx_ensemble** out) {
    if(!h || !out) return set_error(SFX_ERR_ARG, "sfx_ensemble_create: NULL");
    *out = nullptr;
    if(max_meshes <= 0 || max_cells <= 0 || max_faces <= 0 || max_entries <= 0 ||
       arena_cells <= 0 || arena_faces <= 0)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_create: sizes must be > 0");
    if(max_cells + 64 * max_meshes > INT32_MAX || max_faces > (INT32_MAX >> 1) ||
       max_entries > INT32_MAX)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_create: too large for 32-bit indices");
    int rc = bind_device(h);
    if(rc) return rc;
    sfx_ensemble* e = new sfx_ensemble();
    e->h = h;
    e->nd = h->ndim;
    e->NC = h->ndim + 2;
    e->prec = h->precision;
    e->tsz = e->prec == SFX_PREC_FP64 ? 8 : 4;
    e->cap_mesh = max_meshes;
    e->cap_cells = max_cells;
    e->cap_faces = max_faces;
    e->cap_ent = max_entries;
    e->cap_gent = max_entries;
    e->arena_cells = arena_cells;
    e->arena_faces = arena_faces;
    e->cap_alloc = max_cells + 64 * max_meshes;
    const int64_t C = e->cap_alloc, F = max_faces, E = max_entries, nd = e->nd, NC = e->NC;
    const size_t t = e->tsz;
    const int ER = flux::entry_rec(e->nd, t == 4);
    hipError_t err = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if(err == hipSuccess) err = dalloc(e, p, bytes);
    };
    A(&e->W, 8 * C * NC);
    A(&e->Q, 8 * C * NC);
    A(&e->dQ, 8 * C * NC);
    A(&e->mass, 8 * C);
    A(&e->total_area, 8 * C);
    A(&e->dt, 8 * C);
    A(&e->volume, 8 * C);
    A(&e->vel, 8 * C * nd);
    A(&e->active, C);
    A(&e->cstatus, C);
    A(&e->cf_off, 8 * (C + 1));
    A(&e->gather_off, 8 * (C + 1));
    A(&e->left, 4 * F);
    A(&e->right, 4 * F);
    A(&e->twin, 4 * F);
    A(&e->ghost, F);
    A(&e->fstatus, F);
    A(&e->area, t * F);
    A(&e->rcent, t * F);
    A(&e->dL, t * F * nd);
    A(&e->dR, t * F * nd);
    A(&e->rRL, t * F * nd);
    A(&e->mc, t * F * nd);
    A(&e->cf_entry, 4 * E);
    A(&e->gather_entry, 4 * E);
    A(&e->egeo, t * E * ER);  // grown by build_entries when the groups' padding needs it
    A(&e->egrp, 8 * ((C + 63) / 64 + 1));
    A(&e->d_c0, 8 * max_meshes);
    A(&e->d_nc, 8 * max_meshes);
    A(&e->d_f0, 8 * max_meshes);
    A(&e->d_nf, 8 * max_meshes);
    A(&e->grad, t * arena_cells * nd * NC);
    A(&e->dQL, t * arena_faces * NC);
    A(&e->d_batch, sizeof(int32_t) * (1 + ens::kMaxBatch));
    A(&e->d_plan, sizeof(ens::Plan));
    e->egeo_bytes = std::max<size_t>(t * E * ER, 256);
    if(err == hipSuccess) err = hipHostMalloc((void**)&e->h_batch,
                                              sizeof(int32_t) * (1 + ens::kMaxBatch));
    if(err == hipSuccess) err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if(err == hipSuccess) err = hipMemsetAsync(e->dQ, 0, 8 * C * NC, e->stream);
    if(err == hipSuccess) err = hipStreamSynchronize(e->stream);
    int dev = 0;
    if(err == hipSuccess) err = hipGetDevice(&dev);
    if(err == hipSuccess)
        err = hipDeviceGetAttribute(&e->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if(err != hipSuccess) {
        ens_free(e);
        return err == hipErrorOutOfMemory
                       ? set_error(SFX_ERR_NOMEM, "sfx_ensemble_create: out of device memory")
                       : hip_error(err, "sfx_ensemble_create");
    }
    *out = e;
    return SFX_OK;
}

int sfx_ensemble_destroy(sfx_ensemble* e) { return ens_free(e); }

int sfx_ensemble_add_mesh(sfx_ensemble* e, const sfx_mesh* m, const double* Q, int64_t E,
                          int64_t G, int32_t* mesh_id) {
    if(!e || !m || !Q || !mesh_id) return set_error(SFX_ERR_ARG, "sfx_ensemble_add_mesh: NULL");
    if(m->ndim != e->nd) return set_error(SFX_ERR_ARG, "sfx_ensemble_add_mesh: ndim");
    const int64_t C = m->C, F = m->F;
    if(C <= 0 || F < 0 || E < 0 || G < 0)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_add_mesh: bad sizes");
    if(!m->W || !m->vel || !m->mass || !m->total_area || !m->dt || !m->volume || !m->left ||
       !m->right || !m->ghost || !m->area || !m->dL || !m->dR || !m->rRL || !m->mc ||
       !m->rcent || !m->cf_off || !m->cf_entry || !m->gather_off || !m->gather_entry)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_add_mesh: missing array");
    if(e->nmesh >= e->cap_mesh || e->ncells + C > e->cap_cells || e->nfaces + F > e->cap_faces ||
       e->nent + E > e->cap_ent || e->ngent + G > e->cap_gent)
        return set_error(SFX_ERR_NOMEM,
                         "sfx_ensemble_add_mesh: the ensemble is full (%lld meshes, %lld cells, "
                         "%lld faces)",
                         (long long)e->nmesh, (long long)e->ncells, (long long)e->nfaces);
    if(C > e->arena_cells || F > e->arena_faces)
        return set_error(SFX_ERR_ARG,
                         "sfx_ensemble_add_mesh: a mesh of %lld cells / %lld faces never fits "
                         "the scratch arena",
                         (long long)C, (long long)F);
    int rc = bind_device(e->h);
    if(rc) return rc;
    const int64_t c0 = (e->cell_end + 63) / 64 * 64, f0 = e->nfaces, e0 = e->nent,
                  g0 = e->ngent;
    const int nd = e->nd, NC = e->NC;
    const size_t t = e->tsz;
    hipStream_t s = e->stream;
    hipError_t err = hipSuccess;
    // the gap cells cell_end .. c0 - 1 have no entries: their offsets repeat
    // the previous mesh's end (read by the host before the stream syncs below)
    const std::vector<int64_t> gap_e(c0 - e->cell_end, e0), gap_g(c0 - e->cell_end, g0);
    if(c0 > e->cell_end) {
        err = hipMemcpyAsync(e->cf_off + e->cell_end + 1, gap_e.data(), 8 * gap_e.size(),
                             hipMemcpyHostToDevice, s);
        if(err == hipSuccess)
            err = hipMemcpyAsync(e->gather_off + e->cell_end + 1, gap_g.data(),
                                 8 * gap_g.size(), hipMemcpyHostToDevice, s);
    }
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        if(err == hipSuccess && bytes)
            err = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
    };
    cp(e->W + c0 * NC, m->W, 8 * C * NC);
    cp(e->Q + c0 * NC, Q, 8 * C * NC);
    cp((char*)e->vel + c0 * nd * 8, m->vel, 8 * C * nd);
    cp(e->mass + c0, m->mass, 8 * C);
    cp(e->total_area + c0, m->total_area, 8 * C);
    cp(e->dt + c0, m->dt, 8 * C);
    cp(e->volume + c0, m->volume, 8 * C);
    if(m->active)
        cp(e->active + c0, m->active, C);
    else if(err == hipSuccess)
        err = hipMemsetAsync(e->active + c0, 1, C, s);
    if(m->closure && e->prec == SFX_PREC_MIXED) {
        if(!e->closure) {
            if(err == hipSuccess) err = dalloc(e, &e->closure, t * e->cap_alloc * nd);
            if(err == hipSuccess) err = hipMemsetAsync(e->closure, 0, t * e->cap_alloc * nd, s);
            drop_graph(e);  // the captured drains were built without it
        }
        for(int l = 0; l < nd; l++)
            cp((char*)e->closure + ((size_t)l * e->cap_alloc + c0) * t,
               (const char*)m->closure + (size_t)l * C * t, t * C);
    }
    cp(e->ghost + f0, m->ghost, F);
    cp((char*)e->area + f0 * t, m->area, t * F);
    cp((char*)e->rcent + f0 * t, m->rcent, t * F);
    for(int l = 0; l < nd; l++) {
        const size_t so = (size_t)l * F * t, dof = ((size_t)l * e->cap_faces + f0) * t;
        cp((char*)e->dL + dof, (const char*)m->dL + so, t * F);
        cp((char*)e->dR + dof, (const char*)m->dR + so, t * F);
        cp((char*)e->rRL + dof, (const char*)m->rRL + so, t * F);
        cp((char*)e->mc + dof, (const char*)m->mc + so, t * F);
    }
    if(err == hipSuccess) err = ens_launch_offset_i32(m->left, e->left + f0, F, c0, s);
    if(err == hipSuccess) err = ens_launch_offset_i32(m->right, e->right + f0, F, c0, s);
    if(err == hipSuccess) {
        if(m->twin)
            err = ens_launch_offset_i32(m->twin, e->twin + f0, F, f0, s);
        else
            err = hipMemsetAsync(e->twin + f0, 0xff, 4 * F, s);  // -1
    }
    if(err == hipSuccess) err = ens_launch_offset_i64(m->cf_off, e->cf_off + c0, C + 1, e0, s);
    if(err == hipSuccess) err = ens_launch_offset_i32(m->cf_entry, e->cf_entry + e0, E, 2 * f0, s);
    if(err == hipSuccess)
        err = ens_launch_offset_i64(m->gather_off, e->gather_off + c0, C + 1, g0, s);
    if(err == hipSuccess)
        err = ens_launch_offset_i32(m->gather_entry, e->gather_entry + g0, G, 2 * f0, s);
    const int64_t row[4] = {c0, C, f0, F};
    int64_t* dst[4] = {e->d_c0, e->d_nc, e->d_f0, e->d_nf};
    for(int k = 0; k < 4 && err == hipSuccess; k++)
        err = hipMemcpyAsync(dst[k] + e->nmesh, &row[k], 8, hipMemcpyHostToDevice, s);
    if(err == hipSuccess) err = hipStreamSynchronize(s);
    if(err != hipSuccess) return hip_error(err, "sfx_ensemble_add_mesh");
    e->c0.push_back(c0);
    e->nc.push_back(C);
    e->f0.push_back(f0);
    e->nf.push_back(F);
    e->ncells += C;
    e->cell_end = c0 + C;
    e->nfaces += F;
    e->nent += E;
    e->ngent += G;
    *mesh_id = (int32_t)e->nmesh++;
    e->entries_built = false;
    drop_graph(e);
    return SFX_OK;
}

int sfx_ensemble_info(const sfx_ensemble* e, int32_t mesh_id, int64_t* cells, int64_t* faces) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_info: NULL");
    if(mesh_id < 0) {
        if(cells) *cells = e->ncells;
        if(faces) *faces = e->nfaces;
        return (int)e->nmesh;
    }
    if(mesh_id >= e->nmesh) return set_error(SFX_ERR_ARG, "sfx_ensemble_info: mesh id");
    if(cells) *cells = e->nc[mesh_id];
    if(faces) *faces = e->nf[mesh_id];
    return SFX_OK;
}

int sfx_ensemble_step(sfx_ensemble* e, const int32_t* ids, int32_t n) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_step: NULL");
    int rc = check_batch(e, ids, n);
    if(rc) return rc;
    if((rc = bind_device(e->h))) return rc;
    if((rc = build_entries(e))) return rc;
    if(n == 0) return SFX_OK;
    // the previous replay must have consumed the pinned batch list
    hipError_t err = hipStreamSynchronize(e->stream);
    if(err != hipSuccess) return hip_error(err, "sfx_ensemble_step");
    return run_batch(e, ids, n);
}

int sfx_ensemble_step_plain(sfx_ensemble* e, int32_t id) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_step_plain: NULL");
    if(id < 0 || id >= e->nmesh) return set_error(SFX_ERR_ARG, "sfx_ensemble_step_plain: id");
    int rc = bind_device(e->h);
    if(rc) return rc;
    if((rc = build_entries(e))) return rc;
    return run_plain(e, id);
}

int sfx_ensemble_sync(sfx_ensemble* e) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_sync: NULL");
    int rc = bind_device(e->h);
    if(rc) return rc;
    hipError_t err = hipStreamSynchronize(e->stream);
    return err == hipSuccess ? SFX_OK : hip_error(err, "sfx_ensemble_sync");
}

int sfx_ensemble_read(sfx_ensemble* e, int32_t id, int what, double* host_out) {
    if(!e || !host_out) return set_error(SFX_ERR_ARG, "sfx_ensemble_read: NULL");
    if(id < 0 || id >= e->nmesh) return set_error(SFX_ERR_ARG, "sfx_ensemble_read: id");
    const double* src = what == SFX_ENS_W ? e->W : what == SFX_ENS_Q ? e->Q : what == SFX_ENS_DQ
                                                                                  ? e->dQ
                                                                                  : nullptr;
    if(!src) return set_error(SFX_ERR_ARG, "sfx_ensemble_read: what");
    int rc = bind_device(e->h);
    if(rc) return rc;
    hipError_t err = hipStreamSynchronize(e->stream);
    if(err == hipSuccess)
        err = hipMemcpy(host_out, src + e->c0[id] * e->NC, 8 * e->nc[id] * e->NC,
                        hipMemcpyDeviceToHost);
    return err == hipSuccess ? SFX_OK : hip_error(err, "sfx_ensemble_read");
}

int sfx_ensemble_set_launch(sfx_ensemble* e, int block, int blocks_per_cu) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_set_launch: NULL");
    if(block < 64 || block > 256 || (block & 63) || blocks_per_cu < 0 || blocks_per_cu > 32)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_set_launch: block %d, blocks/CU %d", block,
                         blocks_per_cu);
    if(block != e->block || blocks_per_cu != e->per_cu) {
        int rc = bind_device(e->h);
        if(rc) return rc;
        hipStreamSynchronize(e->stream);
        drop_graph(e);
        e->block = block;
        e->per_cu = blocks_per_cu;
    }
    return SFX_OK;
}

int sfx_ensemble_get_launch(const sfx_ensemble* e, int* block, int* blocks_per_cu) {
    if(!e) return set_error(SFX_ERR_ARG, "sfx_ensemble_get_launch: NULL");
    if(block) *block = e->block;
    if(blocks_per_cu) *blocks_per_cu = e->per_cu;
    return SFX_OK;
}

int sfx_ensemble_autotune(sfx_ensemble* e, const int32_t* ids, int32_t n, int reps,
                          double* best_ms) {
    if(!e || reps <= 0) return set_error(SFX_ERR_ARG, "sfx_ensemble_autotune: bad arguments");
    int rc = check_batch(e, ids, n);
    if(rc) return rc;
    if(n == 0) return set_error(SFX_ERR_ARG, "sfx_ensemble_autotune: empty batch");
    if((rc = bind_device(e->h))) return rc;
    if((rc = build_entries(e))) return rc;
    // the trial steps advance the batch's meshes: keep W, Q and the masses
    // (which the update refreshes) to restore
    int64_t cells = 0;
    for(int32_t k = 0; k < n; k++) cells += e->nc[ids[k]];
    hipError_t err = hipStreamSynchronize(e->stream);
    double* save = nullptr;
    if(err == hipSuccess) err = hipMalloc(&save, 8 * (size_t)cells * (2 * e->NC + 1));
    if(err != hipSuccess) return hip_error(err, "sfx_ensemble_autotune");
    auto copy_state = [&](bool back) {
        int64_t o = 0, om = 0;
        for(int32_t k = 0; k < n; k++) {
            const size_t bytes = 8 * (size_t)e->nc[ids[k]] * e->NC;
            double* w = e->W + e->c0[ids[k]] * e->NC;
            double* q = e->Q + e->c0[ids[k]] * e->NC;
            double* m = e->mass + e->c0[ids[k]];
            double* sw = save + o;
            double* sq = save + (size_t)cells * e->NC + o;
            double* sm = save + 2 * (size_t)cells * e->NC + om;
            if(back) {
                hipMemcpyAsync(w, sw, bytes, hipMemcpyDeviceToDevice, e->stream);
                hipMemcpyAsync(q, sq, bytes, hipMemcpyDeviceToDevice, e->stream);
                hipMemcpyAsync(m, sm, 8 * e->nc[ids[k]], hipMemcpyDeviceToDevice, e->stream);
            } else {
                hipMemcpyAsync(sw, w, bytes, hipMemcpyDeviceToDevice, e->stream);
                hipMemcpyAsync(sq, q, bytes, hipMemcpyDeviceToDevice, e->stream);
                hipMemcpyAsync(sm, m, 8 * e->nc[ids[k]], hipMemcpyDeviceToDevice, e->stream);
            }
            o += (int64_t)e->nc[ids[k]] * e->NC;
            om += e->nc[ids[k]];
        }
    };
    copy_state(false);
    hipEvent_t t0, t1;
    hipEventCreate(&t0);
    hipEventCreate(&t1);
    const int blocks[3] = {64, 128, 256};
    const int per_cu[5] = {0, 1, 2, 4, 8};  // 0: one block per tile
    double best = 1e300;
    int bb = e->block, bp = e->per_cu;
    e->h_batch[0] = n;
    std::memcpy(e->h_batch + 1, ids, sizeof(int32_t) * n);
    hipMemcpyAsync(e->d_batch, e->h_batch, sizeof(int32_t) * (1 + n), hipMemcpyHostToDevice,
                   e->stream);
    for(int b : blocks)
        for(int p : per_cu) {
            e->block = b;
            e->per_cu = p;
            launch_batch(e, e->stream);  // warm
            hipEventRecord(t0, e->stream);
            for(int r = 0; r < reps; r++) launch_batch(e, e->stream);
            hipEventRecord(t1, e->stream);
            err = hipEventSynchronize(t1);
            float ms = 0.f;
            if(err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
            if(err != hipSuccess) break;
            if(ms / reps < best) {
                best = ms / reps;
                bb = b;
                bp = p;
            }
        }
    copy_state(true);
    hipStreamSynchronize(e->stream);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    hipFree(save);
    if(err != hipSuccess) return hip_error(err, "sfx_ensemble_autotune");
    e->block = bb;
    e->per_cu = bp;
    drop_graph(e);
    if(best_ms) *best_ms = best;
    return SFX_OK;
}

int sfx_ensemble_serve(sfx_ensemble* e, int mode, int64_t nreq, const double* arrival_s,
                       const int32_t* mesh_id, double* latency_s, int32_t* batch_of,
                       sfx_serve_stats* st) {
    if(!e || nreq < 0 || (nreq > 0 && (!arrival_s || !mesh_id || !latency_s)))
        return set_error(SFX_ERR_ARG, "sfx_ensemble_serve: bad arguments");
    if(mode != SFX_SERVE_BATCHED && mode != SFX_SERVE_PLAIN && mode != SFX_SERVE_QUEUE)
        return set_error(SFX_ERR_ARG, "sfx_ensemble_serve: mode");
    for(int64_t r = 0; r < nreq; r++) {
        if(mesh_id[r] < 0 || mesh_id[r] >= e->nmesh)
            return set_error(SFX_ERR_ARG, "sfx_ensemble_serve: request %lld names mesh %d",
                             (long long)r, mesh_id[r]);
        if(r > 0 && arrival_s[r] < arrival_s[r - 1])
            return set_error(SFX_ERR_ARG, "sfx_ensemble_serve: arrivals must be sorted");
    }
    int rc = bind_device(e->h);
    if(rc) return rc;
    if((rc = build_entries(e))) return rc;
    if(mode == SFX_SERVE_BATCHED && (rc = capture(e))) return rc;
    hipError_t err = hipStreamSynchronize(e->stream);
    if(err != hipSuccess) return hip_error(err, "sfx_ensemble_serve");
    if(mode == SFX_SERVE_QUEUE) {
        std::vector<int32_t> bat(nreq > 0 ? nreq : 1, -1);
        for(int64_t r = 0; r < nreq; r++) latency_s[r] = -1.0;
        int64_t nbatch = 0;
        double wall = 0.0;
        if(nreq > 0 && (rc = serve_queue(e, nreq, arrival_s, mesh_id, latency_s, bat.data(),
                                         &nbatch, &wall)))
            return rc;
        if(batch_of) std::copy(bat.begin(), bat.begin() + nreq, batch_of);
        if(st) {
            std::vector<int64_t> per(nbatch > 0 ? nbatch : 1, 0);
            for(int64_t r = 0; r < nreq; r++)
                if(bat[r] >= 0 && bat[r] < nbatch) per[bat[r]]++;
            std::vector<double> v(latency_s, latency_s + nreq);
            double sum = 0.0;
            for(double x : v) sum += x;
            st->nreq = nreq;
            st->nbatch = nbatch;
            st->wall_s = wall;
            st->busy_s = wall;
            st->p50_s = pct(v, 0.50);
            st->p99_s = pct(v, 0.99);
            st->max_s = v.empty() ? 0.0 : *std::max_element(v.begin(), v.end());
            st->mean_s = v.empty() ? 0.0 : sum / (double)v.size();
            st->mean_batch = nbatch ? (double)nreq / (double)nbatch : 0.0;
            st->max_batch = *std::max_element(per.begin(), per.end());
        }
        return SFX_OK;
    }
    hipEvent_t done;
    if((err = hipEventCreateWithFlags(&done, hipEventDisableTiming)) != hipSuccess)
        return hip_error(err, "sfx_ensemble_serve");
    const int32_t max_batch = mode == SFX_SERVE_BATCHED ? ens::kMaxBatch : 1;
    std::vector<int32_t> pending;     // request indices, FIFO
    std::vector<int32_t> pend_mesh;   // their meshes
    std::vector<int32_t> inflight, adm(ens::kMaxBatch), ids(ens::kMaxBatch);
    std::vector<double> lat;
    lat.reserve(nreq);
    int64_t next = 0, completed = 0, nbatch = 0, maxb = 0, sumb = 0;
    double busy = 0.0, launched_at = 0.0;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    while(completed < nreq) {
        const double tnow = now();
        while(next < nreq && arrival_s[next] <= tnow) {
            pending.push_back((int32_t)next);
            pend_mesh.push_back(mesh_id[next]);
            next++;
        }
        if(!inflight.empty()) {
            err = hipEventQuery(done);
            if(err == hipErrorNotReady) continue;
            if(err != hipSuccess) break;
            const double tdone = now();
            busy += tdone - launched_at;
            for(int32_t r : inflight) latency_s[r] = tdone - arrival_s[r];
            completed += (int64_t)inflight.size();
            inflight.clear();
            continue;
        }
        if(pending.empty()) {
            if(next < nreq) {
                const double wait = arrival_s[next] - now();
                if(wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
            }
            continue;
        }
        int32_t na = 0;
        if((rc = sfx_serve_admit((int64_t)pending.size(), pend_mesh.data(), e->nc.data(),
                                 e->nf.data(), e->arena_cells, e->arena_faces, max_batch,
                                 adm.data(), &na)))
            break;
        if(na == 0) {
            rc = set_error(SFX_ERR_ARG, "sfx_ensemble_serve: a request does not fit the arena");
            break;
        }
        for(int32_t k = 0; k < na; k++) {
            ids[k] = pend_mesh[adm[k]];
            inflight.push_back(pending[adm[k]]);
            if(batch_of) batch_of[pending[adm[k]]] = (int32_t)nbatch;
        }
        // drop the admitted ones from the queue, keeping the order of the rest
        std::vector<uint8_t> gone(pending.size(), 0);
        for(int32_t k = 0; k < na; k++) gone[adm[k]] = 1;
        size_t w = 0;
        for(size_t k = 0; k < pending.size(); k++)
            if(!gone[k]) {
                pending[w] = pending[k];
                pend_mesh[w] = pend_mesh[k];
                w++;
            }
        pending.resize(w);
        pend_mesh.resize(w);
        launched_at = now();
        if(mode == SFX_SERVE_BATCHED) {
            rc = run_batch(e, ids.data(), na);
        } else {
            rc = run_plain(e, ids[0]);
        }
        if(rc) break;
        if((err = hipEventRecord(done, e->stream)) != hipSuccess) break;
        nbatch++;
        sumb += na;
        maxb = std::max<int64_t>(maxb, na);
    }
    const double wall = now();
    hipStreamSynchronize(e->stream);
    hipEventDestroy(done);
    if(rc) return rc;
    if(err != hipSuccess && err != hipErrorNotReady) return hip_error(err, "sfx_ensemble_serve");
    if(st) {
        std::vector<double> v(latency_s, latency_s + nreq);
        st->nreq = nreq;
        st->nbatch = nbatch;
        st->wall_s = wall;
        st->busy_s = busy;
        st->p50_s = pct(v, 0.50);
        st->p99_s = pct(v, 0.99);
        st->max_s = v.empty() ? 0.0 : *std::max_element(v.begin(), v.end());
        double sum = 0.0;
        for(double x : v) sum += x;
        st->mean_s = v.empty() ? 0.0 : sum / (double)v.size();
        st->mean_batch = nbatch ? (double)sumb / (double)nbatch : 0.0;
        st->max_batch = maxb;
    }
    return SFX_OK;
}

}  // extern "C"
