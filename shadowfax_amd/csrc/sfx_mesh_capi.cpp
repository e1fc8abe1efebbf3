// sfx_mesh_capi.cpp -- C ABI for meshes (host tessellation, per-face
// geometry, gather index) and for the face-flux pass kernels.
//
// Host arithmetic here must reproduce the reference's FP64 expressions bit
// for bit; this file is compiled with -ffp-contract=off (no FMA), which the
// x86-64 baseline has no instruction for anyway.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "flux_pass_impl.h"
#include "sfx_capi_internal.h"
#include "sfx_launch.h"
#include "sfx_mesh_host.h"
#include "shadowfax_amd.h"

using namespace sfx;

namespace {

template <typename V>
typename V::value_type* dup(const V& v) {
    using T = typename V::value_type;
    T* p = (T*)std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if(p && !v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

template <typename T>
flux::MeshView<T> view(const sfx_mesh* m) {
    flux::MeshView<T> v;
    v.C = m->C;
    v.F = m->F;
    v.W = m->W;
    v.vel = (const double*)m->vel;
    v.mass = m->mass;
    v.total_area = m->total_area;
    v.dt = m->dt;
    v.volume = m->volume;
    v.active = m->active;
    v.left = m->left;
    v.right = m->right;
    v.ghost = m->ghost;
    v.twin = m->twin;
    v.area = (const T*)m->area;
    v.dL = (const T*)m->dL;
    v.dR = (const T*)m->dR;
    v.rRL = (const T*)m->rRL;
    v.mc = (const T*)m->mc;
    v.rcent = (const T*)m->rcent;
    v.cf_off = m->cf_off;
    v.cf_entry = m->cf_entry;
    v.egeo = (const T*)m->egeo;
    v.egrp = m->egeo_group;
    v.eslot = m->egeo_group ? m->egeo_slot : nullptr;
    if(m->egeo_group && m->egeo_rows > 0)
        v.egeo_bytes = (uint64_t)m->egeo_rows * 64 * flux::entry_rec(m->ndim, sizeof(T) == 4) *
                       sizeof(T);
    v.gather_off = m->gather_off;
    v.gather_entry = m->gather_entry;
    v.closure = (const T*)m->closure;
    if constexpr(sizeof(T) == 4) v.frec = (const float4*)m->frec;
    return v;
}

int check_mesh(const sfx_solver* h, const sfx_mesh* m, const char* who, bool faces,
               bool gradients, bool gather) {
    if(!h || !m) return set_error(SFX_ERR_ARG, "%s: NULL handle or mesh", who);
    if(m->ndim != h->ndim)
        return set_error(SFX_ERR_ARG, "%s: mesh ndim %d != solver ndim %d", who, m->ndim,
                         h->ndim);
    if(m->C < 0 || m->F < 0) return set_error(SFX_ERR_ARG, "%s: negative size", who);
    if(m->C > INT32_MAX || m->F > (INT32_MAX >> 1))
        return set_error(SFX_ERR_ARG, "%s: mesh too large for 32-bit indices", who);
    if(m->W && ((uintptr_t)m->W & 15))
        return set_error(SFX_ERR_ARG, "%s: W must be 16-byte aligned", who);
    if(m->egeo && ((uintptr_t)m->egeo & 15))
        return set_error(SFX_ERR_ARG, "%s: egeo must be 16-byte aligned", who);
    if(m->frec && ((uintptr_t)m->frec & 15))
        return set_error(SFX_ERR_ARG, "%s: frec must be 16-byte aligned", who);
    if(m->egeo_group && !m->cf_off)
        return set_error(SFX_ERR_ARG, "%s: egeo_group needs cf_off", who);
    if(m->egeo_rows < 0) return set_error(SFX_ERR_ARG, "%s: egeo_rows < 0", who);
    if(gradients && (!m->W || !m->volume || !m->cf_off || !m->egeo))
        return set_error(SFX_ERR_ARG, "%s: missing array for the gradient pass (W, volume, "
                         "cf_off, egeo)", who);
    if(faces && (!m->W || !m->vel || !m->mass || !m->total_area || !m->dt || !m->left ||
                 !m->right || !m->ghost || !m->area || !m->dL || !m->dR || !m->rRL ||
                 !m->mc || !m->rcent))
        return set_error(SFX_ERR_ARG, "%s: missing array for the face pass", who);
    if(gather && (!m->gather_off || !m->gather_entry))
        return set_error(SFX_ERR_ARG, "%s: missing gather index", who);
    if(gather && h->precision == SFX_PREC_MIXED && m->twin && !m->ghost)
        return set_error(SFX_ERR_ARG, "%s: twin given without ghost flags", who);
    return SFX_OK;
}

}  // namespace

extern "C" {

int sfx_voronoi_periodic(int ndim, int64_t C, const double* pos, const double* box,
                         int nthreads, sfx_host_mesh* out) {
    if(!pos || !box || !out) return set_error(SFX_ERR_ARG, "sfx_voronoi_periodic: NULL");
    if(C <= 0) return set_error(SFX_ERR_ARG, "sfx_voronoi_periodic: C must be > 0");
    std::memset(out, 0, sizeof *out);
    HostMesh m;
    int rc;
    if(ndim == 2)
        rc = voronoi2d_periodic(C, pos, box[0], box[1], m, nthreads);
    else if(ndim == 3)
        rc = voronoi3d_periodic(C, pos, box[0], box[1], box[2], m, nthreads);
    else
        return set_error(SFX_ERR_ARG, "sfx_voronoi_periodic: ndim must be 2 or 3");
    if(rc != 0)
        return set_error(SFX_ERR_ARG, "sfx_voronoi_periodic: tessellation failed (%d)", rc);
    out->ndim = ndim;
    out->C = m.C;
    out->F = m.F;
    out->pos = dup(m.pos);
    out->volume = dup(m.volume);
    out->centroid = dup(m.centroid);
    out->left = dup(m.left);
    out->right = dup(m.right);
    out->area = dup(m.area);
    out->midpoint = dup(m.midpoint);
    out->shift = dup(m.shift);
    out->ghost = dup(m.ghost);
    out->twin = dup(m.twin);
    out->cf_off = dup(m.cf_off);
    out->cf_face = dup(m.cf_face);
    out->fv_off = dup(m.fv_off);
    out->fverts = dup(m.fverts);
    if(!out->pos || !out->left || !out->cf_face || !out->fv_off || !out->fverts) {
        sfx_host_mesh_free(out);
        return set_error(SFX_ERR_NOMEM, "sfx_voronoi_periodic: out of host memory");
    }
    return SFX_OK;
}

void sfx_host_mesh_free(sfx_host_mesh* m) {
    if(!m) return;
    void* ps[] = {m->pos,      m->volume, m->centroid, m->left,  m->right,
                  m->area,     m->midpoint, m->shift,  m->ghost, m->twin,
                  m->cf_off,   m->cf_face, m->fv_off, m->fverts};
    for(void* p : ps) std::free(p);
    std::memset(m, 0, sizeof *m);
}

int sfx_face_geometry(int ndim, int64_t C, int64_t F, const double* pos, const double* centroid,
                      const int32_t* left, const int32_t* right, const double* midpoint,
                      const double* shift, double* dL, double* dR, double* rRL, double* mc,
                      double* rcent) {
    if(ndim != 2 && ndim != 3) return set_error(SFX_ERR_ARG, "sfx_face_geometry: ndim");
    if(!pos || !centroid || !left || !right || !midpoint || !shift || !dL || !dR || !rRL ||
       !mc || !rcent)
        return set_error(SFX_ERR_ARG, "sfx_face_geometry: NULL array");
    for(int64_t f = 0; f < F; f++) {
        const int64_t L = left[f], R = right[f];
        if(L < 0 || L >= C || R >= C)
            return set_error(SFX_ERR_ARG, "sfx_face_geometry: face %lld index out of range",
                             (long long)f);
        const int64_t base = (R >= 0) ? R : L;
        double xl[3], xr[3];
        for(int d = 0; d < ndim; d++) {
            xl[d] = pos[d * C + L];
            xr[d] = pos[d * C + base] + shift[d * F + f];
        }
        // VorFace::calculate_flux, src/VorFace.cpp:731-734 (right) and
        // :777-780 (left): d = midpoint - centroid + (cell pos - generator pos)
        for(int n = ndim; n--;) {
            dL[n * F + f] = midpoint[n * F + f] - centroid[n * C + L] + (xl[n] - xl[n]);
            if(R >= 0)
                dR[n * F + f] = midpoint[n * F + f] - centroid[n * C + R] + (pos[n * C + R] - xr[n]);
            else
                dR[n * F + f] = 0.0;
            rRL[n * F + f] = xr[n] - xl[n];
            mc[n * F + f] = midpoint[n * F + f] - 0.5 * (xl[n] + xr[n]);
        }
        if(R >= 0) {
            double cd[3];
            for(int d = 0; d < ndim; d++) cd[d] = centroid[d * C + L] - centroid[d * C + R];
            rcent[f] = (ndim == 3) ? std::sqrt(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2])
                                   : std::sqrt(cd[0] * cd[0] + cd[1] * cd[1]);
        } else {
            rcent[f] = 0.0;
        }
    }
    return SFX_OK;
}

int sfx_cell_closure(int ndim, int64_t C, int64_t F, const int64_t* cf_off,
                     const int32_t* cf_entry, const double* area, const double* rRL,
                     double* closure) {
    if(ndim != 2 && ndim != 3) return set_error(SFX_ERR_ARG, "sfx_cell_closure: ndim");
    if(C < 0 || F < 0) return set_error(SFX_ERR_ARG, "sfx_cell_closure: negative size");
    if(!cf_off || !cf_entry || !area || !rRL || !closure)
        return set_error(SFX_ERR_ARG, "sfx_cell_closure: NULL array");
    for(int64_t i = 0; i < C; i++) {
        double S[3] = {0.0, 0.0, 0.0};
        if(cf_off[i] > cf_off[i + 1])
            return set_error(SFX_ERR_ARG, "sfx_cell_closure: cf_off not ascending");
        // the faces in reverse walk order, as the gradient loop visits them
        for(int64_t j = cf_off[i + 1]; j-- > cf_off[i];) {
            const int32_t e = cf_entry[j];
            const int64_t f = e >> 1;
            if(e < 0 || f >= F) return set_error(SFX_ERR_ARG, "sfx_cell_closure: entry");
            const double A = area[f];
            if(!(A != 0.0)) continue;
            double r[3];
            for(int d = 0; d < ndim; d++) r[d] = (e & 1) ? rRL[d * F + f] : -rRL[d * F + f];
            const double rn = (ndim == 3) ? std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2])
                                          : std::sqrt(r[0] * r[0] + r[1] * r[1]);
            for(int d = 0; d < ndim; d++) S[d] += A * r[d] / rn;
        }
        for(int d = 0; d < ndim; d++) closure[d * C + i] = S[d];
    }
    return SFX_OK;
}

int sfx_mesh_cell_entries(int64_t C, int64_t F, const int32_t* left, const int64_t* cf_off,
                          const int32_t* cf_face, int32_t* cf_entry) {
    if(!left || !cf_off || !cf_face || !cf_entry)
        return set_error(SFX_ERR_ARG, "sfx_mesh_cell_entries: NULL array");
    if(F > (INT32_MAX >> 1)) return set_error(SFX_ERR_ARG, "sfx_mesh_cell_entries: F too large");
    for(int64_t i = 0; i < C; i++)
        for(int64_t j = cf_off[i]; j < cf_off[i + 1]; j++) {
            const int32_t f = cf_face[j];
            if(f < 0 || f >= F) return set_error(SFX_ERR_ARG, "sfx_mesh_cell_entries: face");
            cf_entry[j] = 2 * f + (left[f] == i ? 0 : 1);
        }
    return SFX_OK;
}

int sfx_mesh_gather_index(int64_t C, int64_t F, const int32_t* left, const int32_t* right,
                          const uint8_t* ghost, int64_t* gather_off, int32_t* gather_entry) {
    if(!left || !right || !gather_off)
        return set_error(SFX_ERR_ARG, "sfx_mesh_gather_index: NULL array");
    std::vector<int64_t> cnt(C + 1, 0);
    for(int64_t f = 0; f < F; f++) {
        if(left[f] < 0 || left[f] >= C || right[f] >= C)
            return set_error(SFX_ERR_ARG, "sfx_mesh_gather_index: face %lld index out of range",
                             (long long)f);
        cnt[left[f] + 1]++;
        if(right[f] >= 0 && !(ghost && ghost[f])) cnt[right[f] + 1]++;
    }
    for(int64_t i = 0; i < C; i++) cnt[i + 1] += cnt[i];
    std::memcpy(gather_off, cnt.data(), sizeof(int64_t) * (C + 1));
    if(!gather_entry) return SFX_OK;
    // faces in ascending order: every cell's list comes out sorted
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for(int64_t f = 0; f < F; f++) {
        gather_entry[fill[left[f]]++] = (int32_t)(2 * f);
        if(right[f] >= 0 && !(ghost && ghost[f])) gather_entry[fill[right[f]]++] = (int32_t)(2 * f + 1);
    }
    return SFX_OK;
}

int64_t sfx_mesh_entry_groups(int64_t C, const int64_t* cf_off, int64_t* egeo_group) {
    if(C < 0 || !cf_off || !egeo_group) {
        set_error(SFX_ERR_ARG, "sfx_mesh_entry_groups: bad arguments");
        return -1;
    }
    const int64_t ng = (C + 63) / 64;
    int64_t rows = 0;
    for(int64_t g = 0; g < ng; g++) {
        egeo_group[g] = rows;
        int64_t k = 0;
        for(int64_t i = 64 * g; i < std::min(C, 64 * g + 64); i++)
            k = std::max(k, cf_off[i + 1] - cf_off[i]);
        rows += k;
    }
    egeo_group[ng] = rows;
    return rows;
}

int64_t sfx_mesh_entry_groups_sorted(int64_t C, const int64_t* cf_off, int64_t window,
                                     int64_t* egeo_slot, int64_t* egeo_group) {
    if(C < 0 || C > INT32_MAX || !cf_off || !egeo_slot || !egeo_group || window <= 0 ||
       window % 64) {
        set_error(SFX_ERR_ARG, "sfx_mesh_entry_groups_sorted: bad arguments");
        return -1;
    }
    std::vector<int64_t> cells;
    for(int64_t w0 = 0; w0 < C; w0 += window) {
        const int64_t w1 = std::min(C, w0 + window);
        cells.resize(w1 - w0);
        for(int64_t i = w0; i < w1; i++) {
            const int64_t n = cf_off[i + 1] - cf_off[i];
            if(n < 0 || n > INT32_MAX) {
                set_error(SFX_ERR_ARG, "sfx_mesh_entry_groups_sorted: cf_off not ascending");
                return -1;
            }
            cells[i - w0] = i;
        }
        std::stable_sort(cells.begin(), cells.end(), [&](int64_t a, int64_t b) {
            return cf_off[a + 1] - cf_off[a] > cf_off[b + 1] - cf_off[b];
        });
        for(int64_t k = 0; k < w1 - w0; k++) {
            const int64_t c = cells[k];
            egeo_slot[w0 + k] = c | ((cf_off[c + 1] - cf_off[c]) << 32);
        }
    }
    const int64_t ng = (C + 63) / 64;
    int64_t rows = 0;
    for(int64_t g = 0; g < ng; g++) {
        egeo_group[g] = rows;
        int64_t k = 0;
        for(int64_t s = 64 * g; s < std::min(C, 64 * g + 64); s++)
            k = std::max(k, egeo_slot[s] >> 32);
        rows += k;
    }
    egeo_group[ng] = rows;
    return rows;
}

int sfx_mesh_entry_record(int ndim, int precision) {
    if((ndim != 2 && ndim != 3) || (precision != SFX_PREC_FP64 && precision != SFX_PREC_MIXED))
        return -1;
    return sfx::flux::entry_rec(ndim, precision == SFX_PREC_MIXED);
}

int sfx_mesh_entry_geometry(sfx_solver* h, const sfx_mesh* m, int64_t E, void* egeo,
                            void* stream) {
    if(!h || !m) return set_error(SFX_ERR_ARG, "sfx_mesh_entry_geometry: NULL handle or mesh");
    int rc = check_mesh(h, m, "sfx_mesh_entry_geometry", false, false, false);
    if(rc) return rc;
    if(E < 0 || E > INT32_MAX)
        return set_error(SFX_ERR_ARG, "sfx_mesh_entry_geometry: E = %lld", (long long)E);
    if(E == 0) return SFX_OK;
    if(!m->cf_entry || !m->area || !m->mc || !m->rRL || !m->dL || !m->dR || !m->left ||
       !m->right || !egeo)
        return set_error(SFX_ERR_ARG, "sfx_mesh_entry_geometry: NULL array");
    if((rc = bind_device(h))) return rc;
    hipError_t e = (h->precision == SFX_PREC_FP64)
                           ? launch_entry_geometry<double>(h->ndim, view<double>(m), E,
                                                           (double*)egeo, (hipStream_t)stream)
                           : launch_entry_geometry<float>(h->ndim, view<float>(m), E,
                                                          (float*)egeo, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "entry geometry kernel launch");
    return SFX_OK;
}

int sfx_mesh_face_record_planes(int ndim, int precision) {
    if((ndim != 2 && ndim != 3) || precision != SFX_PREC_MIXED) return -1;
    return ndim == 3 ? sfx::flux::face_record_planes<3>() : sfx::flux::face_record_planes<2>();
}

int sfx_mesh_face_records(sfx_solver* h, const sfx_mesh* m, void* frec, void* stream) {
    if(!h || !m) return set_error(SFX_ERR_ARG, "sfx_mesh_face_records: NULL handle or mesh");
    if(h->precision != SFX_PREC_MIXED)
        return set_error(SFX_ERR_ARG, "sfx_mesh_face_records: mixed-precision handles only");
    int rc = check_mesh(h, m, "sfx_mesh_face_records", false, false, false);
    if(rc) return rc;
    if(m->F == 0) return SFX_OK;
    if(!m->left || !m->right || !m->area || !m->rcent || !m->rRL || !m->mc || !m->dL ||
       !m->dR || !frec)
        return set_error(SFX_ERR_ARG, "sfx_mesh_face_records: NULL array");
    if((uintptr_t)frec & 15)
        return set_error(SFX_ERR_ARG, "sfx_mesh_face_records: frec must be 16-byte aligned");
    if((rc = bind_device(h))) return rc;
    hipError_t e = launch_face_records(h->ndim, view<float>(m), (float*)frec, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "face record kernel launch");
    return SFX_OK;
}

int sfx_gradients(sfx_solver* h, const sfx_mesh* m, void* grad, void* stream) {
    int rc = check_mesh(h, m, "sfx_gradients", false, true, false);
    if(rc) return rc;
    if(!grad) return set_error(SFX_ERR_ARG, "sfx_gradients: grad is NULL");
    if((rc = bind_device(h))) return rc;
    hipError_t e = (h->precision == SFX_PREC_FP64)
                           ? launch_gradients<double>(h->ndim, view<double>(m), (double*)grad,
                                                      (hipStream_t)stream)
                           : launch_gradients<float>(h->ndim, view<float>(m), (float*)grad,
                                                     (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "gradient kernel launch");
    return SFX_OK;
}

int sfx_face_fluxes(sfx_solver* h, const sfx_mesh* m, const void* grad, void* dQL,
                    uint8_t* face_status, void* stream) {
    int rc = check_mesh(h, m, "sfx_face_fluxes", true, false, false);
    if(rc) return rc;
    if(!grad || !dQL) return set_error(SFX_ERR_ARG, "sfx_face_fluxes: NULL grad or dQL");
    if((rc = bind_device(h))) return rc;
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    hipError_t e = (h->precision == SFX_PREC_FP64)
                           ? launch_face_fluxes<double>(h->ndim, h->kind, view<double>(m),
                                                        (const double*)grad, h->c,
                                                        (double*)dQL,
                                                        face_status, ctr, (hipStream_t)stream)
                           : launch_face_fluxes<float>(h->ndim, h->kind, view<float>(m),
                                                       (const float*)grad, h->c,
                                                       (float*)dQL,
                                                       face_status, ctr, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "face flux kernel launch");
    return SFX_OK;
}

int sfx_gather_dQ(sfx_solver* h, const sfx_mesh* m, const void* dQL, double* dQ,
                  void* stream) {
    int rc = check_mesh(h, m, "sfx_gather_dQ", false, false, true);
    if(rc) return rc;
    if(!dQL || !dQ) return set_error(SFX_ERR_ARG, "sfx_gather_dQ: NULL dQL or dQ");
    if((rc = bind_device(h))) return rc;
    hipError_t e =
            (h->precision == SFX_PREC_FP64)
                    ? launch_gather_fp64(h->ndim, m->C, m->F, m->gather_off, m->gather_entry,
                                         (const double*)dQL, dQ, (hipStream_t)stream)
                    : launch_gather_mixed(h->ndim, m->C, m->F, m->gather_off, m->gather_entry,
                                          m->ghost, m->twin, (const float*)dQL, dQ,
                                          (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "gather kernel launch");
    return SFX_OK;
}

int sfx_face_flux_pass(sfx_solver* h, const sfx_mesh* m, void* grad, void* dQL, double* dQ,
                       uint8_t* face_status, void* stream) {
    int rc = check_mesh(h, m, "sfx_face_flux_pass", true, true, true);
    if(rc) return rc;
    if((rc = sfx_gradients(h, m, grad, stream))) return rc;
    if((rc = sfx_face_fluxes(h, m, grad, dQL, face_status, stream))) return rc;
    return sfx_gather_dQ(h, m, dQL, dQ, stream);
}

int sfx_update_step_timed(sfx_solver* h, int64_t C, double* Q, double* dQ,
                          const double* volume, double* W, double* mass,
                          const uint64_t* endtime, uint64_t current_time,
                          uint8_t* cell_status, void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_update_step: NULL handle");
    if(C < 0) return set_error(SFX_ERR_ARG, "sfx_update_step: C < 0");
    if(C == 0) return SFX_OK;
    if(!Q || !dQ || !volume || !W) return set_error(SFX_ERR_ARG, "sfx_update_step: NULL array");
    int rc = bind_device(h);
    if(rc) return rc;
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    hipError_t e = launch_update(h->ndim, h->c, C, Q, dQ, volume, W, endtime, current_time,
                                 cell_status, mass, ctr, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "update kernel launch");
    return SFX_OK;
}

int sfx_update_step(sfx_solver* h, int64_t C, double* Q, double* dQ, const double* volume,
                    double* W, double* mass, uint8_t* cell_status, void* stream) {
    return sfx_update_step_timed(h, C, Q, dQ, volume, W, mass, nullptr, 0, cell_status, stream);
}

}  // extern "C"

// ---- generator order -------------------------------------------------

namespace {

// Hilbert index of integer coordinates X[0..n) with b bits each: Skilling's
// transpose method (J. Skilling, "Programming the Hilbert curve", AIP Conf.
// Proc. 707, 381 (2004)): undo the excess work of the curve's inverse, Gray
// encode, then interleave the transposed bits, most significant first.
uint64_t hilbert_index(uint32_t* X, int b, int n) {
    const uint32_t M = 1u << (b - 1);
    for(uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for(int i = 0; i < n; i++) {
            if(X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for(int i = 1; i < n; i++) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for(uint32_t Q = M; Q > 1; Q >>= 1)
        if(X[n - 1] & Q) t ^= Q - 1;
    for(int i = 0; i < n; i++) X[i] ^= t;
    uint64_t key = 0;
    for(int bit = b - 1; bit >= 0; bit--)
        for(int i = 0; i < n; i++) key = (key << 1) | ((X[i] >> bit) & 1u);
    return key;
}

}  // namespace

extern "C" int sfx_hilbert_order(int ndim, int64_t C, const double* pos, const double* box,
                                 int64_t* order) {
    if(ndim != 2 && ndim != 3) return set_error(SFX_ERR_ARG, "sfx_hilbert_order: ndim");
    if(C < 0 || (C > 0 && (!pos || !box || !order)))
        return set_error(SFX_ERR_ARG, "sfx_hilbert_order: NULL array");
    const int b = ndim == 2 ? 31 : 21;
    const double scale = (double)(1u << (b - 1)) * 2.0;
    std::vector<std::pair<uint64_t, int64_t>> key(C);
    for(int64_t i = 0; i < C; i++) {
        uint32_t X[3] = {0, 0, 0};
        for(int d = 0; d < ndim; d++) {
            double u = pos[d * C + i] / box[d];
            if(!(u >= 0.0 && u < 1.0))
                return set_error(SFX_ERR_ARG, "sfx_hilbert_order: point %lld outside the box",
                                 (long long)i);
            double s = std::floor(u * scale);
            X[d] = (uint32_t)std::min(s, scale - 1.0);
        }
        key[i] = {hilbert_index(X, b, ndim), i};
    }
    std::sort(key.begin(), key.end());
    for(int64_t i = 0; i < C; i++) order[i] = key[i].second;
    return SFX_OK;
}
