// sfx_capi.cpp -- implementation of the C ABI declared in
// include/shadowfax_amd.h: solver handles, argument checking, error
// reporting, and dispatch to the HIP kernels.
#include "shadowfax_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "sfx_capi_internal.h"
#include "sfx_launch.h"

namespace {

thread_local std::string g_err;

}  // namespace

namespace sfx {

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_error(hipError_t e, const char* where) {
    return set_error(SFX_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
}

int ensure_scratch(sfx_solver* h, size_t bytes) {
    if(h->scratch_bytes >= bytes) return SFX_OK;
    // the old buffer may still be read by queued work
    hipError_t e = hipDeviceSynchronize();
    if(e != hipSuccess) return hip_error(e, "scratch: synchronize");
    release_scratch(h);
    const size_t want = bytes + bytes / 2;
    e = hipMalloc(&h->d_scratch, want);
    if(e != hipSuccess) {
        h->d_scratch = nullptr;
        return e == hipErrorOutOfMemory ? set_error(SFX_ERR_NOMEM, "scratch: out of memory")
                                        : hip_error(e, "scratch: hipMalloc");
    }
    h->scratch_bytes = want;
    return SFX_OK;
}

void release_scratch(sfx_solver* h) {
    if(h->d_scratch) hipFree(h->d_scratch);
    h->d_scratch = nullptr;
    h->scratch_bytes = 0;
}

int bind_device(const sfx_solver* h) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if(e != hipSuccess) return hip_error(e, "hipGetDevice");
    if(cur != h->device) {
        e = hipSetDevice(h->device);
        if(e != hipSuccess) return hip_error(e, "hipSetDevice");
    }
    return SFX_OK;
}

}  // namespace sfx

using namespace sfx;

extern "C" {

const char* sfx_last_error(void) { return g_err.c_str(); }

const char* sfx_version(void) { return "shadowfax_amd 0.1 gfx950"; }

int sfx_solver_create(int kind, int ndim, double gamma, double tolerance, double cutoff,
                      int precision, sfx_solver** out) {
    if(!out) return set_error(SFX_ERR_ARG, "sfx_solver_create: out is NULL");
    *out = nullptr;
    if(kind != SFX_EXACT && kind != SFX_HLLC)
        return set_error(SFX_ERR_ARG, "sfx_solver_create: unknown kind %d", kind);
    if(ndim != 2 && ndim != 3)
        return set_error(SFX_ERR_ARG, "sfx_solver_create: ndim must be 2 or 3, got %d", ndim);
    if(!(gamma > 1.0) || !std::isfinite(gamma))
        return set_error(SFX_ERR_ARG, "sfx_solver_create: gamma must be > 1, got %g", gamma);
    if(!(tolerance > 0.0))
        return set_error(SFX_ERR_ARG, "sfx_solver_create: tolerance must be > 0");
    if(precision != SFX_PREC_FP64 && precision != SFX_PREC_MIXED)
        return set_error(SFX_ERR_ARG, "sfx_solver_create: unknown precision %d", precision);
    sfx_solver* h = new sfx_solver();
    h->kind = kind;
    h->ndim = ndim;
    h->precision = precision;
    Consts& c = h->c;
    // ExactRiemannSolver.cpp:52-61
    c.gamma = gamma;
    c.gp1d2g = 0.5 * (gamma + 1.) / gamma;
    c.gm1d2g = 0.5 * (gamma - 1.) / gamma;
    c.gm1dgp1 = (gamma - 1.) / (gamma + 1.);
    c.tdgp1 = 2. / (gamma + 1.);
    c.tdgm1 = 2. / (gamma - 1.);
    c.gm1d2 = 0.5 * (gamma - 1.);
    c.tgdgm1 = 2. * gamma / (gamma - 1.);
    c.ginv = 1. / gamma;
    c.tol = tolerance;
    c.cutoff = cutoff;
    c.itercap = 1000;
    // integer-power shortcuts of the mixed path hold when (g-1)/(2g) = 1/5
    c.fast53 = std::fabs(gamma - 5.0 / 3.0) < 1e-12;
    hipError_t e = hipGetDevice(&h->device);
    if(e != hipSuccess) {
        delete h;
        return hip_error(e, "hipGetDevice");
    }
    if(e == hipSuccess)
        e = hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->device);
    e = hipMalloc(&h->d_ctr, sizeof(unsigned long long) * CNT_N * kCtrSlots);
    if(e != hipSuccess) {
        delete h;
        return hip_error(e, "hipMalloc(counters)");
    }
    e = hipMemset(h->d_ctr, 0, sizeof(unsigned long long) * CNT_N * kCtrSlots);
    if(e != hipSuccess) {
        hipFree(h->d_ctr);
        delete h;
        return hip_error(e, "hipMemset(counters)");
    }
    h->stats_on = true;
    *out = h;
    return SFX_OK;
}

int sfx_solver_create_by_name(const char* name, int ndim, double gamma, double tolerance,
                              double cutoff, int precision, sfx_solver** out) {
    if(!name) return set_error(SFX_ERR_ARG, "sfx_solver_create_by_name: name is NULL");
    if(!strcmp(name, "Exact")) return sfx_solver_create(SFX_EXACT, ndim, gamma, tolerance, cutoff, precision, out);
    if(!strcmp(name, "HLLC")) return sfx_solver_create(SFX_HLLC, ndim, gamma, tolerance, cutoff, precision, out);
    if(!strcmp(name, "TRRS") || !strcmp(name, "NoPressure") || !strcmp(name, "Advection"))
        return set_error(SFX_ERR_UNSUPPORTED, "solver type %s has no GPU implementation", name);
    return set_error(SFX_ERR_ARG, "unknown solver type %s", name);
}

int sfx_solver_destroy(sfx_solver* h) {
    if(!h) return SFX_OK;
    release_scratch(h);
    if(h->d_ctr) hipFree(h->d_ctr);
    if(h->h_stage) hipHostFree(h->h_stage);
    if(h->d_stage) hipFree(h->d_stage);
    delete h;
    return SFX_OK;
}

const char* sfx_solver_tag(const sfx_solver* h) {
    if(!h) return "";
    return h->kind == SFX_EXACT ? "EXAC" : "HLLC";
}

double sfx_solver_gamma(const sfx_solver* h) { return h ? h->c.gamma : 0.0; }
int sfx_solver_ndim(const sfx_solver* h) { return h ? h->ndim : 0; }
int sfx_solver_precision(const sfx_solver* h) { return h ? h->precision : -1; }

int sfx_solver_get_stats(sfx_solver* h, sfx_stats* out) {
    if(!h || !out) return set_error(SFX_ERR_ARG, "sfx_solver_get_stats: NULL argument");
    int rc = bind_device(h);
    if(rc) return rc;
    hipError_t e = hipDeviceSynchronize();
    if(e != hipSuccess) return hip_error(e, "hipDeviceSynchronize");
    unsigned long long slots[CNT_N * kCtrSlots];
    e = hipMemcpy(slots, h->d_ctr, sizeof slots, hipMemcpyDeviceToHost);
    if(e != hipSuccess) return hip_error(e, "hipMemcpy(counters)");
    unsigned long long v[CNT_N] = {0, 0, 0, 0, 0};
    for(int j = 0; j < kCtrSlots; j++)
        for(int k = 0; k < CNT_N; k++) v[k] += slots[j * CNT_N + k];
    out->neval = v[CNT_NEVAL];
    out->nbrent = v[CNT_NBRENT];
    out->nvacuum = v[CNT_NVACUUM];
    out->niter = v[CNT_NITER];
    out->nfail = v[CNT_NFAIL];
    return SFX_OK;
}

int sfx_solver_reset_stats(sfx_solver* h) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_solver_reset_stats: NULL handle");
    int rc = bind_device(h);
    if(rc) return rc;
    hipError_t e = hipDeviceSynchronize();
    if(e == hipSuccess) e = hipMemset(h->d_ctr, 0, sizeof(unsigned long long) * CNT_N * kCtrSlots);
    if(e != hipSuccess) return hip_error(e, "reset_stats");
    return SFX_OK;
}

int sfx_solver_enable_stats(sfx_solver* h, int enabled) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_solver_enable_stats: NULL handle");
    h->stats_on = enabled != 0;
    return SFX_OK;
}

int sfx_solver_set_split(sfx_solver* h, int iterations) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_solver_set_split: NULL handle");
    if(iterations < 0 || iterations > 64)
        return set_error(SFX_ERR_ARG, "sfx_solver_set_split: %d (0..64)", iterations);
    h->split = iterations;
    return SFX_OK;
}

int sfx_solver_get_split(const sfx_solver* h) { return h ? h->split : -1; }

int sfx_solve_for_flux_batch(sfx_solver* h, int64_t N, const double* WL, const double* WR,
                             const double* n, const double* v, const uint8_t* reflective,
                             double* F, uint8_t* status, void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch: NULL handle");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch: N < 0");
    if(N == 0) return SFX_OK;
    if(!WL || !WR || !n || !F)
        return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch: NULL array argument");
    int rc = bind_device(h);
    if(rc) return rc;
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    hipError_t e;
    if(h->precision == SFX_PREC_FP64)
        e = launch_solve_for_flux_fp64(h->kind, h->ndim, h->c, N, WL, WR, n, v, reflective, F,
                                       status, ctr, (hipStream_t)stream);
    else if(h->kind == SFX_EXACT && h->split > 0) {
        if((rc = ensure_scratch(h, exact_split_scratch_bytes(N)))) return rc;
        e = launch_exact_mixed_split<double>(h->ndim, h->c, N, WL, WR, n, v, reflective, F,
                                             status, ctr, h->d_scratch, h->split,
                                             4 * std::max(h->ncu, 1), (hipStream_t)stream);
    } else
        e = launch_solve_for_flux_mixed<double>(h->kind, h->ndim, h->c, N, WL, WR, n, v,
                                                reflective, F, status, ctr, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "solve_for_flux kernel launch");
    return SFX_OK;
}

int sfx_solve_for_flux_batch_f32(sfx_solver* h, int64_t N, const float* WL, const float* WR,
                                 const float* n, const float* v, const uint8_t* reflective,
                                 float* F, uint8_t* status, void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch_f32: NULL handle");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch_f32: N < 0");
    if(N == 0) return SFX_OK;
    if(!WL || !WR || !n || !F)
        return set_error(SFX_ERR_ARG, "sfx_solve_for_flux_batch_f32: NULL array argument");
    int rc = bind_device(h);
    if(rc) return rc;
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    hipError_t e;
    if(h->kind == SFX_EXACT && h->split > 0) {
        if((rc = ensure_scratch(h, exact_split_scratch_bytes(N)))) return rc;
        e = launch_exact_mixed_split<float>(h->ndim, h->c, N, WL, WR, n, v, reflective, F,
                                            status, ctr, h->d_scratch, h->split,
                                            4 * std::max(h->ncu, 1), (hipStream_t)stream);
    } else {
        e = launch_solve_for_flux_mixed<float>(h->kind, h->ndim, h->c, N, WL, WR, n, v,
                                               reflective, F, status, ctr, (hipStream_t)stream);
    }
    if(e != hipSuccess) return hip_error(e, "solve_for_flux_f32 kernel launch");
    return SFX_OK;
}

int sfx_exact_solve_batch(sfx_solver* h, int64_t N, const double* WL, const double* WR,
                          const double* n, const uint8_t* reflective, double* Whalf,
                          double* mach, uint8_t* status, void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_exact_solve_batch: NULL handle");
    if(h->kind != SFX_EXACT)
        return set_error(SFX_ERR_UNSUPPORTED,
                         "solve() is only defined for the exact solver (HLLC::solve aborts)");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_exact_solve_batch: N < 0");
    if(N == 0) return SFX_OK;
    if(!WL || !WR || !n || !Whalf)
        return set_error(SFX_ERR_ARG, "sfx_exact_solve_batch: NULL array argument");
    int rc = bind_device(h);
    if(rc) return rc;
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    hipError_t e = launch_exact_solve_fp64(h->ndim, h->c, N, WL, WR, n, reflective, Whalf,
                                           nullptr, mach, status, ctr, (hipStream_t)stream);
    if(e != hipSuccess) return hip_error(e, "exact_solve kernel launch");
    return SFX_OK;
}

// ---- per-cell operations over batches (FP64) -----------------------------

int sfx_time_derivative_batch(sfx_solver* h, int64_t N, const double* W, const double* grad,
                              double* dWdt, void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_time_derivative_batch: NULL handle");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_time_derivative_batch: N < 0");
    if(N == 0) return SFX_OK;
    if(!W || !grad || !dWdt)
        return set_error(SFX_ERR_ARG, "sfx_time_derivative_batch: NULL array argument");
    int rc = bind_device(h);
    if(rc) return rc;
    hipError_t e = launch_cell_op(0, h->kind, h->ndim, h->c.gamma, N, W, grad, dWdt, nullptr,
                                  (hipStream_t)stream);
    return e == hipSuccess ? SFX_OK : hip_error(e, "time derivative kernel launch");
}

int sfx_get_W_batch(sfx_solver* h, int64_t N, const double* volume, double* Q, double* W,
                    void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_get_W_batch: NULL handle");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_get_W_batch: N < 0");
    if(N == 0) return SFX_OK;
    if(!volume || !Q || !W) return set_error(SFX_ERR_ARG, "sfx_get_W_batch: NULL array");
    int rc = bind_device(h);
    if(rc) return rc;
    hipError_t e = launch_cell_op(1, h->kind, h->ndim, h->c.gamma, N, volume, nullptr, W, Q,
                                  (hipStream_t)stream);
    return e == hipSuccess ? SFX_OK : hip_error(e, "get_W kernel launch");
}

int sfx_hydro_timestep(sfx_solver* h, int64_t C, const double* W, const double* vel,
                       const double* volume, const sfx_timeline* tl, uint64_t* starttime,
                       uint64_t* endtime, double* dt, uint8_t* active, uint64_t* result,
                       uint8_t* status, void* stream) {
    if(!h || !tl) return set_error(SFX_ERR_ARG, "sfx_hydro_timestep: NULL handle or timeline");
    if(C < 0) return set_error(SFX_ERR_ARG, "sfx_hydro_timestep: C < 0");
    if(!result) return set_error(SFX_ERR_ARG, "sfx_hydro_timestep: NULL result");
    if(C > 0 && (!W || !vel || !volume || !starttime || !endtime || !dt || !active))
        return set_error(SFX_ERR_ARG, "sfx_hydro_timestep: NULL array");
    if(tl->integer_maxtime == 0 || tl->max_timestep == 0 ||
       tl->min_timestep > tl->max_timestep)
        return set_error(SFX_ERR_ARG,
                         "sfx_hydro_timestep: integer_maxtime and max_timestep must be > 0 "
                         "and min_timestep <= max_timestep");
    int rc = bind_device(h);
    if(rc) return rc;
    TimelineArgs a;
    a.integer_maxtime = tl->integer_maxtime;
    a.current_time = tl->current_time;
    a.timestep = tl->timestep;
    a.max_timestep = tl->max_timestep;
    a.min_timestep = tl->min_timestep;
    a.tottime = tl->tottime;
    a.mintime = tl->mintime;
    a.cfl = tl->cfl;
    a.global_timestep = tl->global_timestep != 0;
    hipError_t e = launch_hydro_timestep(h->kind, h->ndim, h->c.gamma, C, a, W, vel, volume,
                                         starttime, endtime, dt, active, result, status,
                                         (hipStream_t)stream);
    return e == hipSuccess ? SFX_OK : hip_error(e, "time step kernel launch");
}

int sfx_get_Q_batch(sfx_solver* h, int64_t N, const double* volume, const double* W, double* Q,
                    void* stream) {
    if(!h) return set_error(SFX_ERR_ARG, "sfx_get_Q_batch: NULL handle");
    if(N < 0) return set_error(SFX_ERR_ARG, "sfx_get_Q_batch: N < 0");
    if(N == 0) return SFX_OK;
    if(!volume || !Q || !W) return set_error(SFX_ERR_ARG, "sfx_get_Q_batch: NULL array");
    int rc = bind_device(h);
    if(rc) return rc;
    hipError_t e = launch_cell_op(2, h->kind, h->ndim, h->c.gamma, N, volume, W, Q, nullptr,
                                  (hipStream_t)stream);
    return e == hipSuccess ? SFX_OK : hip_error(e, "get_Q kernel launch");
}

// ---- scalar drop-ins: stage one problem through a pinned buffer -----------

static int stage(sfx_solver* h) {
    if(h->d_stage) return SFX_OK;
    hipError_t e = hipMalloc(&h->d_stage, kStageDoubles * sizeof(double));
    if(e != hipSuccess) return hip_error(e, "hipMalloc(stage)");
    e = hipHostMalloc(&h->h_stage, kStageDoubles * sizeof(double), hipHostMallocDefault);
    if(e != hipSuccess) return hip_error(e, "hipHostMalloc(stage)");
    return SFX_OK;
}

// staging layout (doubles): WL[5] WR[5] n[3] v[3] out[5] mach[1] | bytes: refl, status
enum { S_WL = 0, S_WR = 5, S_N = 10, S_V = 13, S_OUT = 16, S_MACH = 21, S_BYTES = 22 };

static int scalar_run(sfx_solver* h, const double* WL, double* WR, const double* n,
                      const double* v, int reflective, double* out, double* mach,
                      uint8_t* status, bool solve_only) {
    int rc = bind_device(h);
    if(rc) return rc;
    rc = stage(h);
    if(rc) return rc;
    const int NC = h->ndim + 2, ND = h->ndim;
    double* hs = h->h_stage;
    for(int k = 0; k < NC; k++) {
        hs[S_WL + k] = WL[k];
        hs[S_WR + k] = WR[k];
    }
    for(int d = 0; d < ND; d++) {
        hs[S_N + d] = n[d];
        hs[S_V + d] = v ? v[d] : 0.0;
    }
    hs[S_MACH] = mach ? *mach : 0.0;
    uint8_t* hb = reinterpret_cast<uint8_t*>(hs + S_BYTES);
    hb[0] = reflective ? 1 : 0;
    hb[1] = 0;
    double* ds = h->d_stage;
    uint8_t* db = reinterpret_cast<uint8_t*>(ds + S_BYTES);
    hipError_t e = hipMemcpy(ds, hs, kStageDoubles * sizeof(double), hipMemcpyHostToDevice);
    if(e != hipSuccess) return hip_error(e, "hipMemcpy(stage in)");
    // SoA with N = 1: component k sits at offset k
    unsigned long long* ctr = h->stats_on ? h->d_ctr : nullptr;
    if(solve_only)
        e = launch_exact_solve_fp64(ND, h->c, 1, ds + S_WL, ds + S_WR, ds + S_N, db,
                                    ds + S_OUT, nullptr, ds + S_MACH, db + 1, ctr, nullptr);
    else if(h->precision == SFX_PREC_FP64)
        e = launch_solve_for_flux_fp64(h->kind, ND, h->c, 1, ds + S_WL, ds + S_WR, ds + S_N,
                                       ds + S_V, db, ds + S_OUT, db + 1, ctr, nullptr);
    else
        e = launch_solve_for_flux_mixed<double>(h->kind, ND, h->c, 1, ds + S_WL, ds + S_WR,
                                                ds + S_N, ds + S_V, db, ds + S_OUT, db + 1, ctr,
                                                nullptr);
    if(e != hipSuccess) return hip_error(e, "scalar kernel launch");
    e = hipMemcpy(hs, ds, kStageDoubles * sizeof(double), hipMemcpyDeviceToHost);
    if(e != hipSuccess) return hip_error(e, "hipMemcpy(stage out)");
    for(int k = 0; k < NC; k++) out[k] = hs[S_OUT + k];
    if(mach) *mach = hs[S_MACH];
    if(status) *status = hb[1];
    if(reflective)  // the reference overwrites the caller's WR (ExactRiemannSolver.cpp:121-124)
        for(int k = 0; k < NC; k++) WR[k] = WL[k];
    return SFX_OK;
}

int sfx_solve(sfx_solver* h, const double* WL, double* WR, const double* n, double* mach,
              int reflective, double* Whalf, uint8_t* status) {
    if(!h || !WL || !WR || !n || !Whalf) return set_error(SFX_ERR_ARG, "sfx_solve: NULL argument");
    if(h->kind != SFX_EXACT)
        return set_error(SFX_ERR_UNSUPPORTED,
                         "solve() is only defined for the exact solver (HLLC::solve aborts)");
    return scalar_run(h, WL, WR, n, nullptr, reflective, Whalf, mach, status, true);
}

int sfx_solve_for_flux(sfx_solver* h, const double* WL, double* WR, const double* n,
                       const double* v, int reflective, double* F, uint8_t* status) {
    if(!h || !WL || !WR || !n || !F)
        return set_error(SFX_ERR_ARG, "sfx_solve_for_flux: NULL argument");
    return scalar_run(h, WL, WR, n, v, reflective, F, nullptr, status, false);
}

}  // extern "C"
