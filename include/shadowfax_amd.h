/*
 * shadowfax_amd.h -- C ABI of the MI355X (gfx950) face-flux engine for
 * Shadowfax.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Every entry point returns SFX_OK (0) or a negative error code; a text for
 * the last error of the calling thread is available from sfx_last_error().
 * Per-problem failures that the reference reports by throwing (NaN or
 * negative Riemann solution, NaN face flux, negative mass/energy after an
 * update) are reported as status bits instead, so no exception crosses the
 * boundary.
 *
 * State layout: W = [rho, v_1..v_ndim, p], Q = [m, p_1..p_ndim, e]
 * (ndim + 2 components, the hydro part of the reference StateVector,
 * src/StateVector.hpp:66-96).  Batched arguments are device pointers in SoA,
 * component-major layout: X[k*N + i].  Scalar arguments are host pointers.
 *
 * Reference interfaces replaced are cited per entry point (paths relative to
 * the Shadowfax source tree).
 */
#ifndef SHADOWFAX_AMD_H
#define SHADOWFAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* solver kinds: RiemannSolverFactory type names "Exact" / "HLLC"
 * (src/riemann/RiemannSolverFactory.hpp:42-45,115-136) */
#define SFX_EXACT 0
#define SFX_HLLC 1

/* precision variants */
#define SFX_PREC_FP64 0  /* plain FP64, reference association (parity path) */
#define SFX_PREC_MIXED 1 /* FP32 bulk, FP64 final Newton step for p*, FP64
                            conserved-quantity accumulation */

/* return codes */
#define SFX_OK 0
#define SFX_ERR_ARG -1
#define SFX_ERR_HIP -2
#define SFX_ERR_NOMEM -3
#define SFX_ERR_UNSUPPORTED -4

/* status bits (uint8 per problem / face / cell) */
#define SFX_ST_SOLVER 1u  /* RiemannSolverException, src/riemann/ExactRiemannSolver.cpp:276-288 */
#define SFX_ST_NANFLUX 2u /* FluxCalculationException, src/VorFace.cpp:956-1022 */
#define SFX_ST_ITERCAP 4u /* iteration cap hit (the reference loops are unbounded) */
#define SFX_ST_UPDATEQ 8u /* UpdateQException, src/utilities/GasParticle.cpp:617-621 */

typedef struct sfx_solver sfx_solver;

/* Replaces the solver's mutable counters _counterTotal, _counterBrent,
 * _vacuum (src/riemann/ExactRiemannSolver.hpp:79-86) and HLLC _neval. */
typedef struct {
    uint64_t neval;   /* Riemann problems solved (get_neval) */
    uint64_t nbrent;  /* exact FP64 solves that used Brent's method */
    uint64_t nvacuum; /* solves that took a vacuum branch */
    uint64_t niter;   /* pressure iterations (Newton + Brent) */
    uint64_t nfail;   /* problems/faces with a nonzero status */
} sfx_stats;

const char* sfx_last_error(void);
const char* sfx_version(void);

/* ---- solver handle -------------------------------------------------- */

/* ExactRiemannSolver(gamma, tolerance, cutoff) / HLLCRiemannSolver(gamma)
 * (src/riemann/ExactRiemannSolver.cpp:48-69, HLLCRiemannSolver.cpp:216-219).
 * ndim is 2 or 3 (the reference's compile-time ndim_).  The handle is bound
 * to the HIP device current at creation. */
int sfx_solver_create(int kind, int ndim, double gamma, double tolerance,
                      double cutoff, int precision, sfx_solver** out);

/* RiemannSolverFactory::generate(type, gamma, tolerance, cutoff)
 * (src/riemann/RiemannSolverFactory.hpp:115-136): name "Exact" or "HLLC";
 * other reference solver names return SFX_ERR_UNSUPPORTED. */
int sfx_solver_create_by_name(const char* name, int ndim, double gamma,
                              double tolerance, double cutoff, int precision,
                              sfx_solver** out);
int sfx_solver_destroy(sfx_solver* h);

/* RiemannSolver::tag (src/riemann/RiemannSolver.hpp:49): "EXAC" / "HLLC" */
const char* sfx_solver_tag(const sfx_solver* h);
/* RiemannSolver::get_gamma (src/riemann/RiemannSolver.hpp:140) */
double sfx_solver_gamma(const sfx_solver* h);
int sfx_solver_ndim(const sfx_solver* h);
int sfx_solver_precision(const sfx_solver* h);
/* RiemannSolver::get_neval (src/riemann/RiemannSolver.hpp:154) and the
 * other counters; synchronizes the handle's device. */
int sfx_solver_get_stats(sfx_solver* h, sfx_stats* out);
int sfx_solver_reset_stats(sfx_solver* h);
/* counters are on by default; off removes their atomics from the kernels */
int sfx_solver_enable_stats(sfx_solver* h, int enabled);

/* ---- scalar drop-ins (host pointers, one problem on the GPU) --------- */

/* RiemannSolver::solve(WL, WR, n, mach, reflective)
 * (src/riemann/RiemannSolver.hpp:62-63, ExactRiemannSolver.cpp:108-290).
 * Exact solver only.  WR is overwritten by WL when reflective, as in the
 * reference; *mach is updated with the maximum shock Mach number. */
int sfx_solve(sfx_solver* h, const double* WL, double* WR, const double* n,
              double* mach, int reflective, double* Whalf, uint8_t* status);

/* RiemannSolver::solve_for_flux(WL, WR, n, v, reflective)
 * (src/riemann/RiemannSolver.hpp:76-77).  v may be NULL (zero). */
int sfx_solve_for_flux(sfx_solver* h, const double* WL, double* WR,
                       const double* n, const double* v, int reflective,
                       double* F, uint8_t* status);

/* ---- batched (device pointers, SoA) ---------------------------------- */

/* A loop of solve_for_flux over N independent problems.
 * WL, WR: [ndim+2][N]; n, v: [ndim][N] (v may be NULL: zero face velocity);
 * reflective: [N] or NULL; F: [ndim+2][N]; status: [N] or NULL.
 * stream: a hipStream_t or NULL.  Asynchronous. */
int sfx_solve_for_flux_batch(sfx_solver* h, int64_t N, const double* WL,
                             const double* WR, const double* n,
                             const double* v, const uint8_t* reflective,
                             double* F, uint8_t* status, void* stream);

/* Same with FP32 storage of inputs and fluxes.  Always runs the mixed
 * kernels (FP32 bulk, FP64 final Newton step for p*). */
int sfx_solve_for_flux_batch_f32(sfx_solver* h, int64_t N, const float* WL,
                                 const float* WR, const float* n,
                                 const float* v, const uint8_t* reflective,
                                 float* F, uint8_t* status, void* stream);

/* A loop of ExactRiemannSolver::solve over N problems (half states at
 * x/t = 0); FP64.  mach: [N] in/out or NULL. */
int sfx_exact_solve_batch(sfx_solver* h, int64_t N, const double* WL,
                          const double* WR, const double* n,
                          const uint8_t* reflective, double* Whalf,
                          double* mach, uint8_t* status, void* stream);

#ifdef __cplusplus
}
#endif
#endif
