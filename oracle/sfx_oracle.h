/*
 * sfx_oracle.h -- CPU oracle for the Shadowfax face-flux path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference algorithm (Shadowfax src/riemann and the
 * VorFace/VorCell flux pass).  It is the checker for the HIP product path and
 * the CPU baseline timed next to it; nothing in the product links or calls it.
 *
 * Build rules that make it bit-reproducible against the reference compiled
 * with g++ -O2 on x86-64: ISO C, -ffp-contract=off, no -march (no FMA), glibc
 * libm pow/sqrt.
 *
 * Data conventions (shared with the product's C-ABI):
 *   ndim in {2,3}; a state has ndim+2 components
 *     W = [rho, vx, vy, (vz), p]   primitives
 *     Q = [m, px, py, (pz), e]     conserved
 *   batch arrays are SoA, component-major: X[k*N + i].
 *
 * Pinning status of each function is stated in the header comment of
 * sfx_oracle.c.
 */
#ifndef SFX_ORACLE_H
#define SFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFXO_EXACT 0
#define SFXO_HLLC 1

/* per-problem / per-face status bits (replace the reference's exceptions) */
#define SFXO_ST_SOLVER 1u    /* RiemannSolverException: NaN/negative solution */
#define SFXO_ST_NANFLUX 2u   /* FluxCalculationException: NaN mass flux      */
#define SFXO_ST_ITERCAP 4u   /* iteration cap reached (reference: unbounded)  */
#define SFXO_ST_UPDATEQ 8u   /* UpdateQException: negative m or e after update */

typedef struct {
    int kind;
    int ndim;
    double gamma;
    /* derived constants, ExactRiemannSolver.cpp:52-61 */
    double gp1d2g, gm1d2g, gm1dgp1, tdgp1, tdgm1, gm1d2, tgdgm1, ginv;
    double tol, cutoff;
} sfxo_solver;

typedef struct {
    uint64_t neval;   /* _counterTotal / _neval */
    uint64_t nbrent;  /* _counterBrent */
    uint64_t nvacuum; /* _vacuum */
    uint64_t niter;   /* Newton + Brent iterations (not in the reference) */
} sfxo_counters;

void sfxo_solver_init(sfxo_solver* s, int kind, int ndim, double gamma,
                      double tol, double cutoff);

/* ExactRiemannSolver::solve (x/t = 0 half state).  WR is overwritten when
 * reflective, as in the reference.  Returns status bits. */
unsigned sfxo_exact_solve(const sfxo_solver* s, double* WL, double* WR,
                          const double* n, double* mach, int reflective,
                          double* Whalf, sfxo_counters* c);

/* RiemannSolver::solve_for_flux for the solver kind in s. v may be NULL (0).*/
unsigned sfxo_solve_for_flux(const sfxo_solver* s, double* WL, double* WR,
                             const double* n, const double* v, int reflective,
                             double* F, sfxo_counters* c);

/* batched solve_for_flux over SoA arrays; v, refl may be NULL; status may be
 * NULL.  Returns OR of all status bits. */
unsigned sfxo_solve_for_flux_batch(const sfxo_solver* s, int64_t N,
                                   const double* WL, const double* WR,
                                   const double* n, const double* v,
                                   const uint8_t* refl, double* F,
                                   uint8_t* status, sfxo_counters* c);

/* batched ExactRiemannSolver::solve (half states); mach: [N] in/out (the
 * reference's running maximum, ExactRiemannSolver.cpp:181-188) or NULL */
unsigned sfxo_exact_solve_batch(const sfxo_solver* s, int64_t N,
                                const double* WL, const double* WR,
                                const double* n, const uint8_t* refl,
                                double* Whalf, double* mach, uint8_t* status,
                                sfxo_counters* c);

/* get_time_derivative; grad is [ndim][ndim+2] (grad[d*(ndim+2)+k]) */
void sfxo_time_derivative(const sfxo_solver* s, const double* W,
                          const double* grad, double* dWdt);

/* VorFace::limit (scalar) and VorFace::reconstruct */
double sfxo_limit(double phimid0, double phiL, double phiR, const double* d,
                  double r, int ndim);
void sfxo_reconstruct(int ndim, const double* W, const double* grad,
                      const double* d, double* out);

/* get_W / get_Q (hydro components; paq not carried).  Returns status. */
unsigned sfxo_get_W(const sfxo_solver* s, double volume, double* Q,
                    double* W);
void sfxo_get_Q(const sfxo_solver* s, double volume, const double* W,
                double* Q);

/*
 * Mesh in SoA form.  C cells, F faces.
 *  cell arrays (component-major, length C per component):
 *    W[(ndim+2)*C], grad[ndim*(ndim+2)*C] laid out grad[(d*(ndim+2)+k)*C+i],
 *    centroid[ndim*C], pos[ndim*C] (generator), vel[ndim*C] (generator
 *    velocity), mass[C] (Q.m), total_area[C], dt[C] (real time of the cell's
 *    timestep), volume[C], active[C] (starttime == current integer time)
 *  face arrays:
 *    left[F], right[F] (-1: reflective mirror), area[F], midpoint[ndim*F],
 *    shift[ndim*F]: right generator position = base + shift, where base is
 *      pos[right] (periodic copy when shift != 0, real neighbour when 0) or,
 *      for a mirror face, pos[left];
 *    ghost[F]: 1 if the right generator is a periodic copy (reference id 1)
 *  per-cell face list in the cell's own walk order (CSR):
 *    cf_off[C+1], cf_face[cf_off[C]]
 */
typedef struct {
    int ndim;
    int64_t C, F;
    const double *W, *grad, *centroid, *pos, *vel, *mass, *total_area, *dt,
            *volume;
    const uint8_t* active;
    const int32_t *left, *right;
    const double *area, *midpoint, *shift;
    const uint8_t* ghost;
    const int64_t* cf_off;
    const int32_t* cf_face;
    /* optional per-face geometry, [ndim][F] (geo_rcent [F]): when geo_rRL is
     * set the face pass takes rRL = xRgen - xL, mc = midpoint - (xL +
     * xRgen)/2, the two sides' d vectors and the centroid distance from these
     * arrays instead of forming them from pos, shift, midpoint and centroid
     * (the same expressions: given the host's FP64 values the results are
     * bit-identical; given values rounded as a pass stores them, the oracle
     * computes exactly on that pass's stored inputs) */
    const double *geo_dL, *geo_dR, *geo_rRL, *geo_mc, *geo_rcent;
} sfxo_mesh;

/* VorCell::estimate_gradient for every active cell; grad_out has the mesh
 * grad layout.  Inactive cells are left untouched (src/Simulation.cpp:
 * 414-420 only re-estimates active cells). */
void sfxo_gradients(const sfxo_solver* s, const sfxo_mesh* m, double* grad_out);
/* the same for cells c0..c1-1 */
void sfxo_gradients_range(const sfxo_solver* s, const sfxo_mesh* m, int64_t c0,
                          int64_t c1, double* grad_out);

/* VorFace::set_v and get_normal per face: vface, normal [ndim][F] */
void sfxo_face_velocities(const sfxo_mesh* m, double* vface, double* normal);

/* VorFace::set_v + VorFace::calculate_flux over faces 0..F-1 in order,
 * accumulating into dQ[(ndim+2)*C] (component-major; caller zeroes it).
 * Optional per-face outputs (may be NULL): dQL_face[(ndim+2)*F],
 * face_status[F].  Returns OR of status bits. */
unsigned sfxo_face_flux_pass(const sfxo_solver* s, const sfxo_mesh* m,
                             double* dQ, double* dQL_face, uint8_t* face_status,
                             sfxo_counters* c);
/* The face-frame left and right states every face hands its Riemann solver
 * (after the predictor; zeros for a skipped face), [(ndim+2)][F] each.
 * Testing aid. */
void sfxo_face_states(const sfxo_solver* s, const sfxo_mesh* m, double* WLo,
                      double* WRo);

/* Faces f0..f1-1 without the scatter (dQL_face required, face_status may be
 * NULL), and the ordered per-cell sums of such fluxes for cells c0..c1-1
 * through the gather index (ascending face order; entries 2 face + side):
 * together the same result as sfxo_face_flux_pass, bit for bit, computable
 * in pieces on several threads. */
unsigned sfxo_face_flux_range(const sfxo_solver* s, const sfxo_mesh* m,
                              int64_t f0, int64_t f1, double* dQL_face,
                              uint8_t* face_status, sfxo_counters* c);
void sfxo_gather(const sfxo_mesh* m, const int64_t* gather_off,
                 const int32_t* gather_entry, const double* dQL_face,
                 int64_t c0, int64_t c1, double* dQ);

/* GasParticle::update_Q for every cell (Q -= dQ, dQ = 0, status on m < 0 or
 * e < 0), then W = get_W(V, Q) for the cells whose step ends now,
 * endtime[i] == tnow (every cell when endtime is NULL).  Returns OR of status
 * bits; cell_status may be NULL. */
unsigned sfxo_update(const sfxo_solver* s, int64_t C, double* Q, double* dQ,
                     const double* volume, double* W, const uint64_t* endtime,
                     uint64_t tnow, uint8_t* cell_status);

/* The integer timeline of TimeLine (src/TimeLine.hpp:52-113): the values
 * calculate_timestep reads. */
typedef struct {
    uint64_t integer_maxtime; /* _integer_maxtime */
    uint64_t current_time;    /* _current_time */
    uint64_t timestep;        /* _timestep (system step) */
    uint64_t max_timestep;    /* _max_timestep */
    uint64_t min_timestep;    /* _min_timestep */
    double tottime;           /* _tottime */
    double mintime;           /* _mintime */
    double cfl;               /* _cfl */
    int32_t global_timestep;  /* ParticleVector::global_timestep() */
    int32_t pad;
} sfxo_timeline;

/* TimeLine::calculate_timestep for the gas cells, no gravity, no tree walk
 * (src/TimeLine.cpp:233-320,418-431): cells with endtime == current_time +
 * timestep are active; each gets the CFL time step h/(|v - v_gen| + c)
 * (h from the volume as VorCell::set_h, src/VorCell.cpp:930-938; c from the
 * solver's get_soundspeed), put on the power-of-two hierarchy of the integer
 * timeline and made to synchronise; start/end times advance
 * (Particle::set_timestep).  W: [C][ndim+2], vel: [C][ndim] (item-major).
 * Outputs: dt[C] = get_realtime(end - start) (the value VorFace reads),
 * active[C], *system_step = the smallest new step (Smax if none), *nactive;
 * cell_status bit SFXO_ST_TIMESTEP where the reference exits on "Timestep 0".
 * Returns the OR of the status bits. */
#define SFXO_ST_TIMESTEP 16u
unsigned sfxo_hydro_timestep(const sfxo_solver* s, int64_t C, const double* W,
                             const double* vel, const double* volume,
                             const sfxo_timeline* tl, uint64_t* starttime,
                             uint64_t* endtime, double* dt, uint8_t* active,
                             uint64_t* system_step, uint64_t* nactive,
                             uint8_t* cell_status);

#ifdef __cplusplus
}
#endif
#endif
