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

/* Wave-level compaction of the mixed exact solver's batched path: every
 * problem gets `iterations` FP32 Newton iterations in the first kernel;
 * the ones still iterating (strong shocks, near vacuum) are queued with
 * their state and finished, densely packed, by a persistent second kernel,
 * so that a few slow lanes do not hold up their wavefront.  0 runs one
 * kernel.  Results are identical either way.  Default 0: on MI355X the
 * bracketed FP32 Newton converges within a few steps even for the 1e6
 * shocks and near-vacuum faces, and no stop count from 1 to 24 made the
 * two kernels faster than one on the 4M-problem mixed batch.  (No
 * reference counterpart: the reference solves one problem at a time.) */
int sfx_solver_set_split(sfx_solver* h, int iterations);
int sfx_solver_get_split(const sfx_solver* h);

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


/* RiemannSolver::get_time_derivative (src/riemann/RiemannSolver.hpp:89-90;
 * ExactRiemannSolver.cpp:892-934, HLLCRiemannSolver.cpp:415-450, including
 * the 2D dWdt[2] that uses gradients[1].vy twice) over N states, FP64.
 * W, dWdt: [ndim+2][N]; grad: [ndim][ndim+2][N] (dW_k/dx_d). */
int sfx_time_derivative_batch(sfx_solver* h, int64_t N, const double* W,
                              const double* grad, double* dWdt, void* stream);

/* The integer timeline of TimeLine (src/TimeLine.hpp:52-113, accessors
 * :145-166): the values calculate_timestep reads.  No cosmology. */
typedef struct sfx_timeline {
    uint64_t integer_maxtime; /* _integer_maxtime */
    uint64_t current_time;    /* _current_time */
    uint64_t timestep;        /* _timestep, the system step */
    uint64_t max_timestep;    /* _max_timestep */
    uint64_t min_timestep;    /* _min_timestep */
    double tottime;           /* _tottime */
    double mintime;           /* _mintime */
    double cfl;               /* _cfl */
    int32_t global_timestep;  /* ParticleVector::global_timestep() */
    int32_t pad;
} sfx_timeline;

#define SFX_ST_TIMESTEP 16u /* the reference exits with "Timestep 0!" */

/* TimeLine::calculate_timestep for the gas cells, no gravity, no tree walk
 * (src/TimeLine.cpp:233-320, global-timestep tail :418-431), over C cells on
 * the device, FP64.  A cell is active when endtime == current_time +
 * timestep; it gets h/(|v - v_gen| + c) (h from the volume as
 * VorCell::set_h, src/VorCell.cpp:930-938; c = the handle's get_soundspeed,
 * ExactRiemannSolver.cpp:681-687 / HLLCRiemannSolver.cpp:458-460) times the
 * CFL number, put on the power-of-two hierarchy of the integer timeline and
 * made to synchronise; its start and end times advance
 * (Particle::set_timestep, src/utilities/Particle.cpp:304-307).  In global
 * mode every cell takes the smallest step.
 *   W [C][ndim+2], vel (generator velocity) [C][ndim], volume [C]: FP64,
 *   cell-major, as the mesh arrays;  starttime, endtime [C]: in/out;
 *   dt [C] out: get_realtime(endtime - starttime), the per-cell value the
 *   face pass reads (VorFace.cpp:695-699);  active [C] out: the cells that
 *   were active, i.e. those whose faces the next pass computes
 *   (VorFace.cpp:675-683);  result: device uint64[2] out, {system step =
 *   the smallest new step (max_timestep if no cell is active), number of
 *   active cells};  status [C] out (may be NULL): SFX_ST_TIMESTEP.
 * Captures into a HIP graph (no memset, no host synchronisation). */
int sfx_hydro_timestep(sfx_solver* h, int64_t C, const double* W, const double* vel,
                       const double* volume, const sfx_timeline* tl, uint64_t* starttime,
                       uint64_t* endtime, double* dt, uint8_t* active, uint64_t* result,
                       uint8_t* status, void* stream);

/* RiemannSolver::get_W(volume, Q, use_energy = true)
 * (src/riemann/RiemannSolver.hpp:120-121; ExactRiemannSolver.cpp:716-766)
 * over N cells, FP64: W, Q [ndim+2][N], volume [N].  Where the pressure
 * floor acts, Q's energy is reset, as the reference does through Q&. */
int sfx_get_W_batch(sfx_solver* h, int64_t N, const double* volume, double* Q,
                    double* W, void* stream);

/* RiemannSolver::get_Q(volume, W) (src/riemann/RiemannSolver.hpp:108;
 * ExactRiemannSolver.cpp:776-804, HLLCRiemannSolver.cpp:471-483) over N
 * cells, FP64. */
int sfx_get_Q_batch(sfx_solver* h, int64_t N, const double* volume,
                    const double* W, double* Q, void* stream);

/* ---- meshes (host side; tessellation stays on the host) -------------- */

/* A Voronoi mesh as the reference's VorTess hands it over (src/VorTess.cpp:
 * 285-396, src/VorFace.cpp:293-403, src/VorCell.cpp:257-411):
 *   cells in generator order; per-cell faces in walk (counter-clockwise)
 *   order; each face created by its first real cell ("left"); a periodic
 *   copy (reference ghost id 1) never owns a face, so periodic boundary faces
 *   appear twice; right = -1 marks a reflective mirror.
 *   shift: right generator position minus the base position (pos[right] for
 *   a periodic copy or a real neighbour, pos[left] for a mirror).
 * Arrays are component-major; owned by the library, freed by
 * sfx_host_mesh_free(). */
typedef struct {
    int ndim;
    int64_t C, F;
    double* pos;        /* [ndim][C] generator positions */
    double* volume;     /* [C] */
    double* centroid;   /* [ndim][C] */
    int32_t* left;      /* [F] */
    int32_t* right;     /* [F] */
    double* area;       /* [F] */
    double* midpoint;   /* [ndim][F] */
    double* shift;      /* [ndim][F] */
    uint8_t* ghost;     /* [F] right generator is a periodic copy */
    int32_t* twin;      /* [F] other copy of a periodic face, or -1 */
    int64_t* cf_off;    /* [C+1] */
    int32_t* cf_face;   /* [cf_off[C]] faces of each cell, walk order */
    int64_t* fv_off;    /* [F+1] each face's vertex list, as stored by its
                           creating cell (VorFace::get_vertices) */
    double* fverts;     /* [ndim][fv_off[F]] */
} sfx_host_mesh;

/* ParticleVector::sort (src/utilities/ParticleVector.cpp:205-283): the
 * reference keeps its particles, and so its cells, in space-filling Hilbert
 * order.  order[k] = index of the k-th generator along a Hilbert curve over
 * the box (31 bits per axis in 2D, 21 in 3D); generators must lie in
 * [0, box[d]).  Cells in this order keep each cell's neighbours close in
 * memory, which the flux-pass gathers rely on. */
int sfx_hilbert_order(int ndim, int64_t C, const double* pos, const double* box,
                      int64_t* order);

/* Periodic Voronoi tessellation of C generators in [0, box[d]) (ndim 2 or
 * 3), built on the host with nthreads threads (0: all cores).  Stands in for
 * VorTess::construct (src/VorTess.cpp:153-396), whose Delaunay predicates
 * need Boost.Multiprecision. */
int sfx_voronoi_periodic(int ndim, int64_t C, const double* pos, const double* box,
                         int nthreads, sfx_host_mesh* out);
void sfx_host_mesh_free(sfx_host_mesh* m);

/* Per-face geometry the flux-pass kernels read, from a raw mesh (host, FP64,
 * the reference's expressions: src/VorFace.cpp:553-660,728-789):
 *   dL = midpoint - centroid_L,
 *   dR = (midpoint - centroid_R) + (pos_R - xRgen)      (right real cell),
 *   rRL = xRgen - pos_L,  mc = midpoint - 0.5 (pos_L + xRgen),
 *   rcent = |centroid_L - centroid_R| (0 for mirrors),
 * with xRgen = base + shift.  All outputs [ndim][F] except rcent [F]. */
int sfx_face_geometry(int ndim, int64_t C, int64_t F, const double* pos,
                      const double* centroid, const int32_t* left,
                      const int32_t* right, const double* midpoint,
                      const double* shift, double* dL, double* dR, double* rRL,
                      double* mc, double* rcent);

/* Closure residual of every cell (host, FP64): closure[d*C + i] = the sum,
 * over cell i's faces with nonzero area in reverse walk order, of
 * area * r_ij[d] / |r_ij|, r_ij = position of i - position of its neighbour
 * (= -rRL for the face's left cell, rRL for its right cell).  The gradient
 * sum of VorCell::estimate_gradient (src/VorCell.cpp:471-500),
 *   sum_j A_ij [(W_j - W_i) c_ij / |r_ij| - (W_i + W_j)/2 r_ij / |r_ij|],
 * equals the difference form sum_j (W_j - W_i) A_ij (c_ij - r_ij/2)/|r_ij|
 * minus W_i * closure_i; the mixed pass sums the difference form (no
 * cancellation of the W_i terms in FP32 weights) and subtracts W_i *
 * closure_i when sfx_mesh.closure is given.  The faces of a Voronoi cell
 * close (sum of A n = 0), so closure is round-off there; a face list that
 * does not close (faces moved or turned into mirrors by hand) needs it.
 * cf_entry: sfx_mesh_cell_entries; rRL: sfx_face_geometry [ndim][F]. */
int sfx_cell_closure(int ndim, int64_t C, int64_t F, const int64_t* cf_off,
                     const int32_t* cf_entry, const double* area, const double* rRL,
                     double* closure);

/* Per-cell face entries in walk order: cf_entry[j] = 2*face + side, side 1
 * when the cell is the face's right (used by the gradient kernel). */
int sfx_mesh_cell_entries(int64_t C, int64_t F, const int32_t* left,
                          const int64_t* cf_off, const int32_t* cf_face,
                          int32_t* cf_entry);

/* Gather index of the deterministic scatter: for every cell, the (face,
 * side) entries it receives in ascending face order -- the order in which
 * VorTessManager::hydro's face loop (src/VorTessManager.cpp:205) applies
 * increase_dQ.  Every face updates its left cell; its right cell only when
 * that is a real local cell (not a periodic copy, not a mirror),
 * src/VorFace.cpp:913-955.  gather_off: [C+1]; gather_entry may be NULL to
 * only size it (gather_off[C] entries). */
int sfx_mesh_gather_index(int64_t C, int64_t F, const int32_t* left,
                          const int32_t* right, const uint8_t* ghost,
                          int64_t* gather_off, int32_t* gather_entry);

/* ---- face-flux pass (device pointers) -------------------------------- */

/* Device mesh for the flux pass.  "T" arrays are double for an
 * SFX_PREC_FP64 handle and float for an SFX_PREC_MIXED handle; the
 * primitive states W are double for both.
 * Layout: arrays the kernels read through a neighbour index are cell- or
 * face-major (W [C][ndim+2], vel [C][ndim], grad [C][ndim][ndim+2],
 * dQ and Q [C][ndim+2], dQL T [F][ndim+2]) so that one gather is one cache
 * line; per-face geometry, read one face per lane, is component-major
 * ([ndim][F]).  Zero-initialise it and set fields by name: optional
 * arrays are NULL when unused. */
typedef struct {
    int ndim;
    int64_t C, F;
    /* cells */
    const double* W;           /* [C][ndim+2] primitives (FP64 in both passes);
                                  16-byte aligned, as egeo */
    const double* vel;         /* [C][ndim] generator velocities (FP64 in
                                  both passes) */
    const double* mass;        /* [C] Q.m */
    const double* total_area;  /* [C] sum of the cell's face areas */
    const double* dt;          /* [C] real time of the cell's time step */
    const double* volume;      /* [C] */
    const uint8_t* active;     /* [C] starttime == current time, or NULL */
    /* faces */
    const int32_t* left;       /* [F] */
    const int32_t* right;      /* [F] -1: mirror */
    const uint8_t* ghost;      /* [F] */
    const int32_t* twin;       /* [F] or NULL: the other copy of a periodic
                                  face, -1 on every other face (mixed pass:
                                  one evaluation per periodic pair when
                                  given) */
    const void* area;          /* T [F] */
    const void* dL;            /* T [ndim][F]  (sfx_face_geometry) */
    const void* dR;            /* T [ndim][F] */
    const void* rRL;           /* T [ndim][F] */
    const void* mc;            /* T [ndim][F] */
    const void* rcent;         /* T [F] */
    /* topology */
    const int64_t* cf_off;     /* [C+1] */
    const int32_t* cf_entry;   /* sfx_mesh_cell_entries */
    const void* egeo;          /* T [cf_off[C]][sfx_mesh_entry_record(...)]
                                  per-entry face records
                                  (sfx_mesh_entry_geometry) */
    const int64_t* gather_off; /* [C+1] */
    const int32_t* gather_entry; /* sfx_mesh_gather_index */
    const int64_t* egeo_group; /* may be NULL: the grouped record layout
                                  (sfx_mesh_entry_groups) -- egeo then holds
                                  rows * 64 records, rows =
                                  egeo_group[ceil(C/64)] */
    int64_t egeo_rows;         /* that rows value, or 0 if not given (with it,
                                  and records under 4 GB, a lane past its
                                  cell's entries reads no padding) */
    const int64_t* egeo_slot;  /* [C] or NULL (grouped layout only): slot s
                                  of the groups (group s / 64, lane s % 64)
                                  holds cell egeo_slot[s] & 0xffffffff, which
                                  has egeo_slot[s] >> 32 entries
                                  (sfx_mesh_entry_groups_sorted); NULL: slot
                                  s holds cell s */
    const void* closure;       /* T [ndim][C] or NULL (mixed pass only):
                                  every cell's closure residual
                                  (sfx_cell_closure); NULL when the cells'
                                  faces close, as those of a Voronoi
                                  tessellation do */
    const void* frec;          /* float4 [sfx_mesh_face_record_planes(...)][F]
                                  or NULL (mixed pass only): the faces' own
                                  values packed by sfx_mesh_face_records,
                                  16-byte aligned; the face kernel then reads
                                  them instead of left, right, area, dL, dR,
                                  rRL, mc and rcent (which stay required) */
} sfx_mesh;

/* Per-entry face records for the gradient pass, built once per
 * tessellation from the face arrays and cf_entry, in each cell's walk order
 * so that the pass streams them; for entry j of cell i (face f seen from
 * side s):
 *   FP64:  [area, mc (ndim), rRL seen from i (ndim), dL or dR (ndim),
 *           the neighbour's index bits (-1: mirror), (3D) pad]
 *          -- the values the gradient loop of VorCell::estimate_gradient
 *          reads per face (src/VorCell.cpp:471-500);
 *   mixed: [area (mc - r/2)/|r| (ndim), the neighbour's index bits
 *          (~f: mirror), (2D) pad | dL or dR (ndim), area, (2D) pad],
 *          r = rRL seen from i -- the weights of the difference form of
 *          the same sum, in two 16-byte halves.
 * E = cf_off[C]; egeo: T [E][sfx_mesh_entry_record(ndim, precision)].
 * Mixed handles read each face's values from m->frec when it is set
 * (sfx_mesh_face_records, built from the same arrays: the same records,
 * fewer loads), else from the face arrays. */
int sfx_mesh_entry_geometry(sfx_solver* h, const sfx_mesh* m, int64_t E, void* egeo,
                            void* stream);

/* Elements of type T per entry record: 8 / 12 (FP64, 2D / 3D), 8 (mixed);
 * -1 for other arguments. */
int sfx_mesh_entry_record(int ndim, int precision);

/* Per-face records of the mixed face pass (no counterpart in the reference:
 * a layout of the per-face values that VorFace::calculate_flux reads,
 * src/VorFace.cpp:670-760), built once per tessellation from the FP32 face
 * arrays of m: plane q of face f at float4 index q F + f,
 *   3D: [left, right, area, rcent] [rRL (3), mc.x] [mc.y, mc.z, dL.x, dL.y]
 *       [dL.z, dR (3)]
 *   2D: [left, right, area, rcent] [rRL (2), mc (2)] [dL (2), dR (2)]
 * with left and right as their int32 bits.  The face kernel then loads a
 * face with 3 or 4 16-byte reads instead of 13 or 17 4-byte ones; the
 * fluxes are bit for bit the same.  frec: float4 [planes][F], 16-byte
 * aligned.  Mixed-precision handles only (SFX_ERR_ARG otherwise). */
int sfx_mesh_face_records(sfx_solver* h, const sfx_mesh* m, void* frec, void* stream);

/* float4 planes per face record: 4 (3D) or 3 (2D) for the mixed precision,
 * -1 for other arguments. */
int sfx_mesh_face_record_planes(int ndim, int precision);

/* The grouped layout of the per-entry records (host): cells in groups of
 * 64 consecutive cells; group g gets record rows egeo_group[g] ..
 * egeo_group[g+1]-1, one per entry of its cell with the most entries, and
 * row r holds every cell's entry r - egeo_group[g] (walk order) side by
 * side as 16-byte planes of the record (plane q of row r, lane l at
 * 16-byte index (r P + q) 64 + l, P = record bytes / 16), zero records
 * padding the shorter cells.  A wavefront of 64 consecutive cells then
 * loads each row with coalesced 1 KB reads.  egeo_group: [ceil(C/64) + 1];
 * returns rows = egeo_group[ceil(C/64)] (egeo then holds rows * 64
 * records), or -1 on bad arguments.  sfx_mesh_entry_geometry fills that
 * layout when sfx_mesh.egeo_group is set. */
int64_t sfx_mesh_entry_groups(int64_t C, const int64_t* cf_off, int64_t* egeo_group);

/* The grouped layout over cells ordered by entry count (host; no
 * counterpart in the reference: a layout of the records above).  Within
 * each window of `window` consecutive cells (a positive multiple of 64; the
 * last window may be shorter) the cells take slots in descending order of
 * their entry counts (ties in cell order), so that a group's 64 cells have
 * nearly equal counts and its rows -- as many as its longest cell has
 * entries -- hold little padding: on a 3D Poisson-Voronoi mesh 16.3 rows
 * per group against 24.2 in cell order, for 15.5 entries per cell.  Each
 * cell's records and their walk order are those of sfx_mesh_entry_groups,
 * so the gradients are bit for bit the same.  egeo_slot: [C], slot s ->
 * cell | (entries << 32); egeo_group: [ceil(C/64) + 1]; returns rows, or -1
 * on bad arguments.  Set sfx_mesh.egeo_slot to the array (on the device)
 * for sfx_mesh_entry_geometry and the gradient pass. */
int64_t sfx_mesh_entry_groups_sorted(int64_t C, const int64_t* cf_off, int64_t window,
                                     int64_t* egeo_slot, int64_t* egeo_group);

/* VorCell::estimate_gradient for every active cell (src/VorCell.cpp:464-530,
 * loop src/Simulation.cpp:414-420).  grad: T [C][ndim][ndim+2]
 * (grad[(i*ndim + d)*(ndim+2) + k] = dW_k/dx_d); inactive cells are
 * untouched. */
int sfx_gradients(sfx_solver* h, const sfx_mesh* m, void* grad, void* stream);

/* VorFace::set_v + calculate_flux for every face (src/VorFace.cpp:586-866):
 * dQL: T [F][ndim+2] the flux exchanged through each face, dt*A*F after the
 * mass-flux limiter, seen from the left cell (the right cell receives -dQL);
 * face_status: [F] or NULL.  Mixed precision with twin: each periodic pair is
 * evaluated once, by its lower-index copy, which also writes the other
 * copy's row (negated: seen from that copy's left cell). */
int sfx_face_fluxes(sfx_solver* h, const sfx_mesh* m, const void* grad,
                    void* dQL, uint8_t* face_status, void* stream);

/* Deterministic scatter: dQ[i*(ndim+2) + k] = sum of the cell's entries in
 * ascending face order (overwrites dQ). */
int sfx_gather_dQ(sfx_solver* h, const sfx_mesh* m, const void* dQL,
                  double* dQ, void* stream);

/* VorTessManager::hydro (src/VorTessManager.cpp:198-212) preceded by the
 * gradient estimation: sfx_gradients + sfx_face_fluxes + sfx_gather_dQ.
 * grad (T [C][ndim][ndim+2]) and dQL (T [F][ndim+2]) are caller-owned work
 * buffers; dQ receives the per-cell change [C][ndim+2]. */
int sfx_face_flux_pass(sfx_solver* h, const sfx_mesh* m, void* grad,
                       void* dQL, double* dQ, uint8_t* face_status,
                       void* stream);

/* No counterpart in the reference.  sfx_face_flux_pass for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the face fluxes dQL, the pass's only
 * scratch, live in a static device array of SFX_STATIC_SCRATCH_BYTES per
 * device instead of a caller workspace; grad stays the caller's (inactive
 * cells keep their gradients from step to step).  Allocates nothing;
 * SFX_ERR_ARG when F * (ndim+2) * sizeof(T) exceeds the scratch.  Every call
 * on one device shares that device's scratch: calls that overlap on one
 * device (other streams or threads) race. */
#define SFX_STATIC_SCRATCH_BYTES (64u << 20)
int sfx_face_flux_pass_static(sfx_solver* h, const sfx_mesh* m, void* grad,
                              double* dQ, uint8_t* face_status, void* stream);

/* GasParticle::update_Q then get_W (src/utilities/GasParticle.cpp:617-624,
 * src/riemann/ExactRiemannSolver.cpp:716-766): Q -= dQ, dQ = 0,
 * W = get_W(volume, Q) (Q, dQ, W [C][ndim+2]); mass [C] (may be NULL) is
 * set to the new Q.m, the sfx_mesh.mass the next pass's mass-flux limiter
 * reads (VorFace::calculate_flux takes the cell's current mass,
 * src/VorFace.cpp:850-851); cell_status bit SFX_ST_UPDATEQ where m or e went
 * negative. */
int sfx_update_step(sfx_solver* h, int64_t C, double* Q, double* dQ,
                    const double* volume, double* W, double* mass,
                    uint8_t* cell_status, void* stream);

/* The same with individual time steps (src/Simulation.cpp:476-483,792-796): Q -= dQ
 * and dQ = 0 for every cell (GasParticle::update_Q, which also raises
 * SFX_ST_UPDATEQ), but W = get_W(volume, Q) only for the cells whose step
 * ends now, endtime[i] == current_time (the time after the step; only those
 * cells have received the fluxes of all their faces); the other cells keep
 * their W, and their Q is not reset by get_W's pressure floor.
 * endtime: device uint64 [C] (sfx_hydro_timestep's), or NULL for every
 * cell (= sfx_update_step). */
int sfx_update_step_timed(sfx_solver* h, int64_t C, double* Q, double* dQ,
                          const double* volume, double* W, double* mass,
                          const uint64_t* endtime, uint64_t current_time,
                          uint8_t* cell_status, void* stream);


/* ---- serving: an ensemble of meshes resident in HBM ------------------ */

/* Many independent meshes held on the device together and stepped in
 * batches (no reference counterpart: the reference runs one mesh per MPI
 * job; the per-mesh step is the same VorTessManager::hydro + update_Q +
 * get_W).  All meshes share one set of arrays with global indices; the
 * per-step scratch (gradients, face fluxes) is an arena of arena_cells
 * cells and arena_faces faces that each batch shares out, so a batch is
 * admitted by its scratch footprint.  A batched step is one replay of a
 * captured HIP graph: the batch list (host to device), a plan kernel and
 * four persistent kernels draining work tiles of all admitted meshes. */
typedef struct sfx_ensemble sfx_ensemble;

typedef struct {
    int64_t nreq, nbatch, max_batch;
    double wall_s;     /* first arrival window to last completion */
    double busy_s;     /* time with a step in flight */
    double p50_s, p99_s, mean_s, max_s;  /* request latency (arrival to done) */
    double mean_batch; /* meshes per step */
} sfx_serve_stats;

#define SFX_ENS_W 0
#define SFX_ENS_Q 1
#define SFX_ENS_DQ 2
#define SFX_SERVE_BATCHED 0 /* batches through the captured graph */
#define SFX_SERVE_PLAIN 1   /* one mesh per step, launches sized to the mesh */
#define SFX_SERVE_QUEUE 2   /* requests posted to a queue on the device, drained,
                               admitted and stepped by one persistent kernel */

/* Capacities are totals over all meshes (entries: cf_off[C] summed, the
 * same bound is used for the gather index).  The handle's precision selects
 * the storage (SFX_PREC_MIXED: FP32 gradients and face geometry). */
int sfx_ensemble_create(sfx_solver* h, int64_t max_meshes, int64_t max_cells,
                        int64_t max_faces, int64_t max_entries,
                        int64_t arena_cells, int64_t arena_faces,
                        sfx_ensemble** out);
int sfx_ensemble_destroy(sfx_ensemble* e);

/* Copies one mesh (device arrays with local indices, the sfx_mesh layout of
 * the handle's precision; egeo, if set, is ignored; closure, if set, is
 * kept by a mixed ensemble) and its conserved
 * variables Q ([C][ndim+2]) into the ensemble.  E = cf_off[C],
 * G = gather_off[C].  Synchronous.  SFX_ERR_NOMEM when the ensemble is
 * full. */
int sfx_ensemble_add_mesh(sfx_ensemble* e, const sfx_mesh* m, const double* Q,
                          int64_t E, int64_t G, int32_t* mesh_id);

/* mesh_id >= 0: its cells and faces; mesh_id < 0: totals, returns the
 * number of meshes. */
int sfx_ensemble_info(const sfx_ensemble* e, int32_t mesh_id, int64_t* cells,
                      int64_t* faces);

/* One step (gradients, face fluxes, gather, update_Q + get_W) of each mesh
 * in ids (no mesh twice; the batch's scratch must fit the arena).
 * Asynchronous on the ensemble's stream. */
int sfx_ensemble_step(sfx_ensemble* e, const int32_t* ids, int32_t n);
/* The same step for one mesh with plain launches sized to it. */
int sfx_ensemble_step_plain(sfx_ensemble* e, int32_t id);
int sfx_ensemble_sync(sfx_ensemble* e);
/* Copies W, Q or dQ (SFX_ENS_*) of one mesh to host_out ([C][ndim+2]). */
int sfx_ensemble_read(sfx_ensemble* e, int32_t id, int what, double* host_out);

/* Launch shape of the batched step's kernels: threads per block (= cells
 * or faces per work tile; 64..256, a multiple of 64) and blocks per CU of
 * the persistent grid (1..32), or 0 for one block per possible tile (the
 * grid then covers the largest tile count the arena allows, and blocks
 * without a tile exit at once). */
int sfx_ensemble_set_launch(sfx_ensemble* e, int block, int blocks_per_cu);
int sfx_ensemble_get_launch(const sfx_ensemble* e, int* block, int* blocks_per_cu);
/* Times reps batched steps of ids for each block size in {64, 128, 256}
 * and blocks per CU in {0 (one block per tile), 1, 2, 4, 8}, keeps the
 * fastest (best_ms: its time per step) and restores the meshes' states
 * afterwards. */
int sfx_ensemble_autotune(sfx_ensemble* e, const int32_t* ids, int32_t n, int reps,
                          double* best_ms);

/* FIFO admission: the longest prefix of the pending requests (meshes
 * pending[k]) whose scratch fits the arena, skipping (and keeping queued)
 * a second request for a mesh already admitted; at most max_batch.
 * admitted[0..*nadmitted) are indices into pending.  Host only. */
int sfx_serve_admit(int64_t npending, const int32_t* pending,
                    const int64_t* mesh_cells, const int64_t* mesh_faces,
                    int64_t arena_cells, int64_t arena_faces, int32_t max_batch,
                    int32_t* admitted, int32_t* nadmitted);

/* Serves an arrival trace in real time, open loop: request r for mesh
 * mesh_id[r] arrives arrival_s[r] seconds after the call (ascending); while
 * the device is busy arrivals queue; when it is free the admitted prefix of
 * the queue is stepped (SFX_SERVE_BATCHED) or the head alone
 * (SFX_SERVE_PLAIN).  SFX_SERVE_QUEUE: the host only posts each request,
 * at its arrival time, to a mailbox in pinned memory; one persistent kernel
 * (launched cooperatively, every block resident) moves posted requests into
 * a queue in HBM, admits them by the same rule, plans and runs each batch's
 * four phases with grid barriers between them, and marks each request done
 * in the mailbox.  latency_s[r] = completion - arrival; batch_of[r] (or
 * NULL) the step that served it. */
int sfx_ensemble_serve(sfx_ensemble* e, int mode, int64_t nreq,
                       const double* arrival_s, const int32_t* mesh_id,
                       double* latency_s, int32_t* batch_of, sfx_serve_stats* st);

/* ---- snapshots: HDF5 files of Gadget-style initial conditions ------- */

/* Replaces the HDF5-library reads of GadgetSnapshotReader::read_snapshot
 * (src/GadgetSnapshotReader.cpp:79-380, through src/io/HDF5tools.hpp) with a
 * self-contained reader of the structures such files use: superblock
 * version 0, symbol-table groups, version 1 object headers, contiguous or
 * compact datasets, attributes.  Chunked or filtered datasets are refused
 * with an error.  Host only; paths are absolute ("/PartType0/Coordinates"). */
typedef struct sfx_h5 sfx_h5;
int sfx_h5_open(const char* path, sfx_h5** out);
int sfx_h5_close(sfx_h5* f);
/* 1 if the object (and, attr != NULL, its attribute) exists, 0 if not */
int sfx_h5_exists(const sfx_h5* f, const char* path, const char* attr);
/* A dataset (attr NULL) or an attribute of the object at path:
 * type_class 0 integer, 1 floating point, 8 enumeration (boolean),
 * -1 other; type_size in bytes; rank and dims[rank] (rank <= 8). */
int sfx_h5_info(const sfx_h5* f, const char* path, const char* attr, int* type_class,
                int* type_size, int* rank, int64_t* dims);
/* All n values (n must equal the element count) converted to double or
 * int64. */
int sfx_h5_read_f64(const sfx_h5* f, const char* path, const char* attr, int64_t n,
                    double* out);
int sfx_h5_read_i64(const sfx_h5* f, const char* path, const char* attr, int64_t n,
                    int64_t* out);
/* The members of the group at path and its attributes ("@name"), one per
 * line, NUL-terminated; returns the bytes needed when cap is too small. */
int sfx_h5_list(const sfx_h5* f, const char* path, char* names, int64_t cap);

/* Writing such files (what GadgetSnapshotWriter writes through the HDF5
 * library, src/GadgetSnapshotWriter.cpp:220-500): build a tree of groups,
 * contiguous datasets and attributes in memory, then save it in the same
 * structures the reader above reads (superblock version 0, symbol-table
 * groups of at most 256 members, version 1 object headers).  Types:
 * SFX_H5_UINT / SFX_H5_INT (1, 2, 4 or 8 bytes), SFX_H5_FLOAT (4 or 8);
 * data little-endian, row-major. */
#define SFX_H5_UINT 0
#define SFX_H5_FLOAT 1
#define SFX_H5_INT 2
typedef struct sfx_h5w sfx_h5w;
int sfx_h5w_create(sfx_h5w** out);
int sfx_h5w_destroy(sfx_h5w* w);
/* a group at an absolute path (its parent must exist) */
int sfx_h5w_group(sfx_h5w* w, const char* path);
int sfx_h5w_dataset(sfx_h5w* w, const char* path, int type, int type_size, int rank,
                    const int64_t* dims, const void* data);
/* an attribute of the group or dataset at path (rank 0: a scalar) */
int sfx_h5w_attr(sfx_h5w* w, const char* path, const char* name, int type, int type_size,
                 int rank, const int64_t* dims, const void* data);
int sfx_h5w_save(sfx_h5w* w, const char* path);

#ifdef __cplusplus
}
#endif
#endif
