// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-callable driver around the reference Shadowfax classes, compiled (by
// oracle/Makefile) together with the reference's own source files, unmodified,
// from the Shadowfax tree.  Built once per dimension (-Dndim_=2 / -Dndim_=3)
// into oracle/_ref/.  Used only to write golden vectors (tests/golden/) and to
// cross-check the C restatement in sfx_oracle.c; it contains no reference
// code, only calls into the reference's public API:
//   RiemannSolver::solve / solve_for_flux / get_time_derivative / get_W /
//   get_Q (src/riemann/RiemannSolver.hpp:62-121),
//   VorCell::estimate_gradient (src/VorCell.cpp:464), VorFace::set_v /
//   get_normal (src/VorFace.cpp:553-660),
// on objects wired through their public setters (VorGen, GasParticle, VorFace,
// VorCell).  Exceptions thrown by the reference are caught and reported as
// status bits.
#include "StateVector.hpp"
#include "Vec.hpp"
#include "VorCell.hpp"
#include "VorFace.hpp"
#include "VorGen.hpp"
#include "riemann/ExactRiemannSolver.hpp"
#include "riemann/HLLCRiemannSolver.hpp"
#include "utilities/GasParticle.hpp"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

const int NC = ndim_ + 2;

RiemannSolver* make_solver(int kind, double gamma, double tol, double cutoff) {
    // never deleted: the reference destructors print statistics to stdout
    if(kind == 1) return new HLLCRiemannSolver(gamma);
    return new ExactRiemannSolver(gamma, tol, cutoff);
}

StateVector load_state(const double* X, int64_t N, int64_t i) {
    StateVector W;
    for(int k = 0; k < NC; k++) W[k] = X[k * N + i];
    return W;
}

Vec load_vec(const double* X, int64_t N, int64_t i) {
    Vec v;
    for(int d = 0; d < ndim_; d++) v[d] = X[d * N + i];
    return v;
}

}  // namespace

extern "C" {

int sfxr_ndim() { return ndim_; }

// solve_for_flux over a SoA batch; status bit 1 = RiemannSolverException.
// counters[0..2] = get_neval, (unused), (unused).
void sfxr_solve_for_flux_batch(int kind, double gamma, double tol,
                               double cutoff, int64_t N, const double* WL,
                               const double* WR, const double* n,
                               const double* v, const uint8_t* refl,
                               double* F, double* WRout, uint8_t* status,
                               uint64_t* counters) {
    RiemannSolver* s = make_solver(kind, gamma, tol, cutoff);
    for(int64_t i = 0; i < N; i++) {
        StateVector wl = load_state(WL, N, i);
        StateVector wr = load_state(WR, N, i);
        Vec nn = load_vec(n, N, i);
        Vec vv;
        if(v) vv = load_vec(v, N, i);
        uint8_t st = 0;
        StateVector f;
        try {
            f = s->solve_for_flux(wl, wr, nn, vv, refl ? refl[i] != 0 : false);
        } catch(...) {
            st = 1;
        }
        for(int k = 0; k < NC; k++) {
            F[k * N + i] = f[k];
            if(WRout) WRout[k * N + i] = wr[k];
        }
        status[i] = st;
    }
    if(counters) counters[0] = s->get_neval();
}

// ExactRiemannSolver::solve (half state at x/t = 0) over a batch.
void sfxr_exact_solve_batch(double gamma, double tol, double cutoff, int64_t N,
                            const double* WL, const double* WR,
                            const double* n, const uint8_t* refl,
                            double* Whalf, double* mach_io, uint8_t* status) {
    RiemannSolver* s = make_solver(0, gamma, tol, cutoff);
    for(int64_t i = 0; i < N; i++) {
        StateVector wl = load_state(WL, N, i);
        StateVector wr = load_state(WR, N, i);
        Vec nn = load_vec(n, N, i);
        double mach = mach_io ? mach_io[i] : 0.;
        uint8_t st = 0;
        StateVector w;
        try {
            w = s->solve(wl, wr, nn, mach, refl ? refl[i] != 0 : false);
        } catch(...) {
            st = 1;
        }
        for(int k = 0; k < NC; k++) Whalf[k * N + i] = w[k];
        if(mach_io) mach_io[i] = mach;
        status[i] = st;
    }
}

// get_time_derivative; grad is [ndim][NC][N].
void sfxr_time_derivative_batch(int kind, double gamma, int64_t N,
                                const double* W, const double* grad,
                                double* dWdt) {
    RiemannSolver* s = make_solver(kind, gamma, 1.e-8, -5.);
    for(int64_t i = 0; i < N; i++) {
        StateVector w = load_state(W, N, i);
        StateVector g[3];
        for(int d = 0; d < ndim_; d++)
            for(int k = 0; k < NC; k++) g[d][k] = grad[(d * NC + k) * N + i];
        StateVector r = s->get_time_derivative(w, g);
        for(int k = 0; k < NC; k++) dWdt[k * N + i] = r[k];
    }
}

// get_W(volume, Q, use_energy=true); Q is updated in place as in the
// reference (pressure floor resets the energy).
void sfxr_get_W_batch(int kind, double gamma, int64_t N, const double* volume,
                      double* Q, double* W) {
    RiemannSolver* s = make_solver(kind, gamma, 1.e-8, -5.);
    for(int64_t i = 0; i < N; i++) {
        StateVector q = load_state(Q, N, i);
        StateVector w = s->get_W(volume[i], q, true);
        for(int k = 0; k < NC; k++) {
            W[k * N + i] = w[k];
            Q[k * N + i] = q[k];
        }
    }
}

void sfxr_get_Q_batch(int kind, double gamma, int64_t N, const double* volume,
                      const double* W, double* Q, uint8_t* status) {
    RiemannSolver* s = make_solver(kind, gamma, 1.e-8, -5.);
    for(int64_t i = 0; i < N; i++) {
        StateVector w = load_state(W, N, i);
        StateVector q;
        uint8_t st = 0;
        try {
            q = s->get_Q(volume[i], w);
        } catch(...) {
            st = 1;
        }
        for(int k = 0; k < NC; k++) Q[k * N + i] = q[k];
        status[i] = st;
    }
}

// Build reference mesh objects from the SoA mesh description used by the
// oracle (see oracle/sfx_oracle.h) and run, for every cell,
// VorCell::estimate_gradient, and for every face VorFace::set_v + get_v +
// get_normal.  grad_out: [ndim][NC][C]; vface, normal: [ndim][F].
void sfxr_mesh_gradients(int64_t C, int64_t F, const double* W,
                         const double* centroid, const double* pos,
                         const double* vel, const double* volume,
                         const int32_t* left, const int32_t* right,
                         const double* area, const double* midpoint,
                         const double* shift, const uint8_t* ghost,
                         const int64_t* cf_off, const int32_t* cf_face,
                         double* grad_out, double* vface, double* normal) {
    std::vector<GasParticle*> parts(C);
    std::vector<VorGen*> gens(C);
    std::vector<VorCell*> cells(C);
    for(int64_t i = 0; i < C; i++) {
        Vec p = load_vec(pos, C, i);
        parts[i] = new GasParticle(p);
        parts[i]->set_W(load_state(W, C, i));
        Vec v = load_vec(vel, C, i);
        parts[i]->set_velocity(v);
        parts[i]->set_centroid(load_vec(centroid, C, i));
#if ndim_ == 3
        gens[i] = new VorGen(p[0], p[1], p[2]);
#else
        gens[i] = new VorGen(p[0], p[1]);
#endif
        gens[i]->set_particle(parts[i]);
        gens[i]->set_id(2);
        cells[i] = new VorCell(gens[i]);
        cells[i]->set_centroid(load_vec(centroid, C, i));
        cells[i]->set_volume(volume[i]);
    }
    std::vector<VorFace*> faces(F);
    std::vector<VorGen*> rgen(F);
    for(int64_t f = 0; f < F; f++) {
        int32_t L = left[f], R = right[f];
        int64_t base = (R >= 0) ? R : L;
        double x[3];
        for(int d = 0; d < ndim_; d++)
            x[d] = pos[d * C + base] + shift[d * F + f];
        if(R >= 0 && !ghost[f]) {
            rgen[f] = gens[R];
        } else {
#if ndim_ == 3
            rgen[f] = new VorGen(x[0], x[1], x[2]);
#else
            rgen[f] = new VorGen(x[0], x[1]);
#endif
            if(R >= 0) {
                rgen[f]->set_particle(parts[R]);
                rgen[f]->set_id(1);
            } else {
                rgen[f]->set_particle(NULL);
                rgen[f]->set_id(1);
            }
        }
        faces[f] = new VorFace((unsigned int)L, gens);
        faces[f]->add_ngb(rgen[f]);
        faces[f]->set_area(area[f]);
        double m[3];
        for(int d = 0; d < ndim_; d++) m[d] = midpoint[d * F + f];
        faces[f]->set_midpoint(m);
    }
    for(int64_t i = 0; i < C; i++) {
        for(int64_t j = cf_off[i]; j < cf_off[i + 1]; j++) {
            int32_t f = cf_face[j];
            cells[i]->add_face(faces[f]);
            if(left[f] == i)
                cells[i]->add_ngb(rgen[f]);
            else
                cells[i]->add_ngb(gens[left[f]]);
        }
    }
    for(int64_t i = 0; i < C; i++) {
        StateVector delta[3];
        cells[i]->estimate_gradient(delta);
        for(int d = 0; d < ndim_; d++)
            for(int k = 0; k < NC; k++)
                grad_out[(d * NC + k) * C + i] = delta[d][k];
    }
    for(int64_t f = 0; f < F; f++) {
        faces[f]->set_v();
        Vec& v = faces[f]->get_v();
        double ang[3];
        faces[f]->get_normal(ang);
        for(int d = 0; d < ndim_; d++) {
            vface[d * F + f] = v[d];
            normal[d * F + f] = ang[d];
        }
    }
    // objects are leaked on purpose: destructors of the reference mesh
    // classes walk global tessellation state this harness does not build
}

// Face and cell geometry with the reference classes: for every face, the
// stored vertex list (fverts [ndim][fv_off[F]]) goes through
// VorFace::add_vertex, then calculate_midpoint and calculate_area in the
// order VorTess::construct calls them (3D: midpoint first, src/VorTess.cpp:
// 266-268; 2D: area first, :336-343, a face with fewer than 3 vertices or a
// 2D face shorter than 1e-13 gets area 0); every cell collects its faces
// (cf_off/cf_face) and runs VorCell::calculate_centroid and
// calculate_volume (src/VorCell.cpp:257-411).
void sfxr_mesh_geometry(int64_t C, int64_t F, const double* pos,
                        const int64_t* fv_off, const double* fverts,
                        const int64_t* cf_off, const int32_t* cf_face,
                        double* area, double* midpoint, double* volume,
                        double* centroid) {
    const int64_t NV = fv_off[F];
    std::vector<VorGen*> gens(C);
    std::vector<VorCell*> cells(C);
    for(int64_t i = 0; i < C; i++) {
#if ndim_ == 3
        gens[i] = new VorGen(pos[i], pos[C + i], pos[2 * C + i]);
#else
        gens[i] = new VorGen(pos[i], pos[C + i]);
#endif
        cells[i] = new VorCell(gens[i]);
    }
    std::vector<VorFace*> faces(F);
    for(int64_t f = 0; f < F; f++) {
        faces[f] = new VorFace(0, gens);
        const int64_t n = fv_off[f + 1] - fv_off[f];
        for(int64_t t = fv_off[f]; t < fv_off[f + 1]; t++) {
#if ndim_ == 3
            faces[f]->add_vertex(new VorGen(fverts[t], fverts[NV + t], fverts[2 * NV + t]));
#else
            faces[f]->add_vertex(new VorGen(fverts[t], fverts[NV + t]));
#endif
        }
#if ndim_ == 3
        if(n < 3) {
            faces[f]->set_area(0.);
        } else {
            faces[f]->calculate_midpoint();
            faces[f]->calculate_area();
        }
#else
        (void)n;
        faces[f]->calculate_area();
        const int64_t t = fv_off[f];
        double dx = fverts[t] - fverts[t + 1], dy = fverts[NV + t] - fverts[NV + t + 1];
        if(sqrt(dx * dx + dy * dy) < 1.e-13) faces[f]->set_area(0.);
        faces[f]->calculate_midpoint();
#endif
        area[f] = faces[f]->get_area();
        Vec& m = faces[f]->get_midpoint();
        for(int d = 0; d < ndim_; d++) midpoint[d * F + f] = m[d];
    }
    for(int64_t i = 0; i < C; i++) {
        for(int64_t j = cf_off[i]; j < cf_off[i + 1]; j++) cells[i]->add_face(faces[cf_face[j]]);
        cells[i]->calculate_centroid();
        cells[i]->calculate_volume();
        volume[i] = cells[i]->get_volume();
        Vec& c = cells[i]->get_centroid();
        for(int d = 0; d < ndim_; d++) centroid[d * C + i] = c[d];
    }
    // leaked on purpose, as above
}

}  // extern "C"
