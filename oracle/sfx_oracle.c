/*
 * sfx_oracle.c -- CPU oracle (FP64 restatement) of the Shadowfax face-flux
 * path.  TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker; never linked into the product.
 *
 * Every function restates the reference expression by expression (same
 * operand order, same association), so that compiled with -ffp-contract=off
 * and glibc libm it reproduces the reference bit for bit.  Reference paths
 * are relative to the Shadowfax source tree.
 *
 * Pinning (how each function is checked, see tests/):
 *   exact solve / solve_for_flux / vacuum / Brent / guess_p  -- golden vectors
 *       from the reference ExactRiemannSolver compiled from source
 *       (oracle/_ref), Toro known answers, Sod star state, invariants.
 *   HLLC solve_for_flux -- golden vectors from the compiled reference
 *       HLLCRiemannSolver, invariants (uniform state, reflective symmetry).
 *   time derivative, get_W, get_Q -- golden vectors from the compiled
 *       reference solvers; closed forms.
 *   gradients (VorCell::estimate_gradient), face velocity (VorFace::set_v),
 *       normal -- golden vectors from the reference VorCell/VorFace compiled
 *       from source on harness-built meshes; linear-field exactness.
 *   limiter, reconstruct, face flux pass (VorFace::calculate_flux) -- parity
 *       with the reference is UNPINNED by a reference run (calculate_flux
 *       needs a TimeLine, whose constructor lives in TimeLine.cpp, which needs
 *       Boost.property_tree, absent here).  Pinned by hand-computed limiter
 *       cases, brute force on tiny meshes, the closed forms of linear
 *       pressure and velocity fields (the predicted face states follow from
 *       the Euler equations in primitive form: a sign error in d, a dropped
 *       1/2 or dt fails them; tests/test_oracle_predictor.py), the vacuum
 *       keep, invariance under a periodic translation, and invariants: dQR
 *       == -dQL bitwise, conservation to round-off, uniform flow gives zero
 *       net dQ, zero gradients reduce the face flux to solve_for_flux
 *       exactly.
 *   hydro time step (TimeLine::calculate_timestep) -- UNPINNED by a reference
 *       run (TimeLine.cpp needs Boost.property_tree).  Pinned by a brute-force
 *       search over all power-of-two fractions of the timeline on small
 *       timelines, hand-worked cases and the invariants of the hierarchy
 *       (the step is the largest admissible one, it divides the remaining
 *       time, it stays within [min, max]).
 */
#include "sfx_oracle.h"

#include <math.h>
#include <string.h>

#ifndef M_PI /* strict C11 <math.h> leaves it out; glibc's value */
#define M_PI 3.14159265358979323846
#endif

/* std::max / std::min semantics (src/StateVector.hpp uses std::max/min):
 * std::max(a,b) = (a < b) ? b : a ; std::min(a,b) = (b < a) ? b : a */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

#define ITERCAP 100000 /* reference loops are unbounded; cap for safety */

/* ExactRiemannSolver ctor, src/riemann/ExactRiemannSolver.cpp:48-63 */
void sfxo_solver_init(sfxo_solver* s, int kind, int ndim, double gamma,
                      double tol, double cutoff) {
    s->kind = kind;
    s->ndim = ndim;
    s->gamma = gamma;
    s->gp1d2g = 0.5 * (gamma + 1.) / gamma;
    s->gm1d2g = 0.5 * (gamma - 1.) / gamma;
    s->gm1dgp1 = (gamma - 1.) / (gamma + 1.);
    s->tdgp1 = 2. / (gamma + 1.);
    s->tdgm1 = 2. / (gamma - 1.);
    s->gm1d2 = 0.5 * (gamma - 1.);
    s->tgdgm1 = 2. * gamma / (gamma - 1.);
    s->ginv = 1. / gamma;
    s->tol = tol;
    s->cutoff = cutoff;
}

/* ---------------------------------------------------------------- exact -- */

/* get_soundspeed, ExactRiemannSolver.cpp:681-687 */
static double ex_soundspeed(const sfxo_solver* s, const double* W) {
    int P = s->ndim + 1;
    if(W[0] != 0.) return sqrt(s->gamma * W[P] / W[0]);
    return 0.;
}

/* fb, ExactRiemannSolver.cpp:556-566 */
static double ex_fb(const sfxo_solver* s, double p, const double* W,
                    double a) {
    int P = s->ndim + 1;
    if(p > W[P]) {
        double A = s->tdgp1 / W[0];
        double B = s->gm1dgp1 * W[P];
        return (p - W[P]) * sqrt(A / (p + B));
    }
    return s->tdgm1 * a * (pow(p / W[P], s->gm1d2g) - 1.);
}

/* gb, ExactRiemannSolver.cpp:575-579 */
static double ex_gb(const sfxo_solver* s, double p, const double* W) {
    int P = s->ndim + 1;
    double A = s->tdgp1 / W[0];
    double B = s->gm1dgp1 * W[P];
    return sqrt(A / (p + B));
}

/* fprimeb, ExactRiemannSolver.cpp:589-599 */
static double ex_fprimeb(const sfxo_solver* s, double p, const double* W,
                         double a) {
    int P = s->ndim + 1;
    if(p > W[P]) {
        double A = s->tdgp1 / W[0];
        double B = s->gm1dgp1 * W[P];
        return (1. - 0.5 * (p - W[P]) / (B + p)) * sqrt(A / (p + B));
    }
    return 1. / (W[0] * a) * pow(p / W[P], -s->gp1d2g);
}

/* f and fprime, ExactRiemannSolver.cpp:611-631 */
static double ex_f(const sfxo_solver* s, double p, const double* WL,
                   const double* WR, double vL, double vR, double aL,
                   double aR) {
    return ex_fb(s, p, WL, aL) + ex_fb(s, p, WR, aR) + (vR - vL);
}
static double ex_fprime(const sfxo_solver* s, double p, const double* WL,
                        const double* WR, double aL, double aR) {
    return ex_fprimeb(s, p, WL, aL) + ex_fprimeb(s, p, WR, aR);
}

/* guess_p, ExactRiemannSolver.cpp:644-673 */
static double ex_guess_p(const sfxo_solver* s, const double* WL,
                         const double* WR, double vL, double vR, double aL,
                         double aR) {
    int P = s->ndim + 1;
    double pguess;
    double pmin = smin(WL[P], WR[P]);
    double pmax = smax(WL[P], WR[P]);
    double qmax = pmax / pmin;
    double ppv = 0.5 * (WL[P] + WR[P]) -
                 0.125 * (vR - vL) * (WL[0] + WR[0]) * (aL + aR);
    ppv = smax(s->tol * 0.5 * (WL[P] + WR[P]), ppv);
    if(qmax <= 2. && pmin <= ppv && ppv <= pmax) {
        pguess = ppv;
    } else if(ppv < pmin) {
        pguess = pow((aL + aR - s->gm1d2 * (vR - vL)) /
                             (aL / pow(WL[P], s->gm1d2g) +
                              aR / pow(WR[P], s->gm1d2g)),
                     s->tgdgm1);
    } else {
        pguess = (ex_gb(s, ppv, WL) * WL[P] + ex_gb(s, ppv, WR) * WR[P] - vR +
                  vL) /
                 (ex_gb(s, ppv, WL) + ex_gb(s, ppv, WR));
    }
    pguess = smax(s->tol * 0.5 * (WL[P] + WR[P]), pguess);
    return pguess;
}

/* BrentsMethodSolve, ExactRiemannSolver.cpp:462-546 */
static double ex_brent(const sfxo_solver* s, double lowerLimit,
                       double upperLimit, double errorTol, const double* WL,
                       const double* WR, double vL, double vR, double aL,
                       double aR, sfxo_counters* c, unsigned* st) {
    double a = lowerLimit;
    double b = upperLimit;
    double cc = 0.;
    double d = 1e230;
    double fa = ex_f(s, a, WL, WR, vL, vR, aL, aR);
    double fb = ex_f(s, b, WL, WR, vL, vR, aL, aR);
    double fc = 0.;
    double ss = 0.;
    double fs = 0.;
    if(fa * fb >= 0) {
        if(fa < fb) return a;
        return b;
    }
    if(fabs(fa) < fabs(fb)) {
        double tmp = a;
        a = b;
        b = tmp;
        tmp = fa;
        fa = fb;
        fb = tmp;
    }
    cc = a;
    fc = fa;
    int mflag = 1;
    long it = 0;
    while(!(fb == 0) && (fabs(a - b) > errorTol * (a + b) * 0.5)) {
        if(++it > ITERCAP) {
            *st |= SFXO_ST_ITERCAP;
            break;
        }
        if((fa != fc) && (fb != fc))
            ss = a * fb * fc / (fa - fb) / (fa - fc) +
                 b * fa * fc / (fb - fa) / (fb - fc) +
                 cc * fa * fb / (fc - fa) / (fc - fb);
        else
            ss = b - fb * (b - a) / (fb - fa);
        double tmp2 = (3. * a + b) / 4.;
        if(!(((ss > tmp2) && (ss < b)) || ((ss < tmp2) && (ss > b))) ||
           (mflag && (fabs(ss - b) >= (fabs(b - cc) / 2.))) ||
           (!mflag && (fabs(ss - b) >= (fabs(cc - d) / 2.))) ||
           (mflag && (fabs(b - cc) < errorTol * (b + cc) * 0.5)) ||
           (!mflag && (fabs(cc - d) < errorTol * (cc + d) * 0.5))) {
            ss = (a + b) / 2.;
            mflag = 1;
        } else {
            mflag = 0;
        }
        fs = ex_f(s, ss, WL, WR, vL, vR, aL, aR);
        d = cc;
        cc = b;
        fc = fb;
        if(fa * fs < 0.) {
            b = ss;
            fb = fs;
        } else {
            a = ss;
            fa = fs;
        }
        if(fabs(fa) < fabs(fb)) {
            double tmp = a;
            a = b;
            b = tmp;
            tmp = fa;
            fa = fb;
            fb = tmp;
        }
    }
    if(c) c->niter += (uint64_t)it;
    return b;
}

/* sanity check after a solve, ExactRiemannSolver.cpp:276-288 / 432-444 */
static unsigned ex_check(int ndim, const double* sol) {
    int P = ndim + 1;
    int bad = (sol[0] < 0.);
    bad |= (sol[0] != sol[0]);
    bad |= (sol[1] != sol[1]);
    bad |= (sol[2] != sol[2]);
    if(ndim == 3) bad |= (sol[3] != sol[3]);
    bad |= (sol[P] < 0.);
    bad |= (sol[P] != sol[P]);
    return bad ? SFXO_ST_SOLVER : 0u;
}

/* rarefaction-fan base, shared by the vacuum and fan branches:
 * 2/(g+1) -+ (g-1)/(g+1)/a*v, ExactRiemannSolver.cpp:214,251,321,353 */

/* solve_vacuum, ExactRiemannSolver.cpp:304-446 */
static unsigned ex_solve_vacuum(const sfxo_solver* s, const double* WL,
                                const double* WR, double vL, double vR,
                                double aL, double aR, const double* n,
                                double* sol, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    double g = s->gamma;
    if(c) c->nvacuum++;
    for(int k = 0; k < NC; k++) sol[k] = 0.;
    if(!(WR[0] != 0.) && !(WL[0] != 0.)) return 0u;
    double vhalf;
    if(!(WR[0] != 0.)) {
        memcpy(sol, WL, sizeof(double) * NC);
        if(vL < aL) {
            double SL = vL + 2. * aL / (g - 1.);
            if(SL > 0.) {
                double base = 2. / (g + 1.) + (g - 1.) / (g + 1.) / aL * vL;
                sol[0] = WL[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (aL + 0.5 * (g - 1.) * vL) - vL;
                sol[P] = WL[P] * pow(base, 2. * g / (g - 1.));
            } else {
                for(int k = 0; k < NC; k++) sol[k] = 0.;
                vhalf = 0.;
            }
        } else {
            vhalf = 0.;
        }
    } else if(!(WL[0] != 0.)) {
        memcpy(sol, WR, sizeof(double) * NC);
        if(-aR < vR) {
            double SR = vR - 2. * aR / (g - 1.);
            if(SR >= 0.) {
                for(int k = 0; k < NC; k++) sol[k] = 0.;
                vhalf = 0.;
            } else {
                double base = 2. / (g + 1.) - (g - 1.) / (g + 1.) / aR * vR;
                sol[0] = WR[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (-aR + 0.5 * (g - 1.) * vR) - vR;
                sol[P] = WR[P] * pow(base, 2. * g / (g - 1.));
            }
        } else {
            vhalf = 0.;
        }
    } else {
        /* vacuum generation */
        double SR = vR - 2. * aR / (g - 1.);
        double SL = vL + 2. * aL / (g - 1.);
        if(SR > 0. && SL < 0.) {
            for(int k = 0; k < NC; k++) sol[k] = 0.;
            vhalf = 0.;
        } else if(SL >= 0.) {
            memcpy(sol, WL, sizeof(double) * NC);
            if(aL > vL) {
                double base = 2. / (g + 1.) + (g - 1.) / (g + 1.) / aL * vL;
                sol[0] = WL[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (aL + 0.5 * (g - 1.) * vL) - vL;
                sol[P] = WL[P] * pow(base, 2. * g / (g - 1.));
            } else {
                vhalf = 0.;
            }
        } else {
            memcpy(sol, WR, sizeof(double) * NC);
            if(-aR < vR) {
                double base = 2. / (g + 1.) - (g - 1.) / (g + 1.) / aR * vR;
                sol[0] = WR[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (-aR + 0.5 * (g - 1.) * vR) - vR;
                sol[P] = WR[P] * pow(base, 2. * g / (g - 1.));
            } else {
                vhalf = 0.;
            }
        }
    }
    for(int d = 0; d < nd; d++) sol[1 + d] += vhalf * n[d];
    return ex_check(nd, sol);
}

/* ExactRiemannSolver::solve, ExactRiemannSolver.cpp:108-290 */
unsigned sfxo_exact_solve(const sfxo_solver* s, double* WL, double* WR,
                          const double* n, double* mach, int reflective,
                          double* sol, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    unsigned st = 0;
    if(c) c->neval++;
    double vL = WL[1] * n[0] + WL[2] * n[1];
    double vR = WR[1] * n[0] + WR[2] * n[1];
    if(nd == 3) {
        vL = WL[1] * n[0] + WL[2] * n[1] + WL[3] * n[2];
        vR = WR[1] * n[0] + WR[2] * n[1] + WR[3] * n[2];
    }
    if(reflective) {
        memcpy(WR, WL, sizeof(double) * NC);
        vR = -vL;
    }
    double aL = ex_soundspeed(s, WL);
    double aR = ex_soundspeed(s, WR);
    if(!(WL[0] != 0.) || !(WR[0] != 0.)) {
        return ex_solve_vacuum(s, WL, WR, vL, vR, aL, aR, n, sol, c);
    }
    if(2. * aL / (s->gamma - 1.) + 2. * aR / (s->gamma - 1.) <= vR - vL) {
        st |= ex_solve_vacuum(s, WL, WR, vL, vR, aL, aR, n, sol, c);
    } else {
        double tolh = s->tol;
        double p = 0.;
        double pguess = ex_guess_p(s, WL, WR, vL, vR, aL, aR);
        double fp = ex_f(s, p, WL, WR, vL, vR, aL, aR);
        double fpguess = ex_f(s, pguess, WL, WR, vL, vR, aL, aR);
        long it = 0;
        if(fp < s->cutoff) {
            while(fabs(p - pguess) > tolh * 0.5 * (p + pguess)) {
                if(++it > ITERCAP) {
                    st |= SFXO_ST_ITERCAP;
                    break;
                }
                p = pguess;
                pguess = pguess - fpguess / ex_fprime(s, pguess, WL, WR, aL, aR);
                if(pguess < 0.) pguess = 0.;
                fpguess = ex_f(s, pguess, WL, WR, vL, vR, aL, aR);
            }
            p = pguess;
        } else {
            if(fp * fpguess >= 0.) {
                while(fabs(p - pguess) > tolh * 0.5 * (p + pguess) &&
                      fpguess < 0.) {
                    if(++it > ITERCAP) {
                        st |= SFXO_ST_ITERCAP;
                        break;
                    }
                    p = pguess;
                    pguess = pguess -
                             fpguess / ex_fprime(s, pguess, WL, WR, aL, aR);
                    fpguess = ex_f(s, pguess, WL, WR, vL, vR, aL, aR);
                }
            }
            if(fabs(p - pguess) > tolh * 0.5 * (p + pguess) && fpguess > 0.) {
                p = 0.;
                if(c) c->nbrent++;
                p = ex_brent(s, p, pguess, s->tol, WL, WR, vL, vR, aL, aR, c,
                             &st);
            } else {
                p = pguess;
            }
        }
        if(c) c->niter += (uint64_t)it;

        double u = 0.5 * (vL + vR) +
                   0.5 * (ex_fb(s, p, WR, aR) - ex_fb(s, p, WL, aL));

        if(mach) {
            if(p > WL[P]) *mach = smax(*mach, sqrt(s->gp1d2g * p / WL[P] + s->gm1d2g));
            if(p > WR[P]) *mach = smax(*mach, sqrt(s->gp1d2g * p / WR[P] + s->gm1d2g));
        }

        double vhalf;
        if(u < 0) {
            memcpy(sol, WR, sizeof(double) * NC);
            double pdpR = p / WR[P];
            if(p > WR[P]) {
                double SR = vR + aR * sqrt(s->gp1d2g * pdpR + s->gm1d2g);
                if(SR > 0) {
                    sol[0] = WR[0] * (pdpR + s->gm1dgp1) /
                             (s->gm1dgp1 * pdpR + 1.);
                    sol[P] = p;
                    vhalf = u - vR;
                } else {
                    vhalf = 0.;
                }
            } else {
                double SHR = vR + aR;
                if(SHR > 0) {
                    double STR = u + aR * pow(pdpR, s->gm1d2g);
                    if(STR <= 0) {
                        sol[0] = WR[0] *
                                 pow(s->tdgp1 - s->gm1dgp1 / aR * vR, s->tdgm1);
                        vhalf = s->tdgp1 * (-aR + s->gm1d2 * vR) - vR;
                        sol[P] = WR[P] * pow(s->tdgp1 - s->gm1dgp1 / aR * vR,
                                             s->tgdgm1);
                    } else {
                        sol[0] = WR[0] * pow(pdpR, s->ginv);
                        sol[P] = p;
                        vhalf = u - vR;
                    }
                } else {
                    vhalf = 0.;
                }
            }
        } else {
            memcpy(sol, WL, sizeof(double) * NC);
            double pdpL = p / WL[P];
            if(p > WL[P]) {
                double SL = vL - aL * sqrt(s->gp1d2g * pdpL + s->gm1d2g);
                if(SL < 0) {
                    sol[0] = WL[0] * (pdpL + s->gm1dgp1) /
                             (s->gm1dgp1 * pdpL + 1.);
                    sol[P] = p;
                    vhalf = u - vL;
                } else {
                    vhalf = 0.;
                }
            } else {
                double SHL = vL - aL;
                if(SHL < 0) {
                    double STL = u - aL * pow(pdpL, s->gm1d2g);
                    if(STL > 0) {
                        sol[0] = WL[0] *
                                 pow(s->tdgp1 + s->gm1dgp1 / aL * vL, s->tdgm1);
                        vhalf = s->tdgp1 * (aL + s->gm1d2 * vL) - vL;
                        sol[P] = WL[P] * pow(s->tdgp1 + s->gm1dgp1 / aL * vL,
                                             s->tgdgm1);
                    } else {
                        sol[0] = WL[0] * pow(pdpL, s->ginv);
                        vhalf = u - vL;
                        sol[P] = p;
                    }
                } else {
                    vhalf = 0.;
                }
            }
        }
        for(int d = 0; d < nd; d++) sol[1 + d] += vhalf * n[d];
    }
    st |= ex_check(nd, sol);
    return st;
}

/* get_energy, ExactRiemannSolver.cpp:695-704 */
static double ex_energy(const sfxo_solver* s, const double* W) {
    if(s->ndim == 3)
        return 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]) +
               W[4] / (s->gamma - 1.);
    return 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2]) + W[3] / (s->gamma - 1.);
}

/* solve_for_flux + get_flux, ExactRiemannSolver.cpp:815-880 */
static unsigned ex_solve_for_flux(const sfxo_solver* s, double* WL, double* WR,
                                  const double* n, const double* v,
                                  int reflective, double* F, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    double W[5];
    double maxmach = 0.;
    unsigned st = sfxo_exact_solve(s, WL, WR, n, &maxmach, reflective, W, c);
    double vz[3] = {0., 0., 0.};
    if(!v) v = vz;
    for(int d = 0; d < nd; d++) W[1 + d] += v[d];
    double Fi[3][5];
    for(int i = 0; i < nd; i++) {
        for(int k = 0; k < NC; k++) Fi[i][k] = 0.;
        if(!(W[0] != 0.)) continue;
        double rel = W[1 + i] - v[i];
        Fi[i][0] = W[0] * rel;
        for(int d = 0; d < nd; d++) Fi[i][1 + d] = W[0] * rel * W[1 + d];
        Fi[i][1 + i] += W[P];
        Fi[i][P] = rel * ex_energy(s, W) + W[P] * W[1 + i];
    }
    for(int k = 0; k < NC; k++) {
        double acc = Fi[0][k] * n[0] + Fi[1][k] * n[1];
        if(nd == 3) acc = acc + Fi[2][k] * n[2];
        F[k] = acc;
    }
    return st;
}

/* ----------------------------------------------------------------- HLLC -- */

/* HLLCRiemannSolver::vacuum_flux, HLLCRiemannSolver.cpp:44-209 */
static void hllc_vacuum_flux(const sfxo_solver* s, const double* WL,
                             const double* WR, const double* n,
                             const double* v, double uL, double uR, double aL,
                             double aR, double* flux) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    double g = s->gamma;
    double sol[5];
    double vhalf;
    if(!(WR[0] != 0.)) {
        memcpy(sol, WL, sizeof(double) * NC);
        if(aL > uL) {
            double SL = uL + 2. * aL / (g - 1.);
            if(SL > 0.) {
                double base = 2. / (g + 1.) + (g - 1.) / (g + 1.) / aL * uL;
                sol[0] = WL[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (aL + 0.5 * (g - 1.) * uL) - uL;
                sol[P] = WL[P] * pow(base, 2. * g / (g - 1.));
            } else {
                for(int k = 0; k < NC; k++) sol[k] = 0.;
                vhalf = 0.;
            }
        } else {
            vhalf = 0.;
        }
    } else if(!(WL[0] != 0.)) {
        memcpy(sol, WR, sizeof(double) * NC);
        if(-uR < aR) {
            double SR = uR - 2. * aR / (g - 1.);
            if(SR >= 0.) {
                for(int k = 0; k < NC; k++) sol[k] = 0.;
                vhalf = 0.;
            } else {
                double base = 2. / (g + 1.) - (g - 1.) / (g + 1.) / aR * uR;
                sol[0] = WR[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (-aR + 0.5 * (g - 1.) * uR) - uR;
                sol[P] = WR[P] * pow(base, 2. * g / (g - 1.));
            }
        } else {
            vhalf = 0.;
        }
    } else {
        double SR = uR - 2. * aR / (g - 1.);
        double SL = uL + 2. * aL / (g - 1.);
        if(SR > 0. && SL < 0.) {
            for(int k = 0; k < NC; k++) sol[k] = 0.;
            vhalf = 0.;
        } else if(SL >= 0.) {
            memcpy(sol, WL, sizeof(double) * NC);
            if(uL < aL) {
                double base = 2. / (g + 1.) + (g - 1.) / (g + 1.) / aL * uL;
                sol[0] = WL[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (aL + 0.5 * (g - 1.) * uL) - uL;
                sol[P] = WL[P] * pow(base, 2. * g / (g - 1.));
            } else {
                vhalf = 0.;
            }
        } else {
            memcpy(sol, WR, sizeof(double) * NC);
            if(-uR < aR) {
                double base = 2. / (g + 1.) - (g - 1.) / (g + 1.) / aR * uR;
                sol[0] = WR[0] * pow(base, 2. / (g - 1.));
                vhalf = 2. / (g + 1.) * (-aR + 0.5 * (g - 1.) * uR) - uR;
                sol[P] = WR[P] * pow(base, 2. * g / (g - 1.));
            } else {
                vhalf = 0.;
            }
        }
    }
    for(int d = 0; d < nd; d++) sol[1 + d] += vhalf * n[d];

    for(int k = 0; k < NC; k++) flux[k] = 0.;
    if(!(sol[0] != 0.)) return;
    double vsol = sol[1] * n[0] + sol[2] * n[1];
    double v2 = sol[1] * sol[1] + sol[2] * sol[2];
    if(nd == 3) {
        vsol += sol[3] * n[2];
        v2 += sol[3] * sol[3];
    }
    flux[0] = sol[0] * vsol;
    for(int d = 0; d < nd; d++)
        flux[1 + d] = sol[0] * sol[1 + d] * vsol + sol[P] * n[d];
    double e = sol[P] / (g - 1.) / sol[0] + 0.5 * v2;
    flux[P] = sol[0] * e * vsol + sol[P] * vsol;
    /* deboost, :198-206 */
    double vv[3] = {0., 0., 0.};
    if(v) {
        for(int d = 0; d < nd; d++) vv[d] = v[d];
    }
    double vnorm2 = vv[0] * vv[0] + vv[1] * vv[1];
    if(nd == 3) vnorm2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
    for(int d = 0; d < nd; d++) flux[1 + d] += vv[d] * flux[0];
    if(nd == 3)
        flux[4] += vv[0] * flux[1] + vv[1] * flux[2] + vv[2] * flux[3] +
                   0.5 * vnorm2 * flux[0];
    else
        flux[3] += vv[0] * flux[1] + vv[1] * flux[2] + 0.5 * vnorm2 * flux[0];
}

/* HLLCRiemannSolver::solve_for_flux, HLLCRiemannSolver.cpp:258-403 */
static unsigned hllc_solve_for_flux(const sfxo_solver* s, double* WL,
                                    double* WR, const double* n,
                                    const double* v, int reflective,
                                    double* flux, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    double g = s->gamma;
    if(c) c->neval++;
    for(int k = 0; k < NC; k++) flux[k] = 0.;
    if(!(WL[0] != 0.) && (reflective || !(WR[0] != 0.))) return 0u;
    double uL = WL[1] * n[0] + WL[2] * n[1];
    double uR = WR[1] * n[0] + WR[2] * n[1];
    if(nd == 3) {
        uL = WL[1] * n[0] + WL[2] * n[1] + WL[3] * n[2];
        uR = WR[1] * n[0] + WR[2] * n[1] + WR[3] * n[2];
    }
    if(reflective) {
        memcpy(WR, WL, sizeof(double) * NC);
        uR = -uL;
    }
    /* get_soundspeed without a rho == 0 guard, :458-460 */
    double aL = sqrt(g * WL[P] / WL[0]);
    double aR = sqrt(g * WR[P] / WR[0]);
    if(!(WL[0] != 0.) || !(WR[0] != 0.)) {
        if(c) c->nvacuum++;
        hllc_vacuum_flux(s, WL, WR, n, v, uL, uR, aL, aR, flux);
        return 0u;
    }
    if(2. * aL / (g - 1.) + 2. * aR / (g - 1.) <= uR - uL) {
        if(c) c->nvacuum++;
        hllc_vacuum_flux(s, WL, WR, n, v, uL, uR, aL, aR, flux);
        return 0u;
    }
    double rhobar = 0.5 * (WL[0] + WR[0]);
    double abar = 0.5 * (aL + aR);
    double pPVRS = 0.5 * (WL[P] + WR[P]) - 0.5 * (uR - uL) * rhobar * abar;
    double pstar = smax(0., pPVRS);
    double qL = 1.;
    if(pstar > WL[P]) qL = sqrt(1. + 0.5 * (g + 1.) / g * (pstar / WL[P] - 1.));
    double qR = 1.;
    if(pstar > WR[P]) qR = sqrt(1. + 0.5 * (g + 1.) / g * (pstar / WR[P] - 1.));
    double SL = uL - aL * qL;
    double SR = uR + aR * qR;
    double Sstar = (WR[P] - WL[P] + WL[0] * uL * (SL - uL) -
                    WR[0] * uR * (SR - uR)) /
                   (WL[0] * (SL - uL) - WR[0] * (SR - uR));
    if(Sstar >= 0.) {
        flux[0] = WL[0] * uL;
        for(int d = 0; d < nd; d++)
            flux[1 + d] = WL[0] * WL[1 + d] * uL + WL[P] * n[d];
        double v2 = WL[1] * WL[1] + WL[2] * WL[2];
        if(nd == 3) v2 += WL[3] * WL[3];
        double eL = WL[P] / (g - 1.) / WL[0] + 0.5 * v2;
        flux[P] = WL[0] * eL * uL + WL[P] * uL;
        if(SL < 0.) {
            double U[5];
            U[0] = 1.;
            for(int d = 0; d < nd; d++) U[1 + d] = WL[1 + d] + (Sstar - uL) * n[d];
            U[P] = eL + (Sstar - uL) * (Sstar + WL[P] / (WL[0] * (SL - uL)));
            double fac = WL[0] * (SL - uL) / (SL - Sstar);
            for(int k = 0; k < NC; k++) U[k] *= fac;
            flux[0] += SL * (U[0] - WL[0]);
            for(int d = 0; d < nd; d++)
                flux[1 + d] += SL * (U[1 + d] - WL[0] * WL[1 + d]);
            flux[P] += SL * (U[P] - WL[0] * eL);
        }
    } else {
        flux[0] = WR[0] * uR;
        for(int d = 0; d < nd; d++)
            flux[1 + d] = WR[0] * WR[1 + d] * uR + WR[P] * n[d];
        double v2 = WR[1] * WR[1] + WR[2] * WR[2];
        if(nd == 3) v2 += WR[3] * WR[3];
        double eR = WR[P] / (g - 1.) / WR[0] + 0.5 * v2;
        flux[P] = WR[0] * eR * uR + WR[P] * uR;
        if(SR > 0.) {
            double U[5];
            U[0] = 1.;
            for(int d = 0; d < nd; d++) U[1 + d] = WR[1 + d] + (Sstar - uR) * n[d];
            U[P] = eR + (Sstar - uR) * (Sstar + WR[P] / (WR[0] * (SR - uR)));
            double fac = WR[0] * (SR - uR) / (SR - Sstar);
            for(int k = 0; k < NC; k++) U[k] *= fac;
            flux[0] += SR * (U[0] - WR[0]);
            for(int d = 0; d < nd; d++)
                flux[1 + d] += SR * (U[1 + d] - WR[0] * WR[1 + d]);
            flux[P] += SR * (U[P] - WR[0] * eR);
        }
    }
    /* deboost, :392-400 */
    double vv[3] = {0., 0., 0.};
    if(v) {
        for(int d = 0; d < nd; d++) vv[d] = v[d];
    }
    double vnorm2 = vv[0] * vv[0] + vv[1] * vv[1];
    if(nd == 3) vnorm2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
    for(int d = 0; d < nd; d++) flux[1 + d] += vv[d] * flux[0];
    if(nd == 3)
        flux[4] += vv[0] * flux[1] + vv[1] * flux[2] + vv[2] * flux[3] +
                   0.5 * vnorm2 * flux[0];
    else
        flux[3] += vv[0] * flux[1] + vv[1] * flux[2] + 0.5 * vnorm2 * flux[0];
    return 0u;
}

unsigned sfxo_solve_for_flux(const sfxo_solver* s, double* WL, double* WR,
                             const double* n, const double* v, int reflective,
                             double* F, sfxo_counters* c) {
    if(s->kind == SFXO_HLLC)
        return hllc_solve_for_flux(s, WL, WR, n, v, reflective, F, c);
    return ex_solve_for_flux(s, WL, WR, n, v, reflective, F, c);
}

unsigned sfxo_solve_for_flux_batch(const sfxo_solver* s, int64_t N,
                                   const double* WL, const double* WR,
                                   const double* n, const double* v,
                                   const uint8_t* refl, double* F,
                                   uint8_t* status, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2;
    unsigned all = 0;
    for(int64_t i = 0; i < N; i++) {
        double wl[5], wr[5], nn[3], vv[3], f[5];
        for(int k = 0; k < NC; k++) {
            wl[k] = WL[k * N + i];
            wr[k] = WR[k * N + i];
        }
        for(int d = 0; d < nd; d++) {
            nn[d] = n[d * N + i];
            vv[d] = v ? v[d * N + i] : 0.;
        }
        unsigned st = sfxo_solve_for_flux(s, wl, wr, nn, v ? vv : NULL,
                                          refl ? refl[i] : 0, f, c);
        for(int k = 0; k < NC; k++) F[k * N + i] = f[k];
        if(status) status[i] = (uint8_t)st;
        all |= st;
    }
    return all;
}

unsigned sfxo_exact_solve_batch(const sfxo_solver* s, int64_t N,
                                const double* WL, const double* WR,
                                const double* n, const uint8_t* refl,
                                double* Whalf, double* mach_io,
                                uint8_t* status, sfxo_counters* c) {
    int nd = s->ndim, NC = nd + 2;
    unsigned all = 0;
    for(int64_t i = 0; i < N; i++) {
        double wl[5], wr[5], nn[3], w[5];
        for(int k = 0; k < NC; k++) {
            wl[k] = WL[k * N + i];
            wr[k] = WR[k * N + i];
        }
        for(int d = 0; d < nd; d++) nn[d] = n[d * N + i];
        double mach = mach_io ? mach_io[i] : 0.;
        unsigned st = sfxo_exact_solve(s, wl, wr, nn, &mach,
                                       refl ? refl[i] : 0, w, c);
        if(mach_io) mach_io[i] = mach;
        for(int k = 0; k < NC; k++) Whalf[k * N + i] = w[k];
        if(status) status[i] = (uint8_t)st;
        all |= st;
    }
    return all;
}

/* ------------------------------------------------- predictor & limiter -- */

/* get_time_derivative: ExactRiemannSolver.cpp:892-934,
 * HLLCRiemannSolver.cpp:415-450 (identical in the hydro components).
 * The 2D dWdt[2] uses gradients[1].vy() twice, as in the reference (:920). */
void sfxo_time_derivative(const sfxo_solver* s, const double* W,
                          const double* grad, double* dWdt) {
    double g = s->gamma;
    if(s->ndim == 3) {
        const double *gx = grad, *gy = grad + 5, *gz = grad + 10;
        dWdt[0] = -W[0] * (gx[1] + gy[2] + gz[3]) - W[1] * gx[0] -
                  W[2] * gy[0] - W[3] * gz[0];
        dWdt[1] = -W[1] * gx[1] - W[2] * gy[1] - W[3] * gz[1] - gx[4] / W[0];
        dWdt[2] = -W[1] * gx[2] - W[2] * gy[2] - W[3] * gz[2] - gy[4] / W[0];
        dWdt[3] = -W[1] * gx[3] - W[2] * gy[3] - W[3] * gz[3] - gz[4] / W[0];
        dWdt[4] = -g * W[4] * (gx[1] + gy[2] + gz[3]) - W[1] * gx[4] -
                  W[2] * gy[4] - W[3] * gz[4];
    } else {
        const double *gx = grad, *gy = grad + 4;
        dWdt[0] = -W[1] * gx[0] - W[2] * gy[0] - W[0] * (gx[1] + gy[2]);
        dWdt[1] = -W[1] * gx[1] - W[2] * gy[1] - gx[3] / W[0];
        dWdt[2] = -W[1] * gy[2] - W[2] * gy[2] - gy[3] / W[0];
        dWdt[3] = -g * W[3] * (gx[1] + gy[2]) - W[1] * gx[3] - W[2] * gy[3];
    }
}

/* VorFace::limit (scalar), src/VorFace.hpp:92-137 */
double sfxo_limit(double phimid0, double phiL, double phiR, const double* d,
                  double r, int ndim) {
    double psi1 = 0.5;
    double psi2 = 0.25;
    double dnrm = (ndim == 3) ? sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
                              : sqrt(d[0] * d[0] + d[1] * d[1]);
    double delta1 = psi1 * fabs(phiL - phiR);
    double delta2 = psi2 * fabs(phiL - phiR);
    double phimin = fmin(phiL, phiR);
    double phimax = fmax(phiL, phiR);
    double phibar = phiL + dnrm / r * (phiR - phiL);
    double phiplus;
    if((phimax + delta1) * phimax > 0.)
        phiplus = phimax + delta1;
    else
        phiplus = phimax / (1. + delta1 / fabs(phimax));
    double phiminus;
    if((phimin - delta1) * phimin > 0.)
        phiminus = phimin - delta1;
    else
        phiminus = phimin / (1. + delta1 / fabs(phimin));
    double phimid;
    if(phiL == phiR) {
        phimid = phiL;
    } else if(phiL < phiR) {
        phimid = fmax(phiminus, fmin(phibar + delta2, phimid0));
    } else {
        phimid = fmin(phiplus, fmax(phibar - delta2, phimid0));
    }
    return phimid;
}

/* VorFace::reconstruct, src/VorFace.hpp:174-196 */
void sfxo_reconstruct(int ndim, const double* W, const double* grad,
                      const double* d, double* out) {
    int NC = ndim + 2;
    for(int k = 0; k < NC; k++) {
        if(ndim == 3)
            out[k] = W[k] + (grad[k] * d[0] + grad[5 + k] * d[1] +
                             grad[10 + k] * d[2]);
        else
            out[k] = W[k] + (grad[k] * d[0] + grad[4 + k] * d[1]);
    }
}

/* ------------------------------------------------------ get_W / get_Q -- */

/* get_W (use_energy = true): Exact ExactRiemannSolver.cpp:716-766, HLLC
 * HLLCRiemannSolver.cpp:495-535; both give the same hydro components. */
unsigned sfxo_get_W(const sfxo_solver* s, double volume, double* Q,
                    double* W) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    double g = s->gamma;
    for(int k = 0; k < NC; k++) W[k] = 0.;
    if(!(Q[0] != 0.)) return 0u;
    W[0] = Q[0] / volume;
    for(int d = 0; d < nd; d++) W[1 + d] = Q[1 + d] / Q[0];
    double p2 = Q[1] * Q[1] + Q[2] * Q[2];
    if(nd == 3) p2 = Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3];
    W[P] = (g - 1.) * (Q[P] - 0.5 * p2 / Q[0]) / volume;
    if(W[P] < 1.e-30) {
        W[P] = 1.e-30;
        Q[P] = 0.5 * p2 / Q[0] + W[P] * volume / (g - 1.);
    }
    if(W[0] < 1.e-30) {
        for(int k = 0; k < NC; k++) W[k] = 0.;
    }
    return 0u;
}

/* get_Q: Exact ExactRiemannSolver.cpp:776-804, HLLC HLLCRiemannSolver.cpp:
 * 471-483 (different energy association). */
void sfxo_get_Q(const sfxo_solver* s, double volume, const double* W,
                double* Q) {
    int nd = s->ndim, P = nd + 1;
    double g = s->gamma;
    Q[0] = W[0] * volume;
    for(int d = 0; d < nd; d++) Q[1 + d] = Q[0] * W[1 + d];
    double v2 = W[1] * W[1] + W[2] * W[2];
    if(nd == 3) v2 = W[1] * W[1] + W[2] * W[2] + W[3] * W[3];
    if(s->kind == SFXO_HLLC)
        Q[P] = volume * W[P] / (g - 1.) + 0.5 * Q[0] * v2;
    else
        Q[P] = (0.5 * W[0] * v2 + W[P] / (g - 1.)) * volume;
}

/* ------------------------------------------------------------ mesh ops -- */

/* VorFace::transform / invtransform, src/VorFace.cpp:417-542 */
static void vf_basis(int nd, const double* left, const double* right,
                     double* nn, double* o, double* p) {
    if(nd == 3) {
        double n4[4];
        n4[0] = right[0] - left[0];
        n4[1] = right[1] - left[1];
        n4[2] = right[2] - left[2];
        n4[3] = sqrt(n4[0] * n4[0] + n4[1] * n4[1] + n4[2] * n4[2]);
        nn[0] = n4[0] / n4[3];
        nn[1] = n4[1] / n4[3];
        nn[2] = n4[2] / n4[3];
        if(nn[0] != 0. || nn[1] != 0.) {
            double o3;
            o[0] = -nn[1];
            o[1] = nn[0];
            o[2] = 0.;
            o3 = sqrt(o[0] * o[0] + o[1] * o[1]);
            o[0] /= o3;
            o[1] /= o3;
            p[0] = -nn[2] * o[1];
            p[1] = nn[2] * o[0];
            p[2] = o3;
        } else {
            o[0] = 0.;
            o[1] = 1.;
            o[2] = 0.;
            p[0] = -1.;
            p[1] = 0.;
            p[2] = 0.;
        }
    }
}

static void vf_mirror_velocity(int nd, double* W, const double* left,
                               const double* right) {
    /* transform, negate the normal component, invtransform */
    if(nd == 3) {
        double nn[3], o[3], p[3];
        vf_basis(nd, left, right, nn, o, p);
        double v0 = W[1], v1 = W[2], v2 = W[3];
        double t1 = v0 * nn[0] + v1 * nn[1] + v2 * nn[2];
        double t2 = v0 * o[0] + v1 * o[1];
        double t3 = v0 * p[0] + v1 * p[1] + v2 * p[2];
        t1 = -t1;
        W[1] = t1 * nn[0] + t2 * o[0] + t3 * p[0];
        W[2] = t1 * nn[1] + t2 * o[1] + t3 * p[1];
        W[3] = t1 * nn[2] + t3 * p[2];
    } else {
        double d[3];
        d[0] = left[0] - right[0];
        d[1] = left[1] - right[1];
        d[2] = sqrt(d[0] * d[0] + d[1] * d[1]);
        double cost = -d[0] / d[2];
        double sint = -d[1] / d[2];
        double W1 = W[1] * cost + W[2] * sint;
        W[2] = -W[1] * sint + W[2] * cost;
        W[1] = W1;
        W[1] = -W[1];
        W1 = W[1] * cost - W[2] * sint;
        W[2] = W[1] * sint + W[2] * cost;
        W[1] = W1;
    }
}

#define CELL(arr, k, i, C) (arr)[(int64_t)(k) * (C) + (i)]

static void right_gen_pos(const sfxo_mesh* m, int64_t f, double* xr) {
    int nd = m->ndim;
    int32_t R = m->right[f];
    int64_t base = (R >= 0) ? R : m->left[f];
    for(int d = 0; d < nd; d++)
        xr[d] = CELL(m->pos, d, base, m->C) + m->shift[d * m->F + f];
}

/* VorCell::estimate_gradient, src/VorCell.cpp:464-530, for cells c0..c1-1 */
void sfxo_gradients_range(const sfxo_solver* s, const sfxo_mesh* m, int64_t c0,
                          int64_t c1, double* grad_out) {
    int nd = m->ndim, NC = nd + 2;
    int64_t C = m->C, F = m->F;
    (void)s;
    for(int64_t i = c0; i < c1; i++) {
        double delta[3][5];
        for(int d = 0; d < 3; d++)
            for(int k = 0; k < 5; k++) delta[d][k] = 0.;
        if(m->active && !m->active[i]) continue;  /* keeps its gradients */
        double W[5], Wmax[5], Wmin[5], xc[3];
        for(int k = 0; k < NC; k++) W[k] = Wmax[k] = Wmin[k] = CELL(m->W, k, i, C);
        for(int d = 0; d < nd; d++) xc[d] = CELL(m->pos, d, i, C);
        int64_t j0 = m->cf_off[i], j1 = m->cf_off[i + 1];
        for(int64_t j = j1; j-- > j0;) {
            int64_t f = m->cf_face[j];
            double Aij = m->area[f];
            if(!(Aij != 0.)) continue;
            double xn[3], Wj[5];
            int32_t L = m->left[f], R = m->right[f];
            if(L == i) {
                right_gen_pos(m, f, xn);
            } else {
                for(int d = 0; d < nd; d++) xn[d] = CELL(m->pos, d, L, C);
            }
            double c[3], rij[3];
            for(int d = 0; d < nd; d++) {
                c[d] = m->midpoint[d * F + f] - 0.5 * (xc[d] + xn[d]);
                rij[d] = xc[d] - xn[d];
            }
            double rnorm = (nd == 3) ? sqrt(rij[0] * rij[0] + rij[1] * rij[1] +
                                            rij[2] * rij[2])
                                     : sqrt(rij[0] * rij[0] + rij[1] * rij[1]);
            if(L == i && R < 0) {
                /* mirror neighbour, :482-489 */
                double xl[3];
                for(int d = 0; d < nd; d++) xl[d] = CELL(m->pos, d, L, C);
                for(int k = 0; k < NC; k++) Wj[k] = W[k];
                vf_mirror_velocity(nd, Wj, xl, xn);
            } else {
                int64_t o = (L == i) ? R : L;
                for(int k = 0; k < NC; k++) Wj[k] = CELL(m->W, k, o, C);
            }
            for(int l = nd; l--;) {
                for(int k = 0; k < NC; k++) {
                    double t = (Wj[k] - W[k]) * c[l] / rnorm -
                               0.5 * (W[k] + Wj[k]) * rij[l] / rnorm;
                    delta[l][k] += Aij * t;
                }
            }
            for(int k = 0; k < NC; k++) {
                Wmax[k] = smax(Wmax[k], Wj[k]);
                Wmin[k] = smin(Wmin[k], Wj[k]);
            }
        }
        double V = m->volume[i];
        for(int l = nd; l--;)
            for(int k = 0; k < NC; k++) delta[l][k] /= V;
        double alpha[5] = {1., 1., 1., 1., 1.};
        for(int64_t j = j1; j-- > j0;) {
            int64_t f = m->cf_face[j];
            if(!(m->area[f] != 0.)) continue;
            double dd[3];
            for(int d = 0; d < nd; d++)
                dd[d] = m->midpoint[d * F + f] - CELL(m->centroid, d, i, C);
            for(int k = 0; k < NC; k++) {
                double dp = delta[0][k] * dd[0] + delta[1][k] * dd[1];
                if(nd == 3) dp += delta[2][k] * dd[2];
                /* StateVector operator/ maps x/0 to 1, StateVector.hpp:607 */
                double q1 = (dp != 0.) ? (Wmax[k] - W[k]) / dp : 1.;
                double q2 = (dp != 0.) ? (Wmin[k] - W[k]) / dp : 1.;
                alpha[k] = smin(alpha[k], smax(q1, q2));
            }
        }
        for(int l = nd; l--;)
            for(int k = 0; k < NC; k++)
                grad_out[(int64_t)(l * NC + k) * C + i] = delta[l][k] * alpha[k];
    }
}

void sfxo_gradients(const sfxo_solver* s, const sfxo_mesh* m,
                    double* grad_out) {
    sfxo_gradients_range(s, m, 0, m->C, grad_out);
}

/* VorFace::set_v, src/VorFace.cpp:586-660; rRL = xr - xl and mc = midpoint -
 * (xl + xr)/2 of the face (face_vectors) */
static void face_velocity(const sfxo_mesh* m, int64_t f, const double* xl,
                          const double* xr, const double* rRL,
                          const double* mc, double* vface) {
    int nd = m->ndim;
    int64_t C = m->C;
    int32_t L = m->left[f], R = m->right[f];
    double vl[3], vr[3];
    for(int d = 0; d < nd; d++) vl[d] = CELL(m->vel, d, L, C);
    if(R >= 0) {
        for(int d = 0; d < nd; d++) vr[d] = CELL(m->vel, d, R, C);
    } else {
        if(nd == 3) {
            vface[0] = vface[1] = vface[2] = 0.;
            return;
        }
        double Wd[4] = {0., vl[0], vl[1], 0.};
        vf_mirror_velocity(nd, Wd, xl, xr);
        vr[0] = Wd[1];
        vr[1] = Wd[2];
    }
    double rRL2 = rRL[0] * rRL[0] + rRL[1] * rRL[1];
    if(nd == 3) rRL2 = rRL[0] * rRL[0] + rRL[1] * rRL[1] + rRL[2] * rRL[2];
    double fac = (vl[0] - vr[0]) * mc[0];
    fac += (vl[1] - vr[1]) * mc[1];
    if(nd == 3) fac += (vl[2] - vr[2]) * mc[2];
    fac /= rRL2;
    for(int d = 0; d < nd; d++) vface[d] = 0.5 * (vl[d] + vr[d]) + fac * rRL[d];
}

/* The generator positions of face f's two sides, rRL = xr - xl and mc =
 * midpoint - (xl + xr)/2 (src/VorFace.cpp:586-660); with the stored
 * geometry, xl = 0 and xr = rRL stand for the positions (only their
 * difference is used). */
static void face_vectors(const sfxo_mesh* m, int64_t f, double* xl, double* xr,
                         double* rRL, double* mc) {
    int nd = m->ndim;
    int64_t C = m->C, F = m->F;
    if(m->geo_rRL) {
        for(int d = 0; d < nd; d++) {
            xl[d] = 0.;
            xr[d] = rRL[d] = m->geo_rRL[d * F + f];
            mc[d] = m->geo_mc[d * F + f];
        }
        return;
    }
    for(int d = 0; d < nd; d++) xl[d] = CELL(m->pos, d, m->left[f], C);
    right_gen_pos(m, f, xr);
    for(int d = 0; d < nd; d++) {
        rRL[d] = xr[d] - xl[d];
        mc[d] = m->midpoint[d * F + f] - 0.5 * (xl[d] + xr[d]);
    }
}

/* VorFace::set_v and VorFace::get_normal for every face (testing aid) */
void sfxo_face_velocities(const sfxo_mesh* m, double* vface, double* normal) {
    int nd = m->ndim;
    int64_t F = m->F;
    for(int64_t f = 0; f < F; f++) {
        double xl[3], xr[3], rRL[3], mc[3], v[3], dd[4];
        face_vectors(m, f, xl, xr, rRL, mc);
        face_velocity(m, f, xl, xr, rRL, mc, v);
        for(int k = 0; k < nd; k++) dd[k] = xl[k] - xr[k];
        dd[3] = (nd == 3) ? sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2])
                          : sqrt(dd[0] * dd[0] + dd[1] * dd[1]);
        for(int k = 0; k < nd; k++) {
            vface[k * F + f] = v[k];
            normal[k * F + f] = -dd[k] / dd[3];
        }
    }
}

/* one side's predictor + reconstruction + limiter, src/VorFace.cpp:714-804.
 * Wraw: boosted state of this side, Wother: boosted state of the other side,
 * limited: whether the limiter runs (the face has a right particle). */
static void predict_side(const sfxo_solver* s, const sfxo_mesh* m,
                         int64_t cell, const double* Wraw, const double* Wother,
                         const double* d, double r, int limited, double dtp,
                         double* out) {
    int nd = m->ndim, NC = nd + 2, P = nd + 1;
    int64_t C = m->C;
    double grad[15], dWdt[5], deltaW[5], Wt[5];
    for(int l = 0; l < nd; l++)
        for(int k = 0; k < NC; k++)
            grad[l * 5 + k] = m->grad[(int64_t)(l * NC + k) * C + cell];
    /* get_time_derivative takes grad rows of stride ndim+2 */
    double gpack[15];
    for(int l = 0; l < nd; l++)
        for(int k = 0; k < NC; k++) gpack[l * NC + k] = grad[l * 5 + k];
    sfxo_time_derivative(s, Wraw, gpack, dWdt);
    for(int k = 0; k < NC; k++) deltaW[k] = 0.5 * dtp * dWdt[k];
    sfxo_reconstruct(nd, Wraw, gpack, d, Wt);
    if(limited) {
        double lim[5];
        for(int k = 0; k < NC; k++)
            lim[k] = sfxo_limit(Wt[k], Wraw[k], Wother[k], d, r, nd);
        for(int k = 0; k < NC; k++) Wt[k] = lim[k];
    }
    if(Wt[0] < 0. || Wt[P] < 0.) {
        for(int k = 0; k < NC; k++) Wt[k] = Wraw[k];
    }
    for(int k = NC; k--;) Wt[k] += deltaW[k];
    if(Wt[0] < 0. || Wt[P] < 0.) {
        for(int k = 0; k < NC; k++) Wt[k] -= deltaW[k];
    }
    for(int k = 0; k < NC; k++) out[k] = Wt[k];
}

/* VorFace::set_v and VorFace::calculate_flux (src/VorFace.cpp:586-866) for
 * face f of a single process (every right generator is local, a periodic
 * copy or a mirror), no gravity: dQLo = the flux exchanged, seen from the
 * left cell (zero when the face is skipped).  Returns the status bits. */
static unsigned face_flux_one(const sfxo_solver* s, const sfxo_mesh* m,
                              int64_t f, double* dQLo, sfxo_counters* c,
                              double* wstate) {
    int nd = m->ndim, NC = nd + 2;
    int64_t C = m->C, F = m->F;
    unsigned st = 0;
    int32_t L = m->left[f], R = m->right[f];
    for(int k = 0; k < NC; k++) dQLo[k] = 0.;
    double xl[3], xr[3], rRL[3], mc[3], vface[3];
    face_vectors(m, f, xl, xr, rRL, mc);
    face_velocity(m, f, xl, xr, rRL, mc, vface);

    /* gating, :672-683 */
    if(!(m->area[f] != 0.)) return 0u;
    int actL = m->active ? m->active[L] : 1;
    int actR = (R >= 0) ? (m->active ? m->active[R] : 1) : 0;
    if(R < 0 && !actL) return 0u;
    if(R >= 0 && !actR && !actL) return 0u;

    /* dt, :695-699 */
    double dt = m->dt[L];
    if(R >= 0) dt = smin(dt, m->dt[R]);

    double WL[5], WR[5];
    for(int k = 0; k < NC; k++) {
        WL[k] = CELL(m->W, k, L, C);
        WR[k] = (R >= 0) ? CELL(m->W, k, R, C) : 0.;
    }
    for(int d = 0; d < nd; d++) {
        WL[1 + d] -= vface[d];
        WR[1 + d] -= vface[d];
    }

    double r = 0.;
    if(R >= 0 && m->geo_rRL) {
        r = m->geo_rcent[f];
    } else if(R >= 0) {
        double cd[3];
        for(int d = 0; d < nd; d++)
            cd[d] = CELL(m->centroid, d, L, C) - CELL(m->centroid, d, R, C);
        r = (nd == 3) ? sqrt(cd[0] * cd[0] + cd[1] * cd[1] + cd[2] * cd[2])
                      : sqrt(cd[0] * cd[0] + cd[1] * cd[1]);
    }
    double WRrec[5] = {0., 0., 0., 0., 0.}, WLrec[5];
    double d[3];
    if(R >= 0) {
        for(int n = nd; n--;) {
            if(m->geo_rRL) {
                d[n] = m->geo_dR[n * F + f];
                continue;
            }
            double correction = CELL(m->pos, n, R, C) - xr[n];
            d[n] = m->midpoint[n * F + f] - CELL(m->centroid, n, R, C) +
                   correction;
        }
        predict_side(s, m, R, WR, WL, d, r, 1, dt, WRrec);
    }
    for(int n = nd; n--;) {
        if(m->geo_rRL) {
            d[n] = m->geo_dL[n * F + f];
            continue;
        }
        double correction = CELL(m->pos, n, L, C) - xl[n];
        d[n] = m->midpoint[n * F + f] - CELL(m->centroid, n, L, C) +
               correction;
    }
    predict_side(s, m, L, WL, WR, d, r, R >= 0, dt, WLrec);
    if(WL[0] != 0.) memcpy(WL, WLrec, sizeof(double) * NC);
    if(WR[0] != 0.) memcpy(WR, WRrec, sizeof(double) * NC);
    if(wstate) {
        memcpy(wstate, WL, sizeof(double) * NC);
        memcpy(wstate + NC, WR, sizeof(double) * NC);
    }

    /* get_normal, :553-571 */
    double nrm[3], dd[4];
    for(int k = 0; k < nd; k++) dd[k] = xl[k] - xr[k];
    dd[3] = (nd == 3) ? sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2])
                      : sqrt(dd[0] * dd[0] + dd[1] * dd[1]);
    for(int k = 0; k < nd; k++) nrm[k] = -dd[k] / dd[3];

    double flux[5];
    st |= sfxo_solve_for_flux(s, WL, WR, nrm, vface, R < 0, flux, c);
    double dQL[5], dQR[5];
    for(int k = 0; k < NC; k++) {
        dQL[k] = dt * m->area[f] * flux[k];
        dQR[k] = -dt * m->area[f] * flux[k];
    }
    /* mass flux limiter, :840-866 */
    double timefac = 1.;
    if(R >= 0) {
        double dtmax = m->dt[L];
        dtmax = smax(dtmax, m->dt[R]);
        timefac = dt / dtmax;
    }
    double areafac = m->area[f] / m->total_area[L];
    areafac *= timefac;
    if(dQL[0] > areafac * m->mass[L]) {
        double corfac = fabs(areafac * m->mass[L] / dQL[0]);
        for(int k = 0; k < NC; k++) dQL[k] *= corfac;
        for(int k = 0; k < NC; k++) dQR[k] = -1. * dQL[k];
    }
    if(R >= 0) {
        double af = m->area[f] / m->total_area[R];
        af *= timefac;
        if(dQR[0] > af * m->mass[R]) {
            double corfac = fabs(af * m->mass[R] / dQR[0]);
            for(int k = 0; k < NC; k++) dQR[k] *= corfac;
            for(int k = 0; k < NC; k++) dQL[k] = -1. * dQR[k];
        }
    }
    if(flux[0] != flux[0]) st |= SFXO_ST_NANFLUX;
    for(int k = 0; k < NC; k++) dQLo[k] = dQL[k];
    return st;
}

/* The scatter of one face (src/VorFace.cpp:913-955, single process): the
 * left cell always takes dQL; the right cell -dQL when it is a local real
 * cell (not a periodic copy, not a mirror).  dQR is bitwise -dQL
 * (:834-835, :855, :864). */
static void scatter_one(const sfxo_mesh* m, int64_t f, const double* dQL,
                        double* dQ) {
    int NC = m->ndim + 2;
    int64_t C = m->C;
    int32_t L = m->left[f], R = m->right[f];
    for(int k = 0; k < NC; k++) CELL(dQ, k, L, C) += dQL[k];
    if(R >= 0 && !(m->ghost && m->ghost[f]))
        for(int k = 0; k < NC; k++) CELL(dQ, k, R, C) += -dQL[k];
}

/* VorTessManager::hydro face loop (src/VorTessManager.cpp:205-210) calling
 * VorFace::set_v and VorFace::calculate_flux for every face in order, each
 * face's flux scattered as soon as it is computed. */
unsigned sfxo_face_flux_pass(const sfxo_solver* s, const sfxo_mesh* m,
                             double* dQ, double* dQL_face, uint8_t* face_status,
                             sfxo_counters* c) {
    int NC = m->ndim + 2;
    int64_t F = m->F;
    unsigned all = 0;
    for(int64_t f = 0; f < F; f++) {
        double dQL[5];
        unsigned st = face_flux_one(s, m, f, dQL, c, NULL);
        scatter_one(m, f, dQL, dQ);
        if(dQL_face)
            for(int k = 0; k < NC; k++) dQL_face[(int64_t)k * F + f] = dQL[k];
        if(face_status) face_status[f] = (uint8_t)st;
        all |= st;
    }
    return all;
}

/* Faces f0..f1-1 without the scatter: dQL_face [(ndim+2)][F] and
 * face_status [F] (may be NULL) of those faces.  For running the pass on
 * several threads; sfxo_gather then applies the fluxes in face order. */
unsigned sfxo_face_flux_range(const sfxo_solver* s, const sfxo_mesh* m,
                              int64_t f0, int64_t f1, double* dQL_face,
                              uint8_t* face_status, sfxo_counters* c) {
    int NC = m->ndim + 2;
    int64_t F = m->F;
    unsigned all = 0;
    for(int64_t f = f0; f < f1; f++) {
        double dQL[5];
        unsigned st = face_flux_one(s, m, f, dQL, c, NULL);
        for(int k = 0; k < NC; k++) dQL_face[(int64_t)k * F + f] = dQL[k];
        if(face_status) face_status[f] = (uint8_t)st;
        all |= st;
    }
    return all;
}

/* The states every face hands its Riemann solver: left and right, in the
 * face frame, after the predictor, limiter and vacuum keep (zeros for a
 * skipped face; a mirror face's right state is zero, the solver mirrors the
 * left one).  WLo, WRo: [(ndim+2)][F].  Testing aid. */
void sfxo_face_states(const sfxo_solver* s, const sfxo_mesh* m, double* WLo,
                      double* WRo) {
    int NC = m->ndim + 2;
    int64_t F = m->F;
    sfxo_counters c;
    memset(&c, 0, sizeof c);
    for(int64_t f = 0; f < F; f++) {
        double dQL[5], w[10];
        memset(w, 0, sizeof w);
        face_flux_one(s, m, f, dQL, &c, w);
        for(int k = 0; k < NC; k++) {
            WLo[(int64_t)k * F + f] = w[k];
            WRo[(int64_t)k * F + f] = w[NC + k];
        }
    }
}

/* dQ of cells c0..c1-1 from per-face fluxes: the cell's gather entries
 * (2 face + side, ascending face order) added in order, -dQL for side 1 --
 * the same additions, in the same order, as the scatter of the face loop
 * (dQ zeroed first). */
void sfxo_gather(const sfxo_mesh* m, const int64_t* gather_off,
                 const int32_t* gather_entry, const double* dQL_face,
                 int64_t c0, int64_t c1, double* dQ) {
    int NC = m->ndim + 2;
    int64_t C = m->C, F = m->F;
    for(int64_t i = c0; i < c1; i++) {
        double acc[5] = {0., 0., 0., 0., 0.};
        for(int64_t j = gather_off[i]; j < gather_off[i + 1]; j++) {
            int32_t e = gather_entry[j];
            int64_t f = e >> 1;
            for(int k = 0; k < NC; k++) {
                double x = dQL_face[(int64_t)k * F + f];
                acc[k] += (e & 1) ? -x : x;
            }
        }
        for(int k = 0; k < NC; k++) CELL(dQ, k, i, C) = acc[k];
    }
}

/* GasParticle::update_Q (src/utilities/GasParticle.cpp:617-624) for every
 * cell (the loop at src/Simulation.cpp:478-483), then get_W for the cells
 * whose step ends at the new current time (src/Simulation.cpp:792-796) */
unsigned sfxo_update(const sfxo_solver* s, int64_t C, double* Q, double* dQ,
                     const double* volume, double* W, const uint64_t* endtime,
                     uint64_t tnow, uint8_t* cell_status) {
    int nd = s->ndim, NC = nd + 2, P = nd + 1;
    unsigned all = 0;
    for(int64_t i = 0; i < C; i++) {
        double q[5], w[5];
        unsigned st = 0;
        for(int k = 0; k < NC; k++) {
            q[k] = CELL(Q, k, i, C) - CELL(dQ, k, i, C);
            CELL(dQ, k, i, C) = 0.;
        }
        if(q[0] < 0. || q[P] < 0.) st |= SFXO_ST_UPDATEQ;
        for(int k = 0; k < NC; k++) CELL(Q, k, i, C) = q[k];
        if(!endtime || endtime[i] == tnow) {
            sfxo_get_W(s, volume[i], q, w);
            for(int k = 0; k < NC; k++) {
                CELL(Q, k, i, C) = q[k];
                CELL(W, k, i, C) = w[k];
            }
        }
        if(cell_status) cell_status[i] = (uint8_t)st;
        all |= st;
    }
    return all;
}

/* ------------------------------------------------------------ time step -- */

/* TimeLine::get_realtime / get_realtime_interval (src/TimeLine.hpp:145-166),
 * no cosmology */
static double tl_realtime(const sfxo_timeline* tl, uint64_t t) {
    double x = (double)t / (double)tl->integer_maxtime;
    return tl->tottime * x + tl->mintime;
}
static double tl_interval(const sfxo_timeline* tl, uint64_t t) {
    double x = (double)t / (double)tl->integer_maxtime;
    return tl->tottime * x;
}

/* TimeLine::calculate_timestep, gas branch (src/TimeLine.cpp:233-320) and the
 * global-timestep tail (:418-431) */
unsigned sfxo_hydro_timestep(const sfxo_solver* s, int64_t C, const double* W,
                             const double* vel, const double* volume,
                             const sfxo_timeline* tl, uint64_t* starttime,
                             uint64_t* endtime, double* dt, uint8_t* active,
                             uint64_t* system_step, uint64_t* nactive,
                             uint8_t* cell_status) {
    const int nd = s->ndim, NC = nd + 2, P = nd + 1;
    const uint64_t Smax = tl->max_timestep;
    const uint64_t now = tl->current_time + tl->timestep;
    uint64_t Smin = Smax, na = 0;
    unsigned st_all = 0;
    for(int64_t i = 0; i < C; i++) {
        uint8_t st = 0;
        const int act = endtime[i] == now;
        active[i] = (uint8_t)act;
        if(act) {
            na++;
            const double* w = W + i * NC;
            /* v = W.v - particle velocity; Vec::norm (src/Vec.hpp:140,378) */
            double vx = w[1] - vel[i * nd], vy = w[2] - vel[i * nd + 1];
            double vi;
            if(nd == 3) {
                double vz = w[3] - vel[i * nd + 2];
                vi = sqrt(vx * vx + vy * vy + vz * vz);
            } else {
                vi = sqrt(vx * vx + vy * vy);
            }
            /* get_soundspeed: ExactRiemannSolver.cpp:681-687 guards rho = 0,
             * HLLCRiemannSolver.cpp:458-460 does not */
            double ci;
            if(s->kind == SFXO_EXACT && !(w[0] != 0.))
                ci = 0.;
            else
                ci = sqrt(s->gamma * w[P] / w[0]);
            /* VorCell::set_h, src/VorCell.cpp:930-938 */
            const double V = volume[i];
            const double h = (nd == 3) ? cbrt(V * 3. / 4. / M_PI) : sqrt(V / M_PI);
            const double rdt = h / (vi + ci);
            const double S = tl->cfl * rdt;
            uint64_t S2;
            if(tl_interval(tl, Smax) < S) {
                S2 = Smax;
            } else {
                S2 = tl->integer_maxtime;
                while(tl_interval(tl, S2) > S) S2 >>= 1;
                if(!S2) {
                    st = SFXO_ST_TIMESTEP;
                } else {
                    while((tl->integer_maxtime - endtime[i]) % S2) S2 >>= 1;
                    S2 = (S2 < tl->min_timestep) ? tl->min_timestep : S2;
                    S2 = (S2 > tl->max_timestep) ? tl->max_timestep : S2;
                }
            }
            if(!st) {
                Smin = (S2 < Smin) ? S2 : Smin;
                if(!tl->global_timestep) {
                    starttime[i] = endtime[i];
                    endtime[i] = starttime[i] + S2;
                }
            }
        }
        if(cell_status) cell_status[i] = st;
        st_all |= st;
    }
    if(tl->global_timestep) {
        for(int64_t i = 0; i < C; i++) {
            starttime[i] = endtime[i];
            endtime[i] = starttime[i] + Smin;
        }
    }
    for(int64_t i = 0; i < C; i++) dt[i] = tl_realtime(tl, endtime[i] - starttime[i]);
    *system_step = Smin;
    *nactive = na;
    return st_all;
}
