// riemann_mixed_impl.h -- mixed-precision Riemann solver device code for
// gfx950 (included by riemann_mixed.hip and flux_pass_mixed.hip).
//
// Storage type T is float (52 B in + 20 B out per 3D problem) or double.
//
// Exact solver (same problem as Shadowfax ExactRiemannSolver::solve_for_flux,
// src/riemann/ExactRiemannSolver.cpp:108-290,815-880): the pressure
// iteration -- the transcendental-heavy bulk of the work -- runs in FP32 on
// the VALU's fast paths (v_log_f32 / v_exp_f32 / v_rcp_f32 / v_sqrt_f32);
// one Newton step in FP64 then gives p* and u* to ~1e-13, and the sampling
// of the solution at x/t = 0 and the flux assembly run in FP64 (a few dozen
// operations; at Mach numbers ~1e3 the flux is too ill-conditioned for an
// FP32 sampled state to meet 1e-6).
//   * f(p) = fL + fR + du is monotone increasing and concave, so Newton from
//     the left of the root converges monotonically; a bracket [lo, hi]
//     catches a first step from the right that leaves it (no serial Brent
//     state machine, a bounded and nearly uniform iteration count per wave);
//   * the reference's pure-Newton scheme (f(0) < cutoff) stops at p* = 0
//     when its first step goes negative; that outcome is reproduced;
//   * pow(x, -(g+1)/(2g)) = pow(x, (g-1)/(2g)) / x: one transcendental per
//     side per step; for g = 5/3 every other power is an integer power of
//     x^(1/5), and the FP64 fifth root is an FP32 seed plus one Newton step.
//
// HLLC (HLLCRiemannSolver::solve_for_flux, src/riemann/HLLCRiemannSolver.cpp:
// 258-403): HBM-bound, so FP32 storage carries the speedup and the
// arithmetic runs in FP64 registers (the reference's expressions, shared
// with the plain path through riemann_fp64_impl.h) at no cost in time.
#pragma once

#include "riemann_fp64_impl.h"
#include "sfx_device.h"

namespace sfx {
namespace mixed {

// FP32 consts
// The FP32 iteration stops once a Newton step moves p by at most this
// fraction of (p_new + p_old); the FP64 step that follows squares the
// remaining error.  Measured on the 4M-problem mixed batch and the vacuum,
// shock and mixed families: the error against FP64 is the same (5.96e-8,
// the FP32 output rounding) for 2e-6, 1e-5, 1e-4 and 1e-3; 1e-4 saves ~8%
// of the solve and leaves margin where Newton converges only linearly
// (p* near 0).
#ifndef SFX_MIXED_STEP_TOL
#define SFX_MIXED_STEP_TOL 1e-4f
#endif

struct CF {
    float gamma, gp1d2g, gm1d2g, gm1dgp1, tdgp1, tdgm1, gm1d2, tgdgm1, ginv, ginvm1;
    float gm1inv;      // 1/(gamma-1)
    float q_fac;       // 0.5*(gamma+1)/gamma (HLLC)
    double gamma_d, tdgp1_d, tdgm1_d, gm1dgp1_d, gm1d2g_d, cutoff_d;
    float halftol;     // 0.5*tol: floor of the initial guess
    int fast53;
    int itmax;
};

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrtf_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float log2f_fast(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float exp2f_fast(float x) { return __builtin_amdgcn_exp2f(x); }
// x^e for x >= 0 (x = 0 gives 0 for e > 0, inf for e < 0)
__device__ __forceinline__ float powf_fast(float x, float e) {
    return exp2f_fast(e * log2f_fast(x));
}
__device__ __forceinline__ float ipow3(float b) { return b * b * b; }
__device__ __forceinline__ float ipow5(float b) {
    float b2 = b * b;
    return b2 * b2 * b;
}

// 1/x in FP64 from the hardware estimate v_rcp_f64 and two Newton-Raphson
// steps r <- r + r (1 - x r) (each step doubles the correct bits): a few
// FP64 FMAs instead of the full division sequence, within an ulp or two of
// 1/x for normal x; 1/0 = inf and 1/inf = 0 as for a division.
__device__ __forceinline__ double rcp_nr(double x) {
    const double r0 = __builtin_amdgcn_rcp(x);
    double r = fma(fma(-x, r0, 1.0), r0, r0);
    r = fma(fma(-x, r, 1.0), r, r);
    // 1/0 and 1/inf: the estimate is exact and the steps would give NaN
    return (r0 == 0.0 || __builtin_isinf(r0)) ? r0 : r;
}

// sqrt(x) in FP64 for x >= 0 from the hardware estimate y = v_rsq_f64(x)
// and one Newton-Raphson step on y = 1/sqrt(x), y <- y (3/2 - x/2 y^2)
// (relative error about 3/2 of the square of the estimate's), then x y.
__device__ __forceinline__ double sqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return (x == 0.0) ? 0.0 : x * y;
}

// x^(1/5) in FP64 for x > 0: exponent split so the FP32 seed stays in range,
// then one Newton step on y^5 = z.
__device__ __forceinline__ double fifth_root_d(double x) {
    int e = __builtin_amdgcn_frexp_exp(x);      // x = m 2^e, m in [0.5, 1)
    double m = __builtin_amdgcn_frexp_mant(x);
    int q = (e >= 0) ? e / 5 : -((-e + 4) / 5);  // floor(e/5)
    int r = e - 5 * q;                           // 0..4
    double z = __builtin_ldexp(m, r);            // [0.5, 16)
    double y = (double)exp2f_fast(0.2f * log2f_fast((float)z));
    double y2 = y * y;
    y = 0.8 * y + 0.2 * z * rcp_nr(y2 * y2);
    return __builtin_ldexp(y, q);
}

// x^gm1d2g in FP64
__device__ __forceinline__ double pow_gm1d2g_d(const CF& c, double x) {
    if(c.fast53) return fifth_root_d(x);
    return pow(x, c.gm1d2g_d);
}

template <int ND, typename T>
struct IO {
    static constexpr int NC = ND + 2;
    static constexpr int P = ND + 1;
};

// One side of the pressure function, FP32: value and derivative.
struct SideF {
    float rho, p, a, A, B, rpinv, tda, ra_inv;
};

__device__ __forceinline__ SideF make_side(const CF& c, float rho, float p, float a) {
    SideF s;
    s.rho = rho;
    s.p = p;
    s.a = a;
    s.A = c.tdgp1 * rcpf(rho);
    s.B = c.gm1dgp1 * p;
    s.rpinv = rcpf(p);
    s.tda = c.tdgm1 * a;
    s.ra_inv = rcpf(rho * a);
    return s;
}

__device__ __forceinline__ void side_f(const CF& c, const SideF& s, float p, float& f, float& df) {
    if(p > s.p) {
        float r = rcpf(p + s.B);
        float sq = sqrtf_fast(s.A * r);
        float dp = p - s.p;
        f = dp * sq;
        df = (1.f - 0.5f * dp * r) * sq;
    } else {
        float x = p * s.rpinv;
        float y = powf_fast(x, c.gm1d2g);
        f = s.tda * (y - 1.f);
        df = s.ra_inv * y * rcpf(x);
    }
}

// FP64 side evaluation at p for the final Newton step
__device__ __forceinline__ void side_f_d(const CF& c, double rho, double pK, double a, double p,
                                         double& f, double& df) {
    if(p > pK) {
        double A = c.tdgp1_d * rcp_nr(rho);
        double B = c.gm1dgp1_d * pK;
        double r = rcp_nr(p + B);
        double sq = sqrt_nr(A * r);
        double dp = p - pK;
        f = dp * sq;
        df = (1.0 - 0.5 * dp * r) * sq;
    } else {
        double x = p * rcp_nr(pK);
        double y = pow_gm1d2g_d(c, x);
        f = c.tdgm1_d * a * (y - 1.0);
        df = y * rcp_nr(rho * a * x);
    }
}

// x^e in FP64 for the sampling: x^(1/5) powers for g = 5/3, pow otherwise
__device__ __forceinline__ double ipow3d(double b) { return b * b * b; }
__device__ __forceinline__ double ipow5d(double b) {
    double b2 = b * b;
    return b2 * b2 * b;
}

// Flux of the sampled face-frame state (rho, s[ND], p) through a face moving
// with v, FP64: F = F(W + v).n with the mass flux relative to the face.
// Same quantity as ExactRiemannSolver::solve_for_flux's per-axis sum
// (ExactRiemannSolver.cpp:863-877), evaluated in projected form.
template <int ND>
__device__ __forceinline__ void state_flux_d(double gm1inv, double rho, const double* s,
                                             double p, const double* n, const double* v,
                                             double* F) {
    const int NC = ND + 2, P = ND + 1;
    if(!(rho != 0.0)) {
#pragma unroll
        for(int k = 0; k < NC; k++) F[k] = 0.0;
        return;
    }
    double w[ND];
    double rel = 0.0, vn = 0.0, w2 = 0.0;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        w[d] = s[d] + v[d];
        rel += s[d] * n[d];
        vn += v[d] * n[d];
        w2 += w[d] * w[d];
    }
    const double F0 = rho * rel;
    F[0] = F0;
#pragma unroll
    for(int d = 0; d < ND; d++) F[1 + d] = F0 * w[d] + p * n[d];
    const double E = 0.5 * rho * w2 + p * gm1inv;
    F[P] = rel * E + p * (rel + vn);
}

// Exact Riemann solver, mixed precision.  WL/WR: FP32 copies for the
// iteration; WL64/WR64/n64/v64: the same values in FP64.  F: FP64 flux.
// The exact solver in three parts so that a batch can pause the FP32
// iteration of its slow problems and resume them elsewhere (wave-level
// compaction): exact_setup (normal velocities, sound speeds, the vacuum
// test, the initial guess), exact_iterate (bracketed Newton from a saved
// state, up to a given iteration count) and exact_finish (the FP64 Newton
// step, sampling and flux).  exact_mixed composes them.

enum : uint8_t { EX_ITER = 0, EX_DONE = 1, EX_STUCK = 2, EX_VACUUM = 3 };

struct ExactCase {
    double uLd, uRd, aLd, aRd;  // FP64 normal velocities and sound speeds
    SideF sL, sR;               // FP32 side data of the pressure function
    float du;                   // uR - uL
    bool pure_newton;           // f(0) < cutoff: the reference's Newton-only scheme
    float p, lo, hi;            // iterate and bracket
    int it;                     // iterations done
    uint8_t phase;
};

template <int ND>
__device__ __forceinline__ void exact_setup(const CF& c, const Consts& cd, const float* WL,
                                            float* WR, const float* n, bool reflective,
                                            const double* WL64, double* WR64,
                                            const double* n64, ExactCase& s) {
    const int NC = ND + 2, P = ND + 1;
    float uL = 0.f, uR = 0.f;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        uL += WL[1 + d] * n[d];
        uR += WR[1 + d] * n[d];
    }
    if(reflective) {
#pragma unroll
        for(int k = 0; k < NC; k++) {
            WR[k] = WL[k];
            WR64[k] = WL64[k];
        }
        uR = -uL;
    }
    const bool vacuum_in = !(WL64[0] != 0.0) || !(WR64[0] != 0.0);
    const float aL = (WL[0] != 0.f) ? sqrtf_fast(c.gamma * WL[P] * rcpf(WL[0])) : 0.f;
    const float aR = (WR[0] != 0.f) ? sqrtf_fast(c.gamma * WR[P] * rcpf(WR[0])) : 0.f;
    // FP64 sound speeds and normal velocities: branch decisions, the final
    // Newton step and the sampling use them
    double uLd = 0.0, uRd = 0.0;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        uLd += WL64[1 + d] * n64[d];
        uRd += WR64[1 + d] * n64[d];
    }
    if(reflective) uRd = -uLd;
    const double aLd = (WL64[0] != 0.0) ? sqrt_nr(cd.gamma * WL64[P] * rcp_nr(WL64[0])) : 0.0;
    const double aRd = (WR64[0] != 0.0) ? sqrt_nr(cd.gamma * WR64[P] * rcp_nr(WR64[0])) : 0.0;
    s.uLd = uLd;
    s.uRd = uRd;
    s.aLd = aLd;
    s.aRd = aRd;
    s.it = 0;
    if(vacuum_in || cd.tdgm1 * aLd + cd.tdgm1 * aRd <= uRd - uLd) {
        s.phase = EX_VACUUM;  // closed forms, no iteration
        return;
    }
    s.phase = EX_ITER;
    s.sL = make_side(c, WL[0], WL[P], aL);
    s.sR = make_side(c, WR[0], WR[P], aR);
    const float du = uR - uL;
    s.du = du;
    // f(0) decides between the reference's two iteration schemes
    // (ExactRiemannSolver.cpp:143-176)
    const double f0d = -(cd.tdgm1 * aLd) + -(cd.tdgm1 * aRd) + (uRd - uLd);
    s.pure_newton = f0d < cd.cutoff;
    // initial guess, ExactRiemannSolver.cpp:644-673 (two-rarefaction /
    // primitive-variable / two-shock)
    const float pmin = fminf(WL[P], WR[P]), pmax = fmaxf(WL[P], WR[P]);
    const float psum = WL[P] + WR[P];
    float ppv = 0.5f * psum - 0.125f * du * (WL[0] + WR[0]) * (aL + aR);
    const float pfloor = c.halftol * psum;
    ppv = fmaxf(pfloor, ppv);
    float p;
    if(pmax <= 2.f * pmin && pmin <= ppv && ppv <= pmax) {
        p = ppv;
    } else if(ppv < pmin) {
        float num = aL + aR - c.gm1d2 * du;
        float den = aL * powf_fast(WL[P], -c.gm1d2g) + aR * powf_fast(WR[P], -c.gm1d2g);
        float z = num * rcpf(den);
        p = c.fast53 ? ipow5(z) : powf_fast(z, c.tgdgm1);
    } else {
        float gL = sqrtf_fast(s.sL.A * rcpf(ppv + s.sL.B));
        float gR = sqrtf_fast(s.sR.A * rcpf(ppv + s.sR.B));
        p = (gL * WL[P] + gR * WR[P] - du) * rcpf(gL + gR);
    }
    s.p = fmaxf(pfloor, p);
    s.lo = 0.f;
    s.hi = __builtin_inff();
}

// Bracketed Newton in FP32 until convergence, the iteration cap, or `stop`
// iterations done (then s.phase stays EX_ITER and the state can be resumed).
// In the reference's pure-Newton scheme a first step below zero is clamped
// to p = 0, where f' is infinite, and the iteration stops there with p* = 0
// (ExactRiemannSolver.cpp:146-154); that outcome is reproduced.
__device__ __forceinline__ void exact_iterate(const CF& c, ExactCase& s, int stop) {
    float p = s.p, lo = s.lo, hi = s.hi;
    int it = s.it;
    uint8_t phase = EX_ITER;
    for(; it < c.itmax; it++) {
        if(it >= stop) break;
        float fl, dl, fr, dr;
        side_f(c, s.sL, p, fl, dl);
        side_f(c, s.sR, p, fr, dr);
        const float fv = fl + fr + s.du;
        if(fv < 0.f) lo = p;
        else hi = p;
        float pn = p - fv * rcpf(dl + dr);
        if(it == 0 && s.pure_newton && fv > 0.f && pn < 0.f) {
            phase = EX_STUCK;
            break;
        }
        if(!(pn >= lo && pn <= hi))
            pn = (hi < __builtin_inff()) ? ((lo > 0.f) ? sqrtf_fast(lo) * sqrtf_fast(hi)
                                                       : 0.01f * hi)
                                         : 2.f * p;
        const bool done = fabsf(pn - p) <= SFX_MIXED_STEP_TOL * (pn + p);
        p = pn;
        if(done) {
            phase = EX_DONE;
            break;
        }
    }
    if(it >= c.itmax) phase = EX_DONE;
    s.p = p;
    s.lo = lo;
    s.hi = hi;
    s.it = it;
    s.phase = phase;
}

// The half state at x/t = 0 of a Riemann problem with vacuum on one side or
// vacuum generated in between: ExactRiemannSolver::solve_vacuum
// (ExactRiemannSolver.cpp:304-446) and the state part of
// HLLCRiemannSolver::vacuum_flux (HLLCRiemannSolver.cpp:44-209) decide the
// same cases in the same order (the two files spell the comparisons with
// their operands swapped).  Mixed path: the case is decided first and the
// rarefaction-fan state formed once, with rcp_nr for the quotients and, for
// gamma = 5/3, the fan powers 2/(g-1) = 3 and 2g/(g-1) = 5 as products
// (pow otherwise).  In a batch with vacuum problems nearly every wavefront
// holds one, so this path's cost is paid by almost every wave.
template <int ND>
__device__ __forceinline__ void vacuum_state_mixed(const CF& c, const Consts& cd,
                                                   const double* WL, const double* WR, double vL,
                                                   double vR, double aL, double aR,
                                                   const double* n, double* sol) {
    constexpr int NC = ND + 2, P = ND + 1;
    // 0: zero state, 1: WL, 2: WR; fan: the rarefaction fan of that side
    int which = 0;
    bool fan = false;
    const bool rl = WL[0] != 0., rr = WR[0] != 0.;
    if(!rr && !rl) {
        which = 0;
    } else if(!rr) {
        which = 1;
        if(vL < aL) {
            fan = vL + cd.tdgm1 * aL > 0.;
            which = fan ? 1 : 0;
        }
    } else if(!rl) {
        which = 2;
        if(-aR < vR) {
            fan = !(vR - cd.tdgm1 * aR >= 0.);
            which = fan ? 2 : 0;
        }
    } else {
        const double SR = vR - cd.tdgm1 * aR, SL = vL + cd.tdgm1 * aL;
        if(SR > 0. && SL < 0.) {
            which = 0;
        } else if(SL >= 0.) {
            which = 1;
            fan = aL > vL;
        } else {
            which = 2;
            fan = -aR < vR;
        }
    }
    const double* WK = (which == 2) ? WR : WL;
    double vhalf = 0.;
#pragma unroll
    for(int k = 0; k < NC; k++) sol[k] = (which == 0) ? 0. : WK[k];
    if(fan) {
        const double sgn = (which == 1) ? 1. : -1.;
        const double aK = (which == 1) ? aL : aR, vK = (which == 1) ? vL : vR;
        const double base = cd.tdgp1 + sgn * cd.gm1dgp1 * rcp_nr(aK) * vK;
        vhalf = cd.tdgp1 * (sgn * aK + cd.gm1d2 * vK) - vK;
        sol[0] = WK[0] * (c.fast53 ? ipow3d(base) : pow(base, cd.tdgm1));
        sol[P] = WK[P] * (c.fast53 ? ipow5d(base) : pow(base, cd.tgdgm1));
    }
#pragma unroll
    for(int d = 0; d < ND; d++) sol[1 + d] += vhalf * n[d];
}

// FP64 final Newton step, sampling at x/t = 0 and flux; returns the status.
template <int ND>
__device__ __forceinline__ uint8_t exact_finish(const CF& c, const Consts& cd, const ExactCase& s,
                                                const double* WL64, const double* WR64,
                                                const double* n64, const double* v64,
                                                double* F, unsigned& iters, unsigned& vac) {
    const int NC = ND + 2, P = ND + 1;
    const double uLd = s.uLd, uRd = s.uRd, aLd = s.aLd, aRd = s.aRd;
    double sol[NC];
    if(s.phase == EX_VACUUM) {
        // closed forms: the FP64 reference expressions
        vac = 1;
        vacuum_state_mixed<ND>(c, cd, WL64, WR64, uLd, uRd, aLd, aRd, n64, sol);
    } else {
        iters += s.it + 1;
        double ps, us;
        if(s.phase == EX_STUCK) {
            // p* = 0: fK(0) = -2 aK/(g-1)
            ps = 0.0;
            us = 0.5 * (uLd + uRd) + 0.5 * (-(cd.tdgm1 * aRd) - -(cd.tdgm1 * aLd));
        } else {
            // FP64 final Newton step for p* (and u*)
            const double pd = (double)s.p;
            double fLd, dLd, fRd, dRd;
            side_f_d(c, WL64[0], WL64[P], aLd, pd, fLd, dLd);
            side_f_d(c, WR64[0], WR64[P], aRd, pd, fRd, dRd);
            const double step = (fLd + fRd + (uRd - uLd)) * rcp_nr(dLd + dRd);
            ps = pd - step;
            if(!(ps > 0.0)) ps = pd;  // keep the FP32 root if the step would leave (0, inf)
            const double delta = ps - pd;
            us = 0.5 * (uLd + uRd) + 0.5 * ((fRd + dRd * delta) - (fLd + dLd * delta));
        }
        // sampling at x/t = 0 in FP64, ExactRiemannSolver.cpp:190-271
        const bool right = us < 0.0;
        const double* WK = right ? WR64 : WL64;
        const double aK = right ? aRd : aLd;
        const double uK = right ? uRd : uLd;
        const double sgn = right ? 1.0 : -1.0;  // right-going (+) or left-going (-) wave
        const double pK = WK[P];
#pragma unroll
        for(int k = 0; k < NC; k++) sol[k] = WK[k];
        double vhalf = 0.0;
        const double pdp = ps * rcp_nr(pK);
        if(ps > pK) {
            const double S = uK + sgn * aK * sqrt_nr(cd.gp1d2g * pdp + cd.gm1d2g);
            if(sgn * S > 0.0) {
                sol[0] = WK[0] * (pdp + cd.gm1dgp1) * rcp_nr(cd.gm1dgp1 * pdp + 1.);
                sol[P] = ps;
                vhalf = us - uK;
            }
        } else {
            const double SH = uK + sgn * aK;
            if(sgn * SH > 0.0) {
                const double y = (pdp > 0.0) ? pow_gm1d2g_d(c, pdp) : 0.0;
                const double ST = us + sgn * aK * y;
                if(right ? (ST <= 0.0) : (ST > 0.0)) {
                    const double base = cd.tdgp1 - sgn * cd.gm1dgp1 * rcp_nr(aK) * uK;
                    sol[0] = WK[0] * (c.fast53 ? ipow3d(base) : pow(base, cd.tdgm1));
                    sol[P] = pK * (c.fast53 ? ipow5d(base) : pow(base, cd.tgdgm1));
                    vhalf = cd.tdgp1 * (-sgn * aK + cd.gm1d2 * uK) - uK;
                } else {
                    sol[0] = WK[0] * (c.fast53 ? ipow3d(y) : pow(pdp, cd.ginv));
                    sol[P] = ps;
                    vhalf = us - uK;
                }
            }
        }
#pragma unroll
        for(int d = 0; d < ND; d++) sol[1 + d] += vhalf * n64[d];
    }
    state_flux_d<ND>(0.5 * cd.tdgm1, sol[0], sol + 1, sol[P], n64, v64, F);
    bool bad = (sol[0] < 0.0) | (sol[0] != sol[0]) | (sol[P] < 0.0) | (sol[P] != sol[P]);
#pragma unroll
    for(int d = 0; d < ND; d++) bad |= (sol[1 + d] != sol[1 + d]);
    return bad ? ST_SOLVER : 0;
}

// HLLCRiemannSolver::solve_for_flux (HLLCRiemannSolver.cpp:258-403) for the
// mixed path.  On this GPU an FP64 add, multiply or FMA issues at the FP32
// rate, while the IEEE FP64 division and square root are long instruction
// sequences, and they are what HLLC spends its time on.  Here every
// quotient and root goes through rcp_nr / sqrt_nr (hardware estimate plus
// Newton-Raphson), close to FP64 accuracy at a fraction of the cost.  FP32
// wave speeds were tried and rejected: with rho and p spread over six
// decades the sound speeds alone cost 9e-7 and the q factors 1.1e-6
// normwise.  The star-state correction SK (U*K - UK) is written through
// r = (S* - uK)/(SK - S*):
//   U*K - UK = rhoK [r, r vK + (1 + r)(S* - uK) n,
//                    r eK + (1 + r)(S* - uK)(S* + pK/(rhoK (SK - uK)))]
// (fac = rhoK (SK - uK)/(SK - S*) = rhoK (1 + r)), so it vanishes with
// S* - uK rather than as a difference of two rounded products (in cold gas,
// a = 4e-3, p = 1e-5, a rounded fac would leave SK rhoK eps, ~4e-5 of the
// pressure flux with FP32 eps).  Vacuum faces take the FP64 closed forms; a
// reflective face has S* = 0 exactly (the FP64 reference's numerator
// vanishes identically there).
// HLLCRiemannSolver::vacuum_flux's flux of the half state
// (HLLCRiemannSolver.cpp:180-208), mixed path
template <int ND>
__device__ __forceinline__ void vacuum_flux_mixed(const CF& cf, const Consts& c, const double* WL,
                                                  const double* WR, const double* n,
                                                  const double* v, double uL, double uR,
                                                  double aL, double aR, double* flux) {
    constexpr int NC = ND + 2, P = ND + 1;
    double sol[NC];
    vacuum_state_mixed<ND>(cf, c, WL, WR, uL, uR, aL, aR, n, sol);
#pragma unroll
    for(int k = 0; k < NC; k++) flux[k] = 0.;
    if(!(sol[0] != 0.)) return;
    double vsol = sol[1] * n[0] + sol[2] * n[1];
    double v2 = sol[1] * sol[1] + sol[2] * sol[2];
    if(ND == 3) {
        vsol += sol[ND] * n[ND - 1];
        v2 += sol[ND] * sol[ND];
    }
    flux[0] = sol[0] * vsol;
#pragma unroll
    for(int d = 0; d < ND; d++) flux[1 + d] = sol[0] * sol[1 + d] * vsol + sol[P] * n[d];
    const double e = 0.5 * c.tdgm1 * sol[P] * rcp_nr(sol[0]) + 0.5 * v2;
    flux[P] = sol[0] * e * vsol + sol[P] * vsol;
    fp64::deboost<ND>(c, v, flux);
}

template <int ND>
__device__ uint8_t hllc_mixed(const CF& cf, const Consts& c, const double* WL, double* WR,
                              const double* n, const double* v, bool reflective, double* flux,
                              unsigned& vac) {
    constexpr int NC = ND + 2, P = ND + 1;
    const double g = c.gamma;
#pragma unroll
    for(int k = 0; k < NC; k++) flux[k] = 0.;
    if(!(WL[0] != 0.) && (reflective || !(WR[0] != 0.))) return 0;
    double uL = WL[1] * n[0] + WL[2] * n[1];
    double uR = WR[1] * n[0] + WR[2] * n[1];
    if(ND == 3) {
        uL += WL[ND] * n[ND - 1];
        uR += WR[ND] * n[ND - 1];
    }
    if(reflective) {
#pragma unroll
        for(int k = 0; k < NC; k++) WR[k] = WL[k];
        uR = -uL;
    }
    // sound speeds: HLLCRiemannSolver::get_soundspeed has no rho = 0 guard;
    // the vacuum cases below read a side's speed only where its rho != 0
    const double aL = sqrt_nr(g * WL[P] * rcp_nr(WL[0]));
    const double aR = sqrt_nr(g * WR[P] * rcp_nr(WR[0]));
    if(!(WL[0] != 0.) || !(WR[0] != 0.) || c.tdgm1 * aL + c.tdgm1 * aR <= uR - uL) {
        vac = 1;
        vacuum_flux_mixed<ND>(cf, c, WL, WR, n, v, uL, uR, aL, aR, flux);
        return 0;
    }
    const double rhobar = 0.5 * (WL[0] + WR[0]);
    const double abar = 0.5 * (aL + aR);
    const double pstar = smax(0., 0.5 * (WL[P] + WR[P]) - 0.5 * (uR - uL) * rhobar * abar);
    const double qc = c.gp1d2g;  // (g + 1)/(2 g)
    double qL = 1., qR = 1.;
    if(pstar > WL[P]) qL = sqrt_nr(1. + qc * ((pstar - WL[P]) * rcp_nr(WL[P])));
    if(pstar > WR[P]) qR = sqrt_nr(1. + qc * ((pstar - WR[P]) * rcp_nr(WR[P])));
    const double SL = uL - aL * qL;
    const double SR = uR + aR * qR;
    // the denominator is -(rhoL aL qL + rhoR aR qR) < 0: no cancellation
    const double num = WR[P] - WL[P] + WL[0] * uL * (SL - uL) - WR[0] * uR * (SR - uR);
    const double den = WL[0] * (SL - uL) - WR[0] * (SR - uR);
    const double Sstar = reflective ? 0.0 : num * rcp_nr(den);
    const bool left = Sstar >= 0.;
    const double* WK = left ? WL : WR;
    const double uK = left ? uL : uR;
    const double SK = left ? SL : SR;
    const double rho = WK[0], pK = WK[P];
    const double rinv = rcp_nr(rho);
    flux[0] = rho * uK;
#pragma unroll
    for(int d = 0; d < ND; d++) flux[1 + d] = rho * WK[1 + d] * uK + pK * n[d];
    double v2 = WK[1] * WK[1] + WK[2] * WK[2];
    if(ND == 3) v2 += WK[ND] * WK[ND];
    const double eK = 0.5 * c.tdgm1 * pK * rinv + 0.5 * v2;
    flux[P] = rho * eK * uK + pK * uK;
    if(left ? (SK < 0.) : (SK > 0.)) {
        // |SK - S*| >= |SK|: S* lies between SL and SR
        const double ds = Sstar - uK;
        const double r = ds * rcp_nr(SK - Sstar);
        const double sr = SK * rho;
        const double t = sr * (1. + r) * ds;
        flux[0] += sr * r;
#pragma unroll
        for(int d = 0; d < ND; d++) flux[1 + d] += sr * r * WK[1 + d] + t * n[d];
        flux[P] += sr * r * eK + t * (Sstar + pK * rinv * rcp_nr(SK - uK));
    }
    fp64::deboost<ND>(c, v, flux);
    return 0;
}

template <int ND>
__device__ uint8_t exact_mixed(const CF& c, const Consts& cd, const float* WL, float* WR,
                               const float* n, bool reflective, const double* WL64,
                               double* WR64, const double* n64, const double* v64, double* F,
                               unsigned& iters, unsigned& vac) {
    ExactCase s;
    exact_setup<ND>(c, cd, WL, WR, n, reflective, WL64, WR64, n64, s);
    if(s.phase == EX_ITER) exact_iterate(c, s, c.itmax);
    return exact_finish<ND>(c, cd, s, WL64, WR64, n64, v64, F, iters, vac);
}

inline CF make_cf(const Consts& c) {
    CF f;
    f.gamma = (float)c.gamma;
    f.gp1d2g = (float)c.gp1d2g;
    f.gm1d2g = (float)c.gm1d2g;
    f.gm1dgp1 = (float)c.gm1dgp1;
    f.tdgp1 = (float)c.tdgp1;
    f.tdgm1 = (float)c.tdgm1;
    f.gm1d2 = (float)c.gm1d2;
    f.tgdgm1 = (float)c.tgdgm1;
    f.ginv = (float)c.ginv;
    f.ginvm1 = (float)(c.ginv - 1.0);
    f.gm1inv = (float)(1.0 / (c.gamma - 1.0));
    f.q_fac = (float)(0.5 * (c.gamma + 1.0) / c.gamma);
    f.gamma_d = c.gamma;
    f.tdgp1_d = c.tdgp1;
    f.tdgm1_d = c.tdgm1;
    f.gm1dgp1_d = c.gm1dgp1;
    f.gm1d2g_d = c.gm1d2g;
    f.cutoff_d = c.cutoff;
    f.halftol = (float)(0.5 * c.tol);
    f.fast53 = c.fast53;
    f.itmax = 24;
    return f;
}

}  // namespace mixed
}  // namespace sfx
