// riemann_mixed.hip -- mixed-precision Riemann solvers for gfx950.
//
// Bulk arithmetic in FP32 on the VALU's fast paths (v_log_f32 / v_exp_f32 /
// v_rcp_f32 / v_sqrt_f32); FP64 only for the last Newton step of the exact
// solver's star pressure p* (and the star velocity u* that comes with it).
// Storage type T is float (52 B in + 20 B out per 3D problem) or double.
//
// Algorithm (same problem as Shadowfax ExactRiemannSolver::solve_for_flux,
// src/riemann/ExactRiemannSolver.cpp:108-290,815-880, and
// HLLCRiemannSolver::solve_for_flux, src/riemann/HLLCRiemannSolver.cpp:
// 258-403), restructured for the GPU:
//   * the pressure function f(p) = fL + fR + du is monotone increasing and
//     concave, so Newton from any point left of the root converges
//     monotonically; a bracket [lo, hi] handles a first step that would
//     leave it (replaces the reference's Brent fallback: no serial state
//     machine, a bounded and uniform iteration count per wave);
//   * pow(x, (g+1)/(-2g)) is evaluated as pow(x, (g-1)/(2g)) / x, one
//     transcendental per side per step; for g = 5/3 all other exponents are
//     integer powers of x^(1/5);
//   * the FP64 step: fifth root seeded in FP32 and refined by one Newton step
//     in FP64 (relative error ~1e-13), giving p* and u* to ~1e-12;
//   * the flux is evaluated in projected form F = F(W).n directly instead of
//     summing per-axis fluxes.
#include "sfx_device.h"
#include "sfx_launch.h"

namespace sfx {
namespace mixed {

// FP32 consts
struct CF {
    float gamma, gp1d2g, gm1d2g, gm1dgp1, tdgp1, tdgm1, gm1d2, tgdgm1, ginv, ginvm1;
    float gm1inv;      // 1/(gamma-1)
    float q_fac;       // 0.5*(gamma+1)/gamma (HLLC)
    double gamma_d, tdgp1_d, tdgm1_d, gm1dgp1_d, gm1d2g_d;
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
    y = 0.8 * y + 0.2 * z / (y2 * y2);
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
        double A = c.tdgp1_d / rho;
        double B = c.gm1dgp1_d * pK;
        double r = 1.0 / (p + B);
        double sq = sqrt(A * r);
        double dp = p - pK;
        f = dp * sq;
        df = (1.0 - 0.5 * dp * r) * sq;
    } else {
        double x = p / pK;
        double y = pow_gm1d2g_d(c, x);
        f = c.tdgm1_d * a * (y - 1.0);
        df = y / (rho * a * x);
    }
}

// Vacuum closed forms in FP32 (ExactRiemannSolver.cpp:304-446 /
// HLLCRiemannSolver.cpp:44-170 have the same branch structure).
template <int ND>
__device__ void vacuum_state(const CF& c, const float* WL, const float* WR, float uL, float uR,
                             float aL, float aR, const float* n, float* sol) {
    const int NC = ND + 2, P = ND + 1;
#pragma unroll
    for(int k = 0; k < NC; k++) sol[k] = 0.f;
    const bool zl = !(WL[0] != 0.f), zr = !(WR[0] != 0.f);
    if(zl && zr) return;
    // which side's fan (if any) covers x/t = 0
    int side = 0;  // 0: vacuum, 1: left state, 2: right state, 3: left fan, 4: right fan
    if(zr) {
        if(uL < aL) side = (uL + c.tdgm1 * aL > 0.f) ? 3 : 0;
        else side = 1;
    } else if(zl) {
        if(-aR < uR) side = (uR - c.tdgm1 * aR >= 0.f) ? 0 : 4;
        else side = 2;
    } else {
        float SR = uR - c.tdgm1 * aR, SL = uL + c.tdgm1 * aL;
        if(SR > 0.f && SL < 0.f) side = 0;
        else if(SL >= 0.f) side = (aL > uL) ? 3 : 1;
        else side = (-aR < uR) ? 4 : 2;
    }
    if(side == 0) return;
    const float* WK = (side == 1 || side == 3) ? WL : WR;
#pragma unroll
    for(int k = 0; k < NC; k++) sol[k] = WK[k];
    if(side >= 3) {
        const bool left = side == 3;
        const float aK = left ? aL : aR, uK = left ? uL : uR;
        const float base = left ? (c.tdgp1 + c.gm1dgp1 * rcpf(aK) * uK)
                                : (c.tdgp1 - c.gm1dgp1 * rcpf(aK) * uK);
        float rf, pf;
        if(c.fast53) {
            rf = ipow3(base);
            pf = ipow5(base);
        } else {
            float lb = log2f_fast(base);
            rf = exp2f_fast(c.tdgm1 * lb);
            pf = exp2f_fast(c.tgdgm1 * lb);
        }
        sol[0] = WK[0] * rf;
        sol[P] = WK[P] * pf;
        const float vhalf = left ? (c.tdgp1 * (aK + c.gm1d2 * uK) - uK)
                                 : (c.tdgp1 * (-aK + c.gm1d2 * uK) - uK);
#pragma unroll
        for(int d = 0; d < ND; d++) sol[1 + d] += vhalf * n[d];
    }
}

// Flux of the sampled state sol (face frame) through a face moving with v:
// F = F(W + v) . n with the mass flux relative to the face.
template <int ND>
__device__ __forceinline__ void state_flux(const CF& c, const float* sol, const float* n,
                                           const float* v, float* F) {
    const int NC = ND + 2, P = ND + 1;
    if(!(sol[0] != 0.f)) {
#pragma unroll
        for(int k = 0; k < NC; k++) F[k] = 0.f;
        return;
    }
    float w[ND];
    float wn = 0.f, vn = 0.f, w2 = 0.f;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        w[d] = sol[1 + d] + v[d];
        wn += w[d] * n[d];
        vn += v[d] * n[d];
        w2 += w[d] * w[d];
    }
    const float rho = sol[0], p = sol[P];
    const float rel = wn - vn;
    const float F0 = rho * rel;
    F[0] = F0;
#pragma unroll
    for(int d = 0; d < ND; d++) F[1 + d] = F0 * w[d] + p * n[d];
    const float E = 0.5f * rho * w2 + p * c.gm1inv;
    F[P] = rel * E + p * wn;
}

// Exact Riemann solver, mixed precision.  Returns status bits.
template <int ND>
__device__ uint8_t exact_mixed(const CF& c, const float* WL, float* WR, const float* n,
                               const float* v, bool reflective, const double* WL64,
                               double* WR64, const double* n64, float* F,
                               unsigned& iters, unsigned& vac) {
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
    const float aL = (WL[0] != 0.f) ? sqrtf_fast(c.gamma * WL[P] * rcpf(WL[0])) : 0.f;
    const float aR = (WR[0] != 0.f) ? sqrtf_fast(c.gamma * WR[P] * rcpf(WR[0])) : 0.f;
    float sol[NC];
    if(!(WL[0] != 0.f) || !(WR[0] != 0.f) || c.tdgm1 * (aL + aR) <= uR - uL) {
        vac = 1;
        vacuum_state<ND>(c, WL, WR, uL, uR, aL, aR, n, sol);
    } else {
        const SideF sL = make_side(c, WL[0], WL[P], aL);
        const SideF sR = make_side(c, WR[0], WR[P], aR);
        const float du = uR - uL;
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
            float gL = sqrtf_fast(sL.A * rcpf(ppv + sL.B));
            float gR = sqrtf_fast(sR.A * rcpf(ppv + sR.B));
            p = (gL * WL[P] + gR * WR[P] - du) * rcpf(gL + gR);
        }
        p = fmaxf(pfloor, p);
        // bracketed Newton in FP32
        float lo = 0.f, hi = __builtin_inff();
        int it = 0;
        for(; it < c.itmax; it++) {
            float fl, dl, fr, dr;
            side_f(c, sL, p, fl, dl);
            side_f(c, sR, p, fr, dr);
            const float fv = fl + fr + du;
            if(fv < 0.f) lo = p;
            else hi = p;
            float pn = p - fv * rcpf(dl + dr);
            if(!(pn > lo && pn < hi)) pn = (lo > 0.f) ? sqrtf_fast(lo) * sqrtf_fast(hi) : 0.01f * hi;
            const bool done = fabsf(pn - p) <= 2e-6f * (pn + p);
            p = pn;
            if(done) break;
        }
        iters += it + 1;
        // FP64 final Newton step for p* (and u*)
        double pd = (double)p;
        const double aLd = sqrt(c.gamma_d * WL64[P] / WL64[0]);
        const double aRd = sqrt(c.gamma_d * WR64[P] / WR64[0]);
        double uLd = 0.0, uRd = 0.0;
#pragma unroll
        for(int d = 0; d < ND; d++) {
            uLd += WL64[1 + d] * n64[d];
            uRd += WR64[1 + d] * n64[d];
        }
        if(reflective) uRd = -uLd;
        double fLd, dLd, fRd, dRd;
        side_f_d(c, WL64[0], WL64[P], aLd, pd, fLd, dLd);
        side_f_d(c, WR64[0], WR64[P], aRd, pd, fRd, dRd);
        const double step = (fLd + fRd + (uRd - uLd)) / (dLd + dRd);
        double pnew = pd - step;
        if(!(pnew > 0.0)) pnew = pd;  // keep the FP32 root if the step would leave (0, inf)
        const double delta = pnew - pd;
        const double ud = 0.5 * (uLd + uRd) + 0.5 * ((fRd + dRd * delta) - (fLd + dLd * delta));
        p = (float)pnew;
        const float u = (float)ud;
        // sampling at x/t = 0, ExactRiemannSolver.cpp:190-271
        const bool right = u < 0.f;
        const float* WK = right ? WR : WL;
        const float aK = right ? aR : aL;
        const float uK = right ? uR : uL;
        const float sgn = right ? 1.f : -1.f;  // right-going (+) or left-going (-) wave
        const float pK = WK[P];
#pragma unroll
        for(int k = 0; k < NC; k++) sol[k] = WK[k];
        float vhalf = 0.f;
        const float pdp = p * rcpf(pK);
        if(p > pK) {
            const float S = uK + sgn * aK * sqrtf_fast(c.gp1d2g * pdp + c.gm1d2g);
            if(sgn * S > 0.f) {
                sol[0] = WK[0] * (pdp + c.gm1dgp1) * rcpf(c.gm1dgp1 * pdp + 1.f);
                sol[P] = p;
                vhalf = u - uK;
            }
        } else {
            const float SH = uK + sgn * aK;
            if(sgn * SH > 0.f) {
                const float y = powf_fast(pdp, c.gm1d2g);
                const float ST = u + sgn * aK * y;
                if(right ? (ST <= 0.f) : (ST > 0.f)) {
                    const float base = c.tdgp1 - sgn * c.gm1dgp1 * rcpf(aK) * uK;
                    float rf, pf;
                    if(c.fast53) {
                        rf = ipow3(base);
                        pf = ipow5(base);
                    } else {
                        const float lb = log2f_fast(base);
                        rf = exp2f_fast(c.tdgm1 * lb);
                        pf = exp2f_fast(c.tgdgm1 * lb);
                    }
                    sol[0] = WK[0] * rf;
                    sol[P] = pK * pf;
                    vhalf = c.tdgp1 * (-sgn * aK + c.gm1d2 * uK) - uK;
                } else {
                    sol[0] = WK[0] * (c.fast53 ? ipow3(y) : powf_fast(pdp, c.ginv));
                    sol[P] = p;
                    vhalf = u - uK;
                }
            }
        }
#pragma unroll
        for(int d = 0; d < ND; d++) sol[1 + d] += vhalf * n[d];
    }
    state_flux<ND>(c, sol, n, v, F);
    uint8_t st = 0;
    bool bad = (sol[0] < 0.f) | (sol[0] != sol[0]) | (sol[P] < 0.f) | (sol[P] != sol[P]);
#pragma unroll
    for(int d = 0; d < ND; d++) bad |= (sol[1 + d] != sol[1 + d]);
    if(bad) st |= ST_SOLVER;
    return st;
}

// HLLC, FP32 (HLLCRiemannSolver.cpp:258-403)
template <int ND>
__device__ uint8_t hllc_mixed(const CF& c, const float* WL, float* WR, const float* n,
                              const float* v, bool reflective, float* flux, unsigned& vac) {
    const int NC = ND + 2, P = ND + 1;
#pragma unroll
    for(int k = 0; k < NC; k++) flux[k] = 0.f;
    if(!(WL[0] != 0.f) && (reflective || !(WR[0] != 0.f))) return 0;
    float uL = 0.f, uR = 0.f;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        uL += WL[1 + d] * n[d];
        uR += WR[1 + d] * n[d];
    }
    if(reflective) {
#pragma unroll
        for(int k = 0; k < NC; k++) WR[k] = WL[k];
        uR = -uL;
    }
    const float aL = sqrtf_fast(c.gamma * WL[P] * rcpf(WL[0]));
    const float aR = sqrtf_fast(c.gamma * WR[P] * rcpf(WR[0]));
    if(!(WL[0] != 0.f) || !(WR[0] != 0.f) || c.tdgm1 * (aL + aR) <= uR - uL) {
        vac = 1;
        float sol[NC];
        vacuum_state<ND>(c, WL, WR, uL, uR, aL, aR, n, sol);
        float zero[ND];
#pragma unroll
        for(int d = 0; d < ND; d++) zero[d] = 0.f;
        // flux in the face frame, then deboost below (HLLCRiemannSolver.cpp:172-206)
        state_flux<ND>(c, sol, n, zero, flux);
    } else {
        const float rhobar = 0.5f * (WL[0] + WR[0]);
        const float abar = 0.5f * (aL + aR);
        const float pPVRS = 0.5f * (WL[P] + WR[P]) - 0.5f * (uR - uL) * rhobar * abar;
        const float pstar = fmaxf(0.f, pPVRS);
        const float qL = (pstar > WL[P]) ? sqrtf_fast(1.f + c.q_fac * (pstar * rcpf(WL[P]) - 1.f)) : 1.f;
        const float qR = (pstar > WR[P]) ? sqrtf_fast(1.f + c.q_fac * (pstar * rcpf(WR[P]) - 1.f)) : 1.f;
        const float SL = uL - aL * qL;
        const float SR = uR + aR * qR;
        const float mL = WL[0] * (SL - uL), mR = WR[0] * (SR - uR);
        const float Sstar = (WR[P] - WL[P] + mL * uL - mR * uR) * rcpf(mL - mR);
        const bool left = Sstar >= 0.f;
        const float* WK = left ? WL : WR;
        const float uK = left ? uL : uR;
        const float SK = left ? SL : SR;
        const float mK = left ? mL : mR;
        float v2 = 0.f;
#pragma unroll
        for(int d = 0; d < ND; d++) v2 += WK[1 + d] * WK[1 + d];
        const float eK = WK[P] * c.gm1inv * rcpf(WK[0]) + 0.5f * v2;
        const float rK = WK[0];
        flux[0] = rK * uK;
#pragma unroll
        for(int d = 0; d < ND; d++) flux[1 + d] = rK * WK[1 + d] * uK + WK[P] * n[d];
        flux[P] = rK * eK * uK + WK[P] * uK;
        if(left ? (SK < 0.f) : (SK > 0.f)) {
            // S_K (U*_K - U_K), with U*_K = m_K/(S_K - S*) [1, w + (S*-u)n,
            // e + (S*-u)(S* + p/m_K)]
            const float fac = mK * rcpf(SK - Sstar);
            const float ds = Sstar - uK;
            flux[0] += SK * (fac - rK);
#pragma unroll
            for(int d = 0; d < ND; d++)
                flux[1 + d] += SK * (fac * (WK[1 + d] + ds * n[d]) - rK * WK[1 + d]);
            flux[P] += SK * (fac * (eK + ds * (Sstar + WK[P] * rcpf(mK))) - rK * eK);
        }
    }
    // deboost to the lab frame (HLLCRiemannSolver.cpp:392-400)
    float vn2 = 0.f, vF = 0.f;
#pragma unroll
    for(int d = 0; d < ND; d++) {
        flux[1 + d] += v[d] * flux[0];
        vn2 += v[d] * v[d];
        vF += v[d] * flux[1 + d];
    }
    flux[P] += vF + 0.5f * vn2 * flux[0];
    return 0;
}

template <int ND, int KIND, typename T>
__global__ __launch_bounds__(256) void solve_for_flux_kernel(
        CF c, int64_t N, const T* __restrict__ WL, const T* __restrict__ WR,
        const T* __restrict__ n, const T* __restrict__ v, const uint8_t* __restrict__ refl,
        T* __restrict__ F, uint8_t* __restrict__ status, unsigned long long* __restrict__ ctr) {
    const int NC = ND + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned iters = 0, vac = 0, fail = 0, valid = 0;
    if(i < N) {
        valid = 1;
        float wl[NC], wr[NC], nn[ND], vv[ND], f[NC];
        double wl64[NC], wr64[NC], n64[ND];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            const T a = WL[k * N + i], b = WR[k * N + i];
            wl[k] = (float)a;
            wr[k] = (float)b;
            wl64[k] = (double)a;
            wr64[k] = (double)b;
        }
#pragma unroll
        for(int d = 0; d < ND; d++) {
            const T a = n[d * N + i];
            nn[d] = (float)a;
            n64[d] = (double)a;
            vv[d] = v ? (float)v[d * N + i] : 0.f;
        }
        const bool r = refl ? refl[i] != 0 : false;
        uint8_t st;
        if(KIND == 0)
            st = exact_mixed<ND>(c, wl, wr, nn, vv, r, wl64, wr64, n64, f, iters, vac);
        else
            st = hllc_mixed<ND>(c, wl, wr, nn, vv, r, f, vac);
#pragma unroll
        for(int k = 0; k < NC; k++) F[k * N + i] = (T)f[k];
        if(status) status[i] = st;
        fail = st != 0;
    }
    if(ctr) {
        wave_add(ctr + CNT_NEVAL, valid);
        wave_add(ctr + CNT_NVACUUM, vac);
        wave_add(ctr + CNT_NITER, iters);
        wave_add(ctr + CNT_NFAIL, fail);
    }
}

static CF make_cf(const Consts& c) {
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
    f.halftol = (float)(0.5 * c.tol);
    f.fast53 = c.fast53;
    f.itmax = 24;
    return f;
}

}  // namespace mixed

template <typename T>
hipError_t launch_solve_for_flux_mixed(int kind, int ndim, const Consts& c, int64_t N,
                                       const T* WL, const T* WR, const T* n, const T* v,
                                       const uint8_t* refl, T* F, uint8_t* status,
                                       unsigned long long* ctr, hipStream_t stream) {
    if(N <= 0) return hipSuccess;
    const mixed::CF cf = mixed::make_cf(c);
    const int bs = 256;
    dim3 grid((unsigned)((N + bs - 1) / bs)), block(bs);
#define SFX_LAUNCH(ND, K)                                                                 \
    hipLaunchKernelGGL((mixed::solve_for_flux_kernel<ND, K, T>), grid, block, 0, stream, cf, \
                       N, WL, WR, n, v, refl, F, status, ctr)
    if(ndim == 3 && kind == 0) SFX_LAUNCH(3, 0);
    else if(ndim == 3) SFX_LAUNCH(3, 1);
    else if(kind == 0) SFX_LAUNCH(2, 0);
    else SFX_LAUNCH(2, 1);
#undef SFX_LAUNCH
    return hipGetLastError();
}

template hipError_t launch_solve_for_flux_mixed<float>(int, int, const Consts&, int64_t,
                                                       const float*, const float*, const float*,
                                                       const float*, const uint8_t*, float*,
                                                       uint8_t*, unsigned long long*,
                                                       hipStream_t);
template hipError_t launch_solve_for_flux_mixed<double>(int, int, const Consts&, int64_t,
                                                        const double*, const double*,
                                                        const double*, const double*,
                                                        const uint8_t*, double*, uint8_t*,
                                                        unsigned long long*, hipStream_t);

}  // namespace sfx
