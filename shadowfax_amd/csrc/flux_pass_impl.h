// flux_pass_impl.h -- the VorCell/VorFace flux pass as gfx950 kernels.
//
//   K2 gradients   one lane per cell: VorCell::estimate_gradient
//                  (Shadowfax src/VorCell.cpp:464-530)
//   K3 face fluxes one lane per face: VorFace::set_v + calculate_flux
//                  (src/VorFace.cpp:586-660,670-866) -> dQL per face
//   K4 gather      one lane per cell: the per-cell sum of the face
//                  contributions in ascending face order, which is the
//                  order VorTessManager::hydro's forward face loop
//                  (src/VorTessManager.cpp:205-210) applies them through
//                  GasParticle::increase_dQ (src/utilities/GasParticle.cpp:550)
//   K5 update      one lane per cell: GasParticle::update_Q + get_W
//                  (src/utilities/GasParticle.cpp:617-624,
//                  src/riemann/ExactRiemannSolver.cpp:716-766)
//
// Template parameters: ND (2, 3), KIND (0 exact, 1 HLLC), T (storage type of
// cell states, gradients and face geometry) and R (arithmetic type of the
// reconstruction).  The FP64 parity pass is T = R = double in a TU compiled
// with -ffp-contract=off, every expression in the reference's association;
// the mixed pass is T = float, R = double with the mixed Riemann solvers and
// FP64 dQ.
//
// Layout: everything a kernel gathers through a neighbour index is
// item-major (one cache line per gather): W [C][NC], vel [C][ND],
// grad [C][ND][NC], dQ [C][NC], Q [C][NC], dQL [F][NC] (T); per-face geometry,
// read one face per lane, is component-major [ND][F]; K2 reads each cell's
// faces through per-entry records (entry_geometry_kernel) stored in the
// cell's walk order, so its loads stream instead of chasing face indices.
// Faces carry geometry precomputed on the host from the tessellation
// with the reference's FP64 expressions (see sfx_mesh_capi.cpp):
//   dL = midpoint - centroid_L,  dR = midpoint - centroid_R + (pos_R - x_Rgen)
//   rRL = x_Rgen - x_L,  mc = midpoint - (x_L + x_Rgen)/2,
//   rcent = |centroid_L - centroid_R| (unshifted centroids, as the reference)
#pragma once

#include "riemann_fp64_impl.h"
#include "riemann_mixed_impl.h"
#include "sfx_device.h"

// Entries loaded ahead per lane in the per-cell loops (gradients, gather).
#ifndef SFX_GRAD_KB
#define SFX_GRAD_KB 2
#endif
#ifndef SFX_GRAD_KB_2D  // FP32-record gradients of 2D meshes (about 6 entries a cell)
#define SFX_GRAD_KB_2D 4
#endif
#ifndef SFX_GRAD_KN_2D  // FP32-record gradients of 2D meshes: neighbour rows in flight
#define SFX_GRAD_KN_2D 2
#endif
#ifndef SFX_REC_BUFFER  // grouped records through a bounded buffer descriptor
#define SFX_REC_BUFFER 1
#endif
#ifndef SFX_GRAD_KL_3D  // FP32-record limiter loop in 3D: records per chunk
#define SFX_GRAD_KL_3D 4
#endif
#ifndef SFX_GRAD_PIPE  // FP32-record gradients: next chunk's records in flight
#define SFX_GRAD_PIPE 1
#endif
#ifndef SFX_GRAD_LDS_K  // 3D FP32-record gradients: limiter d vectors kept in LDS per lane
#define SFX_GRAD_LDS_K 16
#endif
#ifndef SFX_GRAD_LDS_K_2D  // the same for 2D meshes; 0: off (8 measured 17.7 against 15.8 us on
#define SFX_GRAD_LDS_K_2D 0  // 512x512: the first walk would load second halves it otherwise skips)
#endif
#ifndef SFX_ROW_LOADS3  // FP64 rows of 5 values in 3 loads (else 4, all 16-byte aligned)
#define SFX_ROW_LOADS3 1
#endif
#ifndef SFX_GATHER_KB
#define SFX_GATHER_KB 8
#endif
#ifndef SFX_GATHER_WAVES  // gathers of FP32 rows over large meshes: waves per SIMD (0: all that fit)
#define SFX_GATHER_WAVES 2
#endif
#ifndef SFX_GATHER_WAVES_F64  // the same for FP64 rows
#define SFX_GATHER_WAVES_F64 1
#endif
#ifndef SFX_GATHER_BIG  // ... of at least this many cells
#define SFX_GATHER_BIG (1 << 20)
#endif
#ifndef SFX_EGEO_U  // entries a lane of the FP32 record build forms at a time
#define SFX_EGEO_U 4  // (3D Sedov build: 1.49 ms; 2: 1.51, 8: 1.55, one at a time: 1.69)
#endif
#ifndef SFX_EGEO_U64  // the same for FP64 records
#define SFX_EGEO_U64 4  // (3D Sedov build: 2.345 ms; 2: 2.465, one at a time: 2.455)
#endif
#ifndef SFX_EGEO_WAVES  // record builds over large meshes: waves per SIMD (0: all that fit)
#define SFX_EGEO_WAVES 2
#endif
#ifndef SFX_GATHER_KB_F64  // FP64 rows: the same registers as 8 FP32 rows
#define SFX_GATHER_KB_F64 4
#endif
#ifndef SFX_GATHER_KB_BIG  // FP32 rows in flight per lane at the large meshes' bounded occupancy
#define SFX_GATHER_KB_BIG 16  // (3D Sedov gather 251 -> 218 us; 8: 221, 32: 455, spilling;
                              // at 1 wave per SIMD, 32: 318 us, 24: 380)
#endif
#ifndef SFX_GATHER_KB_BIG_F64  // the same for FP64 rows
#define SFX_GATHER_KB_BIG_F64 8  // (409 -> 374 us; 16: 389)
#endif

namespace sfx {
namespace flux {

// topology entries: 2*face + side (side 1: the cell is the face's right)
__device__ __forceinline__ int32_t ent_face(int32_t e) { return e >> 1; }
__device__ __forceinline__ bool ent_right(int32_t e) { return e & 1; }

template <typename T>
struct MeshView {
    int64_t C, F;
    const double* W;      // [C][NC], FP64 in both passes
    const double* vel;    // [C][ND], FP64 in both passes
    const double* mass;   // [C]
    const double* total_area;
    const double* dt;
    const double* volume;
    const uint8_t* active;  // may be null: all active
    const int32_t* left;
    const int32_t* right;
    const uint8_t* ghost;
    const int32_t* twin;
    const T* area;
    const T* dL;     // [ND][F]
    const T* dR;     // [ND][F]
    const T* rRL;    // [ND][F]
    const T* mc;     // [ND][F]
    const T* rcent;  // [F]
    const int64_t* cf_off;
    const int32_t* cf_entry;
    const T* egeo;   // [cf_off[C]][entry_rec<ND, T>()]
    const int64_t* gather_off;
    const int32_t* gather_entry;
    // FP32 records grouped by 64 consecutive cells (see grad_cell_f32), the
    // first record row of every group; null: records in entry order
    const int64_t* egrp = nullptr;
    // bytes of the grouped records, 0 if not known: when known and under
    // 4 GB they are read through a buffer descriptor with that bound
    uint64_t egeo_bytes = 0;
    // grouped records: slot s (group s >> 6, lane s & 63) holds cell
    // eslot[s] & 0xffffffff with eslot[s] >> 32 entries; null: cell s
    const int64_t* eslot = nullptr;
    // [ND][C] every cell's closure residual sum A r_ij/|r_ij| (mixed
    // gradients), or null where the cells' faces close
    const T* closure = nullptr;
    // mixed pass: the faces' own values as 3 (2D) or 4 (3D) planes of F
    // float4 (face_record_kernel), or null: read from the arrays above
    const float4* frec = nullptr;
};

// Per-entry record of a cell's face, in the cell's walk order.  FP64
// storage (the parity pass):
//   [A, c (ND), r_ij (ND), d (ND), neighbour index bits, (3D: pad)]
// with c = mc, r_ij = rRL seen from the cell (-rRL for the left cell),
// d = dL or dR, neighbour = the other cell or -1 (mirror).  FP32 storage
// (the mixed pass, which sums the difference form of the gradient), two
// 16-byte halves:
//   [g (ND), neighbour index bits, (2D: pad) | d (ND), A, (2D: pad)]
// with g = A (c - r_ij/2)/|r_ij| formed in FP64 and rounded once, and the
// neighbour = the other cell or ~f for a mirror (so the face normal can be
// read back); the limiter's second loop reads only the second half.
__host__ __device__ constexpr int entry_rec(int ND, bool f32) {
    return f32 ? 8 : (ND == 2 ? 8 : 12);
}
// field offsets of the FP32 record: g, the neighbour, d, A; a 2D record
// carries A a second time in its first half (A2), so the gradient loop reads
// only first halves
template <int ND>
struct Rec32 {
    static constexpr int G = 0, NB = ND, D = 4, A = 4 + ND, A2 = (ND == 2) ? 3 : A;
};

// Row o of an item-major FP64 array of NC = 4 or 5 values per item
// ([n][NC], 16-byte aligned base) with 16-byte loads.  For NC = 5 the row
// starts at s = 5 o; from the aligned base b = s & ~1 two 16-byte loads
// hold values b .. b + 3 and two 8-byte loads the values b + 4 and (odd
// rows) b + 5, so no load reads past the row's last value s + 4; the row is
// picked out by the parity of s.  No branch: the loads of several rows go
// out back to back and nothing waits for them before their values are used.
template <int NC>
__device__ __forceinline__ void load_row(const double* __restrict__ a, int64_t o, int64_t n,
                                         double* out) {
    static_assert(NC == 4 || NC == 5, "rows of 4 or 5 values");
    if constexpr(NC == 4) {
        const double2* p = reinterpret_cast<const double2*>(a + o * 4);
        const double2 x = p[0], y = p[1];
        out[0] = x.x;
        out[1] = x.y;
        out[2] = y.x;
        out[3] = y.y;
    } else if constexpr(SFX_ROW_LOADS3) {
        // three loads: two 16-byte ones at the row's own 8-byte alignment
        // (global memory takes them unaligned) and one 8-byte one
        (void)n;
        typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
        const double* r = a + o * 5;
        const d2u x = *reinterpret_cast<const d2u*>(r);
        const d2u y = *reinterpret_cast<const d2u*>(r + 2);
        out[0] = x.x;
        out[1] = x.y;
        out[2] = y.x;
        out[3] = y.y;
        out[4] = r[4];
    } else {
        (void)n;
        const int64_t s = o * 5, base = s & ~int64_t(1);
        const bool odd = (s & 1) != 0;
        const double2* p = reinterpret_cast<const double2*>(a + base);
        const double2 x = p[0], y = p[1];
        // (base + 4 + (s & 1): an index, not a choice between two loads)
        const double z0 = a[base + 4], z1 = a[base + 4 + (s & 1)];
        out[0] = odd ? x.y : x.x;
        out[1] = odd ? y.x : x.y;
        out[2] = odd ? y.y : y.x;
        out[3] = odd ? z0 : y.y;
        out[4] = odd ? z1 : z0;
    }
}
template <int ND, typename T>
__host__ __device__ constexpr int entry_rec() {
    return entry_rec(ND, sizeof(T) == 4);
}

__device__ __forceinline__ double index_bits(int32_t i, double) {
    return __longlong_as_double((long long)i);
}
__device__ __forceinline__ float index_bits(int32_t i, float) { return __int_as_float(i); }
__device__ __forceinline__ int32_t bits_index(double v) {
    return (int32_t)__double_as_longlong(v);
}
__device__ __forceinline__ int32_t bits_index(float v) { return __float_as_int(v); }

template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return fmin(a, b); }
template <typename R>
__device__ __forceinline__ R rmax(R a, R b) { return fmax(a, b); }
template <typename R>
__device__ __forceinline__ R smaxr(R a, R b) { return (a < b) ? b : a; }
template <typename R>
__device__ __forceinline__ R sminr(R a, R b) { return (b < a) ? b : a; }

// Mirror of a velocity through the face: transform, negate the normal
// component, invtransform (src/VorFace.cpp:417-542); the basis depends only
// on rRL = right - left.
template <int ND, typename R>
__device__ __forceinline__ void mirror_velocity(R* vel, const R* rRL) {
    if(ND == 2) {
        // d = left - right = -rRL
        R d0 = -rRL[0], d1 = -rRL[1];
        R d2 = sqrt(d0 * d0 + d1 * d1);
        R cost = -d0 / d2;
        R sint = -d1 / d2;
        R W1 = vel[0] * cost + vel[1] * sint;
        R W2 = -vel[0] * sint + vel[1] * cost;
        W1 = -W1;
        R V1 = W1 * cost - W2 * sint;
        R V2 = W1 * sint + W2 * cost;
        vel[0] = V1;
        vel[1] = V2;
    } else {
        R n0 = rRL[0], n1 = rRL[1], n2 = rRL[ND - 1];
        R n3 = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        n0 /= n3;
        n1 /= n3;
        n2 /= n3;
        R o0, o1, p0, p1, p2;
        if(n0 != R(0) || n1 != R(0)) {
            o0 = -n1;
            o1 = n0;
            R o3 = sqrt(o0 * o0 + o1 * o1);
            o0 /= o3;
            o1 /= o3;
            p0 = -n2 * o1;
            p1 = n2 * o0;
            p2 = o3;
        } else {
            o0 = 0;
            o1 = 1;
            p0 = -1;
            p1 = 0;
            p2 = 0;
        }
        R v0 = vel[0], v1 = vel[1], v2 = vel[ND - 1];
        R t1 = v0 * n0 + v1 * n1 + v2 * n2;
        R t2 = v0 * o0 + v1 * o1;
        R t3 = v0 * p0 + v1 * p1 + v2 * p2;
        t1 = -t1;
        vel[0] = t1 * n0 + t2 * o0 + t3 * p0;
        vel[1] = t1 * n1 + t2 * o1 + t3 * p1;
        vel[ND - 1] = t1 * n2 + t3 * p2;
    }
}

// a / b: IEEE division in FP64 (the parity pass); in FP32 (the mixed pass's
// offsets) the hardware reciprocal, about 1 ulp, without the denormal
// scaling the compiler wraps around an FP32 divide
template <typename R>
__device__ __forceinline__ R quot_r(R a, R b) {
    return a / b;
}
template <>
__device__ __forceinline__ float quot_r<float>(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

// get_time_derivative (ExactRiemannSolver.cpp:892-934 / HLLCRiemannSolver.cpp:
// 415-450); g[d][k] = dW_k/dx_d.  The 2D dWdt[2] uses gradients[1].vy()
// twice, as the reference does (:920 / :443).
template <int ND, typename R>
__device__ __forceinline__ void time_derivative(R gamma, const R* W, const R (*g)[ND + 2],
                                                R* dWdt) {
    if(ND == 3) {
        const R* gx = g[0];
        const R* gy = g[1];
        const R* gz = g[ND - 1];
        dWdt[0] = -W[0] * (gx[1] + gy[2] + gz[3]) - W[1] * gx[0] - W[2] * gy[0] - W[3] * gz[0];
        dWdt[1] = -W[1] * gx[1] - W[2] * gy[1] - W[3] * gz[1] - quot_r(gx[4], W[0]);
        dWdt[2] = -W[1] * gx[2] - W[2] * gy[2] - W[3] * gz[2] - quot_r(gy[4], W[0]);
        dWdt[3] = -W[1] * gx[3] - W[2] * gy[3] - W[3] * gz[3] - quot_r(gz[4], W[0]);
        dWdt[ND + 1] = -gamma * W[4] * (gx[1] + gy[2] + gz[3]) - W[1] * gx[4] - W[2] * gy[4] -
                       W[3] * gz[4];
    } else {
        const R* gx = g[0];
        const R* gy = g[1];
        dWdt[0] = -W[1] * gx[0] - W[2] * gy[0] - W[0] * (gx[1] + gy[2]);
        dWdt[1] = -W[1] * gx[1] - W[2] * gy[1] - quot_r(gx[3], W[0]);
        dWdt[2] = -W[1] * gy[2] - W[2] * gy[2] - quot_r(gy[3], W[0]);
        dWdt[ND + 1] = -gamma * W[3] * (gx[1] + gy[2]) - W[1] * gx[3] - W[2] * gy[3];
    }
}

// VorFace::limit, scalar (src/VorFace.hpp:92-137); dnrm = |d| precomputed
template <typename R>
__device__ __forceinline__ R limit1(R phimid0, R phiL, R phiR, R dnrm, R r) {
    const R psi1 = R(0.5), psi2 = R(0.25);
    R delta1 = psi1 * fabs(phiL - phiR);
    R delta2 = psi2 * fabs(phiL - phiR);
    R phimin = rmin(phiL, phiR);
    R phimax = rmax(phiL, phiR);
    R phibar = phiL + dnrm / r * (phiR - phiL);
    R phiplus = ((phimax + delta1) * phimax > R(0)) ? phimax + delta1
                                                    : phimax / (R(1) + delta1 / fabs(phimax));
    R phiminus = ((phimin - delta1) * phimin > R(0)) ? phimin - delta1
                                                     : phimin / (R(1) + delta1 / fabs(phimin));
    if(phiL == phiR) return phiL;
    if(phiL < phiR) return rmax(phiminus, rmin(phibar + delta2, phimid0));
    return rmin(phiplus, rmax(phibar - delta2, phimid0));
}

// One side of VorFace::calculate_flux's predictor (src/VorFace.cpp:714-804)
template <int ND, typename R>
__device__ __forceinline__ void predict_side(R gamma, const R* Wraw, const R* Wother,
                                             const R (*g)[ND + 2], const R* d, R r,
                                             bool limited, R dtp, R* out) {
    const int NC = ND + 2, P = ND + 1;
    R dWdt[NC], deltaW[NC], Wt[NC];
    time_derivative<ND, R>(gamma, Wraw, g, dWdt);
#pragma unroll
    for(int k = 0; k < NC; k++) deltaW[k] = R(0.5) * dtp * dWdt[k];
    // VorFace::reconstruct (src/VorFace.hpp:174-196)
#pragma unroll
    for(int k = 0; k < NC; k++) {
        if(ND == 3)
            Wt[k] = Wraw[k] + (g[0][k] * d[0] + g[1][k] * d[1] + g[ND - 1][k] * d[ND - 1]);
        else
            Wt[k] = Wraw[k] + (g[0][k] * d[0] + g[1][k] * d[1]);
    }
    if(limited) {
        R dnrm = (ND == 3) ? sqrt(d[0] * d[0] + d[1] * d[1] + d[ND - 1] * d[ND - 1])
                           : sqrt(d[0] * d[0] + d[1] * d[1]);
#pragma unroll
        for(int k = 0; k < NC; k++) Wt[k] = limit1<R>(Wt[k], Wraw[k], Wother[k], dnrm, r);
    }
    if(Wt[0] < R(0) || Wt[P] < R(0)) {
#pragma unroll
        for(int k = 0; k < NC; k++) Wt[k] = Wraw[k];
    }
#pragma unroll
    for(int k = NC; k--;) Wt[k] += deltaW[k];
    if(Wt[0] < R(0) || Wt[P] < R(0)) {
#pragma unroll
        for(int k = 0; k < NC; k++) Wt[k] -= deltaW[k];
    }
#pragma unroll
    for(int k = 0; k < NC; k++) out[k] = Wt[k];
}

// predict_side for the mixed pass: the state is the FP64 cell state plus an
// FP32 offset.  Slopes, the time derivative and the pairwise limiter act on
// offsets from phiL = Wraw (phiR - phiL formed in FP64, then rounded), so
// their FP32 rounding is relative to the offsets, not to the state; the two
// sign tests of the limiter's phi+/phi- (discontinuous in phiL) stay in
// FP64.  With D = phiR - phiL, d1 = psi1 |D|:
//   phibar - phiL = |d|/r D
//   phiL < phiR: phi- - phiL = -d1, or -phiL d1/(phiL + d1) for 0 <= phiL <= d1
//   phiL > phiR: phi+ - phiL = +d1, or |phiL| d1/(|phiL| + d1) for -d1 <= phiL <= 0
// (the closed forms of phimin/(1 + d1/|phimin|) - phiL).
template <int ND>
__device__ __forceinline__ void predict_side_offset(float gamma, const double* Wraw,
                                                    const double* Wother,
                                                    const float (*g)[ND + 2], const float* d,
                                                    float r, bool limited, double dtp,
                                                    double* out) {
    constexpr int NC = ND + 2, P = ND + 1;
    float Wf[NC], dWdt[NC], dW[NC], off[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) Wf[k] = (float)Wraw[k];
    time_derivative<ND, float>(gamma, Wf, g, dWdt);
    const float hdt = 0.5f * (float)dtp;
#pragma unroll
    for(int k = 0; k < NC; k++) {
        dW[k] = hdt * dWdt[k];
        off[k] = (ND == 3) ? g[0][k] * d[0] + g[1][k] * d[1] + g[ND - 1][k] * d[ND - 1]
                           : g[0][k] * d[0] + g[1][k] * d[1];
    }
    if(limited) {
        // |d|/r and phibar - phiL in FP64: the limiter's bound is an offset
        // of the size of the jump itself, rounded once
        const double d0 = d[0], d1v = d[1], d2v = (ND == 3) ? (double)d[ND - 1] : 0.0;
        const double t = mixed::sqrt_nr(d0 * d0 + d1v * d1v + d2v * d2v) * mixed::rcp_nr((double)r);
#pragma unroll
        for(int k = 0; k < NC; k++) {
            const double phiL = Wraw[k], phiR = Wother[k];
            if(phiL == phiR) {
                off[k] = 0.f;
                continue;
            }
            const double Dd = phiR - phiL;
            const double d1d = 0.5 * fabs(Dd);
            const float d1 = (float)d1d, d2 = (float)(0.25 * fabs(Dd));
            const float pl = (float)phiL;
            if(phiL < phiR) {
                const float om = ((phiL - d1d) * phiL > 0.0) ? -d1
                                 : (pl == 0.f)               ? 0.f
                                                             : -(d1 * pl) * mixed::rcpf(pl + d1);
                off[k] = fmaxf(om, fminf((float)(t * Dd) + d2, off[k]));
            } else {
                const float op = ((phiL + d1d) * phiL > 0.0) ? d1
                                 : (pl == 0.f)               ? 0.f
                                                             : (d1 * -pl) * mixed::rcpf(d1 - pl);
                off[k] = fminf(op, fmaxf((float)(t * Dd) - d2, off[k]));
            }
        }
    }
    double Wt[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) Wt[k] = Wraw[k] + (double)off[k];
    if(Wt[0] < 0.0 || Wt[P] < 0.0) {
#pragma unroll
        for(int k = 0; k < NC; k++) off[k] = 0.f;
    }
    double Wp[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) Wp[k] = Wraw[k] + (double)(off[k] + dW[k]);
    const bool keep = !(Wp[0] < 0.0 || Wp[P] < 0.0);
#pragma unroll
    for(int k = 0; k < NC; k++) out[k] = keep ? Wp[k] : Wraw[k] + (double)off[k];
}

// ------------------------------------------------------- face records --

// Planes of F float4 holding, for face f, the values the mixed face kernel
// reads of the face itself, so that it loads them with 3 (2D) or 4 (3D)
// 16-byte loads instead of 13 (2D) or 17 (3D) 4-byte ones (the kernel's
// address unit, busy for most of its time, handles a wave's load in about
// the same number of cycles whatever its width):
//   3D: [left, right, area, rcent] [rRL (3), mc.x] [mc.y, mc.z, dL.x, dL.y]
//       [dL.z, dR (3)]
//   2D: [left, right, area, rcent] [rRL (2), mc (2)] [dL (2), dR (2)]
// (indices as their bits).  Built once per tessellation, as the entry
// records are; the values are the FP32 arrays' own, so the face fluxes are
// bit for bit those of the array reads.
template <int ND>
constexpr int face_record_planes() {
    return ND == 3 ? 4 : 3;
}

template <int ND>
__global__ __launch_bounds__(256) void face_record_kernel(MeshView<float> m,
                                                          float4* __restrict__ frec) {
    const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t F = m.F;
    if(f >= F) return;
    float v[4 * face_record_planes<ND>()];
    v[0] = __int_as_float(m.left[f]);
    v[1] = __int_as_float(m.right[f]);
    v[2] = m.area[f];
    v[3] = m.rcent[f];
#pragma unroll
    for(int l = 0; l < ND; l++) {
        v[4 + l] = m.rRL[l * F + f];
        v[4 + ND + l] = m.mc[l * F + f];
        v[4 + 2 * ND + l] = m.dL[l * F + f];
        v[4 + 3 * ND + l] = m.dR[l * F + f];
    }
#pragma unroll
    for(int q = 0; q < face_record_planes<ND>(); q++)
        frec[q * F + f] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

// ------------------------------------------------------------- K2 grads --

// One lane per cell entry: the record of entry j (face f = e >> 1 seen from
// side e & 1) from the face arrays; run once per tessellation.
// A face's own values as the face records hold them (face_record_kernel):
// 3 (2D) or 4 (3D) 16-byte loads.
template <int ND>
struct FaceVals {
    int32_t left, right;
    float area, rcent, rRL[ND], mc[ND], dL[ND], dR[ND];
};
template <int ND>
__device__ __forceinline__ FaceVals<ND> load_face_record(const float4* __restrict__ frec,
                                                         int64_t F, int64_t f) {
    FaceVals<ND> o;
    const float4 p0 = frec[f], p1 = frec[F + f], p2 = frec[2 * F + f];
    o.left = __float_as_int(p0.x);
    o.right = __float_as_int(p0.y);
    o.area = p0.z;
    o.rcent = p0.w;
    if constexpr(ND == 3) {
        const float4 p3 = frec[3 * F + f];
        const float v[12] = {p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w, p3.x, p3.y, p3.z, p3.w};
#pragma unroll
        for(int l = 0; l < ND; l++) {
            o.rRL[l] = v[l];
            o.mc[l] = v[3 + l];
            o.dL[l] = v[6 + l];
            o.dR[l] = v[9 + l];
        }
    } else {
        const float v[8] = {p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w};
#pragma unroll
        for(int l = 0; l < ND; l++) {
            o.rRL[l] = v[l];
            o.mc[l] = v[2 + l];
            o.dL[l] = v[4 + l];
            o.dR[l] = v[6 + l];
        }
    }
    return o;
}

// The FP32 record of cell entry e = 2 face + side (entry_rec(ND, true)).
template <int ND>
__device__ __forceinline__ void entry_record_from_face(const FaceVals<ND>& fv, int32_t e,
                                                       float* rec);

template <int ND>
__device__ __forceinline__ void entry_record_f32(const MeshView<float>& m, int32_t e,
                                                 float* rec) {
    const int64_t F = m.F;
    const int32_t f = ent_face(e);
    const bool isR = ent_right(e);
    // the face's values: from its record (4 loads in 3D) when the face
    // records are built, else from the face arrays (the same values)
    FaceVals<ND> fv;
    if(m.frec) {
        fv = load_face_record<ND>(m.frec, F, f);
    } else {
        fv.left = m.left[f];
        fv.right = m.right[f];
        fv.area = m.area[f];
#pragma unroll
        for(int l = 0; l < ND; l++) {
            fv.rRL[l] = m.rRL[l * F + f];
            fv.mc[l] = m.mc[l * F + f];
            fv.dL[l] = m.dL[l * F + f];
            fv.dR[l] = m.dR[l * F + f];
        }
    }
    entry_record_from_face<ND>(fv, e, rec);
}

// The FP32 record of cell entry e from its face's values fv.
template <int ND>
__device__ __forceinline__ void entry_record_from_face(const FaceVals<ND>& fv, int32_t e,
                                                       float* rec) {
    const int32_t f = ent_face(e);
    const bool isR = ent_right(e);
    // g = A (c - r_ij/2)/|r_ij| in FP64
    double r[ND], r2 = 0.0;
#pragma unroll
    for(int l = 0; l < ND; l++) {
        const double rr = (double)fv.rRL[l];
        r[l] = isR ? rr : -rr;
        r2 += r[l] * r[l];
    }
    const double s = (double)fv.area / sqrt(r2);
    using RL = Rec32<ND>;
#pragma unroll
    for(int k = 0; k < 8; k++) rec[k] = 0.f;
#pragma unroll
    for(int l = 0; l < ND; l++) {
        rec[RL::G + l] = (float)(s * ((double)fv.mc[l] - 0.5 * r[l]));
        rec[RL::D + l] = isR ? fv.dR[l] : fv.dL[l];
    }
    const int32_t nb = isR ? fv.left : fv.right;
    rec[RL::NB] = index_bits(nb >= 0 ? nb : ~f, 0.f);
    rec[RL::A] = fv.area;
    rec[RL::A2] = fv.area;
}

// The FP64 record of cell entry e (entry_rec(ND, false)), in two steps: the
// face's values it needs (its side's d vector and neighbour), then the record.
template <int ND>
struct EntryVals64 {
    double area, mc[ND], rRL[ND], d[ND];
    int32_t nb;
};
template <int ND>
__device__ __forceinline__ EntryVals64<ND> load_entry_vals64(const MeshView<double>& m,
                                                             int32_t e) {
    const int64_t F = m.F;
    const int32_t f = ent_face(e);
    const bool isR = ent_right(e);
    EntryVals64<ND> v;
    v.area = m.area[f];
#pragma unroll
    for(int l = 0; l < ND; l++) {
        v.rRL[l] = m.rRL[l * F + f];
        v.mc[l] = m.mc[l * F + f];
        v.d[l] = (isR ? m.dR : m.dL)[l * F + f];
    }
    v.nb = (isR ? m.left : m.right)[f];
    return v;
}
template <int ND>
__device__ __forceinline__ void entry_record_from_vals64(const EntryVals64<ND>& v, int32_t e,
                                                         double* rec) {
    constexpr int ER = entry_rec(ND, false);
    const bool isR = ent_right(e);
    rec[0] = v.area;
#pragma unroll
    for(int l = 0; l < ND; l++) {
        rec[1 + l] = v.mc[l];
        rec[1 + ND + l] = isR ? v.rRL[l] : -v.rRL[l];
        rec[1 + 2 * ND + l] = v.d[l];
    }
    rec[1 + 3 * ND] = index_bits(v.nb, 0.0);
    if(ER > 2 + 3 * ND) rec[ER - 1] = 0.0;
}
template <int ND>
__device__ __forceinline__ void entry_record_f64(const MeshView<double>& m, int32_t e,
                                                 double* rec) {
    entry_record_from_vals64<ND>(load_entry_vals64<ND>(m, e), e, rec);
}

template <int ND, typename T>
__device__ __forceinline__ void entry_record(const MeshView<T>& m, int32_t e, T* rec) {
    if constexpr(sizeof(T) == 4)
        entry_record_f32<ND>(m, e, rec);
    else
        entry_record_f64<ND>(m, e, rec);
}

template <int ND, typename T>
__global__ __launch_bounds__(256) void entry_geometry_kernel(MeshView<T> m, int64_t E,
                                                             T* __restrict__ egeo) {
    constexpr int ER = entry_rec<ND, T>();
    const int64_t j = xcd_block() * blockDim.x + threadIdx.x;
    if(j >= E) return;
    T rec[ER];
    entry_record<ND, T>(m, m.cf_entry[j], rec);
#pragma unroll
    for(int k = 0; k < ER; k++) egeo[j * ER + k] = rec[k];
}

// The grouped record layout: cells in groups of 64 (cell i: group i / 64,
// lane i % 64); group g owns record rows egrp[g] .. egrp[g + 1] - 1, as
// many as its cell with the most entries; row r holds the k-th entry
// (k = r - egrp[g], walk order) of all 64 cells, lane beside lane, as
// P 16-byte planes of the record (2 for FP32, 4 / 6 for FP64 in 2D / 3D):
// plane q of (row r, lane) at 16-byte index (r P + q) 64 + lane.  Cells
// with fewer entries (and lanes past the last cell) get zero records (area
// 0: skipped).  A wavefront of 64 consecutive cells then reads row after
// row with coalesced 1 KB loads where the entry-ordered layout reads 64
// scattered records.
template <int ND, typename T>
__host__ __device__ constexpr int rec_planes() {
    return entry_rec<ND, T>() * (int)sizeof(T) / 16;
}

// Record at position p of a cell: entry base + p (entry order) or row
// base + p, lane `lane` (grouped).
// The grouped records through a buffer descriptor bounded by their size
// (m.egeo_bytes, at most 2 GB): a lane past its own cell's entries gives the
// offset 2^31, past the bound for all four dwords of its load (no 32-bit
// wrap), and the hardware returns zeros (area 0: skipped) without a memory
// access, so padding rows cost no bandwidth.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const void* egeo, uint64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(egeo), (short)0, (int)(uint32_t)bytes,
                                             0x00020000);
}
// 16-byte plane idx of the records; own: the lane's cell has this entry
__device__ __forceinline__ float4 rec_plane(__amdgpu_buffer_rsrc_t rs, int64_t idx, bool own) {
    const uint32_t off = own ? (uint32_t)(idx * 16) : 0x80000000u;
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}

template <int ND, typename T, bool GROUPED>
__device__ __forceinline__ void load_record(const MeshView<T>& m, int64_t base, int64_t p,
                                            int lane, T* out) {
    constexpr int ER = entry_rec<ND, T>();
    if constexpr(GROUPED) {
        constexpr int P = rec_planes<ND, T>();
        const float4* E = reinterpret_cast<const float4*>(m.egeo);
#pragma unroll
        for(int q = 0; q < P; q++)
            reinterpret_cast<float4*>(out)[q] = E[((base + p) * P + q) * 64 + lane];
    } else {
#pragma unroll
        for(int q = 0; q < ER; q++) out[q] = m.egeo[(base + p) * ER + q];
    }
}

// One block per group of 64 slots (XCD-aware order).  Lane l holds slot
// 64 g + l: cell c_l (the slot itself, or m.eslot's cell) with n_l entries
// from cf_off[c_l]; the block's threads walk the group's entries lane after
// lane, entry after entry (neighbouring threads then read a cell's
// neighbouring entries): walk index j belongs to the last lane l with
// off[l] <= j, off the prefix sums of the n_l (in LDS; a lane without
// entries never matches), as that lane's entry j - off[l], written to row
// egrp[g] + j - off[l].  A second loop writes zeros to the rows past each
// lane's last entry.
template <int ND, typename T>
__global__ __launch_bounds__(256) void entry_geometry_grouped_kernel(MeshView<T> m,
                                                                     T* __restrict__ egeo) {
    constexpr int ER = entry_rec<ND, T>(), P = rec_planes<ND, T>();
    __shared__ int64_t off[65], base[64];
    const int64_t g = xcd_block();
    const int64_t s0 = g * 64;
    const int nc = (int)((m.C - s0) < 64 ? (m.C - s0) : 64);
    if((int)threadIdx.x < 64) {
        const int l = threadIdx.x;
        int64_t n = 0, b = 0;
        if(l < nc) {
            if(m.eslot) {
                const int64_t v = m.eslot[s0 + l];
                b = m.cf_off[v & 0xffffffff];
                n = v >> 32;
            } else {
                b = m.cf_off[s0 + l];
                n = m.cf_off[s0 + l + 1] - b;
            }
        }
        base[l] = b;
        // the lanes' counts summed across the first wave (inclusive scan)
#pragma unroll
        for(int d = 1; d < 64; d <<= 1) {
            const int64_t t = __shfl_up(n, d, 64);
            if(l >= d) n += t;
        }
        off[l + 1] = n;
        if(l == 0) off[0] = 0;
    }
    __syncthreads();
    const int64_t r0 = m.egrp[g], K = m.egrp[g + 1] - r0;
    const int64_t j1 = off[nc];
    float4* out = reinterpret_cast<float4*>(egeo);
    // the group's entry j: lane lo with off[lo] <= j < off[lo + 1], its k-th
    auto find = [&](int64_t j, int& lo_out) {
        int lo = 0, hi = nc;
        while(hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if(off[mid] <= j)
                lo = mid;
            else
                hi = mid;
        }
        lo_out = lo;
        return j - off[lo];
    };
    if constexpr(sizeof(T) == 4) {
        if(m.frec) {
            // FP32 records from the face records: a lane's U entries at a
            // time, their indices, then their faces' records, loaded before
            // any record is formed (entries past the group's last read the
            // last one's, unused), so a lane waits for two loads per U
            // entries instead of two per entry
            constexpr int U = SFX_EGEO_U;
            for(int64_t jb = threadIdx.x; jb < j1; jb += U * blockDim.x) {
                int lo[U];
                int64_t k[U];
                int32_t e[U];
#pragma unroll
                for(int u = 0; u < U; u++) {
                    const int64_t j = jb + u * (int64_t)blockDim.x;
                    k[u] = find(j < j1 ? j : j1 - 1, lo[u]);
                }
#pragma unroll
                for(int u = 0; u < U; u++) e[u] = m.cf_entry[base[lo[u]] + k[u]];
                FaceVals<ND> fv[U];
#pragma unroll
                for(int u = 0; u < U; u++) fv[u] = load_face_record<ND>(m.frec, m.F, ent_face(e[u]));
                // the loads stay here, ahead of the tests below (sunk into
                // them, each entry's would wait for the previous entry's)
                asm volatile("" ::: "memory");
#pragma unroll
                for(int u = 0; u < U; u++) {
                    if(jb + u * (int64_t)blockDim.x >= j1) continue;
                    alignas(16) float rec[ER];
                    entry_record_from_face<ND>(fv[u], e[u], rec);
#pragma unroll
                    for(int q = 0; q < P; q++)
                        out[((r0 + k[u]) * P + q) * 64 + lo[u]] =
                                reinterpret_cast<const float4*>(rec)[q];
                }
            }
            goto pad;
        }
    } else {
        // FP64 records: the same, a lane's U64 entries at a time
        constexpr int U = SFX_EGEO_U64;
        for(int64_t jb = threadIdx.x; jb < j1; jb += U * blockDim.x) {
            int lo[U];
            int64_t k[U];
            int32_t e[U];
#pragma unroll
            for(int u = 0; u < U; u++) {
                const int64_t j = jb + u * (int64_t)blockDim.x;
                k[u] = find(j < j1 ? j : j1 - 1, lo[u]);
            }
#pragma unroll
            for(int u = 0; u < U; u++) e[u] = m.cf_entry[base[lo[u]] + k[u]];
            EntryVals64<ND> v[U];
#pragma unroll
            for(int u = 0; u < U; u++) v[u] = load_entry_vals64<ND>(m, e[u]);
            asm volatile("" ::: "memory");
#pragma unroll
            for(int u = 0; u < U; u++) {
                if(jb + u * (int64_t)blockDim.x >= j1) continue;
                alignas(16) double rec[ER];
                entry_record_from_vals64<ND>(v[u], e[u], rec);
#pragma unroll
                for(int q = 0; q < P; q++)
                    out[((r0 + k[u]) * P + q) * 64 + lo[u]] =
                            reinterpret_cast<const float4*>(rec)[q];
            }
        }
        goto pad;
    }
    for(int64_t j = threadIdx.x; j < j1; j += blockDim.x) {
        int lo;
        const int64_t k = find(j, lo);
        alignas(16) T rec[ER];
        entry_record<ND, T>(m, m.cf_entry[base[lo] + k], rec);
#pragma unroll
        for(int q = 0; q < P; q++)
            out[((r0 + k) * P + q) * 64 + lo] = reinterpret_cast<const float4*>(rec)[q];
    }
pad:
    for(int64_t sl = threadIdx.x; sl < K * 64; sl += blockDim.x) {
        const int64_t k = sl >> 6;
        const int lane = (int)(sl & 63);
        const int64_t n = lane < nc ? off[lane + 1] - off[lane] : 0;
        if(k < n) continue;
#pragma unroll
        for(int q = 0; q < P; q++) out[((r0 + k) * P + q) * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// DIFF selects the form of the Green-Gauss sum.  The reference's
//   sum_f A [(Wj - Wi) c/|r| - (Wi + Wj)/2 r/|r|]
// (DIFF = false, the FP64 parity pass) cancels the Wi sum_f A r/|r| part
// only through the cell's closure sum_f A n_f = 0, which FP32 face areas
// hold to 6e-8 only: the residual is a spurious gradient proportional to W
// itself.  DIFF = true (the mixed pass) sums the algebraically equal
//   sum_f A (Wj - Wi) (c - r/2)/|r|
// that needs no cancellation.
// The gradients of the mixed pass from FP32 records (entry_rec<ND, float>):
//   delta[l][k] = sum_f (Wj_k - Wi_k) g_l / V
// with the differences and sums in FP64, and the cell-wise limiter
// (src/VorCell.cpp:505-521) without a division per face: since Wmax >= W >=
// Wmin, a face's candidate max((Wmax - W)/dp, (Wmin - W)/dp) is a/b with
// a = Wmax - W, b = dp for dp > 0 and a = W - Wmin, b = -dp for dp < 0
// (dp = 0 gives 1, which never lowers alpha <= 1); the smallest a/b is kept
// as a pair, compared by cross products, and divided out once per component.
// LK > 0: the block's LDS array dcache ([LK][ND][256] floats) keeps the d
// vectors of the cell's positions 0 .. LK - 1 from the first walk, so the
// limiter reads those from LDS instead of loading the records' second
// halves again (zeros where the area is 0, which the limiter skips either
// way: d = 0 gives dp = 0, or NaN against a non-finite gradient, and neither
// changes alpha).
template <int ND, bool GROUPED, bool ACT, bool BUF, int LK = 0>
__device__ __forceinline__ void grad_cell_f32(const MeshView<float>& m, int64_t s,
                                              float* __restrict__ grad, int64_t gshift,
                                              float* __restrict__ dcache = nullptr) {
    constexpr int NC = ND + 2, ER = entry_rec<ND, float>();
    // Slot s: grouped, the cell of lane s & 63 of group s >> 6 (m.eslot's,
    // or cell s); else cell s.  The loads that depend on nothing go out
    // together, before any wait: the record range, the slot's cell and
    // entry count, then the activity flag (ACT: m.active is set), the
    // cell's state and volume.  Inactive cells (which keep their gradients)
    // walk no records and store nothing; a branch on the flag here would
    // put its load alone in front of all the others.
    int64_t i = s, j0, n, own = 0;
    if constexpr(GROUPED) {
        const int64_t g = s >> 6;
        j0 = m.egrp[g];
        n = m.egrp[g + 1] - j0;
        if(m.eslot) {
            const int64_t v = m.eslot[s];
            i = v & 0xffffffff;
            own = v >> 32;  // the cell's own rows
        } else if constexpr(BUF) {
            own = m.cf_off[s + 1] - m.cf_off[s];
        }
    } else {
        j0 = m.cf_off[i];
        n = m.cf_off[i + 1] - j0;
    }
    const bool act = !ACT || m.active[i] != 0;
    const __amdgpu_buffer_rsrc_t rs = rec_rsrc(m.egeo, BUF ? m.egeo_bytes : 0);
    // the cell's records, visited in reverse walk order (src/VorCell.cpp:471)
    // as positions n - 1 .. 0: entry j0 + p (entry order), or row j0 + p of
    // the cell's group, lane i % 64 (grouped; rows past the cell's own
    // entries hold area 0 and are skipped, so every lane of a group walks
    // the same rows)
    const float4* E4 = reinterpret_cast<const float4*>(m.egeo);
    const int lane = (int)(s & 63);
    auto half = [&](int64_t p, int h) -> float4 {
        if constexpr(GROUPED && BUF)
            return rec_plane(rs, (2 * (j0 + p) + h) * 64 + lane, p < own);
        else if constexpr(GROUPED)
            return E4[(2 * (j0 + p) + h) * 64 + lane];
        else
            return E4[(j0 + p) * (ER / 4) + h];
    };
    constexpr int KB = (ND == 2) ? SFX_GRAD_KB_2D : SFX_GRAD_KB;
    // positions pb - 1 .. pb - KB; those below 0 read position 0 (n > 0)
    // and are skipped where used (u >= pb), not masked here: a register
    // write right after the load would wait for it
    auto load_chunk = [&](int64_t pb, float4* L, float4* H) {
#pragma unroll
        for(int u = 0; u < KB; u++) {
            const int64_t p = pb - 1 - u;
            L[u] = half(p >= 0 ? p : 0, 0);
            H[u] = half(p >= 0 ? p : 0, 1);
        }
    };
    // grouped records through the buffer descriptor: the first chunk's
    // loads go out first, without a test on n or the activity flag (a row
    // the lane's cell does not have reads as zeros without a fetch; an
    // inactive cell walks its rows and stores nothing), and the memory
    // clobber keeps them ahead of the cell's own loads, so both are in
    // flight together; sunk behind the loop's entry test, they went out
    // only after the cell's state had arrived
    constexpr bool EAGER = GROUPED && BUF;
    float4 lo[KB], hi[KB];
    if constexpr(EAGER) {
        load_chunk(n, lo, hi);
        asm volatile("" ::: "memory");
    }
    double W[NC], Wmax[NC], Wmin[NC], delta[ND][NC];
    load_row<NC>(m.W, i, m.C, W);
    const double Vinv = mixed::rcp_nr(m.volume[i]);
    if(!EAGER && !act) n = 0;
#pragma unroll
    for(int k = 0; k < NC; k++) {
        Wmax[k] = Wmin[k] = W[k];
#pragma unroll
        for(int l = 0; l < ND; l++) delta[l][k] = 0.0;
    }
    // 2D: chunks of 4 records (a cell has about 6), each chunk's neighbour
    // rows two at a time -- 126 VGPRs, so four waves per SIMD and the 2D
    // benchmark mesh's 4,096 waves in one round (chunks of 8, rows four at a
    // time: 152 VGPRs, three waves, 20.3 against 18.1 us); 3D: chunk after
    // chunk, the next chunk's records in flight while this chunk's
    // neighbour states load
    constexpr bool PIPE = ND == 3 && SFX_GRAD_PIPE;
    if constexpr(!EAGER) {
        if(n > 0) load_chunk(n, lo, hi);
    }
    for(int64_t pb = n; pb > 0; pb -= KB) {
        // the neighbour rows in batches of KN (all KB at once in 3D)
        constexpr int KN = (ND == 2) ? SFX_GRAD_KN_2D : KB;
        float4 nlo[PIPE ? KB : 1], nhi[PIPE ? KB : 1];
        if constexpr(LK > 0) {
            // the chunk's d vectors go to LDS as soon as its records are in,
            // so their registers are free again before the neighbour rows
#pragma unroll
            for(int u = 0; u < KB; u++) {
                const int64_t p = pb - 1 - u;
                if(p >= 0 && p < LK) {
                    // the limiter's area test: the second half's w (3D), z (2D)
                    const bool z = !(((ND == 3) ? hi[u].w : hi[u].z) != 0.f);
                    float* c = dcache + p * (ND * 256) + threadIdx.x;
                    c[0] = z ? 0.f : hi[u].x;
                    c[256] = z ? 0.f : hi[u].y;
                    if(ND == 3) c[512] = z ? 0.f : hi[u].z;
                }
            }
        }
#pragma unroll
        for(int u0 = 0; u0 < KB; u0 += KN) {
            // every neighbour row is loaded unconditionally (a mirror
            // neighbour is the cell's own row, reflected below), so the
            // loads go out back to back
            double Wj[KN][NC];
            int32_t nbo[KN];
#pragma unroll
            for(int v = 0; v < KN; v++) {
                const int u = u0 + v;
                nbo[v] = bits_index((ND == 3) ? lo[u].w : lo[u].z);
                load_row<NC>(m.W, nbo[v] >= 0 ? (int64_t)nbo[v] : i, m.C, Wj[v]);
            }
            if constexpr(PIPE) {
                if(u0 == 0) load_chunk(pb - KB, nlo, nhi);
            }
            // mirror neighbours (src/VorCell.cpp:482-489) of face ~o: only a
            // left cell has one, so rRL is seen from it
#pragma unroll
            for(int v = 0; v < KN; v++) {
                if(nbo[v] < 0) {
                    const int32_t f = ~nbo[v];
                    double rr[ND];
#pragma unroll
                    for(int l = 0; l < ND; l++) rr[l] = (double)m.rRL[l * m.F + f];
                    mirror_velocity<ND, double>(Wj[v] + 1, rr);
                }
            }
#pragma unroll
            for(int v = 0; v < KN; v++) {
                const int u = u0 + v;
                // A (0: a padding row) from the first half in 2D
                const float A = (ND == 3) ? hi[u].w : lo[u].w;
                if(u >= pb || !(A != 0.f)) continue;
                const double gl[3] = {(double)lo[u].x, (double)lo[u].y,
                                      (ND == 3) ? (double)lo[u].z : 0.0};
#pragma unroll
                for(int k = 0; k < NC; k++) {
                    const double dk = Wj[v][k] - W[k];
#pragma unroll
                    for(int l = 0; l < ND; l++) delta[l][k] += dk * gl[l];
                    Wmax[k] = smaxr<double>(Wmax[k], Wj[v][k]);
                    Wmin[k] = sminr<double>(Wmin[k], Wj[v][k]);
                }
            }
            // the next batch's loads stay behind this batch's use
            if constexpr(KN < KB) asm volatile("" ::: "memory");
        }
        if constexpr(PIPE) {
#pragma unroll
            for(int u = 0; u < KB; u++) {
                lo[u] = nlo[u];
                hi[u] = nhi[u];
            }
        } else if(pb > KB) {
            load_chunk(pb - KB, lo, hi);
        }
    }
    // the reference sum (src/VorCell.cpp:493-494) is the difference form
    // minus W_i times the cell's closure residual, which vanishes when the
    // cell's faces close; a face list that does not close brings it
    if(m.closure) {
        double S[ND];
#pragma unroll
        for(int l = 0; l < ND; l++) S[l] = (double)m.closure[l * m.C + i];
#pragma unroll
        for(int l = 0; l < ND; l++)
#pragma unroll
            for(int k = 0; k < NC; k++) delta[l][k] -= W[k] * S[l];
    }
#pragma unroll
    for(int l = 0; l < ND; l++)
#pragma unroll
        for(int k = 0; k < NC; k++) delta[l][k] *= Vinv;
    double na[NC], nb[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) na[k] = nb[k] = 1.0;
    // the limiter over the records' second halves (d and A)
    auto limit = [&](const float4& h, bool valid) {
        const float A = (ND == 3) ? h.w : h.z;
        if(!valid || !(A != 0.f)) return;
        const double dd[3] = {(double)h.x, (double)h.y, (ND == 3) ? (double)h.z : 0.0};
#pragma unroll
        for(int k = 0; k < NC; k++) {
            double dp = delta[0][k] * dd[0] + delta[1][k] * dd[1];
            if(ND == 3) dp += delta[ND - 1][k] * dd[ND - 1];
            const double a = (dp > 0.0) ? Wmax[k] - W[k] : W[k] - Wmin[k];
            const double b = fabs(dp);
            if(dp != 0.0 && a * nb[k] < na[k] * b) {
                na[k] = a;
                nb[k] = b;
            }
        }
    };
    {
        // KL halves a time, their loads out together (a pipelined loop, the
        // next chunk loading while this one is used, saves little here: the
        // work per chunk is short next to a load's latency)
        constexpr int KL = (ND == 2) ? 8 : SFX_GRAD_KL_3D;
        // positions nc .. n - 1 from the records, 0 .. nc - 1 (LK > 0) from
        // LDS, in the same descending order
        const int64_t nc = (LK > 0) ? (n < LK ? n : (int64_t)LK) : 0;
        for(int64_t pb = n; pb > nc; pb -= KL) {
            float4 h[KL];
#pragma unroll
            for(int u = 0; u < KL; u++) {
                const int64_t p = pb - 1 - u;
                h[u] = half(p >= 0 ? p : 0, 1);
            }
#pragma unroll
            for(int u = 0; u < KL; u++) limit(h[u], pb - 1 - u >= nc);
        }
        if constexpr(LK > 0) {
            for(int64_t pb = nc; pb > 0; pb -= 4) {
                float4 h[4];
#pragma unroll
                for(int u = 0; u < 4; u++) {
                    const int64_t p = pb - 1 - u;
                    const float* c = dcache + (p >= 0 ? p : 0) * (ND * 256) + threadIdx.x;
                    h[u] = make_float4(c[0], c[256], (ND == 3) ? c[512] : 1.f, 1.f);
                }
#pragma unroll
                for(int u = 0; u < 4; u++) limit(h[u], u < pb);
            }
        }
    }
    if(!act) return;
#pragma unroll
    for(int k = 0; k < NC; k++) {
        const double alpha = na[k] * mixed::rcp_nr(nb[k]);
#pragma unroll
        for(int l = 0; l < ND; l++)
            grad[((i + gshift) * ND + l) * NC + k] = (float)(delta[l][k] * alpha);
    }
}

// Cell i; its gradients go to grad[(i + gshift) * ND * NC ...] (gshift: 0
// for a whole mesh, the arena offset of the cell's mesh in an ensemble).
template <int ND, typename T, typename R, bool DIFF, bool GROUPED, bool ACT>
__device__ __forceinline__ void grad_cell_f64(const MeshView<T>& m, int64_t s,
                                              T* __restrict__ grad, int64_t gshift) {
    constexpr int NC = ND + 2, ER = entry_rec<ND, T>();
    // slot s (as in grad_cell_f32); the loads that depend on nothing go out
    // together: the record range and the slot's cell, then the activity
    // flag (ACT: m.active is set), the cell's state and volume; inactive
    // cells (which keep their gradients) walk no records and store nothing
    int64_t i = s, base, n;
    if constexpr(GROUPED) {
        base = m.egrp[s >> 6];
        n = m.egrp[(s >> 6) + 1] - base;
        if(m.eslot) i = m.eslot[s] & 0xffffffff;
    } else {
        base = m.cf_off[i];
        n = m.cf_off[i + 1] - base;
    }
    const bool act = !ACT || m.active[i] != 0;
    R W[NC], Wmax[NC], Wmin[NC], delta[ND][NC];
#pragma unroll
    for(int k = 0; k < NC; k++) {
        W[k] = Wmax[k] = Wmin[k] = (R)m.W[i * NC + k];
#pragma unroll
        for(int l = 0; l < ND; l++) delta[l][k] = R(0);
    }
    const R V = (R)m.volume[i];
    if(!act) n = 0;
    // the cell's records as positions n - 1 .. 0 (reverse walk order,
    // src/VorCell.cpp:471): entries base + p, or rows base + p of the
    // cell's group (grouped; rows past the cell's own entries have area 0);
    // positions below 0 of a chunk read position 0 and are skipped where
    // used (u >= pb)
    const int lane = (int)(s & 63);
    // KB records at a time: the records, then the neighbour states of a
    // chunk are loaded before any is used (every neighbour row
    // unconditionally: a mirror neighbour is the cell's own row, reflected
    // afterwards), so a lane has KB independent loads in flight
    constexpr int KB = SFX_GRAD_KB;
    for(int64_t pb = n; pb > 0; pb -= KB) {
        R A[KB], c[KB][ND], rij[KB][ND], Wj[KB][NC];
        int32_t o[KB];
#pragma unroll
        for(int u = 0; u < KB; u++) {
            const int64_t p = pb - 1 - u;
            alignas(16) T rec[ER];
            load_record<ND, T, GROUPED>(m, base, p >= 0 ? p : 0, lane, rec);
            A[u] = (R)rec[0];
#pragma unroll
            for(int l = 0; l < ND; l++) {
                c[u][l] = (R)rec[1 + l];
                rij[u][l] = (R)rec[1 + ND + l];
            }
            o[u] = bits_index(rec[1 + 3 * ND]);
        }
#pragma unroll
        for(int u = 0; u < KB; u++) {
            const int64_t src = o[u] >= 0 ? (int64_t)o[u] : i;
#pragma unroll
            for(int k = 0; k < NC; k++) Wj[u][k] = (R)m.W[src * NC + k];
        }
#pragma unroll
        for(int u = 0; u < KB; u++) {
            if(o[u] < 0) {
                // mirror neighbour (:482-489); only a left cell has one, so
                // rRL = -r_ij
                R rr[ND];
#pragma unroll
                for(int l = 0; l < ND; l++) rr[l] = -rij[u][l];
                mirror_velocity<ND, R>(Wj[u] + 1, rr);
            }
        }
#pragma unroll
        for(int u = 0; u < KB; u++) {
            if(u >= pb || !(A[u] != R(0))) continue;  // also the padding
            const R rnorm =
                    (ND == 3) ? sqrt(rij[u][0] * rij[u][0] + rij[u][1] * rij[u][1] +
                                     rij[u][ND - 1] * rij[u][ND - 1])
                              : sqrt(rij[u][0] * rij[u][0] + rij[u][1] * rij[u][1]);
#pragma unroll
            for(int l = ND; l--;) {
#pragma unroll
                for(int k = 0; k < NC; k++) {
                    R t;
                    if(DIFF)
                        t = (Wj[u][k] - W[k]) * ((c[u][l] - R(0.5) * rij[u][l]) / rnorm);
                    else
                        t = (Wj[u][k] - W[k]) * c[u][l] / rnorm -
                            R(0.5) * (W[k] + Wj[u][k]) * rij[u][l] / rnorm;
                    delta[l][k] += A[u] * t;
                }
            }
#pragma unroll
            for(int k = 0; k < NC; k++) {
                Wmax[k] = smaxr<R>(Wmax[k], Wj[u][k]);
                Wmin[k] = sminr<R>(Wmin[k], Wj[u][k]);
            }
        }
    }
#pragma unroll
    for(int l = ND; l--;)
#pragma unroll
        for(int k = 0; k < NC; k++) delta[l][k] /= V;
    R alpha[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) alpha[k] = R(1);
    // the limiter, KL records a time (their loads out together), same order
    constexpr int KL = 4;
    for(int64_t pb = n; pb > 0; pb -= KL) {
        R Ah[KL], dd[KL][ND];
#pragma unroll
        for(int u = 0; u < KL; u++) {
            const int64_t p = pb - 1 - u;
            alignas(16) T rec[ER];
            load_record<ND, T, GROUPED>(m, base, p >= 0 ? p : 0, lane, rec);
            Ah[u] = (R)rec[0];
#pragma unroll
            for(int l = 0; l < ND; l++) dd[u][l] = (R)rec[1 + 2 * ND + l];
        }
#pragma unroll
        for(int u = 0; u < KL; u++) {
            if(u >= pb || !(Ah[u] != R(0))) continue;
#pragma unroll
            for(int k = 0; k < NC; k++) {
                R dp = delta[0][k] * dd[u][0] + delta[1][k] * dd[u][1];
                if(ND == 3) dp += delta[ND - 1][k] * dd[u][ND - 1];
                R q1 = (dp != R(0)) ? (Wmax[k] - W[k]) / dp : R(1);
                R q2 = (dp != R(0)) ? (Wmin[k] - W[k]) / dp : R(1);
                alpha[k] = sminr<R>(alpha[k], smaxr<R>(q1, q2));
            }
        }
    }
    if(!act) return;
#pragma unroll
    for(int l = 0; l < ND; l++)
#pragma unroll
        for(int k = 0; k < NC; k++)
            grad[((i + gshift) * ND + l) * NC + k] = (T)(delta[l][k] * alpha[k]);
}

// the grouped FP32 records are read through a buffer descriptor when their
// size is known and at most 2 GB (below the out-of-range offset 2^31); the
// FP64 records keep plain loads (measured: the 2D FP64 gradients took
// 64 µs through the descriptor against 50 µs)
template <typename T>
__device__ __forceinline__ bool rec_buffer(const MeshView<T>& m) {
    return SFX_REC_BUFFER && sizeof(T) == 4 && m.egeo_bytes > 0 &&
           m.egeo_bytes <= 0x7FFFFFF0ull;
}

// Slot s: a cell of the grouped layout (see grad_cell_f32), or cell s.
template <int ND, typename T, typename R, bool DIFF>
__device__ __forceinline__ void grad_cell(const MeshView<T>& m, int64_t s, T* __restrict__ grad,
                                          int64_t gshift) {
    if constexpr(sizeof(T) == 4) {
        static_assert(DIFF, "FP32 records carry the difference-form weights");
        if(m.egrp && rec_buffer(m)) {
            if(m.active)
                grad_cell_f32<ND, true, true, true>(m, s, grad, gshift);
            else
                grad_cell_f32<ND, true, false, true>(m, s, grad, gshift);
        } else if(m.egrp) {
            if(m.active)
                grad_cell_f32<ND, true, true, false>(m, s, grad, gshift);
            else
                grad_cell_f32<ND, true, false, false>(m, s, grad, gshift);
        } else {
            if(m.active)
                grad_cell_f32<ND, false, true, false>(m, s, grad, gshift);
            else
                grad_cell_f32<ND, false, false, false>(m, s, grad, gshift);
        }
    } else {
        if(m.egrp) {
            if(m.active)
                grad_cell_f64<ND, T, R, DIFF, true, true>(m, s, grad, gshift);
            else
                grad_cell_f64<ND, T, R, DIFF, true, false>(m, s, grad, gshift);
        } else {
            if(m.active)
                grad_cell_f64<ND, T, R, DIFF, false, true>(m, s, grad, gshift);
            else
                grad_cell_f64<ND, T, R, DIFF, false, false>(m, s, grad, gshift);
        }
    }
}

template <int ND, typename T, typename R, bool DIFF>
__global__ __launch_bounds__(256) void gradient_kernel(MeshView<T> m, T* __restrict__ grad) {
    const int64_t s = xcd_block() * blockDim.x + threadIdx.x;  // slot
    if(s < m.C) grad_cell<ND, T, R, DIFF>(m, s, grad, 0);
}

// LDS positions per lane and waves per SIMD of gradient_lds_kernel: 3D
// keeps 16 positions (48 KB a block) at 3 waves per SIMD (at most 168
// VGPRs), as gradient_kernel runs; 2D (when SFX_GRAD_LDS_K_2D is set) at 4
// waves per SIMD (at most 128 VGPRs).
template <int ND>
constexpr int grad_lds_k() {
    return ND == 3 ? SFX_GRAD_LDS_K : SFX_GRAD_LDS_K_2D;
}
#ifndef SFX_GRAD_WAVES_3D  // (3D Sedov gradients: 382 us; 2 waves with 4 records a chunk: 394)
#define SFX_GRAD_WAVES_3D 3
#endif
template <int ND>
constexpr int grad_lds_waves() {
    return ND == 3 ? SFX_GRAD_WAVES_3D : 4;
}

// FP32 records, grouped and read through the buffer descriptor (the host
// launches this only then, in blocks of 256): the limiter's d vectors of
// each cell's first grad_lds_k positions stay in LDS.
template <int ND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grad_lds_waves<ND>()))) void
gradient_lds_kernel(MeshView<float> m, float* __restrict__ grad) {
    constexpr int LK = grad_lds_k<ND>() > 0 ? grad_lds_k<ND>() : 1;
    __shared__ float dcache[LK * ND * 256];
    const int64_t s = xcd_block() * blockDim.x + threadIdx.x;  // slot
    if(s >= m.C) return;
    if(m.active)
        grad_cell_f32<ND, true, true, true, LK>(m, s, grad, 0, dcache);
    else
        grad_cell_f32<ND, true, false, true, LK>(m, s, grad, 0, dcache);
}

// ------------------------------------------------------- K3 face fluxes --

template <int ND, typename T, typename R>
__device__ __forceinline__ void load_grad(const T* grad, int64_t C, int64_t cell,
                                          R (*g)[ND + 2]) {
    const int NC = ND + 2;
    (void)C;
#pragma unroll
    for(int l = 0; l < ND; l++)
#pragma unroll
        for(int k = 0; k < NC; k++) g[l][k] = (R)grad[(cell * ND + l) * NC + k];
}

// The activity flag read for every cell of a mesh without an activity array.
static __device__ const uint8_t kAllActive = 1;

// MIXED: 0 = FP64 parity solvers, 1 = mixed solvers (and, for a periodic
// face whose other copy has a lower index, no evaluation: the gather applies
// that copy's flux with the opposite sign).
// Face f: its flux goes to dQL[(f + qshift) * NC ...]; the gradients of its
// cells are read at grad[(cell + gshift) * ND * NC ...]; the solver counters
// are added to cnt.
template <int ND, int KIND, int MIXED, typename T, typename R, bool FREC = false>
__device__ __forceinline__ void face_item(const MeshView<T>& m, int64_t f,
                                          const T* __restrict__ grad, int64_t gshift,
                                          const Consts& cd, const mixed::CF& cf,
                                          T* __restrict__ dQL, int64_t qshift,
                                          uint8_t* __restrict__ fstatus,
                                          unsigned (&cnt)[CNT_N]) {
    const int NC = ND + 2, P = ND + 1;
    unsigned iters = 0, brent = 0, vac = 0, fail = 0, solved = 0;
    {
        const int64_t C = m.C, F = m.F;
        (void)C;
        // Every load goes out in two waves before the arithmetic: the face's
        // own values, then its two cells' (a face without a right cell reads
        // its left cell's in their place, unused).  A branch between loads
        // (on activity, on a missing right cell) would make each later load
        // wait for the earlier ones.
        int32_t L, Rc;
        R A, rr[ND], mcv[ND], dLf[ND], dRf[ND];
        T rcent_f;
        if constexpr(FREC) {
            // the face's record (face_record_kernel)
            const FaceVals<ND> fv = load_face_record<ND>(m.frec, F, f);
            L = fv.left;
            Rc = fv.right;
            A = (R)fv.area;
            rcent_f = fv.rcent;
#pragma unroll
            for(int l = 0; l < ND; l++) {
                rr[l] = (R)fv.rRL[l];
                mcv[l] = (R)fv.mc[l];
                dLf[l] = (R)fv.dL[l];
                dRf[l] = (R)fv.dR[l];
            }
        } else {
            L = m.left[f];
            Rc = m.right[f];
            A = (R)m.area[f];
#pragma unroll
            for(int l = 0; l < ND; l++) {
                rr[l] = (R)m.rRL[l * F + f];
                mcv[l] = (R)m.mc[l * F + f];
                dLf[l] = (R)m.dL[l * F + f];
                dRf[l] = (R)m.dR[l * F + f];
            }
            rcent_f = m.rcent[f];
        }
        const bool has_r = Rc >= 0;
        const int64_t Rs = has_r ? (int64_t)Rc : (int64_t)L;
        // twin >= 0 marks a periodic copy (-1 on every other face); the left
        // indices stand in for a missing twin array (never used)
        const int32_t twin_f = (m.twin ? m.twin : m.left)[f];
        // global loads (a select between two pointers would make them flat
        // loads, which also wait with the scalar loads)
        typedef const __attribute__((address_space(1))) uint8_t gu8;
        const gu8* actp = (const gu8*)(m.active ? m.active : &kAllActive);
        const uint8_t actLb = actp[m.active ? (int64_t)L : 0];
        const uint8_t actRb = actp[m.active ? Rs : 0];
        R vl[ND], vr[ND];
#pragma unroll
        for(int l = 0; l < ND; l++) {
            vl[l] = (R)m.vel[(int64_t)L * ND + l];
            vr[l] = (R)m.vel[Rs * ND + l];
        }
        const double dtL = m.dt[L], dtR = m.dt[Rs];
        // 3D: the cells' masses and total areas load with the rest; 2D: after
        // the solve (fewer registers held through it: 2D mixed faces
        // 26.9 -> 25.3 us, while 3D lost 2% that way).  The 3D pin below
        // waits for these four loads before the states and gradients go
        // out; pinned after those instead, the 3D mixed face kernel took
        // 493 against 483 us (and with the states also pinned before the
        // gradients' loads, 491 against 480 us).
        constexpr bool LATE = ND == 2;
        double massL = 0.0, massR = 0.0, areaL = 0.0, areaR = 0.0;
        if constexpr(!LATE) {
            massL = m.mass[L];
            massR = m.mass[Rs];
            areaL = m.total_area[L];
            areaR = m.total_area[Rs];
            asm volatile("" ::"v"(massL), "v"(massR), "v"(areaL), "v"(areaR));
        }
        R WL[NC], WR[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            WL[k] = (R)m.W[(int64_t)L * NC + k];
            WR[k] = (R)m.W[Rs * NC + k];
        }
        using GT = std::conditional_t<MIXED != 0, float, R>;
        GT gL[ND][NC], gR[ND][NC];
        load_grad<ND, T, GT>(grad, C, L + gshift, gL);
        load_grad<ND, T, GT>(grad, C, Rs + gshift, gR);
        // the skip test below needs only the flags, but a branch on them
        // would let the compiler move the other loads behind it
#pragma unroll
        for(int k = 0; k < NC; k++) asm volatile("" ::"v"(WL[k]), "v"(WR[k]));
#pragma unroll
        for(int l = 0; l < ND; l++)
#pragma unroll
            for(int k = 0; k < NC; k++) asm volatile("" ::"v"(gL[l][k]), "v"(gR[l][k]));
        asm volatile("" ::"v"(dtL), "v"(dtR));
        uint8_t st = 0;
        double out[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) out[k] = 0.0;
        const bool actL = actLb != 0;
        const bool actR = has_r && actRb != 0;
        bool skip = !(A != R(0)) || (!has_r && !actL) || (has_r && !actR && !actL);
        // the lower-index copy of a periodic pair is evaluated and writes
        // both rows (the other one negated: that copy's left cell is this
        // one's right cell); the higher-index copy writes nothing
        const bool twin_skip = MIXED && m.twin && twin_f >= 0 && twin_f < f;
        const bool twin_write = MIXED && m.twin && twin_f > f;
        if(twin_skip) skip = true;
        if(!skip) {
            solved = 1;
            R vface[ND];
            if(!has_r) {
#pragma unroll
                for(int l = 0; l < ND; l++) vr[l] = R(0);
            }
            // VorFace::set_v (src/VorFace.cpp:586-660)
            if(has_r || ND == 2) {
                if(!has_r) {
#pragma unroll
                    for(int l = 0; l < ND; l++) vr[l] = vl[l];
                    mirror_velocity<ND, R>(vr, rr);
                }
                R rRL2 = rr[0] * rr[0] + rr[1] * rr[1];
                if(ND == 3) rRL2 = rr[0] * rr[0] + rr[1] * rr[1] + rr[ND - 1] * rr[ND - 1];
                R fac = (vl[0] - vr[0]) * mcv[0];
                fac += (vl[1] - vr[1]) * mcv[1];
                if(ND == 3) fac += (vl[ND - 1] - vr[ND - 1]) * mcv[ND - 1];
                if constexpr(MIXED)
                    fac *= mixed::rcp_nr(rRL2);
                else
                    fac /= rRL2;
#pragma unroll
                for(int l = 0; l < ND; l++) vface[l] = R(0.5) * (vl[l] + vr[l]) + fac * rr[l];
            } else {
#pragma unroll
                for(int l = 0; l < ND; l++) vface[l] = R(0);
            }
            // dt, :695-699
            double dt = dtL;
            if(has_r) dt = smin(dt, dtR);
            if(!has_r) {
#pragma unroll
                for(int k = 0; k < NC; k++) WR[k] = R(0);
            }
#pragma unroll
            for(int l = 0; l < ND; l++) {
                WL[1 + l] -= vface[l];
                WR[1 + l] -= vface[l];
            }
            const R rc = has_r ? (R)rcent_f : R(0);
            R WRrec[NC], WLrec[NC];
            const R gam = (R)cd.gamma;
            if constexpr(MIXED) {
                const float gf = (float)cd.gamma, rcf = (float)rcent_f;
                float dRs[ND], dLs[ND];
#pragma unroll
                for(int l = 0; l < ND; l++) {
                    dRs[l] = (float)dRf[l];
                    dLs[l] = (float)dLf[l];
                }
                if(has_r) predict_side_offset<ND>(gf, WR, WL, gR, dRs, rcf, true, dt, WRrec);
                predict_side_offset<ND>(gf, WL, WR, gL, dLs, rcf, has_r, dt, WLrec);
            } else {
                if(has_r) predict_side<ND, R>(gam, WR, WL, gR, dRf, rc, true, (R)dt, WRrec);
                predict_side<ND, R>(gam, WL, WR, gL, dLf, rc, has_r, (R)dt, WLrec);
            }
            if(WL[0] != R(0)) {
#pragma unroll
                for(int k = 0; k < NC; k++) WL[k] = WLrec[k];
            }
            if(WR[0] != R(0)) {
#pragma unroll
                for(int k = 0; k < NC; k++) WR[k] = WRrec[k];
            }
            // VorFace::get_normal (src/VorFace.cpp:553-571): n = -(xL - xR)/|xL - xR|
            R nrm = (ND == 3) ? sqrt(rr[0] * rr[0] + rr[1] * rr[1] + rr[ND - 1] * rr[ND - 1])
                              : sqrt(rr[0] * rr[0] + rr[1] * rr[1]);
            double flux[NC];
            if(!MIXED) {
                double n64[ND], v64[ND], wl[NC], wr[NC];
#pragma unroll
                for(int l = 0; l < ND; l++) {
                    n64[l] = (double)(rr[l] / nrm);
                    v64[l] = (double)vface[l];
                }
#pragma unroll
                for(int k = 0; k < NC; k++) {
                    wl[k] = (double)WL[k];
                    wr[k] = (double)WR[k];
                }
                if(KIND == 0)
                    st |= fp64::ex_solve_for_flux<ND>(cd, wl, wr, n64, v64, !has_r, flux, iters,
                                                      brent, vac);
                else
                    st |= fp64::hllc_solve_for_flux<ND>(cd, wl, wr, n64, v64, !has_r, flux, vac);
            } else {
                double n64[ND], v64[ND], wl64[NC], wr64[NC];
                const double ninv = mixed::rcp_nr(mixed::sqrt_nr(
                        (double)rr[0] * (double)rr[0] + (double)rr[1] * (double)rr[1] +
                        (ND == 3 ? (double)rr[ND - 1] * (double)rr[ND - 1] : 0.0)));
#pragma unroll
                for(int l = 0; l < ND; l++) {
                    n64[l] = (double)rr[l] * ninv;
                    v64[l] = (double)vface[l];
                }
#pragma unroll
                for(int k = 0; k < NC; k++) {
                    wl64[k] = (double)WL[k];
                    wr64[k] = (double)WR[k];
                }
                if(KIND == 0) {
                    float wlf[NC], wrf[NC], nf[ND];
#pragma unroll
                    for(int k = 0; k < NC; k++) {
                        wlf[k] = (float)WL[k];
                        wrf[k] = (float)WR[k];
                    }
#pragma unroll
                    for(int l = 0; l < ND; l++) nf[l] = (float)n64[l];
                    st |= mixed::exact_mixed<ND>(cf, cd, wlf, wrf, nf, !has_r, wl64, wr64, n64,
                                                 v64, flux, iters, vac);
                } else {
                    st |= mixed::hllc_mixed<ND>(cf, cd, wl64, wr64, n64, v64, !has_r, flux,
                                                vac);
                }
            }
            // dQL = (dt A) F, dQR = -dQL (src/VorFace.cpp:834-835)
            const double dA = dt * (double)A;
            double dQ[NC];
#pragma unroll
            for(int k = 0; k < NC; k++) dQ[k] = dA * flux[k];
            // mass flux limiter, :840-866; dQR stays -dQL bitwise
            if constexpr(LATE) {
                massL = m.mass[L];
                massR = m.mass[Rs];
                areaL = m.total_area[L];
                areaR = m.total_area[Rs];
            }
            // the mixed pass takes its quotients through rcp_nr
            auto quot = [](double a, double b) {
                if constexpr(MIXED)
                    return a * mixed::rcp_nr(b);
                else
                    return a / b;
            };
            double timefac = 1.;
            if(has_r) {
                double dtmax = dtL;
                dtmax = smax(dtmax, dtR);
                timefac = quot(dt, dtmax);
            }
            double areafac = quot((double)A, areaL);
            areafac *= timefac;
            if(dQ[0] > areafac * massL) {
                double corfac = fabs(quot(areafac * massL, dQ[0]));
#pragma unroll
                for(int k = 0; k < NC; k++) dQ[k] *= corfac;
            }
            if(has_r) {
                double af = quot((double)A, areaR);
                af *= timefac;
                // dQR[0] = -dQL[0]
                if(-dQ[0] > af * massR) {
                    double corfac = fabs(quot(af * massR, -dQ[0]));
#pragma unroll
                    for(int k = 0; k < NC; k++) dQ[k] = -(-dQ[k] * corfac);
                }
            }
            if(flux[0] != flux[0]) st |= ST_NANFLUX;
#pragma unroll
            for(int k = 0; k < NC; k++) out[k] = dQ[k];
            fail = st != 0;
        }
        if(!twin_skip) {
#pragma unroll
            for(int k = 0; k < NC; k++) dQL[(f + qshift) * NC + k] = (T)out[k];
        }
        if(twin_write) {
#pragma unroll
            for(int k = 0; k < NC; k++) dQL[((int64_t)twin_f + qshift) * NC + k] = (T)(-out[k]);
        }
        if(fstatus) fstatus[f] = st;
    }
    cnt[CNT_NEVAL] += solved;
    cnt[CNT_NBRENT] += brent;
    cnt[CNT_NVACUUM] += vac;
    cnt[CNT_NITER] += iters;
    cnt[CNT_NFAIL] += fail;
}

template <int ND, int KIND, int MIXED, typename T, typename R, bool FREC = false>
__global__ __launch_bounds__(256) void face_flux_kernel(MeshView<T> m, const T* __restrict__ grad,
                                                        Consts cd, mixed::CF cf,
                                                        T* __restrict__ dQL,
                                                        uint8_t* __restrict__ fstatus,
                                                        unsigned long long* __restrict__ ctr) {
    const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
    unsigned cnt[CNT_N] = {0u, 0u, 0u, 0u, 0u};
    if(f < m.F)
        face_item<ND, KIND, MIXED, T, R, FREC>(m, f, grad, 0, cd, cf, dQL, 0, fstatus, cnt);
    block_add_counters(ctr, cnt);
}

// ------------------------------------------------------------ K4 gather --

// Cell i: dQ[i] = its entries' dQL[(f + qshift)] in ascending face order,
// summed in FP64 (T: storage of dQL).
template <int ND, int MIXED, typename T, int KB0 = 0>
__device__ __forceinline__ void gather_cell(int64_t i, const int64_t* __restrict__ goff,
                                            const int32_t* __restrict__ gent,
                                            const uint8_t* __restrict__ ghost,
                                            const int32_t* __restrict__ twin,
                                            const T* __restrict__ dQL, int64_t qshift,
                                            double* __restrict__ dQ) {
    const int NC = ND + 2;
    // every row of dQL holds its face's flux seen from the left cell (the
    // mixed face pass writes both copies of a periodic pair): no twins here
    (void)ghost;
    (void)twin;
    double acc[NC];
#pragma unroll
    for(int k = 0; k < NC; k++) acc[k] = 0.0;
    const int64_t j0 = goff[i], j1 = goff[i + 1];
    // KB entries at a time (KB0, or SFX_GATHER_KB / _F64): their indices,
    // then their rows, are loaded before any is added, so a lane has that
    // many independent loads in flight; the additions keep ascending face
    // order
    constexpr int KB = KB0 > 0 ? KB0 : (sizeof(T) == 8 ? SFX_GATHER_KB_F64 : SFX_GATHER_KB);
    for(int64_t jb = j0; jb < j1; jb += KB) {
        int32_t f[KB];
        bool neg[KB];
#pragma unroll
        for(int u = 0; u < KB; u++) {
            const int32_t e = gent[(jb + u < j1) ? jb + u : j0];
            f[u] = ent_face(e);
            neg[u] = ent_right(e);
        }
        double x[KB][NC];
#pragma unroll
        for(int u = 0; u < KB; u++)
#pragma unroll
            for(int k = 0; k < NC; k++) x[u][k] = (double)dQL[((int64_t)f[u] + qshift) * NC + k];
        // entries past the cell's last keep acc by a select, not a branch:
        // a branch here would let each row's load sink behind the previous
        // row's addition
#pragma unroll
        for(int u = 0; u < KB; u++) {
            const bool valid = jb + u < j1;
#pragma unroll
            for(int k = 0; k < NC; k++) {
                const double sum = acc[k] + (neg[u] ? -x[u][k] : x[u][k]);
                acc[k] = valid ? sum : acc[k];
            }
        }
    }
#pragma unroll
    for(int k = 0; k < NC; k++) dQ[i * NC + k] = acc[k];
}

// LDS bytes a launch over a mesh of `items` cells reserves per block of 256
// threads (one wave per SIMD, static_lds bytes of its own), so that at most
// `waves` blocks share a CU's 160 KB; meshes below `big` cells keep every
// block that fits.  Some kernels here move far fewer bytes at low occupancy,
// measured on the 3D Sedov mesh (2,097,152 cells, 16.5M faces):
//   * the gather reads every face's row twice, from its left and its right
//     cell; the right cell's read hits the XCD's 4 MB L2 only while the row
//     is still there, and with every block the CU can hold in flight (8
//     waves per SIMD: 65,536 cells an XCD) the rows between the two reads no
//     longer fit.  FP32 rows at 2 waves per SIMD: 526 MB read instead of 765,
//     9% less time; FP64 rows at 1: 906 MB instead of 2,348 (2 waves: 1,325
//     MB, 10% less time; 1 wave: 3.5% more than that);
//   * the record build's scattered face reads: 2 waves per SIMD read 4.4 GB
//     instead of 6.7 (FP64) and 1.7 instead of 2.9 (FP32), 8% and 3% faster.
// The gradients and the face fluxes lost time at 2 or 3 waves per SIMD and
// moved the same bytes.  Small meshes, a few waves per SIMD in all, keep
// their full occupancy: there the loads' latency is what counts.
inline unsigned occupancy_lds_bytes(int waves, int64_t items, int64_t big, unsigned static_lds) {
    if(waves <= 0 || items < big) return 0;
    const unsigned per = ((163840u / (unsigned)waves) & ~1023u) - 1024u;
    return per > static_lds ? per - static_lds : 0u;
}
template <typename T>
inline unsigned gather_lds_bytes(int64_t C) {
    return occupancy_lds_bytes(sizeof(T) == 8 ? SFX_GATHER_WAVES_F64 : SFX_GATHER_WAVES, C,
                               SFX_GATHER_BIG, 0);
}

// waves per SIMD a gather kernel is compiled for: KB0 > 0 is the variant for
// large meshes, which run at the bounded occupancy of gather_lds_bytes, so
// the compiler may spend that occupancy's registers on rows in flight
template <typename T, int KB0>
constexpr int gather_waves() {
    return KB0 > 0 ? (sizeof(T) == 8 ? SFX_GATHER_WAVES_F64 : SFX_GATHER_WAVES) : 0;
}

template <int ND, int MIXED, typename T, int KB0 = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
        gather_waves<T, KB0>() > 0 ? gather_waves<T, KB0>() : 1,
        gather_waves<T, KB0>() > 0 ? gather_waves<T, KB0>() : 10))) void
gather_kernel(int64_t C, int64_t F,
                                                     const int64_t* __restrict__ goff,
                                                     const int32_t* __restrict__ gent,
                                                     const uint8_t* __restrict__ ghost,
                                                     const int32_t* __restrict__ twin,
                                                     const T* __restrict__ dQL,
                                                     double* __restrict__ dQ) {
    (void)F;
    const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
    if(i < C) gather_cell<ND, MIXED, T, KB0>(i, goff, gent, ghost, twin, dQL, 0, dQ);
}

// ------------------------------------------------------------ K5 update --

// GasParticle::update_Q then get_W (use_energy = true).
// get_W with use_energy (ExactRiemannSolver.cpp:716-766, the same
// expressions as HLLCRiemannSolver.cpp:495-535): primitives w of a cell of
// volume V from its conserved q; where the pressure floor acts, q's energy is
// reset to match, as the reference does through its Q reference.
template <int ND>
__device__ __forceinline__ void get_W_state(double g, double V, double* q, double* w) {
    const int NC = ND + 2, P = ND + 1;
#pragma unroll
    for(int k = 0; k < NC; k++) w[k] = 0.;
    if(q[0] != 0.) {
        w[0] = q[0] / V;
#pragma unroll
        for(int d = 0; d < ND; d++) w[1 + d] = q[1 + d] / q[0];
        double p2 = q[1] * q[1] + q[2] * q[2];
        if(ND == 3) p2 = q[1] * q[1] + q[2] * q[2] + q[ND] * q[ND];
        w[P] = (g - 1.) * (q[P] - 0.5 * p2 / q[0]) / V;
        if(w[P] < 1.e-30) {
            w[P] = 1.e-30;
            q[P] = 0.5 * p2 / q[0] + w[P] * V / (g - 1.);
        }
        if(w[0] < 1.e-30) {
#pragma unroll
            for(int k = 0; k < NC; k++) w[k] = 0.;
        }
    }
}

// get_Q (ExactRiemannSolver.cpp:776-804 / HLLCRiemannSolver.cpp:471-483:
// the two associate the energy differently).
template <int ND, int KIND>
__device__ __forceinline__ void get_Q_state(double g, double V, const double* w, double* q) {
    const int P = ND + 1;
    q[0] = w[0] * V;
#pragma unroll
    for(int d = 0; d < ND; d++) q[1 + d] = q[0] * w[1 + d];
    double v2 = w[1] * w[1] + w[2] * w[2];
    if(ND == 3) v2 = w[1] * w[1] + w[2] * w[2] + w[3] * w[3];
    if(KIND == 1)
        q[P] = V * w[P] / (g - 1.) + 0.5 * q[0] * v2;
    else
        q[P] = (0.5 * w[0] * v2 + w[P] / (g - 1.)) * V;
}

// Cell i; returns 1 when m or e went negative.  Q -= dQ and dQ = 0 for every
// cell (GasParticle::update_Q, src/Simulation.cpp:476-483); W = get_W(V, Q)
// only where `rebuild` (the cells whose time step ends now, endtime ==
// currentTime, src/Simulation.cpp:792-796: only they have received the
// fluxes of all their faces); the others keep their W, and their Q is not
// touched by get_W's pressure floor.  mass (may be null) takes the new Q.m:
// the next pass's mass-flux limiter reads the cell's current mass
// (src/VorFace.cpp:850-851).
template <int ND>
__device__ __forceinline__ unsigned update_cell(const Consts& cd, int64_t i,
                                                double* __restrict__ Q, double* __restrict__ dQ,
                                                const double* __restrict__ volume,
                                                double* __restrict__ Wout,
                                                uint8_t* __restrict__ cstatus,
                                                double* __restrict__ mass,
                                                bool rebuild = true) {
    const int NC = ND + 2, P = ND + 1;
    unsigned fail = 0;
    {
        double q[NC], w[NC];
#pragma unroll
        for(int k = 0; k < NC; k++) {
            q[k] = Q[i * NC + k] - dQ[i * NC + k];
            dQ[i * NC + k] = 0.0;
        }
        uint8_t st = (q[0] < 0. || q[P] < 0.) ? ST_UPDATEQ : 0;
        if(rebuild) get_W_state<ND>(cd.gamma, volume[i], q, w);
#pragma unroll
        for(int k = 0; k < NC; k++) Q[i * NC + k] = q[k];
        if(mass) mass[i] = q[0];
        if(rebuild) {
#pragma unroll
            for(int k = 0; k < NC; k++) Wout[i * NC + k] = w[k];
        }
        if(cstatus) cstatus[i] = st;
        fail = st != 0;
    }
    return fail;
}

// endtime null: every cell's W is rebuilt
template <int ND>
__global__ __launch_bounds__(256) void update_kernel(Consts cd, int64_t C, double* __restrict__ Q,
                                                     double* __restrict__ dQ,
                                                     const double* __restrict__ volume,
                                                     double* __restrict__ Wout,
                                                     const uint64_t* __restrict__ endtime,
                                                     uint64_t tnow,
                                                     uint8_t* __restrict__ cstatus,
                                                     double* __restrict__ mass,
                                                     unsigned long long* __restrict__ ctr) {
    const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
    unsigned fail = 0;
    if(i < C) {
        const bool rebuild = endtime ? endtime[i] == tnow : true;
        fail = update_cell<ND>(cd, i, Q, dQ, volume, Wout, cstatus, mass, rebuild);
    }
    const unsigned cv[CNT_N] = {0u, 0u, 0u, 0u, fail};
    block_add_counters(ctr, cv);
}

}  // namespace flux
}  // namespace sfx
