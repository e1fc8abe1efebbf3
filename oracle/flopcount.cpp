// flopcount.cpp -- TEST INFRASTRUCTURE ONLY (measurement).
//
// The oracle (sfx_oracle.c, the FP64 restatement of the reference path)
// compiled as C++ with every `double` replaced by a counting scalar, so that
// running it over a workload counts the floating-point operations of the
// reference algorithm on those exact inputs.  Counting rule:
//   +, -, *, / (binary, and their compound assignments) and sqrt: 1 each;
//   pow, cbrt, exp, log: 1 each, also tallied apart as transcendentals;
//   comparisons, negation, fabs, fmin, fmax, copies: 0.
// The oracle is compiled without FMA contraction, so it has no fused
// multiply-adds to count.  Everything is FP64 (the reference computes in
// FP64 throughout).
//
// Built by oracle/Makefile into oracle/libflops.so; it exports the oracle's
// sfxo_* entry points (same C ABI: the scalar is one double, passed and
// returned like one) and sfxf_tally / sfxf_reset.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

namespace {
unsigned long long g_flops = 0, g_trans = 0;
}

struct fdouble {
    double v;
    fdouble() = default;
    constexpr fdouble(double x) : v(x) {}
    fdouble& operator+=(fdouble o) { g_flops++; v += o.v; return *this; }
    fdouble& operator-=(fdouble o) { g_flops++; v -= o.v; return *this; }
    fdouble& operator*=(fdouble o) { g_flops++; v *= o.v; return *this; }
    fdouble& operator/=(fdouble o) { g_flops++; v /= o.v; return *this; }
};
static_assert(sizeof(fdouble) == sizeof(double), "one double");

static inline fdouble operator+(fdouble a, fdouble b) { g_flops++; return fdouble(a.v + b.v); }
static inline fdouble operator-(fdouble a, fdouble b) { g_flops++; return fdouble(a.v - b.v); }
static inline fdouble operator*(fdouble a, fdouble b) { g_flops++; return fdouble(a.v * b.v); }
static inline fdouble operator/(fdouble a, fdouble b) { g_flops++; return fdouble(a.v / b.v); }
static inline fdouble operator-(fdouble a) { return fdouble(-a.v); }
static inline bool operator<(fdouble a, fdouble b) { return a.v < b.v; }
static inline bool operator>(fdouble a, fdouble b) { return a.v > b.v; }
static inline bool operator<=(fdouble a, fdouble b) { return a.v <= b.v; }
static inline bool operator>=(fdouble a, fdouble b) { return a.v >= b.v; }
static inline bool operator==(fdouble a, fdouble b) { return a.v == b.v; }
static inline bool operator!=(fdouble a, fdouble b) { return a.v != b.v; }
static inline fdouble sqrt(fdouble a) { g_flops++; return fdouble(::sqrt(a.v)); }
static inline fdouble pow(fdouble a, fdouble b) {
    g_flops++;
    g_trans++;
    return fdouble(::pow(a.v, b.v));
}
static inline fdouble cbrt(fdouble a) {
    g_flops++;
    g_trans++;
    return fdouble(::cbrt(a.v));
}
static inline fdouble fabs(fdouble a) { return fdouble(::fabs(a.v)); }
static inline fdouble fmin(fdouble a, fdouble b) { return fdouble(::fmin(a.v, b.v)); }
static inline fdouble fmax(fdouble a, fdouble b) { return fdouble(::fmax(a.v, b.v)); }

#define double fdouble
#include "sfx_oracle.c"
#undef double

extern "C" {

// out[0] = operations (transcendentals included), out[1] = transcendentals
void sfxf_tally(unsigned long long* out) {
    out[0] = g_flops;
    out[1] = g_trans;
}

void sfxf_reset(void) {
    g_flops = 0;
    g_trans = 0;
}

}  // extern "C"
