// sfx_capi_internal.h -- the solver handle behind the opaque sfx_solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>

#include "sfx_device.h"

constexpr int kStageDoubles = 24;

struct sfx_solver {
    int kind = 0;
    int ndim = 3;
    int precision = 0;
    int device = 0;
    sfx::Consts c{};
    bool stats_on = true;
    unsigned long long* d_ctr = nullptr;  // CNT_N device counters
    double* d_stage = nullptr;            // scalar entry points: device staging
    double* h_stage = nullptr;            // pinned host staging
    // work buffer (the straggler queue of the two-phase exact solve), grown
    // on demand
    void* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    int split = 0;  // FP32 iterations before a problem is queued (0: one kernel)
    int ncu = 0;    // compute units of the device
};

namespace sfx {
int set_error(int code, const char* fmt, ...);
int hip_error(hipError_t e, const char* where);
int bind_device(const sfx_solver* h);
void release_scratch(sfx_solver* h);
int ensure_scratch(sfx_solver* h, size_t bytes);
}  // namespace sfx
