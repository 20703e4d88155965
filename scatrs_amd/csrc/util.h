// Small fill kernels (each translation unit keeps its own copy: no relocatable device code).
// Buffers are cleared with kernels rather than hipMemsetAsync so a call stays capturable in a
// HIP graph.
#pragma once

#include <hip/hip_runtime.h>

#include "context.h"

namespace scatrs {

static __global__ void fill_u64(unsigned long long *p, unsigned long long v, int64_t count) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

static __global__ void fill_u32(unsigned int *p, unsigned int v, int64_t count) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

static __global__ void fill_f64(double *p, double v, int64_t count) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// Status: everything 0, then first_bad and first_outside at "none".
static inline void zero_status(scatrs_ctx *ctx, Status *st) {
    static_assert(sizeof(Status) % 8 == 0, "status is u64-sized");
    fill_u64<<<1, 64, 0, ctx->stream>>>(reinterpret_cast<unsigned long long *>(st), 0,
                                        sizeof(Status) / 8);
    fill_u64<<<1, 1, 0, ctx->stream>>>(&st->first_bad, ~0ull, 1);
    fill_u64<<<1, 1, 0, ctx->stream>>>(&st->first_outside, ~0ull, 1);
}

}  // namespace scatrs
