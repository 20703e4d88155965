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

// Per-block partial status of a prep kernel: the sum and maximum of N over valid items and the
// first invalid and first outside-domain item. Reduced by reduce_status into Status, so no
// address takes more than one atomic per block.
struct Partial {
    unsigned long long sum, bad, outside;
    int max, pad;
};

// Block reduction of (n, ok, bad index, outside index) into partial[blockIdx.x]; blockDim is a
// multiple of 64 and at most 1024.
static __device__ inline void block_partial(int n, bool ok, unsigned long long bad,
                                            unsigned long long outside, Partial *partial) {
    __shared__ Partial wp[16];
    unsigned long long sum = ok ? (unsigned long long)n : 0ull;
    int mx = ok ? n : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
        const int o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
        const unsigned long long b = __shfl_xor(bad, off, 64), u = __shfl_xor(outside, off, 64);
        bad = b < bad ? b : bad;
        outside = u < outside ? u : outside;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wp[w] = Partial{sum, bad, outside, mx, 0};
    __syncthreads();
    if (threadIdx.x == 0) {
        Partial p = wp[0];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            p.sum += wp[i].sum;
            p.max = wp[i].max > p.max ? wp[i].max : p.max;
            p.bad = wp[i].bad < p.bad ? wp[i].bad : p.bad;
            p.outside = wp[i].outside < p.outside ? wp[i].outside : p.outside;
        }
        partial[blockIdx.x] = p;
    }
}

// One block of 1024 threads folds `count` partials into the status block.
static __global__ void __launch_bounds__(1024) reduce_status(const Partial *partial,
                                                             int64_t count, Status *st) {
    __shared__ Partial sp[1024];
    Partial p{0ull, ~0ull, ~0ull, 0, 0};
    for (int64_t i = threadIdx.x; i < count; i += 1024) {
        const Partial q = partial[i];
        p.sum += q.sum;
        p.max = q.max > p.max ? q.max : p.max;
        p.bad = q.bad < p.bad ? q.bad : p.bad;
        p.outside = q.outside < p.outside ? q.outside : p.outside;
    }
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            Partial &a = sp[threadIdx.x];
            const Partial b = sp[threadIdx.x + off];
            a.sum += b.sum;
            a.max = b.max > a.max ? b.max : a.max;
            a.bad = b.bad < a.bad ? b.bad : a.bad;
            a.outside = b.outside < a.outside ? b.outside : a.outside;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st->total_terms = sp[0].sum;
        st->n_max = sp[0].max;
        st->first_bad = sp[0].bad;
        st->first_outside = sp[0].outside;
        st->code_invalid = 0;
    }
}

}  // namespace scatrs
