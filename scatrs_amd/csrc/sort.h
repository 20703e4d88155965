// Counting sort of work items by their number of terms, largest first, so that the 64 lanes
// of a wavefront get items of about the same N and the longest ones start earliest. The
// histogram is filled by the prep kernels; a one-block scan turns it into bucket offsets and
// a scatter writes the permutation (the order inside a bucket does not matter: every item's
// result is independent of where it runs).
#pragma once

#include <hip/hip_runtime.h>

namespace scatrs {

constexpr int kBuckets = 4096;

// Bucket of an item with n terms: descending n, items above 4094 terms share bucket 0.
__host__ __device__ __forceinline__ int bucket_of(int n) {
    const int c = n < kBuckets - 2 ? n : kBuckets - 2;
    return kBuckets - 2 - c;  // n = 0 (invalid) lands in kBuckets - 2; kBuckets - 1 unused
}

// Exclusive scan of kBuckets counts in one block of 1024 threads (4 buckets each).
__global__ void __launch_bounds__(1024) scan_buckets(const unsigned *hist, unsigned *offsets) {
    __shared__ unsigned part[1024];
    const int t = threadIdx.x;
    unsigned v[4];
    unsigned local = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        v[j] = hist[t * 4 + j];
        local += v[j];
    }
    part[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    unsigned run = part[t] - local;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        offsets[t * 4 + j] = run;
        run += v[j];
    }
}

__global__ void scatter_by_key(const int *keys, unsigned *offsets, int *perm, int64_t len) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= len) return;
    const unsigned pos = atomicAdd(&offsets[keys[i]], 1u);
    perm[pos] = (int)i;
}

// hist has 2 * kBuckets entries: the counts, then room for the offsets.
inline void sort_by_keys(const int *keys, unsigned *hist, int *perm, int64_t len,
                         hipStream_t s) {
    unsigned *offsets = hist + kBuckets;
    scan_buckets<<<1, 1024, 0, s>>>(hist, offsets);
    scatter_by_key<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(keys, offsets, perm, len);
}

}  // namespace scatrs
