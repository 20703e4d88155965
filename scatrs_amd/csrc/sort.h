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
static __global__ void __launch_bounds__(1024) scan_buckets(const unsigned *hist, unsigned *offsets) {
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

// Elements per thread of the histogram and scatter kernels (256 threads per block): 4096 per
// block, so that each block's pass over the 4096 LDS buckets (zeroing, and one global atomic
// per non-empty bucket) serves as many elements as it has buckets.
constexpr int kSortPerThread = 16;
constexpr int kSortBlock = 256 * kSortPerThread;

// Block-local histogram in LDS, then one global atomic per non-empty bucket of the block.
static __device__ inline void block_histogram(const int *keys_local, int nkeys, unsigned *lcount,
                                       unsigned *hist) {
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x) lcount[b] = 0u;
    __syncthreads();
    for (int j = 0; j < nkeys; ++j) atomicAdd(&lcount[keys_local[j]], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x)
        if (lcount[b]) atomicAdd(&hist[b], lcount[b]);
}

// Histogram of keys[0, len) (keys already written).
static __global__ void __launch_bounds__(256) histogram_keys(const int *keys, int64_t len,
                                                      unsigned *hist) {
    __shared__ unsigned lcount[kBuckets];
    int local[kSortPerThread];
    int nk = 0;
    for (int j = 0; j < kSortPerThread; ++j) {
        const int64_t i = (int64_t)blockIdx.x * kSortBlock + j * 256 + threadIdx.x;
        if (i < len) local[nk++] = keys[i];
    }
    block_histogram(local, nk, lcount, hist);
}

// Each block ranks its elements per bucket in LDS, reserves one range per non-empty bucket
// from the global offsets, and writes the permutation.
static __global__ void __launch_bounds__(256) scatter_by_key(const int *keys, unsigned *offsets,
                                                      int *perm, int64_t len) {
    __shared__ unsigned lcount[kBuckets];
    __shared__ unsigned lbase[kBuckets];
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x) lcount[b] = 0u;
    __syncthreads();
    int key[kSortPerThread];
    unsigned rank[kSortPerThread];
    for (int j = 0; j < kSortPerThread; ++j) {
        const int64_t i = (int64_t)blockIdx.x * kSortBlock + j * 256 + threadIdx.x;
        key[j] = -1;
        if (i < len) {
            key[j] = keys[i];
            rank[j] = atomicAdd(&lcount[key[j]], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kBuckets; b += blockDim.x)
        if (lcount[b]) lbase[b] = atomicAdd(&offsets[b], lcount[b]);
    __syncthreads();
    for (int j = 0; j < kSortPerThread; ++j) {
        const int64_t i = (int64_t)blockIdx.x * kSortBlock + j * 256 + threadIdx.x;
        if (key[j] >= 0) perm[lbase[key[j]] + rank[j]] = (int)i;
    }
}

// hist has 2 * kBuckets entries: the counts, then room for the offsets.
static inline void sort_by_keys(const int *keys, unsigned *hist, int *perm, int64_t len,
                         hipStream_t s) {
    unsigned *offsets = hist + kBuckets;
    scan_buckets<<<1, 1024, 0, s>>>(hist, offsets);
    scatter_by_key<<<(unsigned)((len + kSortBlock - 1) / kSortBlock), 256, 0, s>>>(keys, offsets,
                                                                                  perm, len);
}

}  // namespace scatrs
