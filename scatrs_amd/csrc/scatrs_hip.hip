// scatrs_hip: context management, input checks and the batched-sphere path
// (scatrs_mie_efficiencies) of the MI355X library. The ensemble path is in ensemble.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scatrs_hip.h"
#include "checks.h"
#include "context.h"
#include "ensemble_common.h"
#include "mie_device.h"
#include "mie_wave.h"
#include "sort.h"
#include "util.h"

namespace scatrs {

hipError_t ensure(scatrs_ctx *ctx, int which, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->cap[which] >= bytes) return hipSuccess;
    if (ctx->buf[which]) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
        e = hipFree(ctx->buf[which]);
        if (e != hipSuccess) return e;
        ctx->buf[which] = nullptr;
        ctx->cap[which] = 0;
    }
    size_t want = std::max(bytes, ctx->cap[which] + ctx->cap[which] / 4);
    want = (want + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&ctx->buf[which], want);
    if (e != hipSuccess) {
        // retry with the exact size
        e = hipMalloc(&ctx->buf[which], bytes);
        if (e != hipSuccess) return e;
        want = bytes;
    }
    ctx->cap[which] = want;
    return hipSuccess;
}

int fail(scatrs_ctx *ctx, int code, const std::string &msg) {
    ctx->error = msg;
    return code;
}

void mark_begin(scatrs_ctx *ctx, hipStream_t stream) {
    if (!ctx->profiling || ctx->ev_count >= scatrs_ctx::kMaxEvents) return;
    const int i = ctx->ev_count;
    if (!ctx->ev_begin[i]) {
        (void)hipEventCreate(&ctx->ev_begin[i]);
        (void)hipEventCreate(&ctx->ev_end[i]);
    }
    (void)hipEventRecord(ctx->ev_begin[i], stream);
}

void mark_end(scatrs_ctx *ctx, hipStream_t stream) {
    if (!ctx->profiling || ctx->ev_count >= scatrs_ctx::kMaxEvents) return;
    (void)hipEventRecord(ctx->ev_end[ctx->ev_count], stream);
    ctx->ev_count++;
}

void mark_begin(scatrs_ctx *ctx) { mark_begin(ctx, ctx->stream); }
void mark_end(scatrs_ctx *ctx) { mark_end(ctx, ctx->stream); }

int check_hip(scatrs_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return SCATRS_OK;
    return fail(ctx, SCATRS_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

// Per sphere: the checks of Mie::compute (scatrs/src/mie/mod.rs:274-320), the number of terms
// and its N bucket for the largest-first order. The first failing element is found with one
// atomic per wavefront (a wavefront minimum first).
__global__ void __launch_bounds__(256) mie_prep(const double *x, const double *m, int64_t len,
                                                int64_t n_fixed, int allow_unvalidated,
                                                int *keys, Partial *partial) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int n = 0;
    bool ok = false;
    unsigned long long bad = ~0ull, outside = ~0ull;
    if (i < len) {
        const double xi = x[i], mr = m[2 * i], mi = m[2 * i + 1];
        long long nn = 0;
        const int code = sphere_code(xi, mr, mi, n_fixed, allow_unvalidated, &nn);
        if (code == kInvalid) bad = (unsigned long long)i;
        if (code == kOutside) outside = (unsigned long long)i;
        ok = code == kValid;
        n = ok ? (int)nn : 0;
        keys[i] = bucket_of(n);
    }
    block_partial(n, ok, bad, outside, partial);
}

// The persistent solve: each wavefront takes 64 spheres at a time, largest N first, from a
// shared counter; each lane solves one sphere. The downward u_k(mx) pass keeps one checkpoint
// per kSphereBlock indices in the wavefront's global slab; the order loop regenerates each
// block into the wavefront's LDS tile. u_k(x) above floor(x) (a few dozen orders) stays in the
// slab.
template <bool DERIVS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mie_batch(
    const double *__restrict__ x, const double *__restrict__ m, int64_t len, int64_t n_fixed,
    const int *__restrict__ perm, unsigned long long *chunk_counter, cx *ck_base,
    double *ux_base, int64_t rows_ck, int64_t rows, int plain, double *__restrict__ values,
    double *__restrict__ jac) {
    __shared__ cx ztile[4][kSphereBlock * 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *ck = ck_base + wave * rows_ck * 64 + lane;
    double *ux = ux_base + wave * rows * 64 + lane;
    cx *tile = &ztile[threadIdx.x >> 6][lane];
    while (true) {
        unsigned long long chunk = 0;
        if (lane == 0) chunk = atomicAdd(chunk_counter, 1ull);
        chunk = __shfl(chunk, 0, 64);
        const int64_t idx = (int64_t)chunk * 64 + lane;
        if ((int64_t)chunk * 64 >= len) break;
        if (idx < len) {
            const int i = perm[idx];
            const double xi = x[i];
            const cx mi = mk(m[2 * (int64_t)i], m[2 * (int64_t)i + 1]);
            const int n = (int)(n_fixed > 0 ? n_fixed : default_n_max(xi));
            double q[5], j[15];
            sphere_efficiencies<kSphereBlock, DERIVS>(xi, mi, n, ck, ux, tile, plain != 0, q, j);
#pragma unroll
            for (int k = 0; k < 5; ++k) values[k * len + i] = q[k];
            if constexpr (DERIVS) {
#pragma unroll
                for (int k = 0; k < 15; ++k) jac[k * len + i] = j[k];
            }
        }
    }
}

// The solve of a small batch, one sphere per wavefront (blocks of one wavefront), largest N
// first from the shared counter: the wavefront's columns (wave_bytes each from cols) hold its
// sphere's u_k(m x) (rows complex) and u_k(x) (rows doubles). DERIVS: the Jacobian too.
template <bool DERIVS>
__global__ __launch_bounds__(64) void mie_wave(const double *__restrict__ x,
                                               const double *__restrict__ m, int64_t len,
                                               int64_t n_fixed, const int *__restrict__ perm,
                                               unsigned long long *counter, char *cols,
                                               int64_t rows, int64_t wave_bytes,
                                               double *__restrict__ values,
                                               double *__restrict__ jac) {
    const int lane = threadIdx.x;
    char *base = cols + (int64_t)blockIdx.x * wave_bytes;
    cx *uz = reinterpret_cast<cx *>(base);
    double *ux = reinterpret_cast<double *>(base + rows * (int64_t)sizeof(cx));
    while (true) {
        unsigned long long idx = 0;
        if (lane == 0) idx = atomicAdd(counter, 1ull);
        idx = __shfl(idx, 0, 64);
        if ((int64_t)idx >= len) break;
        const int i = perm[idx];
        const double xi = x[i];
        const cx mi = mk(m[2 * (int64_t)i], m[2 * (int64_t)i + 1]);
        const int n = (int)(n_fixed > 0 ? n_fixed : default_n_max(xi));
        double q[5], j[15];
        sphere_efficiencies_wave<DERIVS>(xi, mi, n, uz, ux, lane, q, NoStage{}, nullptr, nullptr,
                                         j);
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 5; ++k) values[k * len + i] = q[k];
            if constexpr (DERIVS) {
#pragma unroll
                for (int k = 0; k < 15; ++k) jac[k * len + i] = j[k];
            }
        }
        // The columns are rewritten by the next sphere only after every lane has read them.
        __syncthreads();
    }
}

// Whether a batch takes mie_wave (with or without the Jacobian): the rule of wave_rule.
bool use_wave(const scatrs_ctx *ctx, int64_t terms, int64_t nmax) {
    return wave_rule(ctx->tune[SCATRS_TUNE_SPHERE_WAVE], ctx->cus, terms, nmax);
}

// ---- the batched-sphere path, shared by the allocating and the no-allocation entry points ----

constexpr size_t kStatusBytes = sizeof(Status) + 64;  // status block, then the work counter
constexpr size_t kHistBytes = 2 * kBuckets * sizeof(unsigned);

// The device memory of one sphere batch beyond its inputs and outputs.
struct MieWork {
    void *status = nullptr;  // Status, then the persistent kernel's work counter
    unsigned *hist = nullptr;
    int *keys = nullptr;
    int *perm = nullptr;
    Partial *partial = nullptr;
    void *slab = nullptr;  // the wavefronts' checkpoint and u_k(x) slabs
};

// Bytes of one wavefront's slab for spheres of up to nmax terms.
size_t mie_slab_per_wave(int64_t nmax) {
    const int64_t rows = nmax + 2;
    const int64_t rows_ck = nmax / kSphereBlock + 2;
    return (size_t)64 * (rows_ck * sizeof(cx) + rows * sizeof(double));
}

// Argument checks common to both entry points.
int mie_begin(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
              double *d_values, int64_t *total_terms) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (total_terms) *total_terms = 0;
    if (len < 0) return fail(ctx, SCATRS_INVALID_INPUT, "the number of spheres must be >= 0");
    if (len == 0) return SCATRS_OK;
    if (len > (int64_t)INT32_MAX)
        return fail(ctx, SCATRS_INVALID_INPUT, "at most 2^31 - 1 spheres per call");
    if (!d_x || !d_m || !d_values)
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    return check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
}

// The checks of every sphere, the N histogram and the N-ranked permutation; waits for the
// status (the one synchronization of a call) and maps the first failing element to the
// reference's message.
int mie_validate(scatrs_ctx *ctx, const MieWork &w, const double *d_x, const double *d_m,
                 int64_t len, int64_t n_max, int allow_unvalidated, int64_t *nmax_out,
                 int64_t *total_terms) {
    int rc;
    const int threads = 256;
    const int64_t blocks = (len + threads - 1) / threads;
    Status *st = static_cast<Status *>(w.status);
    unsigned long long *chunk_counter = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(w.status) + sizeof(Status));
    hipStream_t s = ctx->stream;
    fill_u64<<<1, 1, 0, s>>>(chunk_counter, 0ull, 1);
    fill_u32<<<(2 * kBuckets + 255) / 256, 256, 0, s>>>(w.hist, 0u, 2 * kBuckets);
    mie_prep<<<(unsigned)blocks, threads, 0, s>>>(d_x, d_m, len, n_max, allow_unvalidated, w.keys,
                                                    w.partial);
    reduce_status<<<1, 1024, 0, s>>>(w.partial, blocks, st);
    histogram_keys<<<(unsigned)((len + kSortBlock - 1) / kSortBlock), 256, 0, s>>>(w.keys, len,
                                                                                   w.hist);
    // The N-ranked permutation needs only the keys: it runs while the host waits for the status
    // (and is unused if a sphere fails its checks).
    sort_by_keys(w.keys, w.hist, w.perm, len, s);
    Status *hst = static_cast<Status *>(ctx->host_pinned);
    if ((rc = check_hip(ctx, hipMemcpyAsync(hst, st, sizeof(Status), hipMemcpyDeviceToHost, s),
                        "status copy")))
        return rc;
    if ((rc = check_hip(ctx, hipStreamSynchronize(s), "prep"))) return rc;
    const unsigned long long bad = std::min(hst->first_bad, hst->first_outside);
    if (bad != ~0ull) {
        double h[3];
        (void)hipMemcpy(&h[0], d_x + bad, sizeof(double), hipMemcpyDeviceToHost);
        (void)hipMemcpy(&h[1], d_m + 2 * bad, 2 * sizeof(double), hipMemcpyDeviceToHost);
        std::string msg;
        const int code = sphere_message(h[0], h[1], h[2], n_max, allow_unvalidated, &msg);
        return fail(ctx, code == kOutside ? SCATRS_OUTSIDE_VALIDATED_DOMAIN : SCATRS_INVALID_INPUT,
                    "element " + std::to_string(bad) + ": " + msg);
    }
    *nmax_out = hst->n_max;
    if (total_terms) *total_terms = (int64_t)hst->total_terms;
    return SCATRS_OK;
}

// The persistent solve over the N-ranked permutation (mie_validate's) on `waves` wavefronts (a
// multiple of 4) whose slabs start at w.slab.
int mie_solve(scatrs_ctx *ctx, const MieWork &w, const double *d_x, const double *d_m,
              int64_t len, int64_t n_max, int64_t nmax, int64_t waves, double *d_values,
              double *d_jacobian) {
    hipStream_t s = ctx->stream;
    const int64_t rows = nmax + 2;
    const int64_t rows_ck = nmax / kSphereBlock + 2;
    unsigned long long *chunk_counter = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(w.status) + sizeof(Status));
    cx *ck = static_cast<cx *>(w.slab);
    double *ux = reinterpret_cast<double *>(ck + (size_t)waves * rows_ck * 64);
    mark_begin(ctx);
    if (d_jacobian)
        mie_batch<true><<<(unsigned)(waves / 4), 256, 0, s>>>(d_x, d_m, len, n_max, w.perm,
                                                             chunk_counter, ck, ux, rows_ck,
                                                             rows, ctx->tune[SCATRS_TUNE_SPHERE_PLAIN],
                                                             d_values, d_jacobian);
    else
        mie_batch<false><<<(unsigned)(waves / 4), 256, 0, s>>>(d_x, d_m, len, n_max, w.perm,
                                                              chunk_counter, ck, ux, rows_ck,
                                                              rows,
                                                              ctx->tune[SCATRS_TUNE_SPHERE_PLAIN],
                                                              d_values, nullptr);
    mark_end(ctx);
    return check_hip(ctx, hipGetLastError(), "mie_batch launch");
}

// The wavefront-per-sphere solve over the N-ranked permutation on `waves` wavefronts whose
// columns (mie_wave_bytes(nmax) each) start at w.slab.
int mie_solve_wave(scatrs_ctx *ctx, const MieWork &w, const double *d_x, const double *d_m,
                   int64_t len, int64_t n_max, int64_t nmax, int64_t waves, double *d_values,
                   double *d_jacobian) {
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(w.status) + sizeof(Status));
    char *cols = static_cast<char *>(w.slab);
    mark_begin(ctx);
    if (d_jacobian)
        mie_wave<true><<<(unsigned)waves, 64, 0, ctx->stream>>>(
            d_x, d_m, len, n_max, w.perm, counter, cols, nmax + 2, mie_wave_bytes(nmax), d_values,
            d_jacobian);
    else
        mie_wave<false><<<(unsigned)waves, 64, 0, ctx->stream>>>(
            d_x, d_m, len, n_max, w.perm, counter, cols, nmax + 2, mie_wave_bytes(nmax), d_values,
            nullptr);
    mark_end(ctx);
    return check_hip(ctx, hipGetLastError(), "mie_wave launch");
}

// The wavefront-per-sphere solve of another entry point's checked and ranked batch (the layered
// entry points' one-layer case, which gives the sphere batch's results bit for bit): *taken says
// whether the sphere batch's rule takes it, and then the solve is queued. The batch's status
// block (kCounters) holds the work counter, perm its largest-first ranking.
int mie_wave_for(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                 int64_t n_fixed, const int *perm, int64_t terms, int64_t nmax, double *d_values,
                 double *d_jacobian, bool *taken) {
    *taken = use_wave(ctx, terms, nmax);
    if (!*taken) return SCATRS_OK;
    int rc;
    const int64_t waves = std::min<int64_t>(len, (int64_t)ctx->cus * 8);
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch,
                                    (size_t)(mie_wave_bytes(nmax) * waves)),
                        "scratch")))
        return rc;
    MieWork w;
    w.status = ctx->buf[scatrs_ctx::kCounters];
    w.perm = const_cast<int *>(perm);
    w.slab = ctx->buf[scatrs_ctx::kScratch];
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(w.status) + sizeof(Status));
    fill_u64<<<1, 1, 0, ctx->stream>>>(counter, 0ull, 1);
    return mie_solve_wave(ctx, w, d_x, d_m, len, n_fixed, nmax, waves, d_values, d_jacobian);
}

// ---- the no-allocation scratch ----

// Device memory of scatrs_mie_efficiencies_noalloc: every call on a device carves its status,
// histogram, keys, permutation, partials and slabs from this array, which the device holds
// from the library's load until the process ends. Calls that overlap on one device race on
// it. In host code the name is only a handle for hipGetSymbolAddress: the device address of
// each device's copy comes from scratch_device_ptr.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int64_t kNoAllocMaxSpheres = int64_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void *scratch_ptr[kMaxDevices] = {};
bool scratch_tried[kMaxDevices] = {};

// Looks up the scratch array's address on `device` (current) once per device.
void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_tried[device]) return;
    scratch_tried[device] = true;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) scratch_ptr[device] = p;
}

void *scratch_device_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    return scratch_ptr[device];
}

}  // namespace scatrs

using namespace scatrs;

extern "C" {

int scatrs_abi_version(void) { return SCATRS_ABI_VERSION; }

int scatrs_create(int device, void *stream, scatrs_ctx **out) {
    if (!out) return SCATRS_INVALID_INPUT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return SCATRS_DEVICE_ERROR;
    if (hipSetDevice(device) != hipSuccess) return SCATRS_DEVICE_ERROR;
    scratch_init(device);
    scatrs_ctx *ctx = new scatrs_ctx();
    ctx->device = device;
    ctx->stream = static_cast<hipStream_t>(stream);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        ctx->cus = cus;
    if (hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return SCATRS_DEVICE_ERROR;
    }
    *out = ctx;
    return SCATRS_OK;
}

int scatrs_destroy(scatrs_ctx *ctx) {
    if (!ctx) return SCATRS_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    else (void)hipDeviceSynchronize();
    for (int i = 0; i < scatrs_ctx::kNumBufs; ++i)
        if (ctx->buf[i]) (void)hipFree(ctx->buf[i]);
    for (auto &r : ctx->angle_rules) {
        if (r.d_atab) (void)hipFree(r.d_atab);
        if (r.d_dtab) (void)hipFree(r.d_dtab);
    }
    if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
    if (ctx->host_io) (void)hipHostFree(ctx->host_io);
    for (int i = 0; i < scatrs_ctx::kMaxEvents; ++i) {
        if (ctx->ev_begin[i]) (void)hipEventDestroy(ctx->ev_begin[i]);
        if (ctx->ev_end[i]) (void)hipEventDestroy(ctx->ev_end[i]);
    }
    delete ctx;
    return SCATRS_OK;
}

int scatrs_set_stream(scatrs_ctx *ctx, void *stream) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->stream = static_cast<hipStream_t>(stream);
    return SCATRS_OK;
}

const char *scatrs_last_error(const scatrs_ctx *ctx) { return ctx ? ctx->error.c_str() : ""; }

int scatrs_set_profiling(scatrs_ctx *ctx, int on) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->profiling = on != 0;
    ctx->ev_count = 0;
    return SCATRS_OK;
}

int scatrs_set_tuning(scatrs_ctx *ctx, int key, int64_t value) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    static const int64_t lo[6] = {1, 1, 1, 1, 0, 0}, hi[6] = {32, 32, 64, 64, 1, 4096};
    if (key < 0 || key > 5)
        return fail(ctx, SCATRS_INVALID_INPUT, "unknown tuning key " + std::to_string(key));
    if (value < lo[key] || value > hi[key])
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "tuning key " + std::to_string(key) + " takes values in [" +
                        std::to_string(lo[key]) + ", " + std::to_string(hi[key]) + "], got " +
                        std::to_string(value));
    ctx->tune[key] = (int)value;
    return SCATRS_OK;
}

int64_t scatrs_get_tuning(const scatrs_ctx *ctx, int key) {
    if (!ctx || key < 0 || key > 5) return -1;
    return ctx->tune[key];
}

int scatrs_set_angular_kernel(scatrs_ctx *ctx, int kind) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    if (kind < 0 || kind > 2)
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "the angular kernel must be 0 (FP64 matrix cores), 1 (FP64 vector) or 2 "
                    "(fp16 matrix cores, operands split in two slices)");
    ctx->angular_kernel = kind;
    return SCATRS_OK;
}

int scatrs_set_phase_cosines(scatrs_ctx *ctx, const double *mu, int64_t n_mu) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (n_mu < 0 || n_mu > kMaxPhaseCosines || (n_mu > 0 && !mu))
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "the phase cosines need 0 to " + std::to_string(kMaxPhaseCosines) +
                        " values, got " + std::to_string(n_mu));
    for (int64_t j = 0; j < n_mu; ++j)
        if (!(mu[j] >= -1.0 && mu[j] <= 1.0))
            return fail(ctx, SCATRS_INVALID_INPUT,
                        "mu must be in [-1, 1], got " + fmt_e(mu[j]) + " at " + std::to_string(j));
    ctx->phase_mu.assign(mu, mu + n_mu);
    // The distinct |mu|, ascending, and each cosine's place among the 2 h cosines -|mu| (at
    // h - 1 - i for the i-th |mu|) and +|mu| (at h + i).
    std::vector<double> a(n_mu);
    for (int64_t j = 0; j < n_mu; ++j) a[j] = std::fabs(mu[j]);
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    const int h = (int)a.size();
    ctx->phase_amu = a;
    ctx->phase_idx.resize(n_mu);
    for (int64_t j = 0; j < n_mu; ++j) {
        const int i = (int)(std::lower_bound(a.begin(), a.end(), std::fabs(mu[j])) - a.begin());
        ctx->phase_idx[j] = mu[j] >= 0.0 ? h + i : h - 1 - i;
    }
    // The cosines at each place, in order (the reverse pass sums their cotangents there).
    ctx->phase_qptr.assign(2 * h + 1, 0);
    for (int64_t j = 0; j < n_mu; ++j) ++ctx->phase_qptr[ctx->phase_idx[j] + 1];
    for (int q = 0; q < 2 * h; ++q) ctx->phase_qptr[q + 1] += ctx->phase_qptr[q];
    ctx->phase_qlist.assign(n_mu, 0);
    std::vector<int> fill(ctx->phase_qptr.begin(), ctx->phase_qptr.end() - 1);
    for (int64_t j = 0; j < n_mu; ++j) ctx->phase_qlist[fill[ctx->phase_idx[j]]++] = (int)j;
    return SCATRS_OK;
}

int scatrs_kernel_times(scatrs_ctx *ctx, double *ms_out, int cap) {
    if (!ctx) return 0;
    int n = 0;
    for (int i = 0; i < ctx->ev_count && n < cap; ++i) {
        if (hipEventSynchronize(ctx->ev_end[i]) != hipSuccess) break;
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, ctx->ev_begin[i], ctx->ev_end[i]) != hipSuccess) break;
        ms_out[n++] = (double)ms;
    }
    ctx->ev_count = 0;
    return n;
}

int64_t scatrs_default_n_max(double x) { return (int64_t)default_n_max(x); }

int scatrs_size_range(int32_t kind, const double *p, double *t_lo, double *t_hi) {
    if (!p || !t_lo || !t_hi || !scatrs::dist_valid(kind, p)) return SCATRS_INVALID_INPUT;
    scatrs::radius_range(kind, p, t_lo, t_hi);
    return SCATRS_OK;
}

int scatrs_size_log_density(int64_t n, const int32_t *kind, const double *p, const double *r,
                            double *ln_u, double *d_ln_u) {
    if (n < 0 || (n > 0 && (!kind || !p || !r || !ln_u))) return SCATRS_INVALID_INPUT;
    auto run = [=](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            if (!scatrs::dist_valid(kind[i], p + 3 * i)) return false;
            ln_u[i] = scatrs::log_density(kind[i], p + 3 * i, r[i], d_ln_u ? d_ln_u + 3 * i : nullptr);
        }
        return true;
    };
    // The radii are independent: slices of at least 32768 on up to 16 host threads (the
    // adaptive engine asks for millions at once). A thread that cannot start leaves its slice
    // to the calling thread.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<int64_t>({(n + 32767) / 32768, (int64_t)hw, 16});
    if (nt <= 1) return run(0, n) ? SCATRS_OK : SCATRS_INVALID_INPUT;
    std::vector<std::thread> pool;
    std::vector<char> ok(nt, 1), started(nt, 0);
    for (int t = 1; t < nt; ++t) {
        try {
            pool.emplace_back([&ok, &run, t, nt, n] { ok[t] = run(t * n / nt, (t + 1) * n / nt); });
            started[t] = 1;
        } catch (...) {
        }
    }
    ok[0] = run(0, n / nt);
    for (int t = 1; t < nt; ++t)
        if (!started[t]) ok[t] = run(t * n / nt, (t + 1) * n / nt);
    for (auto &th : pool) th.join();
    for (char v : ok)
        if (!v) return SCATRS_INVALID_INPUT;
    return SCATRS_OK;
}

int scatrs_mie_efficiencies(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                            int64_t n_max, int allow_unvalidated, double *d_values,
                            double *d_jacobian, int64_t *total_terms) {
    int rc = mie_begin(ctx, d_x, d_m, len, d_values, total_terms);
    if (rc != SCATRS_OK || len == 0) return rc;
    const int64_t blocks = (len + 255) / 256;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kCounters, kStatusBytes), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kHist, kHistBytes), "alloc"))) return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kKeys, (size_t)len * sizeof(int)), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kPerm, (size_t)len * sizeof(int)), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kNodes, (size_t)blocks * sizeof(Partial)),
                        "alloc")))
        return rc;
    MieWork w;
    w.status = ctx->buf[scatrs_ctx::kCounters];
    w.hist = static_cast<unsigned *>(ctx->buf[scatrs_ctx::kHist]);
    w.keys = static_cast<int *>(ctx->buf[scatrs_ctx::kKeys]);
    w.perm = static_cast<int *>(ctx->buf[scatrs_ctx::kPerm]);
    w.partial = static_cast<Partial *>(ctx->buf[scatrs_ctx::kNodes]);
    int64_t nmax = 0, terms = 0;
    if ((rc = mie_validate(ctx, w, d_x, d_m, len, n_max, allow_unvalidated, &nmax, &terms)))
        return rc;
    if (total_terms) *total_terms = terms;
    if (use_wave(ctx, terms, nmax)) {
        const int64_t waves = std::min<int64_t>(len, (int64_t)ctx->cus * 8);
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch,
                                        (size_t)(mie_wave_bytes(nmax) * waves)),
                            "scratch")))
            return rc;
        w.slab = ctx->buf[scatrs_ctx::kScratch];
        return mie_solve_wave(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jacobian);
    }
    // Persistent wavefronts: the tuned number per CU (default 8: two per SIMD, which the
    // kernel's registers allow; 4-wave blocks share the CU's LDS tiles of kSphereBlock orders),
    // fewer if the slabs would pass 8 GiB.
    const size_t per_wave = mie_slab_per_wave(nmax);
    int64_t waves = std::min<int64_t>((int64_t)ctx->cus * ctx->tune[SCATRS_TUNE_SPHERE_WAVES],
                                      (len + 63) / 64);
    waves = std::min<int64_t>(waves, std::max<int64_t>(4, (int64_t)((8ull << 30) / per_wave)));
    waves = (waves + 3) / 4 * 4;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch, per_wave * (size_t)waves), "scratch")))
        return rc;
    w.slab = ctx->buf[scatrs_ctx::kScratch];
    return mie_solve(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jacobian);
}

int scatrs_mie_efficiencies_host(scatrs_ctx *ctx, const double *x, const double *m, int64_t len,
                                 int64_t n_max, int allow_unvalidated, double *values,
                                 double *jacobian, int64_t *total_terms) {
    int rc = mie_begin(ctx, x, m, len, values, total_terms);
    if (rc != SCATRS_OK || len == 0) return rc;
    // Up to kHostChecks spheres: the checks of every sphere, its N and the largest-first
    // ranking on the host (the first failing element reported as the device path reports it);
    // larger batches are checked and ranked on the device, as scatrs_mie_efficiencies does.
    constexpr int64_t kHostChecks = 65536;
    const bool host_checks = len <= kHostChecks;
    std::vector<int> nn(host_checks ? (size_t)len : 0);
    std::vector<unsigned> count(kBuckets + 1, 0u);
    int64_t nmax = 0, terms = 0;
    for (int64_t i = 0; host_checks && i < len; ++i) {
        long long n = 0;
        const int code = sphere_code(x[i], m[2 * i], m[2 * i + 1], n_max, allow_unvalidated, &n);
        if (code != kValid) {
            std::string msg;
            const int c = sphere_message(x[i], m[2 * i], m[2 * i + 1], n_max, allow_unvalidated,
                                         &msg);
            return fail(ctx, c == kOutside ? SCATRS_OUTSIDE_VALIDATED_DOMAIN : SCATRS_INVALID_INPUT,
                        "element " + std::to_string(i) + ": " + msg);
        }
        nn[(size_t)i] = (int)n;
        nmax = std::max<int64_t>(nmax, n);
        terms += n;
        ++count[(size_t)bucket_of((int)n) + 1];
    }
    for (int b = 0; b < kBuckets; ++b) count[(size_t)b + 1] += count[(size_t)b];
    // Pinned staging: x, m (interleaved), then the ranking; the outputs come back through it.
    const bool derivs = jacobian != nullptr;
    const size_t in_bytes = (size_t)len * 3 * sizeof(double) + (size_t)len * sizeof(int);
    const size_t out_bytes = (size_t)len * (derivs ? 20 : 5) * sizeof(double);
    const size_t io_bytes = std::max(in_bytes, out_bytes);
    if (ctx->host_io_cap < io_bytes) {
        if (ctx->host_io) {
            if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "host staging")))
                return rc;
            (void)hipHostFree(ctx->host_io);
            ctx->host_io = nullptr;
            ctx->host_io_cap = 0;
        }
        const size_t want = std::max(io_bytes, ctx->host_io_cap + ctx->host_io_cap / 4);
        if ((rc = check_hip(ctx, hipHostMalloc(&ctx->host_io, want, hipHostMallocDefault),
                            "host staging")))
            return rc;
        ctx->host_io_cap = want;
    }
    char *hio = static_cast<char *>(ctx->host_io);
    std::memcpy(hio, x, (size_t)len * sizeof(double));
    std::memcpy(hio + (size_t)len * sizeof(double), m, (size_t)len * 2 * sizeof(double));
    int *hperm = reinterpret_cast<int *>(hio + (size_t)len * 3 * sizeof(double));
    if (host_checks)
        for (int64_t i = 0; i < len; ++i) hperm[count[(size_t)bucket_of(nn[(size_t)i])]++] = (int)i;
    // Device: the inputs and ranking, then the outputs.
    const size_t out_off = (in_bytes + 255) & ~size_t(255);
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kHostIO, out_off + out_bytes), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kCounters, kStatusBytes), "alloc")))
        return rc;
    char *dio = static_cast<char *>(ctx->buf[scatrs_ctx::kHostIO]);
    const double *d_x = reinterpret_cast<const double *>(dio);
    const double *d_m = d_x + len;
    double *d_values = reinterpret_cast<double *>(dio + out_off);
    double *d_jac = derivs ? d_values + 5 * len : nullptr;
    hipStream_t s = ctx->stream;
    if ((rc = check_hip(ctx, hipMemcpyAsync(dio, hio, host_checks ? in_bytes : 3 * (size_t)len *
                                                                             sizeof(double),
                                            hipMemcpyHostToDevice, s),
                        "input copy")))
        return rc;
    MieWork w;
    w.status = ctx->buf[scatrs_ctx::kCounters];
    w.perm = reinterpret_cast<int *>(dio + (size_t)len * 3 * sizeof(double));
    if (host_checks) {
        unsigned long long *counter = reinterpret_cast<unsigned long long *>(
            static_cast<char *>(w.status) + sizeof(Status));
        fill_u64<<<1, 1, 0, s>>>(counter, 0ull, 1);
    } else {
        const int64_t blocks = (len + 255) / 256;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kHist, kHistBytes), "alloc"))) return rc;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kKeys, (size_t)len * sizeof(int)), "alloc")))
            return rc;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kNodes, (size_t)blocks * sizeof(Partial)),
                            "alloc")))
            return rc;
        w.hist = static_cast<unsigned *>(ctx->buf[scatrs_ctx::kHist]);
        w.keys = static_cast<int *>(ctx->buf[scatrs_ctx::kKeys]);
        w.partial = static_cast<Partial *>(ctx->buf[scatrs_ctx::kNodes]);
        if ((rc = mie_validate(ctx, w, d_x, d_m, len, n_max, allow_unvalidated, &nmax, &terms)))
            return rc;
    }
    if (use_wave(ctx, terms, nmax)) {
        const int64_t waves = std::min<int64_t>(len, (int64_t)ctx->cus * 8);
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch,
                                        (size_t)(mie_wave_bytes(nmax) * waves)),
                            "scratch")))
            return rc;
        w.slab = ctx->buf[scatrs_ctx::kScratch];
        rc = mie_solve_wave(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jac);
    } else {
        const size_t per_wave = mie_slab_per_wave(nmax);
        int64_t waves = std::min<int64_t>((int64_t)ctx->cus * ctx->tune[SCATRS_TUNE_SPHERE_WAVES],
                                          (len + 63) / 64);
        waves = std::min<int64_t>(waves, std::max<int64_t>(4, (int64_t)((8ull << 30) / per_wave)));
        waves = (waves + 3) / 4 * 4;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch, per_wave * (size_t)waves),
                            "scratch")))
            return rc;
        w.slab = ctx->buf[scatrs_ctx::kScratch];
        rc = mie_solve(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jac);
    }
    if (rc) return rc;
    if ((rc = check_hip(ctx, hipMemcpyAsync(hio, d_values, out_bytes, hipMemcpyDeviceToHost, s),
                        "output copy")))
        return rc;
    if ((rc = check_hip(ctx, hipStreamSynchronize(s), "sphere batch"))) return rc;
    std::memcpy(values, hio, (size_t)len * 5 * sizeof(double));
    if (derivs) std::memcpy(jacobian, hio + (size_t)len * 5 * sizeof(double),
                            (size_t)len * 15 * sizeof(double));
    if (total_terms) *total_terms = terms;
    return SCATRS_OK;
}

int scatrs_mie_efficiencies_noalloc(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                    int64_t len, int64_t n_max, int allow_unvalidated,
                                    double *d_values, double *d_jacobian, int64_t *total_terms) {
    int rc = mie_begin(ctx, d_x, d_m, len, d_values, total_terms);
    if (rc != SCATRS_OK || len == 0) return rc;
    if (len > kNoAllocMaxSpheres)
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "the no-allocation entry point takes at most " +
                        std::to_string(kNoAllocMaxSpheres) + " spheres per call, got " +
                        std::to_string(len));
    char *base = static_cast<char *>(scratch_device_ptr(ctx->device));
    if (!base)
        return fail(ctx, SCATRS_DEVICE_ERROR,
                    "the no-allocation scratch of device " + std::to_string(ctx->device) +
                        " was not set up (scatrs_create could not find its address)");
    // The scratch array, carved in 256-byte aligned pieces: status, histogram, keys,
    // permutation, per-block partials, then the wavefronts' slabs in what is left.
    const int64_t blocks = (len + 255) / 256;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base + off;
        off += (bytes + 255) & ~size_t(255);
        return p;
    };
    MieWork w;
    w.status = take(kStatusBytes);
    w.hist = reinterpret_cast<unsigned *>(take(kHistBytes));
    w.keys = reinterpret_cast<int *>(take((size_t)len * sizeof(int)));
    w.perm = reinterpret_cast<int *>(take((size_t)len * sizeof(int)));
    w.partial = reinterpret_cast<Partial *>(take((size_t)blocks * sizeof(Partial)));
    w.slab = base + off;
    const size_t slab_bytes = kScratchBytes - off;
    int64_t nmax = 0, terms = 0;
    if ((rc = mie_validate(ctx, w, d_x, d_m, len, n_max, allow_unvalidated, &nmax, &terms)))
        return rc;
    if (total_terms) *total_terms = terms;
    if (use_wave(ctx, terms, nmax)) {
        // As many wavefronts as the scratch holds columns for, up to two per SIMD.
        const int64_t wb = mie_wave_bytes(nmax);
        const int64_t waves = std::min<int64_t>({len, (int64_t)ctx->cus * 8,
                                                 (int64_t)(slab_bytes / (size_t)wb)});
        if (waves < 1) {
            if (total_terms) *total_terms = 0;
            return fail(ctx, SCATRS_INVALID_INPUT,
                        "N = " + std::to_string(nmax) +
                            " terms need more column memory than the no-allocation scratch "
                            "holds (" + std::to_string(kScratchBytes >> 20) + " MiB)");
        }
        return mie_solve_wave(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jacobian);
    }
    // As many persistent wavefronts as the scratch holds slabs for, up to the tuned number per
    // CU, in whole blocks of 4.
    const size_t per_wave = mie_slab_per_wave(nmax);
    int64_t waves = std::min<int64_t>((int64_t)ctx->cus * ctx->tune[SCATRS_TUNE_SPHERE_WAVES],
                                      (len + 63) / 64);
    waves = (waves + 3) / 4 * 4;
    waves = std::min<int64_t>(waves, (int64_t)(slab_bytes / per_wave));
    waves = waves / 4 * 4;
    if (waves < 4) {
        if (total_terms) *total_terms = 0;
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "N = " + std::to_string(nmax) +
                        " terms need more slab memory than the no-allocation scratch holds (" +
                        std::to_string(kScratchBytes >> 20) + " MiB)");
    }
    return mie_solve(ctx, w, d_x, d_m, len, n_max, nmax, waves, d_values, d_jacobian);
}

}  // extern "C"
