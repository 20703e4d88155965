// Host-side context of the library: the device, the stream, grow-only device buffers and the
// last error message.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

struct scatrs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    std::string error;
    // Grow-only device buffers, indexed by purpose.
    enum Buf {
        kScratch = 0,
        kPerm,
        kKeys,
        kCounters,
        kHist,
        kStaging,
        kNodes,
        kTables,
        kAccum,
        kMisc,
        kMisc2,
        kMisc3,
        kAdjStaging,  // reverse mode: coefficient cotangents, laid out like kStaging
        kAdjNodes,    // reverse mode: per-node table sums, backscatter sums, gradients
        kAdjEns,      // reverse mode: per-ensemble cotangents of the sums and tables
        kAdjSlabs,    // reverse mode: per-workgroup angle-by-column panels
        kOrder,       // ensembles: per-node sort keys and the nodes ranked by N, histogram
        kTables16,    // ensembles: the w pi_n, w tau_n table split in fp16 slices
        kGreekFac,    // ensembles: per-ensemble factor of the Greek coefficients
        kStaging32,   // reverse mode: the fp32 form of the staged coefficients (kernel 2)
        kPhaseMu,     // phase at the caller's cosines: |mu| ascending, then each cosine's index
        kTablesMu,    // ... the w pi_n, w tau_n table at |mu| (and its fp16 slices)
        kAccumMu,     // ... the summed tables of a chunk, and each ensemble's phase factor
        kGeom,        // ensembles: every node's radius and quadrature weight (ens_prep)
        kHostIO,      // host-array sphere batch: its inputs and ranking, then its outputs
        kNumBufs
    };
    void *buf[kNumBufs] = {};
    size_t cap[kNumBufs] = {};
    void *host_pinned = nullptr;  // small pinned status block
    void *host_io = nullptr;      // pinned staging of the host-array sphere batch (grow-only)
    size_t host_io_cap = 0;
    // Ensemble setup kept between calls: the Gauss-Legendre rule in radius (device copy in
    // kMisc) of the last size, and the angle rules of the last sizes, each with its device
    // table (weights, then the non-negative cosines) and its Wigner d table for one order.
    int64_t rule_nr = -1;
    std::vector<double> rr_nodes, rr_weights;
    struct AngleRule {
        int64_t n = 0;
        std::vector<double> nodes, weights;
        double *d_atab = nullptr;
        int64_t order = -1;
        double *d_dtab = nullptr;
        size_t dtab_cap = 0;
        uint64_t used = 0;
    };
    static constexpr int kMaxAngleRules = 64;
    std::vector<AngleRule> angle_rules;
    uint64_t use_clock = 0;
    // Launch parameters the project's autotuner sets (scatrs_set_tuning; scatrs_amd/tune.py):
    // persistent wavefronts per CU of the sphere batch and of the ensembles' sphere node solves,
    // the share of free device memory one chunk of ensembles stages (1/share), and the blocks per
    // CU a small batch's angular products aim for when they split node groups; whether the sphere
    // batch tries the plain path first; when a sphere batch is solved one sphere per wavefront.
    int tune[6] = {16, 8, 4, 2, 1, 8};
    bool profiling = false;
    int angular_kernel = 0;  // 0: FP64 MFMA, 1: FP64 vector (plain), 2: fp16-split MFMA
    std::vector<double> phase_mu;  // the caller's phase cosines (empty: the angle rule's)
    std::vector<double> phase_amu;  // their distinct |mu|, ascending
    std::vector<int> phase_idx;     // each cosine's place among -|mu|, +|mu| (2 h of them)
    std::vector<int> phase_qptr, phase_qlist;  // the cosines at each place, gathered
    // Record and replay of a forward ensemble call for HIP graph capture
    // (scatrs_ensemble_fixed_graph): the first run records every value the host reads back
    // (each ensemble's number of terms, the largest N, the chunking, each chunk's node-launch
    // plan); the captured run takes them from the record instead, so it issues no host
    // synchronization, copy to the host or allocation.
    struct Replay {
        int mode = 0;  // 0 off, 1 recording, 2 capturing from the record
        std::vector<int64_t> terms;
        int n_max = 0;
        int64_t total_terms = 0;
        int64_t max_terms = 0;
        std::vector<std::vector<int64_t>> plans;  // per chunk: (start, end, rows, waves) ...
        std::vector<size_t> scratch;             // per chunk: the node launches' slab bytes
        size_t next_plan = 0;
        int *node_n_rec = nullptr;  // device: every node's N at recording, for the guard
        int *guard = nullptr;       // device: set when a replay's inputs change some node's N
    };
    Replay replay;
    static constexpr int kMaxEvents = 1024;
    hipEvent_t ev_begin[kMaxEvents] = {}, ev_end[kMaxEvents] = {};
    int ev_count = 0;
};

namespace scatrs {

// A device buffer of at least `bytes`, grown (never shrunk) on demand.
hipError_t ensure(scatrs_ctx *ctx, int which, size_t bytes);

int fail(scatrs_ctx *ctx, int code, const std::string &msg);

// Events around the dominant kernel of a call, when profiling is on, on the context's stream or
// on `stream`.
void mark_begin(scatrs_ctx *ctx);
void mark_end(scatrs_ctx *ctx);
void mark_begin(scatrs_ctx *ctx, hipStream_t stream);
void mark_end(scatrs_ctx *ctx, hipStream_t stream);
int check_hip(scatrs_ctx *ctx, hipError_t e, const char *what);

// Validation codes written by the prep kernels.
enum : int { kValid = 0, kInvalid = 1, kOutside = 2 };

// The status block the prep kernels fill and the host reads back: the first invalid element
// (atomicMin of its index), its code, the largest N and the sum of N.
struct Status {
    unsigned long long first_bad;
    int code_invalid;  // bitmask of codes seen at first_bad is resolved on the host
    int n_max;
    unsigned long long total_terms;
    unsigned long long first_outside;
    int pad[2];
};

}  // namespace scatrs
